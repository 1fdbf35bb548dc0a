// Sanitizer harness for the host-side parsers (built by
// tests/test_sanitizers.py with -fsanitize=address,undefined, never shipped):
// for every file argument, parses it as JSON (and, when that succeeds,
// checks that dump -> parse -> dump is a fixed point) and loads it as a scene
// file when its name ends in ".json" (which imports its OBJ / PLY). Parse
// errors are expected for the malformed inputs; any sanitizer report ends
// the process with a nonzero status.
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>

#include "json.h"
#include "log.h"
#include "scene_loader.h"

int main(int argc, char **argv) {
    koishi::log_enabled() = false;
    int parsed = 0, loaded = 0, rejected = 0;
    for (int i = 1; i < argc; ++i) {
        std::string path = argv[i];
        std::ifstream f(path, std::ios::binary);
        std::ostringstream ss;
        ss << f.rdbuf();
        const std::string text = ss.str();
        try {
            koishi::Json j = koishi::Json::parse(text);
            std::string a = j.dump();
            std::string b = koishi::Json::parse(a).dump();
            if (a != b) {
                std::cerr << path << ": dump is not a fixed point\n";
                return 2;
            }
            ++parsed;
        } catch (const std::exception &) {
            ++rejected;
        }
        if (path.size() > 5 && path.compare(path.size() - 5, 5, ".json") == 0) {
            try {
                koishi::LoadedScene s = koishi::load_scene_file(path);
                ++loaded;
            } catch (const std::exception &) {
                ++rejected;
            }
        }
    }
    std::cout << "ok parsed " << parsed << " loaded " << loaded << " rejected " << rejected << "\n";
    return 0;
}
