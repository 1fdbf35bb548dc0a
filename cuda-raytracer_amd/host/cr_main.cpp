// cr_main.cpp -- the drop-in `cr` command line (src/cr.cu:9-68).
//
//   cr <scene.json> [-o a.png] [-s spp] [--resolution WxH] [-t tracer]
//      [-k kernel] [-a allocator] [-l] [-h] [-v]
//      [--seed N] [--dump-raw file] [--device N] [--bvh-width W]
//
// Same options and defaults as the reference (cxxopts, src/cr.cu:11-21);
// unlike the reference, -t and -k are honoured (they override the scene
// file's "tracer" entry) and a fatal error exits with status 1.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "log.h"
#include "renderer.h"

namespace {

const char *kHelp =
    "Ray tracer for heterogeneous systems, by KoishiChan~ (MI355X / gfx950 build)\n"
    "Usage:\n"
    "  cr <scene.json> [OPTION...]\n\n"
    "  -v, --visualize               Visualize BVH using openGL. (not provided)\n"
    "  -o arg                        Place the output into <file>. (default: a.png)\n"
    "  -s, --sample-per-pixel arg    Number of sample points per pixel. (default: 1)\n"
    "  -l, --list                    List all valid renderers.\n"
    "  -h, --help                    Show help message.\n"
    "  -t, --tracer arg              Specify target tracer. (default: the scene's)\n"
    "  -k, --kernel arg              Specify kernel function. (default: Radiance)\n"
    "  -a, --allocator arg           Specify allocator. (accepted, unused)\n"
    "      --resolution arg          Specify target resolution. (default: 1024x768)\n"
    "      --seed arg                RNG seed. (default: 1)\n"
    "      --dump-raw arg            Also write the FP64 image (KRRAW1 format).\n"
    "      --device arg              HIP device ordinal. (default: 0)\n"
    "      --bvh-width arg           BVH width 2, 4 or 8. (default: 4)\n";

struct Opts {
    std::vector<std::string> positional;
    std::string out = "a.png", tracer, kernel, allocator = "HybridAllocator", resolution = "1024x768", dump_raw;
    unsigned spp = 1, device = 0, bvh_width = 0;
    unsigned long long seed = 1;
    bool help = false, list = false, visualize = false;
};

Opts parse(int argc, char **argv) {
    Opts o;
    auto need = [&](int &i, const std::string &name, const std::string &inl) -> std::string {
        if (!inl.empty()) return inl;
        if (i + 1 >= argc) throw std::logic_error("Option '" + name + "' is missing an argument");
        return argv[++i];
    };
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        std::string name, inl;
        if (a.rfind("--", 0) == 0) {
            size_t eq = a.find('=');
            name = a.substr(2, eq == std::string::npos ? std::string::npos : eq - 2);
            if (eq != std::string::npos) inl = a.substr(eq + 1);
        } else if (a.size() >= 2 && a[0] == '-') {
            name = a.substr(1, 1);
            inl = a.substr(2);
        } else {
            o.positional.push_back(a);
            continue;
        }
        auto num = [&](const std::string &v) {
            char *end = nullptr;
            unsigned long long x = std::strtoull(v.c_str(), &end, 10);
            if (!end || *end) throw std::logic_error("Argument '" + v + "' failed to parse");
            return x;
        };
        if (name == "o") o.out = need(i, name, inl);
        else if (name == "s" || name == "sample-per-pixel") o.spp = unsigned(num(need(i, name, inl)));
        else if (name == "t" || name == "tracer") o.tracer = need(i, name, inl);
        else if (name == "k" || name == "kernel") o.kernel = need(i, name, inl);
        else if (name == "a" || name == "allocator") o.allocator = need(i, name, inl);
        else if (name == "resolution") o.resolution = need(i, name, inl);
        else if (name == "seed") o.seed = num(need(i, name, inl));
        else if (name == "dump-raw") o.dump_raw = need(i, name, inl);
        else if (name == "device") o.device = unsigned(num(need(i, name, inl)));
        else if (name == "bvh-width") o.bvh_width = unsigned(num(need(i, name, inl)));
        else if (name == "h" || name == "help") o.help = true;
        else if (name == "l" || name == "list") o.list = true;
        else if (name == "v" || name == "visualize") o.visualize = true;
        else throw std::logic_error("Option '" + name + "' does not exist");
    }
    return o;
}

}  // namespace

int main(int argc, char **argv) {
    try {
        Opts o = parse(argc, argv);
        if (o.help) {
            std::cout << kHelp;
        } else if (o.visualize) {
            KINFO("fatal", "the OpenGL BVH viewer is not provided by this build");
            return 1;
        } else if (o.list) {
            KLOG("tracers:");
            for (auto &n : koishi::core::tracer_names()) KLOG(" ", n);
            KLOG("kernels:");
            for (auto &n : koishi::core::kernel_names()) KLOG(" ", n);
        } else {
            if (o.positional.empty()) throw std::logic_error("no scene file given");
            unsigned w = 1024, h = 768;
            std::sscanf(o.resolution.c_str(), "%ux%u", &w, &h);
            KLOG("Sample", o.spp, "points per pixel");
            koishi::core::RenderOptions ro;
            ro.tracer = o.tracer;
            ro.kernel = o.kernel;
            ro.seed = o.seed;
            ro.dump_raw = o.dump_raw;
            ro.device = int(o.device);
            ro.bvh_width = o.bvh_width;
            koishi::core::Renderer(w, h).render(o.positional[0], o.out, o.spp, ro);
        }
    } catch (const std::exception &err) {
        KINFO("fatal", err.what());
        return 1;
    }
    return 0;
}
