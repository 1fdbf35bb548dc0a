// ThreadSanitizer harness for the CPU oracle's multithreaded render (built
// by tests/test_sanitizers.py with -fsanitize=thread together with
// oracle/oracle_capi.cpp, never shipped): loads a scene file with the host
// loader and renders it with the CPUMulticore thread scheme (rows
// interleaved over threads, per-thread statistics merged after the joins)
// on 4 threads, twice; the two images must be identical.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "kr_api.h"
#include "log.h"
#include "scene_loader.h"

extern "C" {
int oracle_scene_create_f64(const kr_scene_desc *d, void **out);
void oracle_scene_free_f64(void *s);
double oracle_render_f64(void *sp, const kr_camera *cam, const kr_params *prm, uint32_t spp_begin,
                         uint32_t spp_count, int nthreads, const uint32_t *pixels, uint32_t npix, double *rgb,
                         uint32_t *valid, uint64_t *counters);
}

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    koishi::log_enabled() = false;
    koishi::LoadedScene s = koishi::load_scene_file(argv[1]);
    kr_scene_desc d = s.desc();
    void *sc = nullptr;
    if (oracle_scene_create_f64(&d, &sc) != KR_OK) return 3;
    kr_params p{};
    p.width = 40;
    p.height = 24;
    p.spp = 4;
    p.max_bounce = s.max_bounce;
    p.seed = 9;
    std::vector<double> a(3 * 40 * 24), b(3 * 40 * 24);
    std::vector<uint32_t> va(40 * 24), vb(40 * 24);
    uint64_t ca[4] = {0, 0, 0, 0}, cb[4] = {0, 0, 0, 0};
    oracle_render_f64(sc, &s.cameras.at(0), &p, 0, 4, 4, nullptr, 40 * 24, a.data(), va.data(), ca);
    oracle_render_f64(sc, &s.cameras.at(0), &p, 0, 4, 4, nullptr, 40 * 24, b.data(), vb.data(), cb);
    oracle_scene_free_f64(sc);
    uint64_t nv = 0;
    for (size_t i = 0; i < va.size(); ++i) nv += va[i];
    if (a != b || va != vb || ca[0] != cb[0] || nv == 0) return 4;
    std::printf("ok %llu valid samples, %llu closest-hit queries\n", (unsigned long long)nv, (unsigned long long)ca[0]);
    return 0;
}
