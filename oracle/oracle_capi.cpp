// oracle_capi.cpp -- C entry points of the CPU oracle (TEST INFRASTRUCTURE).
//
// Compiled twice into liboracle.so: with ORACLE_REAL=double and strict IEEE
// flags (suffix _f64, the FP64 oracle) and with ORACLE_REAL=float and the
// reference's release flags -O3 -ffast-math -fno-finite-math-only
// -funroll-loops (CMakeLists.txt:11; suffix _f32, the faithful CPU baseline).
// Callers: tests/, __graft_entry__.smoke(), bench.py's cpu_baseline leg.
#include "koishi_oracle.hpp"

#ifndef ORACLE_REAL
#error "define ORACLE_REAL"
#endif
#define OR_CAT2(a, b) a##b
#define OR_CAT(a, b) OR_CAT2(a, b)
#define OR_FN(name) OR_CAT(name, ORACLE_SUFFIX)

using namespace koishi_oracle;
using R = ORACLE_REAL;

namespace {
inline V3<R> dv(const double *p) { return {R(p[0]), R(p[1]), R(p[2])}; }
inline void put(double *p, V3<R> v) {
    p[0] = double(v.x);
    p[1] = double(v.y);
    p[2] = double(v.z);
}
}  // namespace

extern "C" {

int OR_FN(oracle_scene_create)(const kr_scene_desc *d, void **out) {
    auto *s = new Scene<R>();
    if (!s->build(*d)) {
        delete s;
        return KR_E_INVALID;
    }
    *out = s;
    return KR_OK;
}

void OR_FN(oracle_scene_free)(void *s) { delete static_cast<Scene<R> *>(s); }

void OR_FN(oracle_scene_set_traversal)(void *s, int mode) { static_cast<Scene<R> *>(s)->traversal = mode; }

// counts[0..3] = lights, prims, reference-BVH nodes (all meshes), padded faces
void OR_FN(oracle_scene_counts)(void *sp, uint64_t *counts) {
    auto *s = static_cast<Scene<R> *>(sp);
    counts[0] = s->lights.size();
    counts[1] = s->prims.size();
    counts[2] = 0;
    counts[3] = 0;
    for (auto &m : s->meshes) {
        counts[2] += m.bvh.size();
        counts[3] += m.faces.size();
    }
}

// Closest hit of n rays; stats (optional, 2 entries) receive node and
// triangle visits summed over all rays.
int OR_FN(oracle_trace_closest)(void *sp, const float *o, const float *d, uint32_t n, double *t, uint32_t *prim,
                                uint32_t *tri, double *u, double *v, uint64_t *stats) {
    auto *s = static_cast<Scene<R> *>(sp);
    Counters cnt;
    for (uint32_t i = 0; i < n; ++i) {
        Ray<R> r{mk<R>(o + 3 * size_t(i)), mk<R>(d + 3 * size_t(i))};
        auto c = s->closest(r, stats ? &cnt : nullptr);
        t[i] = c.hit ? double(c.h.t) : INFINITY;
        prim[i] = c.hit ? c.prim : KR_MISS;
        tri[i] = c.hit ? c.tri : 0;
        u[i] = c.hit ? double(c.h.u) : 0;
        v[i] = c.hit ? double(c.h.v) : 0;
        if (c.hit && s->prims[c.prim].type == KR_PRIM_SPHERE) u[i] = v[i] = 0;
    }
    if (stats) {
        stats[0] = cnt.nodes;
        stats[1] = cnt.tris;
    }
    return KR_OK;
}

int OR_FN(oracle_trace_shadow)(void *sp, const float *o, const float *d, const float *tmax, uint32_t n, uint8_t *occ) {
    auto *s = static_cast<Scene<R> *>(sp);
    for (uint32_t i = 0; i < n; ++i) {
        Seg<R> sg{mk<R>(o + 3 * size_t(i)), mk<R>(d + 3 * size_t(i)), R(tmax[i])};
        occ[i] = s->occluded(sg) ? 1 : 0;
    }
    return KR_OK;
}

// Renders samples [spp_begin, spp_begin + spp_count) and ADDS into rgb/valid.
// Returns the wall time of the trace loop in seconds. counters (optional, 4
// entries): closest-hit queries, shadow queries, node visits, triangle tests.
double OR_FN(oracle_render)(void *sp, const kr_camera *cam, const kr_params *prm, uint32_t spp_begin,
                            uint32_t spp_count, int nthreads, const uint32_t *pixels, uint32_t npix, double *rgb,
                            uint32_t *valid, uint64_t *counters) {
    auto *s = static_cast<Scene<R> *>(sp);
    PathStats st;
    double sec = render(*s, *cam, *prm, spp_begin, spp_count, nthreads, pixels, npix, rgb, valid,
                        counters ? &st : nullptr);
    if (counters) {
        counters[0] = st.closest;
        counters[1] = st.shadow;
        counters[2] = st.trav.nodes;
        counters[3] = st.trav.tris;
    }
    return sec;
}

// Radiance of one path (pixel index, sample index) from an explicit ray.
void OR_FN(oracle_path_radiance)(void *sp, const double *o, const double *d, uint32_t pixel, uint32_t sample,
                                 uint64_t seed, uint32_t max_bounce, double *L) {
    auto *s = static_cast<Scene<R> *>(sp);
    PathRng rng(pixel, sample, seed);
    put(L, radiance(*s, Ray<R>{dv(o), dv(d)}, rng, max_bounce, nullptr));
}

// The profile slice of path (x, y, k) of a frame (Radiance kernel): rec
// receives `depth` records of 14 values (ray o, d, f, L, theta, shader),
// zero beyond the recorded ones; L the path radiance; *bounce the loop count.
void OR_FN(oracle_profile_path)(void *sp, const kr_camera *cam, const kr_params *prm, uint32_t x, uint32_t y,
                                uint32_t k, uint32_t depth, double *rec, double *L, uint32_t *bounce) {
    auto *s = static_cast<Scene<R> *>(sp);
    Lens<R> lens(*cam, prm->width, prm->height, prm->spp);
    PathRng rng(y * prm->width + x, k, prm->seed);
    ProfSlice<R> sl;
    sl.bounces.assign(depth, ProfRecord<R>{Ray<R>{{0, 0, 0}, {0, 0, 0}}, {0, 0, 0}, {0, 0, 0}, R(0), 0u});
    radiance(*s, lens.sample(x, y, k), rng, prm->max_bounce, nullptr, &sl);
    for (uint32_t i = 0; i < depth; ++i) {
        const ProfRecord<R> &b = sl.bounces[i];
        double *o = rec + 14 * size_t(i);
        put(o, b.ray.o);
        put(o + 3, b.ray.d);
        put(o + 6, b.f);
        put(o + 9, b.L);
        o[12] = double(b.theta);
        o[13] = double(b.shader);
    }
    put(L, sl.L);
    *bounce = sl.bounce;
}

int OR_FN(oracle_light_samples)(void *sp, uint32_t light, double *p_n, double *pdf) {
    auto *s = static_cast<Scene<R> *>(sp);
    if (light >= s->lights.size()) return KR_E_INVALID;
    const Prim &p = s->prims[s->lights[light].prim];
    if (p.type != KR_PRIM_MESH) return KR_E_INVALID;
    const Mesh<R> &m = s->meshes[p.index];
    for (int i = 0; i < 1024; ++i) {
        put(p_n + 6 * i, m.samples[i].p);
        put(p_n + 6 * i + 3, m.samples[i].n);
    }
    *pdf = double(m.pdf);
    return KR_OK;
}

void OR_FN(oracle_lens_ray)(const kr_camera *cam, uint32_t w, uint32_t h, uint32_t spp, uint32_t x, uint32_t y,
                            uint32_t k, double *o, double *d) {
    Lens<R> lens(*cam, w, h, spp);
    Ray<R> r = lens.sample(x, y, k);
    put(o, r.o);
    put(d, r.d);
}

// ---- component functions (known-answer tests) ----
int OR_FN(oracle_ray_bbox)(const double *o, const double *d, const double *mn, const double *mx) {
    Ray<R> r{dv(o), dv(d)};
    return ray_bbox(r, dv(mn), dv(mx)) ? 1 : 0;
}
int OR_FN(oracle_ray_triangle)(const double *o, const double *d, const double *v0, const double *e1, const double *e2,
                               double *tuv) {
    Ray<R> r{dv(o), dv(d)};
    Hit<R> h;
    int ok = ray_triangle(r, dv(v0), dv(e1), dv(e2), h) ? 1 : 0;
    tuv[0] = double(h.t);
    tuv[1] = double(h.u);
    tuv[2] = double(h.v);
    return ok;
}
int OR_FN(oracle_seg_bbox)(const double *o, const double *d, double t, const double *mn, const double *mx) {
    Seg<R> s{dv(o), dv(d), R(t)};
    return seg_bbox(s, dv(mn), dv(mx)) ? 1 : 0;
}
int OR_FN(oracle_seg_triangle)(const double *o, const double *d, double t, const double *v0, const double *e1,
                               const double *e2) {
    Seg<R> s{dv(o), dv(d), R(t)};
    return seg_triangle(s, dv(v0), dv(e1), dv(e2)) ? 1 : 0;
}
int OR_FN(oracle_sphere_intersect)(const double *c, double r, const double *o, const double *d, double *t, double *n) {
    Sphere<R> s{dv(c), R(r) * R(r)};
    Hit<R> h;
    int ok = Scene<R>::sphere_intersect(s, Ray<R>{dv(o), dv(d)}, h) ? 1 : 0;
    *t = double(h.t);
    put(n, h.data);
    return ok;
}
void OR_FN(oracle_sphere_sample)(const double *c, double r, double u0, double u1, double *p, double *n, double *pdf) {
    R r2 = R(r) * R(r);
    V3<R> w = from_euler(R(u0) * Const<R>::pi(), R(u1) * Const<R>::pi() * R(2));
    put(p, dv(c) + std::sqrt(r2) * w);
    put(n, w);
    *pdf = double(R(1) / (R(4) * Const<R>::pi() * r2));
}
void OR_FN(oracle_sample_cos)(double u0, double u1, double *w) { put(w, sample_cos(R(u0), R(u1))); }
void OR_FN(oracle_beckmann_sample)(double u0, double u1, double alpha2, double *w) {
    put(w, beckmann_sample(R(u0), R(u1), R(alpha2)));
}
double OR_FN(oracle_beckmann_f)(const double *w, double alpha2) { return double(beckmann_f(dv(w), R(alpha2))); }
void OR_FN(oracle_reflect)(const double *i, const double *n, double *out) { put(out, reflect(dv(i), dv(n))); }

// Shader sample/eval in the local frame: f and emission out.
// Leaf shaders only; uc is the Glass branch uniform. wi_in is the wi a
// failed refraction leaves in place.
void OR_FN(oracle_shader_sample)(const kr_shader_desc *sh, const double *wo, double u0, double u1, double u2,
                                 double uc, const double *wi_in, double *wi, double *f) {
    Vary<R> vy;
    vy.wo = dv(wo);
    if (wi_in) vy.wi = dv(wi_in);
    shader_sample(*sh, vy, R(u0), R(u1), R(u2), R(uc));
    put(wi, vy.wi);
    put(f, vy.f);
}
void OR_FN(oracle_shader_eval)(const kr_shader_desc *sh, const double *wo, const double *wi, double uc, double *f,
                               double *emission) {
    Vary<R> vy;
    vy.wo = dv(wo);
    vy.wi = dv(wi);
    shader_eval(*sh, vy, R(uc));
    put(f, vy.f);
    put(emission, vy.emission);
}

double OR_FN(oracle_fresnel)(double ior, const double *wo) {
    Vary<R> vy;
    vy.wo = dv(wo);
    return double(fresnel(R(ior), vy));
}
int OR_FN(oracle_refract)(const double *i, const double *n, double eta, double *out) {
    V3<R> o{0, 0, 0};
    bool ok = refract(dv(i), o, dv(n), R(eta));
    put(out, o);
    return ok ? 1 : 0;
}

// Frame of a hit: u = normalize(n x (0,1,0)), v = normalize(n x u); local/global.
void OR_FN(oracle_frame)(const double *n, const double *w, double *u, double *v, double *local, double *global) {
    Vary<R> vy;
    vy.n = dv(n);
    vy.u = normalize(cross(vy.n, V3<R>{0, 1, 0}));
    vy.v = normalize(cross(vy.n, vy.u));
    put(u, vy.u);
    put(v, vy.v);
    put(local, vy.local(dv(w)));
    put(global, vy.global(dv(w)));
}

void OR_FN(oracle_emit_seg)(const double *p, const double *q, double *o, double *d, double *t) {
    Vary<R> vy;
    vy.p = dv(p);
    Seg<R> s = vy.emit_seg(dv(q));
    put(o, s.o);
    put(d, s.d);
    *t = double(s.t);
}

#ifdef ORACLE_WITH_PHILOX
void oracle_philox(const uint32_t *ctr, const uint32_t *key, uint32_t *out) { philox4x32_10(ctr, key, out); }
float oracle_u01(uint32_t x) { return u01(x); }
#endif

}  // extern "C"
