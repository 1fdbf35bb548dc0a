// ref_shim.cpp -- C entry points over the reference's own hot-path headers.
//
// TEST INFRASTRUCTURE. Built by `make -C oracle ref` against the reference
// sources where they lie (REF=/root/reference by default, -Iinclude -Isrc),
// into oracle/_ref/libkoishi_ref.so. It contains no reference source: it
// includes the reference's headers and forwards fixed inputs to them, so the
// oracle's restatement can be checked against the reference's own arithmetic.
// Only the header-only part of the path builds here: mesh.cu and scene.cu
// need assimp, which is absent, so BVH build/traversal and Scene are pinned by
// other means (brute force, invariants).
#include <core/basic/ray.hpp>
#include <core/meta/varyings.hpp>
#include <core/misc/lens.hpp>
#include <core/primitive/sphere.hpp>
#include <ext/distribution/beckmann.hpp>
#include <ext/distribution/cosine.hpp>
#include <ext/scala/fresnel.hpp>
#include <util/hemisphere.hpp>

using namespace koishi;
using namespace koishi::core;

namespace {
inline float3 f3(const double *p) { return float3{float(p[0]), float(p[1]), float(p[2])}; }
inline void put(double *p, const float3 &v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}
}  // namespace

extern "C" {

int ref_ray_bbox(const double *o, const double *d, const double *mn, const double *mx) {
    Ray r;
    r.o = f3(o);
    r.d = normalized_float3(f3(d));
    return r.intersect_bbox(f3(mn), f3(mx)) ? 1 : 0;
}

int ref_ray_triangle(const double *o, const double *d, const double *v0, const double *e1, const double *e2,
                     double *tuv) {
    Ray r;
    r.o = f3(o);
    r.d = normalized_float3(f3(d));
    Hit h;
    int ok = r.intersect_triangle(f3(v0), f3(e1), f3(e2), h) ? 1 : 0;
    tuv[0] = h.t;
    tuv[1] = h.uv.x;
    tuv[2] = h.uv.y;
    return ok;
}

int ref_seg_bbox(const double *o, const double *d, double t, const double *mn, const double *mx) {
    Seg s;
    s.o = f3(o);
    s.d = normalized_float3(f3(d));
    s.t = float(t);
    return s.intersect_bbox(f3(mn), f3(mx)) ? 1 : 0;
}

int ref_seg_triangle(const double *o, const double *d, double t, const double *v0, const double *e1,
                     const double *e2) {
    Seg s;
    s.o = f3(o);
    s.d = normalized_float3(f3(d));
    s.t = float(t);
    return s.intersect_triangle(f3(v0), f3(e1), f3(e2)) ? 1 : 0;
}

int ref_sphere_intersect(const double *c, double r, const double *o, const double *d, double *t, double *n) {
    Sphere s(f3(c), float(r), 0);
    char pool_mem[16];
    HybridAllocator pool(pool_mem, sizeof(pool_mem));
    Ray ray;
    ray.o = f3(o);
    ray.d = normalized_float3(f3(d));
    Hit h;
    int ok = s.intersect(ray, h, pool) ? 1 : 0;
    *t = h.t;
    put(n, h.data);
    return ok;
}

void ref_sphere_sample(const double *c, double r, double u0, double u1, double *p, double *n, double *pdf) {
    Sphere s(f3(c), float(r), 0);
    float fp;
    LocalVaryings lv = s.sample(float2{float(u0), float(u1)}, fp);
    put(p, lv.p);
    put(n, lv.n);
    *pdf = fp;
}

void ref_sample_cos(double u0, double u1, double *w) { put(w, H::sampleCos(float2{float(u0), float(u1)})); }

void ref_beckmann_sample(double u0, double u1, double roughness, double *w, double *pdf) {
    Properties props;
    props["roughness"] = float(roughness);
    ext::BeckmannDistribution dist(props);
    float fp;
    put(w, dist.sample(float3{float(u0), float(u1), 0.f}, fp));
    *pdf = fp;
}

double ref_beckmann_f(const double *w, double roughness) {
    Properties props;
    props["roughness"] = float(roughness);
    ext::BeckmannDistribution dist(props);
    return dist.f(solid(f3(w))).x;
}

double ref_cos_f(const double *w) {
    Properties props;
    ext::CosDistribution dist(props);
    return dist.f(solid(f3(w))).x;
}

int ref_refract(const double *i, const double *n, double eta, double *out) {
    solid r(float3{0, 0, 1});
    bool ok = refract(solid(f3(i)), r, solid(f3(n)), float(eta));
    put(out, ok ? float3(r) : float3{0, 0, 0});
    return ok ? 1 : 0;
}

double ref_fresnel(double ior, const double *wo) {
    char buf[32];
    HybridAllocator al(buf, sizeof buf);
    Varyings vy;
    vy.wo = solid(f3(wo));
    ext::Fresnel fr(Properties{{"ior", float(ior)}});
    return fr.compute(vy, al);
}

void ref_reflect(const double *i, const double *n, double *out) {
    put(out, reflect(f3(i), normalized_float3(f3(n))));
}

void ref_lens_ray(int lens, double fovx, const double *pos, const double *target, const double *up, unsigned w,
                  unsigned h, unsigned spp, unsigned x, unsigned y, unsigned k, double *o, double *d) {
    CameraConfig cc;
    cc.lens = lens;
    cc.fovx = float(fovx);
    cc.position = f3(pos);
    cc.target = f3(target);
    cc.upaxis = f3(up);
    Ray r;
    if (lens == 0) {
        PinholeLens l(cc, w, h, spp);
        r = l.sample(x, y, k);
    } else {
        OrthographicLens l(cc, w, h, spp);
        r = l.sample(x, y, k);
    }
    put(o, r.o);
    put(d, r.d);
}

// Frame of Scene::intersect (src/core/meta/scene.hpp:48-50) and the local /
// global transforms of LocalVaryings.
void ref_frame(const double *n, const double *w, double *u, double *v, double *local, double *global) {
    LocalVaryings lv;
    lv.n = normalized_float3(f3(n));
    lv.u = normalize(cross(lv.n, float3{0, 1, 0}));
    lv.v = normalize(cross(lv.n, lv.u));
    put(u, lv.u);
    put(v, lv.v);
    put(local, lv.local(normalized_float3(f3(w))));
    put(global, lv.global(solid(f3(w))));
}

void ref_emit_seg(const double *p, const double *q, double *o, double *d, double *t) {
    LocalVaryings lv;
    lv.p = f3(p);
    Seg s = lv.emitSeg(f3(q));
    put(o, s.o);
    put(d, s.d);
    *t = s.t;
}

void ref_emit_ray(const double *p, const double *w, double *o, double *d) {
    LocalVaryings lv;
    lv.p = f3(p);
    Ray r = lv.emitRay(solid(f3(w)));
    put(o, r.o);
    put(d, r.d);
}

}  // extern "C"
