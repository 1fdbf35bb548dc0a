// koishi_oracle.hpp -- CPU restatement of the reference Radiance path.
//
// TEST INFRASTRUCTURE. This is the oracle that the HIP path is checked
// against. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it; the product library never links or calls it.
//
// It restates, templated on the scalar type R, the reference's CPU path:
//   RadianceKernel::execute   src/core/kernel/radiance.cu:14-81
//   Scene::intersect          src/core/meta/scene.hpp:35-67
//   LocalVaryings             src/core/meta/varyings.hpp:9-50
//   Ray / Seg / Hit           src/core/basic/ray.hpp:11-98
//   Sphere                    src/core/primitive/sphere.hpp:9-59
//   Mesh + BVH build          src/core/primitive/mesh.hpp:61-98, mesh.cu:23-330
//   AreaLight                 src/core/light/areaLight.hpp:19-31
//   Primitive::sample(p,u)    src/core/primitive/primitive.hpp:25-33
//   Lambert/Glossy/Emission   src/ext/shader/{lambert,glossy,emission}.hpp
//   Cosine/Beckmann           src/ext/distribution/{cosine,beckmann}.hpp
//   H:: helpers               src/util/hemisphere.hpp:14-95
//   Pinhole/Orthographic lens src/core/misc/lens.hpp:13-66
//   CPUMulticoreTracer        src/core/tracer/cpuMulticore.cu:14-122
// at R = double (the FP64 oracle) and R = float (the faithful CPU baseline,
// built with the reference's release flags).
//
// Deliberate replacement: the reference's Sampler is a time-seeded LCG shared
// by all threads (src/core/misc/sampler.hpp:22-36), so no run is
// reproducible. Here every random number is Philox4x32-10 (Salmon, Moraes,
// Dror, Shaw, SC'11) of a counter (pixel, sample, bounce, stream) under the
// key `seed`, and the GPU path draws the very same numbers. The order of
// draws per bounce is the reference's: light pick (1), light point (2), BSDF
// sample3 (3), Russian roulette (1).
//
// Pins (what checks each part): tests/test_oracle_*.py.
#pragma once

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

#include "kr_api.h"

namespace koishi_oracle {
namespace {  // internal linkage: each TU (f32 fast-math, f64 strict) keeps its own copy

// ---------------------------------------------------------------- vectors --
template <class R>
struct V3 {
    R x, y, z;
};
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> inline V3<R> operator*(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline V3<R> operator/(V3<R> a, V3<R> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class R> inline V3<R> &operator+=(V3<R> &a, V3<R> b) { a = a + b; return a; }
template <class R> inline V3<R> &operator*=(V3<R> &a, V3<R> b) { a = a * b; return a; }
template <class R> inline V3<R> &operator/=(V3<R> &a, R s) { a = a / s; return a; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// normalize(a) = a * rsqrt(|a|^2), rsqrt = 1/sqrt  (src/vec/vmath.hpp:48-52,648-652)
template <class R> inline V3<R> normalize(V3<R> a) { return a * (R(1) / std::sqrt(dot(a, a))); }
template <class R> inline R length(V3<R> a) { return std::sqrt(dot(a, a)); }
// __func::min/max are `a < b ? a : b` / `a > b ? a : b` (src/vec/vmath.hpp:21-31);
// the NaN behaviour of the reference depends on this exact form.
template <class R> inline R smin(R a, R b) { return a < b ? a : b; }
template <class R> inline R smax(R a, R b) { return a > b ? a : b; }
template <class R> inline V3<R> vmin(V3<R> a, V3<R> b) { return {smin(a.x, b.x), smin(a.y, b.y), smin(a.z, b.z)}; }
template <class R> inline V3<R> vmax(V3<R> a, V3<R> b) { return {smax(a.x, b.x), smax(a.y, b.y), smax(a.z, b.z)}; }
template <class R> inline V3<R> vabs(V3<R> a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
template <class R> inline bool hasnan(V3<R> a) { return a.x != a.x || a.y != a.y || a.z != a.z; }
template <class R> inline bool hasinf(V3<R> a) { return std::isinf(a.x) || std::isinf(a.y) || std::isinf(a.z); }
template <class R> inline V3<R> mk(const float *p) { return {R(p[0]), R(p[1]), R(p[2])}; }

template <class R> struct Const {
    static constexpr R pi() { return R(3.14159265358979323846); }
    static constexpr R inv_pi() { return R(1) / R(3.14159265358979323846); }
    static constexpr R spawn_eps() { return R(4e-3f); }  // src/core/meta/varyings.hpp:22,33
};

// ------------------------------------------------------------------- RNG --
// Philox4x32-10: 10 rounds of the published bijection with multipliers
// 0xD2511F53 / 0xCD9E8D57 and Weyl key increments 0x9E3779B9 / 0xBB67AE85.
inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = uint64_t(0xD2511F53u) * c0;
        uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
        uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
        uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
// 24 high bits -> [0, 1), exact in both float and double.
inline float u01(uint32_t x) { return float(x >> 8) * (1.0f / 16777216.0f); }

// Stream tags (counter word 3).
constexpr uint32_t kStreamPath = 0u;
constexpr uint32_t kStreamLight = 0x4C495445u;  // light sample tables

constexpr uint32_t kStreamChoice = 1u;          // shader choices (Mixed, Glass)

// Per-path random numbers: dimension d < 8 of bounce b is word (d & 3) of
// Philox(ctr = {pixel, sample, 2*b + (d >> 2), kStreamPath}, key = seed);
// d = 8..15 is word (d & 3) of Philox({pixel, sample, 2*b + ((d - 8) >> 2),
// kStreamChoice}). Dimensions: 0 light pick, 1-2 light point, 3-5 BSDF
// sample3, 6 roulette, 8-11 the shader choices (Mixed levels, then Glass's
// Fresnel pick) of compute_f_by_wi_wo, 12-15 those of sample_wi_f_by_wo.
struct PathRng {
    uint32_t pixel, sample;
    uint32_t key[2];
    int cached = -1;
    uint32_t buf[4];
    PathRng(uint32_t px, uint32_t s, uint64_t seed) : pixel(px), sample(s) {
        key[0] = uint32_t(seed);
        key[1] = uint32_t(seed >> 32);
    }
    float get(int bounce, int dim) {
        bool choice = dim >= 8;
        int d = choice ? dim - 8 : dim;
        int block = (2 * bounce + (d >> 2)) * 2 + (choice ? 1 : 0);
        if (block != cached) {
            uint32_t ctr[4] = {pixel, sample, uint32_t(2 * bounce + (d >> 2)), choice ? kStreamChoice : kStreamPath};
            philox4x32_10(ctr, key, buf);
            cached = block;
        }
        return u01(buf[d & 3]);
    }
};

// ------------------------------------------------------------- geometry --
template <class R> struct Ray { V3<R> o, d; };
template <class R> struct Seg { V3<R> o, d; R t; };
template <class R> struct Hit {
    uint32_t id = 0;
    R t = std::numeric_limits<R>::infinity();
    R u = 0, v = 0;
    V3<R> data{0, 0, 0};
    explicit operator bool() const { return t != std::numeric_limits<R>::infinity(); }
};

// Ray::intersect_bbox (src/core/basic/ray.hpp:29-36): slab test, no tmax.
template <class R> inline bool ray_bbox(const Ray<R> &r, V3<R> vmin_, V3<R> vmax_, R *t0out = nullptr) {
    V3<R> invv = {R(1) / r.d.x, R(1) / r.d.y, R(1) / r.d.z};
    V3<R> a = (vmin_ - r.o) * invv, b = (vmax_ - r.o) * invv;
    V3<R> tmn = vmin(a, b), tmx = vmax(a, b);
    R t0 = smax(tmn.x, smax(tmn.y, tmn.z)), t1 = smin(tmx.x, smin(tmx.y, tmx.z));
    if (t0out) *t0out = t0;
    return t0 <= t1 && t1 >= 0;
}

// Ray::intersect_triangle (src/core/basic/ray.hpp:37-52): Moller-Trumbore,
// two-sided, no det guard, hit iff t > 0, u >= 0, v >= 0, u + v <= 1.
template <class R> inline bool ray_triangle(const Ray<R> &r, V3<R> v0, V3<R> e1, V3<R> e2, Hit<R> &hit) {
    V3<R> T = r.o - v0;
    V3<R> P = cross(r.d, e2);
    V3<R> Q = cross(T, e1);
    R det = dot(e1, P);
    R u = dot(T, P), v = dot(r.d, Q), t = dot(e2, Q);
    R invdet = R(1) / det;
    t *= invdet;
    u *= invdet;
    v *= invdet;
    hit.t = t;
    hit.u = u;
    hit.v = v;
    return t > R(0) && u >= R(0) && v >= R(0) && u + v <= R(1);
}

// Seg::intersect_bbox (src/core/basic/ray.hpp:59-79): segment/box
// separating-axis test.
template <class R> inline bool seg_bbox(const Seg<R> &s, V3<R> vmin_, V3<R> vmax_) {
    V3<R> c = (vmin_ + vmax_) * R(.5f);
    V3<R> l = vmax_ - c;
    V3<R> p0 = s.o - c, p1 = p0 + s.d * s.t;
    V3<R> m = (p0 + p1) * R(.5f);
    V3<R> w = m - p0;
    V3<R> W = vabs(w);
    V3<R> x = vabs(m) - W - l;
    if (x.x > 0 || x.y > 0 || x.z > 0) return false;
    if (std::fabs(-w.z * m.y + w.y * m.z) > l.y * W.z + l.z * W.y ||
        std::fabs(-w.z * m.x + w.x * m.z) > l.x * W.z + l.z * W.x ||
        std::fabs(-w.y * m.x + w.x * m.y) > l.x * W.y + l.y * W.x)
        return false;
    return true;
}

// Seg::intersect_triangle (src/core/basic/ray.hpp:81-90): the end points
// must not lie strictly on one side of the plane, then the ray test.
template <class R> inline bool seg_triangle(const Seg<R> &s, V3<R> v0, V3<R> d1, V3<R> d2) {
    V3<R> n = cross(d1, d2);
    if (dot(s.o - v0, n) * dot(s.o + s.d * s.t - v0, n) > 0) return false;
    Hit<R> h;
    Ray<R> r{s.o, s.d};
    return ray_triangle(r, v0, d1, d2, h);
}

// ------------------------------------------------------------ hemisphere --
// H::fromEular (src/util/hemisphere.hpp:66-74): x = sin(th) sin(phi),
// y = sin(th) cos(phi), z = cos(th).
template <class R> inline V3<R> from_euler(R sinth, R costh, R phi) {
    return {sinth * std::sin(phi), sinth * std::cos(phi), costh};
}
template <class R> inline V3<R> from_euler(R th, R phi) { return from_euler(std::sin(th), std::cos(th), phi); }

// H::sampleCos (src/util/hemisphere.hpp:76-95): Shirley-Chiu concentric disk
// lifted to the hemisphere.
template <class R> inline V3<R> sample_cos(R rn0, R rn1) {
    R ux = R(2) * rn0 - R(1), uy = R(2) * rn1 - R(1);
    if (ux == 0 && uy == 0) return {0, 0, 1};
    R r, phi;
    const R pi = Const<R>::pi();
    if (ux * ux > uy * uy) {
        r = ux;
        phi = (pi / R(4)) * (uy / ux);
    } else {
        r = uy;
        phi = (pi / R(2)) - ((pi / R(4)) * (ux / uy));
    }
    return {r * std::cos(phi), r * std::sin(phi), std::sqrt(smax(R(0), R(1) - r * r))};
}

// reflect(I, N) = N * dot(N, I) * 2 - I (src/vec/vmath.hpp:470-473)
template <class R> inline V3<R> reflect(V3<R> I, V3<R> N) { return N * (dot(N, I) * R(2)) - I; }

// BeckmannDistribution::f (src/ext/distribution/beckmann.hpp:19-24),
// returned as the scalar multiplying (1,1,1).
template <class R> inline R beckmann_f(V3<R> w, R alpha2) {
    R cos2 = w.z * w.z;
    R sin2 = smax(R(0), R(1) - cos2);
    R tg2th = sin2 / cos2;
    R cos4th = cos2 * cos2;
    return std::exp(-tg2th / alpha2) / (Const<R>::pi() * alpha2 * cos4th) * std::fabs(w.z);
}
// BeckmannDistribution::sample (src/ext/distribution/beckmann.hpp:26-38)
template <class R> inline V3<R> beckmann_sample(R u0, R u1, R alpha2) {
    R logu = std::log(u0);
    R tg2th = -alpha2 * logu;
    R phi = u1 * R(2) * Const<R>::pi();
    R costh = R(1) / std::sqrt(R(1) + tg2th);
    R sinth = std::sqrt(smax(R(0), R(1) - costh * costh));
    return from_euler(sinth, costh, phi);
}
// CosDistribution::f (src/ext/distribution/cosine.hpp:16-19)
template <class R> inline R cos_f(V3<R> w) { return w.z >= 0 ? Const<R>::inv_pi() : R(0); }

// ------------------------------------------------------------------ lens --
// Lens ctor + PinholeLens/OrthographicLens::sample (src/core/misc/lens.hpp:13-66).
template <class R> struct Lens {
    uint32_t lens, w, h, spp, pw;
    V3<R> o, n, u, v, du, dv;
    Lens(const kr_camera &c, uint32_t w_, uint32_t h_, uint32_t spp_) : lens(c.lens), w(w_), h(h_), spp(spp_) {
        pw = uint32_t(std::ceil(std::sqrt(double(spp))));
        o = mk<R>(c.position);
        double half = double(c.fovx) * .5;
        double den = 2 * std::tan(3.14159265358979323846 * half / 180.);
        n = (normalize(mk<R>(c.target)) * R(float(w))) / R(den);
        u = normalize(cross(n, mk<R>(c.upaxis)));
        v = normalize(cross(n, u));
        du = u / R(pw);
        dv = v / R(pw);
    }
    Ray<R> sample(uint32_t x, uint32_t y, uint32_t k) const {
        R fx = R(R(float(x)) - R(double(w) * .5));
        R fy = R(R(float(y)) - R(double(h) * .5));
        if (lens == KR_LENS_ORTHOGRAPHIC) {
            V3<R> d = u * fx + v * fy;
            d += du * R(float(k % pw)) + dv * R(float(k / pw));
            return {o + d * R(4.f) / R(w), normalize(n)};
        }
        V3<R> t = n + u * fx + v * fy;
        return {o, normalize(t + du * R(float(k % pw)) + dv * R(float(k / pw)))};
    }
};

// ------------------------------------------------------------ reference BVH --
// BVHNode {vmax, vmin, begin, end, offset} (src/core/primitive/mesh.hpp:17-21).
template <class R> struct BNode {
    V3<R> vmax, vmin;
    uint32_t begin, end, offset;
};
template <class R> struct TriInfo {
    uint32_t face;  // index into the padded input face list
    V3<R> vmax, vmin;
};
// area() is the box volume (src/core/primitive/mesh.cu:23-27).
template <class R> inline R box_volume(V3<R> vmin_, V3<R> vmax_) {
    V3<R> e = vmax_ - vmin_;
    return e.x * e.y * e.z;
}

template <class R> struct BvhBuilder {
    std::vector<TriInfo<R>> &info;
    std::vector<BNode<R>> &out;
    // doCreateBVH + traversal/plainlizeBVH (src/core/primitive/mesh.cu:35-133),
    // emitting nodes directly in DFS order: left child = i + 1, offset = distance
    // to the right child, offset 0 = leaf. Leaves hold <= 16 triangles, splits
    // fall on multiples of 4, the first split whose prefix volume reaches the
    // suffix volume wins, else the median.
    void build(size_t b, size_t e) {
        BNode<R> node;
        node.vmax = info[b].vmax;
        node.vmin = info[b].vmin;
        for (size_t i = b; i != e; ++i) {
            node.vmax = vmax(node.vmax, info[i].vmax);
            node.vmin = vmin(node.vmin, info[i].vmin);
        }
        size_t len = e - b;
        node.offset = len > 16;
        node.begin = uint32_t(b);
        node.end = uint32_t(e);
        size_t me = out.size();
        out.push_back(node);
        if (!node.offset) return;
        V3<R> w = node.vmax - node.vmin;
        auto it0 = info.begin() + b, it1 = info.begin() + e;
        if (w.x >= w.y && w.x >= w.z)
            std::sort(it0, it1, [](const TriInfo<R> &a, const TriInfo<R> &c) { return a.vmin.x < c.vmin.x; });
        else if (w.y >= w.x && w.y >= w.z)
            std::sort(it0, it1, [](const TriInfo<R> &a, const TriInfo<R> &c) { return a.vmin.y < c.vmin.y; });
        else
            std::sort(it0, it1, [](const TriInfo<R> &a, const TriInfo<R> &c) { return a.vmin.z < c.vmin.z; });
        std::vector<R> low(len), high(len);
        const R inf = std::numeric_limits<R>::infinity();
        V3<R> mn{inf, inf, inf}, mx{-inf, -inf, -inf};
        for (size_t i = 0; i < len; ++i) {
            mn = vmin(info[b + i].vmin, mn);
            mx = vmax(info[b + i].vmax, mx);
            low[i] = box_volume(mn, mx);
        }
        mn = {inf, inf, inf};
        mx = {-inf, -inf, -inf};
        for (size_t i = len; i-- > 0;) {
            mn = vmin(info[b + i].vmin, mn);
            mx = vmax(info[b + i].vmax, mx);
            high[i] = box_volume(mn, mx);
        }
        size_t split = len / 2;
        for (size_t i = 4; i < len; i += 4) {
            if (low[i - 1] - high[i] >= R(0)) {
                split = i;
                break;
            }
        }
        build(b, b + split);
        out[me].offset = uint32_t(out.size() - me);
        build(b + split, e);
    }
};

// ----------------------------------------------------------------- scene --
template <class R> struct Face { V3<R> o, d1, d2; };
template <class R> struct Nrm { V3<R> n0, n1, n2; };
template <class R> struct LV { V3<R> p, n; };  // the part of LocalVaryings a light sample fills

template <class R> struct Mesh {
    std::vector<Face<R>> faces;     // BVH order, padded to a multiple of 4
    std::vector<Nrm<R>> normals;
    std::vector<uint32_t> face_input;  // BVH slot -> triangle index within the primitive
    std::vector<BNode<R>> bvh;
    std::vector<Face<R>> pad_faces;    // input order + padding (light table source)
    std::vector<Nrm<R>> pad_normals;
    std::vector<LV<R>> samples;        // 1024-entry light table (emissive meshes only)
    R pdf = 1;
};
template <class R> struct Sphere { V3<R> o; R r2; };
struct Prim {
    uint32_t type, index, shader;
};
template <class R> struct Light {
    uint32_t prim;
    V3<R> emission;
    bool point = false;  // PointLight (src/core/light/pointLight.hpp:10-36)
    V3<R> pos{0, 0, 0};
};

struct Counters {
    uint64_t nodes = 0, tris = 0;
};

// The DFS stack of a mesh traversal: a fixed array on the thread's stack,
// as the reference's Stack is a fixed block of its per-thread pool
// (src/core/primitive/mesh.cu:204), instead of a heap vector per query. A
// DFS holds at most one pending right child per tree level; kr_scene_desc
// meshes build trees far shallower than the capacity, which is checked.
struct NodeStack {
    static constexpr int kCap = 256;
    uint32_t v[kCap];
    int top = 0;
    bool empty() const { return top == 0; }
    void push(uint32_t x) {
        if (top == kCap) {
            std::fprintf(stderr, "oracle: BVH traversal stack overflow\n");
            std::abort();
        }
        v[top++] = x;
    }
    uint32_t pop() { return v[--top]; }
};

template <class R> struct Scene {
    std::vector<Mesh<R>> meshes;
    std::vector<Sphere<R>> spheres;
    std::vector<Prim> prims;
    std::vector<kr_shader_desc> shaders;
    std::vector<Light<R>> lights;
    int traversal = 0;  // 0: reference order, no culling; 1: same tree, t-culled

    // Builds one reference Mesh per mesh primitive (src/core/primitive/mesh.cu:135-198):
    // pad the triangle list to a multiple of 4 with copies of the last one,
    // build, and store faces as {o, d1 = v1 - o, d2 = v2 - o} in BVH order.
    void add_mesh(const kr_scene_desc &d, const kr_prim_desc &p) {
        Mesh<R> m;
        std::vector<TriInfo<R>> info;
        uint32_t n = p.tri_count;
        uint32_t padded = (n + 3) / 4 * 4;
        auto vert = [&](uint32_t vi) { return mk<R>(d.positions + 3 * size_t(vi)); };
        auto nrm = [&](uint32_t vi) { return mk<R>(d.normals + 3 * size_t(vi)); };
        for (uint32_t j = 0; j < padded; ++j) {
            uint32_t t = p.tri_begin + (j < n ? j : n - 1);
            const uint32_t *ix = d.indices + 3 * size_t(t);
            V3<R> a = vert(ix[0]), b = vert(ix[1]), c = vert(ix[2]);
            m.pad_faces.push_back({a, b - a, c - a});
            m.pad_normals.push_back({nrm(ix[0]), nrm(ix[1]), nrm(ix[2])});
            info.push_back({j, vmax(a, vmax(b, c)), vmin(a, vmin(b, c))});
        }
        BvhBuilder<R> bb{info, m.bvh};
        bb.build(0, info.size());
        for (auto &ti : info) {
            m.faces.push_back(m.pad_faces[ti.face]);
            m.normals.push_back(m.pad_normals[ti.face]);
            m.face_input.push_back(ti.face < n ? ti.face : n - 1);
        }
        prims.push_back({KR_PRIM_MESH, uint32_t(meshes.size()), p.shader});
        meshes.push_back(std::move(m));
    }

    bool build(const kr_scene_desc &d) {
        shaders.assign(d.shaders, d.shaders + d.num_shaders);
        // "pointLight" assets join the light list while the assets are read,
        // before area lights are promoted (src/core/meta/scene.cu:158-168)
        for (uint32_t i = 0; i < d.num_point_lights; ++i) {
            Light<R> L;
            L.prim = ~0u;
            L.point = true;
            L.pos = mk<R>(d.point_lights[i].position);
            L.emission = mk<R>(d.point_lights[i].emission);
            lights.push_back(L);
        }
        for (uint32_t i = 0; i < d.num_prims; ++i) {
            const kr_prim_desc &p = d.prims[i];
            if (p.shader >= d.num_shaders) return false;
            if (p.type == KR_PRIM_MESH) {
                if (p.tri_count == 0) return false;
                add_mesh(d, p);
            } else {
                Sphere<R> s;
                s.o = mk<R>(p.center);
                // Sphere ctor stores r2 = r * r in float (src/core/primitive/sphere.hpp:51-56)
                s.r2 = R(p.radius) * R(p.radius);
                prims.push_back({KR_PRIM_SPHERE, uint32_t(spheres.size()), p.shader});
                spheres.push_back(s);
            }
        }
        // Light promotion (src/core/meta/scene.cu:200-216): every primitive with an
        // Emission shader becomes an AreaLight with E = color * strength
        // (src/ext/shader/emission.hpp:12-19), in primitive order.
        for (uint32_t i = 0; i < prims.size() && !(d.flags & KR_SCENE_NO_AREA_LIGHTS); ++i) {
            const kr_shader_desc &sh = shaders[prims[i].shader];
            if (sh.type != KR_SHADER_EMISSION) continue;
            V3<R> e = mk<R>(sh.color) * R(sh.strength);
            Light<R> L;
            L.prim = i;
            L.emission = e;
            lights.push_back(L);
            if (prims[i].type == KR_PRIM_MESH) gen_samples(meshes[prims[i].index], uint32_t(lights.size() - 1), d.light_seed);
        }
        return true;
    }

    // Mesh::generateSamples (src/core/primitive/mesh.hpp:61-85): 1024 points,
    // triangle picked uniformly by index over the padded face list, uv =
    // (min, max) of two uniforms, p = o + d1 uv.x + d2 uv.y (may leave the
    // triangle), n = normalize(interplot(n0, n1, n2, uv)); pdf = 1 / sum of
    // the padded faces' areas. The uniforms are Philox(ctr = {light, i, 0,
    // kStreamLight}, key = light_seed) words 0, 1, 2.
    void gen_samples(Mesh<R> &m, uint32_t light, uint64_t seed) {
        uint32_t key[2] = {uint32_t(seed), uint32_t(seed >> 32)};
        size_t F = m.pad_faces.size();
        m.samples.resize(1024);
        for (uint32_t i = 0; i < 1024; ++i) {
            uint32_t ctr[4] = {light, i, 0, kStreamLight}, o[4];
            philox4x32_10(ctr, key, o);
            R xi = R(u01(o[0]));
            uint32_t idx = smin(uint32_t(std::floor(xi * R(F))), uint32_t(F - 1));
            R a = R(u01(o[1])), b = R(u01(o[2]));
            R ux = smin(a, b), uy = smax(a, b);
            const Face<R> &f = m.pad_faces[idx];
            const Nrm<R> &nn = m.pad_normals[idx];
            m.samples[i].p = f.o + f.d1 * ux + f.d2 * uy;
            m.samples[i].n = normalize(nn.n0 * (R(1) - ux - uy) + nn.n1 * ux + nn.n2 * uy);
        }
        R s = 0;
        for (auto &f : m.pad_faces) s += length(cross(f.d1, f.d2)) * R(.5);
        m.pdf = R(1) / s;
    }

    // Mesh::intersect(Ray) (src/core/primitive/mesh.cu:200-267): DFS, both
    // children tested, left first, right pushed when both hit; leaf faces keep
    // the strictly closer hit. traversal == 1 adds a conservative t-cull, which
    // only skips boxes that cannot hold a strictly closer hit.
    bool mesh_intersect(const Mesh<R> &m, const Ray<R> &ray, Hit<R> &hit, Counters *cnt) const {
        hit.t = std::numeric_limits<R>::infinity();
        NodeStack Q;
        Q.push(0);
        const R slack = R(1) + R(64) * std::numeric_limits<R>::epsilon();
        while (!Q.empty()) {
            uint32_t i = Q.pop();
            bool dead = false;
            while (uint32_t off = m.bvh[i].offset) {
                R t0l, t0r;
                int left = ray_bbox(ray, m.bvh[i + 1].vmin, m.bvh[i + 1].vmax, &t0l);
                int right = ray_bbox(ray, m.bvh[i + off].vmin, m.bvh[i + off].vmax, &t0r);
                if (cnt) cnt->nodes += 2;
                if (traversal == 1) {
                    if (left && t0l > hit.t * slack) left = 0;
                    if (right && t0r > hit.t * slack) right = 0;
                }
                if (!left && !right) {
                    dead = true;
                    break;
                }
                if (left && right) Q.push(i + off);
                i = left ? i + 1 : i + off;
            }
            if (dead) continue;
            for (uint32_t j = m.bvh[i].begin; j < m.bvh[i].end; ++j) {
                Hit<R> h1;
                const Face<R> &f = m.faces[j];
                if (cnt) cnt->tris++;
                if (ray_triangle(ray, f.o, f.d1, f.d2, h1) && h1.t < hit.t) {
                    hit = h1;
                    hit.id = j;
                }
            }
        }
        return hit.t != std::numeric_limits<R>::infinity();
    }

    // Mesh::intersect(Seg) (src/core/primitive/mesh.cu:269-330): any hit.
    bool mesh_occluded(const Mesh<R> &m, const Seg<R> &seg) const {
        NodeStack Q;
        Q.push(0);
        while (!Q.empty()) {
            uint32_t i = Q.pop();
            bool dead = false;
            while (uint32_t off = m.bvh[i].offset) {
                int left = seg_bbox(seg, m.bvh[i + 1].vmin, m.bvh[i + 1].vmax);
                int right = seg_bbox(seg, m.bvh[i + off].vmin, m.bvh[i + off].vmax);
                if (!left && !right) {
                    dead = true;
                    break;
                }
                if (left && right) Q.push(i + off);
                i = left ? i + 1 : i + off;
            }
            if (dead) continue;
            for (uint32_t j = m.bvh[i].begin; j < m.bvh[i].end; ++j) {
                const Face<R> &f = m.faces[j];
                if (seg_triangle(seg, f.o, f.d1, f.d2)) return true;
            }
        }
        return false;
    }

    // Sphere::intersect (src/core/primitive/sphere.hpp:31-54).
    static bool sphere_intersect(const Sphere<R> &s, const Ray<R> &ray, Hit<R> &hit) {
        V3<R> oc = s.o - ray.o;
        R poc = dot(ray.d, oc);
        R oc2 = dot(oc, oc);
        R d2 = oc2 - poc * poc;
        R det = s.r2 - d2;
        if (det < 0) return false;
        det = std::sqrt(det);
        if ((hit.t = poc - det) < 0) hit.t = poc + det;
        hit.data = normalize(ray.o + hit.t * ray.d - s.o);
        return hit.t >= 0;
    }

    struct Closest {
        bool hit = false;
        uint32_t prim = KR_MISS, tri = 0;
        Hit<R> h;
    };

    // Scene::intersect(Ray) primitive loop (src/core/meta/scene.hpp:35-47):
    // the strictly closer primitive wins.
    Closest closest(const Ray<R> &r, Counters *cnt = nullptr) const {
        Closest c;
        for (uint32_t pi = 0; pi < prims.size(); ++pi) {
            Hit<R> h1;
            bool ok;
            if (prims[pi].type == KR_PRIM_MESH)
                ok = mesh_intersect(meshes[prims[pi].index], r, h1, cnt);
            else
                ok = sphere_intersect(spheres[prims[pi].index], r, h1);
            if (ok && h1.t < c.h.t) {
                c.h = h1;
                c.prim = pi;
                c.hit = true;
            }
        }
        if (c.hit && prims[c.prim].type == KR_PRIM_MESH) c.tri = meshes[prims[c.prim].index].face_input[c.h.id];
        return c;
    }

    // Scene::intersect(Seg) (src/core/meta/scene.hpp:57-67); spheres use the
    // default Primitive::intersect(Seg): ray hit with t <= seg.t
    // (src/core/primitive/primitive.hpp:16-21).
    bool occluded(const Seg<R> &s) const {
        for (auto &p : prims) {
            if (p.type == KR_PRIM_MESH) {
                if (mesh_occluded(meshes[p.index], s)) return true;
            } else {
                Hit<R> h;
                Ray<R> r{s.o, s.d};
                if (sphere_intersect(spheres[p.index], r, h) && h.t <= s.t) return true;
            }
        }
        return false;
    }
};

// Varyings (src/core/meta/varyings.hpp:9-50) with the frame of
// Scene::intersect (src/core/meta/scene.hpp:48-53).
template <class R> struct Vary {
    V3<R> u, v, n, p, wo, wi{0, 0, 0}, f{0, 0, 0}, emission{0, 0, 0};
    uint32_t shader = 0;
    Ray<R> emit_ray(V3<R> w) const { return {p + w * Const<R>::spawn_eps(), w}; }
    Seg<R> emit_seg(V3<R> q) const {
        Seg<R> s;
        V3<R> d = q - p;
        s.d = normalize(d);
        V3<R> nwo = normalize(s.d);
        V3<R> diff = nwo * Const<R>::spawn_eps();
        s.o = p + diff;
        s.t = length(d - R(2) * diff);
        return s;
    }
    V3<R> local(V3<R> w) const { return normalize(V3<R>{dot(w, u), dot(w, v), dot(w, n)}); }
    V3<R> global(V3<R> w) const { return normalize(w.x * u + w.y * v + w.z * n); }
};

template <class R> inline bool scene_intersect(const Scene<R> &s, const Ray<R> &r, Vary<R> &vy, Counters *cnt) {
    auto c = s.closest(r, cnt);
    if (!c.hit) return false;
    const Prim &p = s.prims[c.prim];
    if (p.type == KR_PRIM_MESH) {
        // Mesh::normal + interplot (src/core/primitive/mesh.hpp:85-91, ray.hpp:93-98)
        const Nrm<R> &nn = s.meshes[p.index].normals[c.h.id];
        vy.n = normalize(nn.n0 * (R(1) - c.h.u - c.h.v) + nn.n1 * c.h.u + nn.n2 * c.h.v);
    } else {
        vy.n = c.h.data;
    }
    vy.u = normalize(cross(vy.n, V3<R>{0, 1, 0}));
    vy.v = normalize(cross(vy.n, vy.u));
    vy.p = r.o + r.d * c.h.t;
    vy.wo = vy.local(-r.d);
    vy.shader = p.shader;
    return true;
}

// Scala<float>::compute for Mixed's fac: Constant (constant.hpp:15-20) or
// Fresnel (fresnel.hpp:15-30), the dielectric reflectance of wo.
template <class R> inline R fresnel(R ior, const Vary<R> &vy) {
    R cosi = vy.wo.z;
    R e = cosi > R(0) ? R(1) / ior : ior;
    cosi = std::fabs(cosi);
    R sini = std::sqrt(smax(R(0), R(1) - cosi * cosi));
    R sinr = e * sini;
    if (sinr >= R(1)) return R(1);
    R cosr = std::sqrt(smax(R(0), R(1) - sinr * sinr));
    R cosior = cosi / cosr;
    R rparl = (cosior - e) / (cosior + e);
    R rperp = (e * cosior - R(1)) / (e * cosior + R(1));
    return R(.5f) * (rparl * rparl + rperp * rperp);
}
template <class R> inline R scala(const kr_shader_desc &sh, const Vary<R> &vy) {
    return sh.fac_type == KR_SCALA_FRESNEL ? fresnel(R(sh.fac_value), vy) : R(sh.fac_value);
}

// Mixed::execute (src/ext/shader/mixed.hpp:22-27) down to a leaf shader:
// one draw per Mixed level, b when u < fac. `dim` is the next choice
// dimension of this execute call and advances with every draw.
template <class R>
inline uint32_t resolve_shader(const std::vector<kr_shader_desc> &shaders, uint32_t idx, const Vary<R> &vy,
                               PathRng &rng, int bounce, int &dim) {
    while (shaders[idx].type == KR_SHADER_MIXED) {
        const kr_shader_desc &m = shaders[idx];
        R f = scala(m, vy);
        R u = R(rng.get(bounce, dim++));
        idx = u < f ? m.child[1] : m.child[0];
    }
    return idx;
}

template <class R> inline R dist_f(const kr_shader_desc &sh, V3<R> w) {
    R a2 = R(sh.roughness) * R(sh.roughness);
    return sh.distribution == KR_DIST_COSINE ? cos_f(w) : beckmann_f(w, a2);
}
template <class R> inline V3<R> dist_sample(const kr_shader_desc &sh, R u0, R u1) {
    R a2 = R(sh.roughness) * R(sh.roughness);
    return sh.distribution == KR_DIST_COSINE ? sample_cos(u0, u1) : beckmann_sample(u0, u1, a2);
}

// refract (src/vec/vmath.hpp:495-503): Snell through the microfacet normal N.
template <class R> inline bool refract(V3<R> I, V3<R> &out, V3<R> N, R eta) {
    R costh = dot(N, I);
    R sinth2 = smax(R(0), R(1) - costh * costh);
    R sinphi2 = sinth2 * eta * eta;
    if (sinphi2 >= R(1)) return false;
    R cosphi = std::sqrt(R(1) - sinphi2);
    out = normalize(eta * (-I) + (eta * costh - cosphi) * N);
    return true;
}

// Refraction branch of RefractionBSDF / GlassBSDF sample_wi_f_by_wo
// (refraction.hpp:24-34, glass.hpp:37-46); wi is left as it was when the
// refraction fails (f = 0).
template <class R> inline void refraction_sample(const kr_shader_desc &sh, Vary<R> &vy, R u0, R u1) {
    V3<R> h = dist_sample(sh, u0, u1);
    R ior = R(sh.ior);
    R e = vy.wo.z > R(0) ? R(1) / ior : ior;
    if (vy.wo.z < R(0)) h = -h;
    if (refract(vy.wo, vy.wi, h, e))
        vy.f = mk<R>(sh.color) / (e < R(1) ? e * e : R(1));
    else
        vy.f = V3<R>{0, 0, 0};
}

// Shader::execute(compute_f_by_wi_wo) of a leaf: Lambert (lambert.hpp:31-33),
// Glossy (glossy.hpp:28-30), Emission = emission + white Lambert
// (emission.hpp:22-27), GlassBSDF (glass.hpp:27-34, 48-51: reflection when
// u < Fresnel(ior)), RefractionBSDF / Transparent / Blackhole: f = 0
// (refraction.hpp:35-38, transparent.hpp:24-26, blackhole.hpp:18-21).
// uc: the Glass branch uniform (its draw follows the Mixed draws).
template <class R> inline void shader_eval(const kr_shader_desc &sh, Vary<R> &vy, R uc) {
    if (sh.type == KR_SHADER_EMISSION) {
        vy.emission = mk<R>(sh.color) * R(sh.strength);
        vy.f = V3<R>{1, 1, 1} * cos_f(vy.wi);
    } else if (sh.type == KR_SHADER_LAMBERT) {
        vy.f = mk<R>(sh.color) * mk<R>(sh.R) * cos_f(vy.wi);
    } else if (sh.type == KR_SHADER_GLOSSY) {
        vy.f = mk<R>(sh.color) * dist_f(sh, normalize(vy.wo + vy.wi));
    } else if (sh.type == KR_SHADER_GLASS) {
        if (uc < fresnel(R(sh.ior), vy))
            vy.f = mk<R>(sh.color) * dist_f(sh, normalize(vy.wo + vy.wi));
        else
            vy.f = mk<R>(sh.color) * dist_f(sh, normalize(vy.wo + vy.wi * R(sh.ior)));
    } else {
        vy.f = V3<R>{0, 0, 0};
    }
}

// Shader::execute(sample_wi_f_by_wo) of a leaf with u = sample3(). Lambert:
// wi from CosDistribution with pdf 1/pi, f = color R f(wi) / pdf
// (lambert.hpp:26-30). Glossy: wi = reflect(wo, h), f = color when wo, wi lie
// on one side (glossy.hpp:24-27, H::isSame hemisphere.hpp:14-17). GlassBSDF:
// that reflection when u < Fresnel(ior), else the refraction branch
// (glass.hpp:22-46). Transparent: wi = -wo, f = color (transparent.hpp:20-23).
// Blackhole: wi = (0, 0, 1), f = 0 (blackhole.hpp:14-21).
template <class R>
inline void shader_sample(const kr_shader_desc &sh, Vary<R> &vy, R u0, R u1, R u2, R uc) {
    (void)u2;
    if (sh.type == KR_SHADER_EMISSION || sh.type == KR_SHADER_LAMBERT) {
        V3<R> c = sh.type == KR_SHADER_EMISSION ? V3<R>{1, 1, 1} : mk<R>(sh.color) * mk<R>(sh.R);
        if (sh.type == KR_SHADER_EMISSION) vy.emission = mk<R>(sh.color) * R(sh.strength);
        vy.wi = sample_cos(u0, u1);
        R pdf = Const<R>::inv_pi();
        vy.f = c * cos_f(vy.wi) / pdf;
    } else if (sh.type == KR_SHADER_GLOSSY) {
        V3<R> h = dist_sample(sh, u0, u1);
        vy.wi = reflect(vy.wo, h);
        vy.f = vy.wo.z * vy.wi.z > 0 ? mk<R>(sh.color) : V3<R>{0, 0, 0};
    } else if (sh.type == KR_SHADER_GLASS) {
        if (uc < fresnel(R(sh.ior), vy)) {
            V3<R> h = dist_sample(sh, u0, u1);
            vy.wi = reflect(vy.wo, h);
            vy.f = vy.wo.z * vy.wi.z > 0 ? mk<R>(sh.color) : V3<R>{0, 0, 0};
        } else {
            refraction_sample(sh, vy, u0, u1);
        }
    } else if (sh.type == KR_SHADER_REFRACTION) {
        refraction_sample(sh, vy, u0, u1);
    } else if (sh.type == KR_SHADER_TRANSPARENT) {
        vy.wi = -vy.wo;
        vy.f = mk<R>(sh.color);
    } else {
        vy.wi = V3<R>{0, 0, 1};
        vy.f = V3<R>{0, 0, 0};
    }
}

// AreaLight::sample (src/core/light/areaLight.hpp:19-31) over
// Primitive::sample(p, u, pdf) (src/core/primitive/primitive.hpp:25-33).
template <class R>
inline V3<R> light_sample(const Scene<R> &s, const Light<R> &L, const Vary<R> &res, R u0, R u1, V3<R> &li,
                          uint64_t *shadow_count) {
    if (L.point) {
        // PointLight::sample (src/core/light/pointLight.hpp:20-27): no pdf, no falloff
        Seg<R> seg = res.emit_seg(L.pos);
        if (shadow_count) ++*shadow_count;
        li = s.occluded(seg) ? V3<R>{0, 0, 0} : L.emission;
        return res.local(normalize(L.pos - res.p));
    }
    const Prim &p = s.prims[L.prim];
    LV<R> lv;
    R pdf;
    if (p.type == KR_PRIM_MESH) {
        // Mesh::sample(u, pdf): table entry floor(u.x * 1024) (mesh.hpp:92-98)
        const Mesh<R> &m = s.meshes[p.index];
        uint32_t idx = smin(uint32_t(std::floor(u0 * R(1024))), uint32_t(1023));
        pdf = m.pdf;
        lv = m.samples[idx];
    } else {
        // Sphere::sample (sphere.hpp:22-30): theta-uniform, pdf = 1/(4 pi r^2)
        const Sphere<R> &sp = s.spheres[p.index];
        V3<R> w = from_euler(u0 * Const<R>::pi(), u1 * Const<R>::pi() * R(2));
        lv.p = sp.o + std::sqrt(sp.r2) * w;
        lv.n = w;
        pdf = R(1) / (R(4) * Const<R>::pi() * sp.r2);
    }
    V3<R> wi = res.p - lv.p;
    R d2 = dot(wi, wi);
    pdf *= d2 / std::fabs(dot(lv.n, normalize(wi)));
    Seg<R> seg = res.emit_seg(lv.p);
    if (shadow_count) ++*shadow_count;
    li = (s.occluded(seg) ? V3<R>{0, 0, 0} : L.emission) / pdf;
    return res.local(normalize(lv.p - res.p));
}

struct PathStats {
    uint64_t closest = 0, shadow = 0;
    Counters trav;
};

// NormalKernel / VisibilityKernel / HalfVectorKernel (src/core/kernel/
// normal.hpp:17-28, visibility.hpp:17-35, halfVector.hpp:17-37): one hit,
// light 0 sampled with dimensions 1-2 of bounce 0. With no light the
// reference indexes an empty list; here the pixel stays black.
template <class R>
inline V3<R> debug_kernel(uint32_t kernel, const Scene<R> &s, Ray<R> ray, PathRng &rng, PathStats *st) {
    Vary<R> vy;
    if (st) st->closest++;
    if (!scene_intersect(s, ray, vy, st ? &st->trav : nullptr)) return V3<R>{0, 0, 0};
    if (kernel == KR_KERNEL_NORMAL) return vy.n * R(.5f) + V3<R>{R(.5f), R(.5f), R(.5f)};
    if (s.lights.empty()) return V3<R>{0, 0, 0};
    V3<R> li;
    vy.wi = light_sample(s, s.lights[0], vy, R(rng.get(0, 1)), R(rng.get(0, 2)), li, st ? &st->shadow : nullptr);
    if (kernel == KR_KERNEL_VISIBILITY)
        return (li.x == R(0) && li.y == R(0) && li.z == R(0)) ? V3<R>{0, 0, 0} : V3<R>{1, 1, 1};
    R v = std::acos(normalize(vy.wi + vy.wo).z) * R(2) / Const<R>::pi();
    return V3<R>{v, v, v};
}

// RadianceKernel::Slice (src/core/kernel/radiance.hpp:17-55): per vertex
// below `depth`, the ray that reached it, the BSDF value of the light sample,
// theta = acos(wo . reflect(wi, n)) in the local frame, the path radiance
// after the vertex's NEE and emission, and the shader id; at the end, the
// loop counter and the path radiance.
template <class R> struct ProfRecord {
    Ray<R> ray;
    V3<R> f, L;
    R theta;
    uint32_t shader;
};
template <class R> struct ProfSlice {
    std::vector<ProfRecord<R>> bounces;  // size = depth
    uint32_t bounce = 0;
    V3<R> L{0, 0, 0};
};

// RadianceKernel::execute (src/core/kernel/radiance.cu:14-81), recording
// into `prof` when given (radiance.cu:38-50, 74-78).
template <class R>
inline V3<R> radiance(const Scene<R> &s, Ray<R> ray, PathRng &rng, uint32_t max_bounce, PathStats *st,
                      ProfSlice<R> *prof = nullptr) {
    V3<R> L{0, 0, 0}, beta{1, 1, 1};
    Vary<R> vy;
    uint32_t nL = uint32_t(s.lights.size());
    uint32_t bounce = 0;
    for (;; ++bounce) {
        if (st) st->closest++;
        if (!(scene_intersect(s, ray, vy, st ? &st->trav : nullptr) && bounce <= max_bounce)) break;
        int b = int(bounce);
        if (nL) {
            V3<R> li;
            R xi = R(rng.get(b, 0));
            uint32_t idx = smin(uint32_t(std::floor(xi * R(nL))), nL - 1);
            R lpdf = R(1) / R(nL);
            vy.wi = light_sample(s, s.lights[idx], vy, R(rng.get(b, 1)), R(rng.get(b, 2)), li,
                                 st ? &st->shadow : nullptr);
            int dim = 8;
            const kr_shader_desc &sh = s.shaders[resolve_shader(s.shaders, vy.shader, vy, rng, b, dim)];
            R uc = sh.type == KR_SHADER_GLASS ? R(rng.get(b, dim++)) : R(0);
            shader_eval(sh, vy, uc);
            L += beta * vy.f * li * std::fabs(dot(vy.wi, V3<R>{0, 0, 1})) / lpdf;
        }
        L += beta * vy.emission;
        if (prof && bounce < prof->bounces.size()) {
            // reflect(wi, n) = 2 (n . wi) n - wi with n = (0, 0, 1) (vmath.hpp:470)
            ProfRecord<R> &b = prof->bounces[bounce];
            b.ray = ray;
            b.theta = std::acos(dot(vy.wo, V3<R>{-vy.wi.x, -vy.wi.y, vy.wi.z}));
            b.f = vy.f;
            b.L = L;
            b.shader = vy.shader;
        }
        int dim = 12;
        const kr_shader_desc &sh = s.shaders[resolve_shader(s.shaders, vy.shader, vy, rng, b, dim)];
        R uc = sh.type == KR_SHADER_GLASS ? R(rng.get(b, dim++)) : R(0);
        shader_sample(sh, vy, R(rng.get(b, 3)), R(rng.get(b, 4)), R(rng.get(b, 5)), uc);
        beta *= vy.f * std::fabs(dot(vy.wi, V3<R>{0, 0, 1}));
        ray = vy.emit_ray(vy.global(vy.wi));
        R rr = smax(beta.x, smax(beta.y, beta.z));
        if (rr < R(1) && bounce > 3) {
            R q = smax(R(.05f), R(1) - rr);
            if (R(rng.get(int(bounce), 6)) < q) break;
            beta /= R(1) - q;
        }
        vy = Vary<R>();
    }
    if (prof) {
        prof->bounce = bounce;
        prof->L = L;
    }
    return L;
}

// CPUMulticoreTracer::execute (src/core/tracer/cpuMulticore.cu:14-122):
// rows j = id, id + nthreads, ...; per pixel the NaN/Inf-free samples are
// summed and counted. Sums are kept in double here and returned unnormalised
// (pixel value = rgb / valid). `pixels` optionally restricts the work to a
// list of pixel indices (y * w + x); out arrays are indexed like `pixels`.
template <class R>
inline double render(const Scene<R> &s, const kr_camera &cam, const kr_params &prm, uint32_t spp_begin,
                     uint32_t spp_count, int nthreads, const uint32_t *pixels, uint32_t npix, double *rgb,
                     uint32_t *valid, PathStats *total) {
    Lens<R> lens(cam, prm.width, prm.height, prm.spp);
    uint32_t W = prm.width, H = prm.height;
    bool subset = pixels != nullptr;
    uint32_t rows = subset ? npix : H;
    if (nthreads <= 0) nthreads = std::max(1, int(std::thread::hardware_concurrency()) - 1);
    std::vector<PathStats> stats(nthreads);
    auto worker = [&](int id) {
        PathStats *st = total ? &stats[id] : nullptr;
        for (uint32_t j = uint32_t(id); j < rows; j += uint32_t(nthreads)) {
            uint32_t xb = 0, xe = W;
            if (subset) xb = 0, xe = 1;
            for (uint32_t i = xb; i != xe; ++i) {
                uint32_t px = subset ? pixels[j] : j * W + i;
                uint32_t x = px % W, y = px / W;
                uint32_t out = subset ? j : px;
                double acc[3] = {0, 0, 0};
                uint32_t nv = 0;
                for (uint32_t k = spp_begin; k != spp_begin + spp_count; ++k) {
                    PathRng rng(px, k, prm.seed);
                    V3<R> r = prm.kernel == KR_KERNEL_RADIANCE
                                  ? radiance(s, lens.sample(x, y, k), rng, prm.max_bounce, st)
                                  : debug_kernel(prm.kernel, s, lens.sample(x, y, k), rng, st);
                    if (!hasnan(r) && !hasinf(r)) {
                        nv++;
                        acc[0] += double(r.x);
                        acc[1] += double(r.y);
                        acc[2] += double(r.z);
                    }
                }
                rgb[3 * size_t(out) + 0] += acc[0];
                rgb[3 * size_t(out) + 1] += acc[1];
                rgb[3 * size_t(out) + 2] += acc[2];
                valid[out] += nv;
            }
        }
    };
    auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> ts;
    for (int id = 0; id < nthreads; ++id) ts.emplace_back(worker, id);
    for (auto &t : ts) t.join();
    double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (total)
        for (auto &p : stats) {
            total->closest += p.closest;
            total->shadow += p.shadow;
            total->trav.nodes += p.trav.nodes;
            total->trav.tris += p.trav.tris;
        }
    return sec;
}

}  // namespace
}  // namespace koishi_oracle
