// kr_device.h -- gfx950 device code: vector math, Philox, wide-BVH traversal,
// watertight triangles, spheres and the Radiance shading rules.
//
// Semantics follow the reference Radiance path (cited per function); the
// arithmetic is FP32 except the sphere quadratic, which is FP64 (the
// reference's r = 1000 ground sphere loses ~4 digits of t in FP32; gfx950
// runs FP64 VALU at half the FP32 rate and there are only a few spheres).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kr_api.h"
#include "kr_layout.h"

namespace kr {

// ---------------------------------------------------------------- vectors --
struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// normalize = a * (1 / sqrt(|a|^2)) (src/vec/vmath.hpp:648-652)
__device__ __forceinline__ f3 normalize(f3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
__device__ __forceinline__ float length(f3 a) { return sqrtf(dot(a, a)); }
// `a > b ? a : b` as in __func::max (src/vec/vmath.hpp:28-31)
__device__ __forceinline__ float smax(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float smin(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ bool finite3(f3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = float(1.0 / 3.14159265358979323846);
constexpr float kSpawnEps = 4e-3f;  // src/core/meta/varyings.hpp:22,33

// ------------------------------------------------------------------- RNG --
// Philox4x32-10 (Salmon et al., SC'11). Same counter layout as the oracle:
// ctr = {pixel, sample, 2*bounce + dim/4, stream}, key = seed.
struct Philox4 {
    uint32_t v[4];
};
__device__ __forceinline__ Philox4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                          uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return {{c0, c1, c2, c3}};
}
__device__ __forceinline__ float u01(uint32_t x) { return float(x >> 8) * (1.0f / 16777216.0f); }

// ------------------------------------------------------------- scene view --
struct DevScene {
    const float4 *nodes;
    const float4 *tris;       // 3 float4 per triangle, leaf order
    const float *tri_nrm;     // 9 floats per global triangle
    const uint32_t *tri_prim; // primitive of each global triangle
    const double *spheres;    // 4 doubles per sphere: centre, r2 (r2 = float(r)*float(r))
    const uint32_t *sphere_prim;
    const PrimRec *prims;
    const ShaderRec *shaders;
    const LightRec *lights;
    const float *light_table; // 6 floats per entry
    uint32_t nspheres, nlights, ntris, stack_cap;
};

// --------------------------------------------------------- ray intersection --
// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2(1) 2013) on a
// per-ray shear. Two-sided; accepts 0 < t < tmax. Barycentrics b1, b2 are
// the weights of v1 and v2, the (u, v) of the reference's Moller-Trumbore
// test (src/core/basic/ray.hpp:37-52).
struct ShearRay {
    f3 o;
    f3 mx, my, mz;  // rows of the shear: A' = M (v - o)
};

__device__ __forceinline__ ShearRay make_shear(f3 o, f3 d) {
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    auto comp = [&](int k) { return k == 0 ? d.x : (k == 1 ? d.y : d.z); };
    float dz = comp(kz);
    if (dz < 0.f) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    float Sx = comp(kx) / dz, Sy = comp(ky) / dz, Sz = 1.0f / dz;
    auto e = [](int k) { return f3{k == 0 ? 1.f : 0.f, k == 1 ? 1.f : 0.f, k == 2 ? 1.f : 0.f}; };
    ShearRay r;
    r.o = o;
    r.mx = e(kx) - e(kz) * Sx;
    r.my = e(ky) - e(kz) * Sy;
    r.mz = e(kz) * Sz;
    return r;
}

// LE: accept t == tmax as well (closest hit breaks exact ties itself).
template <bool LE = false>
__device__ __forceinline__ bool tri_hit(const ShearRay &r, f3 v0, f3 v1, f3 v2, float tmax, float &t, float &b1,
                                        float &b2) {
    f3 A = v0 - r.o, B = v1 - r.o, C = v2 - r.o;
    float Ax = dot(r.mx, A), Ay = dot(r.my, A);
    float Bx = dot(r.mx, B), By = dot(r.my, B);
    float Cx = dot(r.mx, C), Cy = dot(r.my, C);
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.f || V == 0.f || W == 0.f) {
        U = float(double(Cx) * double(By) - double(Cy) * double(Bx));
        V = float(double(Ax) * double(Cy) - double(Ay) * double(Cx));
        W = float(double(Bx) * double(Ay) - double(By) * double(Ax));
    }
    if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
    float det = U + V + W;
    if (det == 0.f) return false;
    float T = U * dot(r.mz, A) + V * dot(r.mz, B) + W * dot(r.mz, C);
    float inv = __builtin_amdgcn_rcpf(det);  // v_rcp_f32, 1 ulp
    float tt = T * inv;
    if (!(tt > 0.f && (LE ? tt <= tmax : tt < tmax))) return false;
    t = tt;
    b1 = V * inv;
    b2 = W * inv;
    return true;
}

// Sphere::intersect (src/core/primitive/sphere.hpp:31-54) in FP64: near root,
// else far root, hit iff t >= 0.
__device__ __forceinline__ bool sphere_hit(const double *s, f3 o, f3 d, double &t) {
    double ocx = s[0] - double(o.x), ocy = s[1] - double(o.y), ocz = s[2] - double(o.z);
    double dx = d.x, dy = d.y, dz = d.z;
    double poc = dx * ocx + dy * ocy + dz * ocz;
    double oc2 = ocx * ocx + ocy * ocy + ocz * ocz;
    double det = s[3] - (oc2 - poc * poc);
    if (det < 0) return false;
    det = sqrt(det);
    t = poc - det;
    if (t < 0) t = poc + det;
    return t >= 0;
}

// ------------------------------------------------------------- BVH walk --
// Stack: the top LDS_DEPTH entries of each lane live in LDS, deeper ones
// spill to a per-lane global area. The spill capacity comes from the tree
// depth, so the stack cannot overflow. LDS holds two [LDS_DEPTH][LANES] word
// arrays, refs then entry distances: lane l of every level sits in bank
// l mod 32, so the lanes of a wave never conflict whatever their depths.
// A lane's whole state is one byte offset, off = (sp * LANES + lane) * 4:
// a push or pop is one ds_write2 / ds_read2 at off (the distance at the
// constant offset LDS_DEPTH * LANES * 4) and an add of LANES * 4, with no
// lane index to keep in a register (off mod LANES * 4 is 4 * lane).
template <int LDS_DEPTH, int LANES = 256>
struct Stack {
    static constexpr uint32_t kRow = LANES * 4u;  // bytes of one level
    static constexpr uint32_t kLds = LDS_DEPTH * kRow;
    static_assert((kRow & (kRow - 1)) == 0, "a stack level spans a power-of-two number of bytes");
    char *lds;         // [2][LDS_DEPTH][LANES] words of the block's __shared__ stack
    uint2 *spill;      // the block's spill area: entry k of lane l at spill[k * gstride + l]
    uint32_t gstride;  // lanes of the whole launch
    uint32_t off;      // (sp * LANES + lane) * 4
    __device__ __forceinline__ void reset() { off &= kRow - 1u; }
    __device__ __forceinline__ bool empty() const { return off < kRow; }
    // the spill entry at the current depth (beyond LDS_DEPTH); lds, spill and
    // gstride are the same for the whole block, so the lane's only register
    // is off
    __device__ __forceinline__ uint2 *spilled() const {
        return spill + size_t(off / kRow - LDS_DEPTH) * gstride + (off & (kRow - 1u)) / 4u;
    }
    __device__ __forceinline__ void push(uint32_t ref, float t) {
        if (off < kLds) {
            *reinterpret_cast<uint32_t *>(lds + off) = ref;
            *reinterpret_cast<float *>(lds + kLds + off) = t;
        } else {
            *spilled() = make_uint2(ref, __float_as_uint(t));
        }
        off += kRow;
    }
    __device__ __forceinline__ void pop(uint32_t &ref, float &t) {
        off -= kRow;
        if (off < kLds) {
            ref = *reinterpret_cast<const uint32_t *>(lds + off);
            t = *reinterpret_cast<const float *>(lds + kLds + off);
        } else {
            uint2 e = *spilled();
            ref = e.x;
            t = __uint_as_float(e.y);
        }
    }
    // entries on the stack
    __device__ __forceinline__ uint32_t depth() const { return off / kRow; }
    // entry k counted from the bottom (k < depth()), as (ref, entry distance bits)
    __device__ __forceinline__ uint2 entry(uint32_t k) const {
        uint32_t o = k * kRow + (off & (kRow - 1u));
        if (o < kLds)
            return make_uint2(*reinterpret_cast<const uint32_t *>(lds + o),
                              *reinterpret_cast<const uint32_t *>(lds + kLds + o));
        return spill[size_t(k - LDS_DEPTH) * gstride + (off & (kRow - 1u)) / 4u];
    }
};

struct RayBox {
    f3 o, inv, oinv;
};
__device__ __forceinline__ RayBox make_raybox(f3 o, f3 d) {
    auto fix = [](float v) { return fabsf(v) < 1e-30f ? copysignf(1e-30f, v) : v; };
    RayBox rb;
    rb.o = o;
    rb.inv = {1.0f / fix(d.x), 1.0f / fix(d.y), 1.0f / fix(d.z)};
    rb.oinv = o * rb.inv;
    return rb;
}

// Slab test of one child with the ray interval [0, tmax]; conservative by
// (1 + 2 gamma_3) on the far distance (Ize, JCGT 2(2) 2013).
__device__ __forceinline__ float box_enter(const RayBox &rb, float lx, float hx, float ly, float hy, float lz, float hz,
                                           float tmax) {
    float tx0 = fmaf(lx, rb.inv.x, -rb.oinv.x), tx1 = fmaf(hx, rb.inv.x, -rb.oinv.x);
    float ty0 = fmaf(ly, rb.inv.y, -rb.oinv.y), ty1 = fmaf(hy, rb.inv.y, -rb.oinv.y);
    float tz0 = fmaf(lz, rb.inv.z, -rb.oinv.z), tz1 = fmaf(hz, rb.inv.z, -rb.oinv.z);
    float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.f));
    float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
    tf *= 1.00000024f;
    return tn <= tf ? tn : __builtin_inff();
}

// The same test with the entering and leaving planes of every axis already
// known (the ray's direction signs pick them: see load_node_nf); identical
// results without the per-axis min / max.
__device__ __forceinline__ float box_enter_nf(const RayBox &rb, float nx, float fx, float ny, float fy, float nz,
                                              float fz, float tmax) {
    float tn = fmaxf(fmaxf(fmaf(nx, rb.inv.x, -rb.oinv.x), fmaf(ny, rb.inv.y, -rb.oinv.y)),
                     fmaxf(fmaf(nz, rb.inv.z, -rb.oinv.z), 0.f));
    float tf = fminf(fminf(fmaf(fx, rb.inv.x, -rb.oinv.x), fmaf(fy, rb.inv.y, -rb.oinv.y)),
                     fminf(fmaf(fz, rb.inv.z, -rb.oinv.z), tmax));
    tf *= 1.00000024f;
    return tn <= tf ? tn : __builtin_inff();
}

struct HitRec {
    float t, u, v;
    uint32_t code;  // global triangle index, kSphereBit|sphere, or kMissCode
    uint32_t mat;   // triangle hits: the material bucket stored with the triangle
};

template <int W>
struct NodeView {
    float lo[3][W], hi[3][W];
    uint32_t ref[W];
};

// A child ref of an inner node is the node's offset in 16-byte units from
// the start of the node array (kr_scene_upload rewrites the builder's node
// indices so), so that its address is a shift, not a multiply.
template <int W>
__device__ __forceinline__ void load_node(const float4 *nodes, uint32_t ref, NodeView<W> &nv) {
    constexpr int NF4 = int(node_floats(W)) / 4;
    const float4 *p = nodes + ref;
    float f[NF4 * 4];
#pragma unroll
    for (int i = 0; i < NF4; ++i) {
        float4 q = p[i];
        f[4 * i + 0] = q.x;
        f[4 * i + 1] = q.y;
        f[4 * i + 2] = q.z;
        f[4 * i + 3] = q.w;
    }
#pragma unroll
    for (int c = 0; c < W; ++c) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            nv.lo[a][c] = f[(2 * a) * W + c];
            nv.hi[a][c] = f[(2 * a + 1) * W + c];
        }
        nv.ref[c] = __float_as_uint(f[6 * W + c]);
    }
}

// Compressed node (qnode_bytes(W) bytes, layout in kr_layout.h): decoded
// with the same fma(q, s, p) the encoder used to make the boxes conservative.
template <int W>
__device__ __forceinline__ void load_qnode(const float4 *nodes, uint32_t ref, NodeView<W> &nv) {
    constexpr int NU4 = int(qnode_bytes(W)) / 16;
    const uint4 *p = reinterpret_cast<const uint4 *>(nodes) + ref;
    uint32_t u[NU4 * 4];
#pragma unroll
    for (int i = 0; i < NU4; ++i) {
        uint4 q = p[i];
        u[4 * i + 0] = q.x;
        u[4 * i + 1] = q.y;
        u[4 * i + 2] = q.z;
        u[4 * i + 3] = q.w;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float org = __uint_as_float(u[a]);
        int e = int(int8_t(uint8_t(u[3] >> (8 * a))));
        float sc = __uint_as_float(uint32_t(e + 127) << 23);
#pragma unroll
        for (int c = 0; c < W; ++c) {
            constexpr int QB = 4 * (4 + W);  // first quantised byte
            int bl = QB + (2 * a) * W + c, bh = QB + (2 * a + 1) * W + c;
            uint32_t ql = (u[bl >> 2] >> (8 * (bl & 3))) & 0xFFu;
            uint32_t qh = (u[bh >> 2] >> (8 * (bh & 3))) & 0xFFu;
            nv.lo[a][c] = fmaf(float(ql), sc, org);
            nv.hi[a][c] = fmaf(float(qh), sc, org);
        }
    }
#pragma unroll
    for (int c = 0; c < W; ++c) nv.ref[c] = u[4 + c];
}

// FP32 node with each axis' planes in the ray's order: nv.lo = the plane the
// ray enters through, nv.hi = the one it leaves through. Along an axis with a
// negative inverse direction (sign bit set, as -1e30 for d = -0) that is hi
// before lo, so the lane loads the other array of the pair (sel = its byte
// distance, W * 4 or 0) instead of loading both and ordering them with
// min / max per child. The selectors are derived from the sign bits per node
// (the empty asm keeps the compiler from holding them in registers across
// the traversal loop, where they would spill).
template <int W>
__device__ __forceinline__ void load_node_nf(const float4 *nodes, uint32_t ref, const RayBox &rb, NodeView<W> &nv) {
    // 32-bit byte offsets from the array base (kr_scene_upload refuses node
    // arrays of 4 GB or more), so the loads address from a scalar base plus
    // one VGPR
    const char *base = reinterpret_cast<const char *>(nodes);
    const uint32_t nb = ref << 4;
    constexpr uint32_t AB = W * 4;  // bytes of one plane array
    uint32_t bits[3] = {__float_as_uint(rb.inv.x), __float_as_uint(rb.inv.y), __float_as_uint(rb.inv.z)};
    asm volatile("" : "+v"(bits[0]), "+v"(bits[1]), "+v"(bits[2]));
    uint32_t sel[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) sel[a] = (bits[a] >> 31) * AB;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float *pn = reinterpret_cast<const float *>(base + (nb + 2 * a * AB + sel[a]));
        const float *pf = reinterpret_cast<const float *>(base + (nb + (2 * a + 1) * AB - sel[a]));
        if (W == 2) {
            float2 n = *reinterpret_cast<const float2 *>(pn), f = *reinterpret_cast<const float2 *>(pf);
            nv.lo[a][0] = n.x;
            nv.lo[a][1] = n.y;
            nv.hi[a][0] = f.x;
            nv.hi[a][1] = f.y;
        } else {
#pragma unroll
            for (int q = 0; q < W / 4; ++q) {
                float4 n = reinterpret_cast<const float4 *>(pn)[q], f = reinterpret_cast<const float4 *>(pf)[q];
                nv.lo[a][4 * q] = n.x;
                nv.lo[a][4 * q + 1] = n.y;
                nv.lo[a][4 * q + 2] = n.z;
                nv.lo[a][4 * q + 3] = n.w;
                nv.hi[a][4 * q] = f.x;
                nv.hi[a][4 * q + 1] = f.y;
                nv.hi[a][4 * q + 2] = f.z;
                nv.hi[a][4 * q + 3] = f.w;
            }
        }
    }
    const uint32_t *r = reinterpret_cast<const uint32_t *>(base + (nb + 6 * AB));
#pragma unroll
    for (int c = 0; c < W; ++c) nv.ref[c] = r[c];
}

// Entry distances of the children of a compressed node (W = 4 or 8): key[c]
// = the distance at which the ray enters child c before tmax, +inf for a
// missed or empty child. The decode is folded into the slab test: along axis
// a a plane is p = q s_a + o_node (s_a a power of two), so its distance is
//   (p - o) inv_a = q (s_a inv_a) + (o_node inv_a - o inv_a),
// one conversion of the byte and one fma per plane, with s_a inv_a exact and
// the second term computed once per node and axis. As for the FP32 nodes, the
// ray's direction signs pick each axis' entering and leaving planes (whole
// words of W bytes), and the leaving distance is widened by 1 + 2^-22.
template <int W>
__device__ __forceinline__ void qnode_enter(const float4 *nodes, uint32_t ref, const RayBox &rb, float tmax,
                                            float (&key)[W], uint32_t (&cref)[W]) {
    static_assert(W == 4 || W == 8, "byte planes of whole words");
    constexpr int NU4 = int(qnode_bytes(W)) / 16;
    constexpr int QW = 4 + W;  // first word of the planes
    constexpr int AW = W / 4;  // words per plane array
    const uint4 *p = reinterpret_cast<const uint4 *>(nodes) + ref;
    uint32_t u[NU4 * 4];
#pragma unroll
    for (int i = 0; i < NU4; ++i) {
        uint4 q = p[i];
        u[4 * i + 0] = q.x;
        u[4 * i + 1] = q.y;
        u[4 * i + 2] = q.z;
        u[4 * i + 3] = q.w;
    }
    const float inv[3] = {rb.inv.x, rb.inv.y, rb.inv.z}, oinv[3] = {rb.oinv.x, rb.oinv.y, rb.oinv.z};
    float A[3], B[3];
    uint32_t nw[3][AW], fw[3][AW];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float s;
        if (W == 4) {
            // scales stored as floats in the node's padding (s_x, s_y as upper halves)
            s = __uint_as_float(a == 0 ? u[14] << 16 : (a == 1 ? u[14] & 0xFFFF0000u : u[15]));
        } else {
            const int e = int(int8_t(uint8_t(u[3] >> (8 * a))));
            s = __uint_as_float(uint32_t(e + 127) << 23);
        }
        A[a] = s * inv[a];
        B[a] = fmaf(__uint_as_float(u[a]), inv[a], -oinv[a]);
        const bool neg = __float_as_uint(inv[a]) >> 31;
#pragma unroll
        for (int w = 0; w < AW; ++w) {
            const uint32_t lo = u[QW + 2 * a * AW + w], hi = u[QW + (2 * a + 1) * AW + w];
            nw[a][w] = neg ? hi : lo;
            fw[a][w] = neg ? lo : hi;
        }
    }
#pragma unroll
    for (int c = 0; c < W; ++c) {
        float tn3[3], tf3[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float qn = float((nw[a][c / 4] >> (8 * (c % 4))) & 0xFFu);
            const float qf = float((fw[a][c / 4] >> (8 * (c % 4))) & 0xFFu);
            tn3[a] = fmaf(qn, A[a], B[a]);
            tf3[a] = fmaf(qf, A[a], B[a]);
        }
        float tn = fmaxf(fmaxf(tn3[0], tn3[1]), fmaxf(tn3[2], 0.f));
        float tf = fminf(fminf(tf3[0], tf3[1]), fminf(tf3[2], tmax)) * 1.00000024f;
        cref[c] = u[4 + c];
        key[c] = (tn <= tf && cref[c] != kEmpty) ? tn : __builtin_inff();
    }
}

template <int W, bool Q>
__device__ __forceinline__ void fetch_node(const float4 *nodes, uint32_t ref, NodeView<W> &nv) {
    if (Q)
        load_qnode<W>(nodes, ref, nv);
    else
        load_node<W>(nodes, ref, nv);
}

// Sorts the children's (entry distance, ref) pairs ascending by distance with
// an optimal comparator network: 1, 5 and 19 comparators for W = 2, 4, 8.
template <int W>
__device__ __forceinline__ void sort_children(float (&key)[W], uint32_t (&ref)[W]) {
    auto ce = [&](int i, int j) {
        bool sw = key[j] < key[i];
        float ki = sw ? key[j] : key[i], kj = sw ? key[i] : key[j];
        uint32_t ri = sw ? ref[j] : ref[i], rj = sw ? ref[i] : ref[j];
        key[i] = ki;
        key[j] = kj;
        ref[i] = ri;
        ref[j] = rj;
    };
    if (W == 2) {
        ce(0, 1);
    } else if (W == 4) {
        ce(0, 1), ce(2, 3), ce(0, 2), ce(1, 3), ce(1, 2);
    } else {
        ce(0, 2), ce(1, 3), ce(4, 6), ce(5, 7), ce(0, 4), ce(1, 5), ce(2, 6), ce(3, 7), ce(0, 1), ce(2, 3);
        ce(4, 5), ce(6, 7), ce(2, 4), ce(3, 5), ce(1, 4), ce(3, 6), ce(1, 2), ce(3, 4), ce(5, 6);
    }
}

// Closest hit over the wide BVH and the sphere list: ordered traversal,
// nearest child first, far children pushed with their entry distance and
// culled on pop. Equivalent to the reference's primitive loop + per-mesh DFS
// (src/core/meta/scene.hpp:35-47, src/core/primitive/mesh.cu:200-267) up to
// exact ties. The walk is cut into the three pieces below so that the
// refilling traversal kernel runs exactly the same steps as trace_closest.

// The ray's hit against the sphere list (linear, FP64), the starting record
// of the walk.
__device__ __forceinline__ HitRec closest_spheres(const DevScene &S, f3 o, f3 d) {
    HitRec h;
    h.t = __builtin_inff();
    h.u = h.v = 0.f;
    h.code = kMissCode;
    h.mat = 0;
    for (uint32_t s = 0; s < S.nspheres; ++s) {
        double ts;
        if (sphere_hit(S.spheres + 4 * s, o, d, ts)) {
            float tf = float(ts);
            if (tf < h.t) {
                h.t = tf;
                h.code = kSphereBit | s;
            }
        }
    }
    return h;
}

// The children of inner node `node`: key[c] = the distance at which the ray
// enters child c before tmax, +inf for a missed or empty child; ref[c] its ref.
template <int W, bool Q>
__device__ __forceinline__ void node_keys(const float4 *nodes, uint32_t node, const RayBox &rb, float tmax,
                                          float (&key)[W], uint32_t (&ref)[W]) {
    if constexpr (Q && W != 2) {
        qnode_enter<W>(nodes, node, rb, tmax, key, ref);
        return;
    }
    NodeView<W> nv;
    if (Q)
        fetch_node<W, Q>(nodes, node, nv);
    else
        load_node_nf<W>(nodes, node, rb, nv);
#pragma unroll
    for (int c = 0; c < W; ++c) {
        if (Q) {
            float tn = box_enter(rb, nv.lo[0][c], nv.hi[0][c], nv.lo[1][c], nv.hi[1][c], nv.lo[2][c], nv.hi[2][c],
                                 tmax);
            bool ok = nv.ref[c] != kEmpty && tn != __builtin_inff();
            key[c] = ok ? tn : __builtin_inff();
        } else {
            // an empty slot holds the box (+inf, -inf), whose entering
            // plane lies at +inf along every axis: box_enter_nf returns +inf
            // for it like for any missed child
            key[c] = box_enter_nf(rb, nv.lo[0][c], nv.hi[0][c], nv.lo[1][c], nv.hi[1][c], nv.lo[2][c], nv.hi[2][c],
                                  tmax);
        }
        ref[c] = nv.ref[c];
    }
}

// One inner node: returns the nearest child whose box the ray enters before
// tmax (kEmpty if none) and pushes the other entered children, farthest
// first, with their entry distances.
template <int W, int LDS_DEPTH, bool Q>
__device__ __forceinline__ uint32_t closest_node(const DevScene &S, const RayBox &rb, uint32_t node, float tmax,
                                                 Stack<LDS_DEPTH> &st) {
    float key[W];
    uint32_t ref[W];
    node_keys<W, Q>(S.nodes, node, rb, tmax, key, ref);
    sort_children<W>(key, ref);
    // missed children sort last, with key +inf
#pragma unroll
    for (int c = W - 1; c >= 1; --c)
        if (key[c] != __builtin_inff()) st.push(ref[c], key[c]);
    return key[0] != __builtin_inff() ? ref[0] : kEmpty;
}

// One leaf: the triangles of `leaf` against the closest hit so far. Hits at
// exactly the same t (coplanar, overlapping triangles, such as two emitters
// in one plane) go to the lowest global triangle index, i.e. to the first
// primitive in scene order, as the reference's loop over primitives with a
// strict `<` decides them (src/core/meta/scene.hpp:39-46), whatever order
// the tree visits the leaves in.
__device__ __forceinline__ uint32_t closest_leaf(const DevScene &S, const ShearRay &sr, uint32_t leaf, HitRec &h) {
    uint32_t first = leaf & kLeafFirstMask;
    uint32_t cnt = ((leaf >> kLeafCountShift) & 0xFu) + 1u;
    for (uint32_t k = 0; k < cnt; ++k) {
        const float4 *tp = S.tris + 3 * size_t(first + k);
        float4 a = tp[0], b = tp[1], c = tp[2];
        float tt, b1, b2;
        const uint32_t code = __float_as_uint(b.w);
        if (tri_hit<true>(sr, f3{a.x, a.y, a.z}, f3{b.x, b.y, b.z}, f3{c.x, c.y, c.z}, h.t, tt, b1, b2) &&
            (tt < h.t || code < h.code)) {
            h.t = tt;
            h.u = b1;
            h.v = b2;
            h.code = code;
            h.mat = __float_as_uint(c.w);
        }
    }
    return cnt;
}

// Pops the next subtree still in front of the closest hit t into `cur`;
// false when none is left (the walk is done).
template <int LDS_DEPTH>
__device__ __forceinline__ bool closest_pop(Stack<LDS_DEPTH> &st, float t, uint32_t &cur) {
    while (!st.empty()) {
        float tn;
        st.pop(cur, tn);
        if (tn <= t) return true;
    }
    return false;
}

// One iteration of the closest-hit walk: the node step if `cur` is an inner
// node, then the leaf if `cur` is (or the node step just entered) a leaf,
// then the pop if the node entered nothing or the leaf was tested; true when
// the walk is over. Taking an entered leaf in the node's iteration keeps the
// lane's sequence of nodes and leaves, so the hit is the same, but the lanes
// that descend and the lanes at a leaf share iterations: with one triangle
// per leaf most node steps enter a leaf.
template <int W, int LDS_DEPTH, bool Q>
__device__ __forceinline__ bool closest_iter(const DevScene &S, const RayBox &rb, const ShearRay &sr,
                                             Stack<LDS_DEPTH> &st, uint32_t &cur, HitRec &h, uint32_t &nvis,
                                             uint32_t &ntri) {
    bool pop = false;
    if (!(cur & kLeafBit)) {
        ++nvis;
        uint32_t next = closest_node<W, LDS_DEPTH, Q>(S, rb, cur, h.t, st);
        pop = next == kEmpty;
        if (!pop) cur = next;
    }
    if (!pop && (cur & kLeafBit)) {
        ntri += closest_leaf(S, sr, cur, h);
        pop = true;
    }
    return pop && !closest_pop(st, h.t, cur);
}

template <int W, int LDS_DEPTH, bool Q = false>
__device__ __forceinline__ HitRec trace_closest(const DevScene &S, f3 o, f3 d, Stack<LDS_DEPTH> &st,
                                                uint64_t *counts = nullptr) {
    HitRec h = closest_spheres(S, o, d);
    if (S.ntris == 0) return h;
    RayBox rb = make_raybox(o, d);
    ShearRay sr = make_shear(o, d);
    uint32_t cur = 0;  // root is always an inner node
    uint32_t nvis = 0, ntri = 0;
    st.reset();
    // One walk step per iteration, or two when a node step enters a leaf
    // (see closest_iter). A lane that reaches a leaf tests it in the same
    // iteration in which the wave's other lanes take their next node,
    // instead of waiting at the leaf until every lane's descent is over,
    // which is what a `continue` into the node branch compiles to.
    while (!closest_iter<W, LDS_DEPTH, Q>(S, rb, sr, st, cur, h, nvis, ntri)) {
    }
    if (counts) {
        counts[0] = nvis;
        counts[1] = ntri;
    }
    return h;
}

// trace_closest's walk resumed from a state (the next ref `cur`, the hit so
// far `h` and the stack), for at most `steps` iterations of closest_iter;
// true when the walk is over. A walk run in such pieces visits the same nodes
// and leaves in the same order as one uninterrupted walk, so its hit is the
// same. Visited nodes and tested triangles are added to nvis and ntri.
template <int W, int LDS_DEPTH, bool Q>
__device__ __forceinline__ bool closest_walk(const DevScene &S, const RayBox &rb, const ShearRay &sr,
                                             Stack<LDS_DEPTH> &st, uint32_t &cur, HitRec &h, uint32_t steps,
                                             uint32_t &nvis, uint32_t &ntri) {
    for (uint32_t k = 0; k < steps; ++k)
        if (closest_iter<W, LDS_DEPTH, Q>(S, rb, sr, st, cur, h, nvis, ntri)) return true;
    return false;
}

// The children an any-hit walk enters at one node (key = entry distance, +inf
// for a missed child): the farthest goes to `cur`, the others onto the stack
// in sorted order; false when at least one child is entered. Entering in slot
// order without the sort costs more node steps than the sort saves.
template <int W, int LDS_DEPTH>
__device__ __forceinline__ bool occluded_children(float (&key)[W], uint32_t (&ref)[W], Stack<LDS_DEPTH> &st,
                                                  uint32_t &cur) {
#pragma unroll
    for (int c = 0; c < W; ++c) key[c] = key[c] != __builtin_inff() ? -key[c] : __builtin_inff();
    sort_children<W>(key, ref);
#pragma unroll
    for (int c = W - 1; c >= 1; --c)
        if (key[c] != __builtin_inff()) st.push(ref[c], -key[c]);
    if (key[0] == __builtin_inff()) return true;
    cur = ref[0];
    return false;
}

// One iteration of the any-hit walk, paired like closest_iter: the node
// step, then the leaf the node step entered (or `cur` if it is a leaf), then
// the pop. Returns 1 when a triangle blocks the segment, 0 when the walk
// ends without one, -1 when it goes on.
template <int W, int LDS_DEPTH, bool Q>
__device__ __forceinline__ int occluded_iter(const DevScene &S, const RayBox &rb, const ShearRay &sr, float limit,
                                             float tmax, Stack<LDS_DEPTH> &st, uint32_t &cur) {
    bool pop = false;
    if (!(cur & kLeafBit)) {
        // sort key: minus the entry distance (farthest first), +inf for a
        // child the segment misses (last)
        float key[W];
        uint32_t ref[W];
        node_keys<W, Q>(S.nodes, cur, rb, limit, key, ref);
        pop = occluded_children<W, LDS_DEPTH>(key, ref, st, cur);
    }
    if (!pop && (cur & kLeafBit)) {
        uint32_t first = cur & kLeafFirstMask;
        uint32_t cnt = ((cur >> kLeafCountShift) & 0xFu) + 1u;
        for (uint32_t k = 0; k < cnt; ++k) {
            const float4 *tp = S.tris + 3 * size_t(first + k);
            float4 a = tp[0], b = tp[1], c = tp[2];
            float tt, b1, b2;
            if (tri_hit(sr, f3{a.x, a.y, a.z}, f3{b.x, b.y, b.z}, f3{c.x, c.y, c.z}, limit, tt, b1, b2) &&
                tt <= tmax)
                return 1;
        }
        pop = true;
    }
    if (pop) {
        if (st.empty()) return 0;
        float tn;
        st.pop(cur, tn);
    }
    return -1;
}

// Any hit on the segment (o, d, tmax): Scene::intersect(Seg)
// (src/core/meta/scene.hpp:57-67). Triangles block for 0 < t <= tmax,
// spheres for a ray hit with t <= tmax (src/core/primitive/primitive.hpp:16-21).
// Scheduled "while-while" with postponed leaves (Aila and Laine, HPG 2009):
// a lane that reaches a leaf parks it and keeps descending until every lane
// of the wave holds a leaf, then all lanes test triangles together, so the
// node and leaf code run in separate wave-wide loops instead of alternating
// under divergence (measured faster than the alternating loop for shadow
// segments; the closest-hit traversal keeps the alternating loop, which
// measured faster there).
template <int W, int LDS_DEPTH, bool Q = false>
__device__ __forceinline__ bool trace_occluded(const DevScene &S, f3 o, f3 d, float tmax, Stack<LDS_DEPTH> &st) {
    for (uint32_t s = 0; s < S.nspheres; ++s) {
        double ts;
        if (sphere_hit(S.spheres + 4 * s, o, d, ts) && ts <= double(tmax)) return true;
    }
    if (S.ntris == 0) return false;
    RayBox rb = make_raybox(o, d);
    ShearRay sr = make_shear(o, d);
    // tri_hit accepts t < limit; the segment includes its end point.
    float limit = tmax * 1.00000012f;
    st.reset();
    auto pop_next = [&]() -> uint32_t {
        if (st.empty()) return kEmpty;
        uint32_t r;
        float tn;
        st.pop(r, tn);
        return r;
    };
    uint32_t cur = 0, leaf = kEmpty;
    while (true) {
        while (cur != kEmpty && !(cur & kLeafBit)) {
            NodeView<W> nv;
            if (Q)
                fetch_node<W, Q>(S.nodes, cur, nv);
            else
                load_node_nf<W>(S.nodes, cur, rb, nv);
            uint32_t next = kEmpty;
#pragma unroll
            for (int c = 0; c < W; ++c) {
                float tn = Q ? box_enter(rb, nv.lo[0][c], nv.hi[0][c], nv.lo[1][c], nv.hi[1][c], nv.lo[2][c],
                                         nv.hi[2][c], limit)
                             : box_enter_nf(rb, nv.lo[0][c], nv.hi[0][c], nv.lo[1][c], nv.hi[1][c], nv.lo[2][c],
                                            nv.hi[2][c], limit);
                // (box_enter_nf never enters an empty slot's (+inf, -inf) box)
                if ((Q ? nv.ref[c] != kEmpty : true) && tn != __builtin_inff()) {
                    if (next == kEmpty)
                        next = nv.ref[c];
                    else
                        st.push(nv.ref[c], tn);
                }
            }
            cur = next != kEmpty ? next : pop_next();
            if (cur != kEmpty && (cur & kLeafBit) && leaf == kEmpty) {
                leaf = cur;
                cur = pop_next();
            }
            if (!__any(leaf == kEmpty)) break;  // every lane still descending holds a leaf
        }
        while (leaf != kEmpty || (cur != kEmpty && (cur & kLeafBit))) {
            uint32_t lf;
            if (leaf != kEmpty) {
                lf = leaf;
                leaf = kEmpty;
            } else {
                lf = cur;
                cur = pop_next();
            }
            uint32_t first = lf & kLeafFirstMask;
            uint32_t cnt = ((lf >> kLeafCountShift) & 0xFu) + 1u;
            for (uint32_t k = 0; k < cnt; ++k) {
                const float4 *tp = S.tris + 3 * size_t(first + k);
                float4 a = tp[0], b = tp[1], c = tp[2];
                float tt, b1, b2;
                if (tri_hit(sr, f3{a.x, a.y, a.z}, f3{b.x, b.y, b.z}, f3{c.x, c.y, c.z}, limit, tt, b1, b2) &&
                    tt <= tmax)
                    return true;
            }
        }
        if (cur == kEmpty) break;
    }
    return false;
}

// The same any-hit walk in iterations of occluded_iter, without postponed
// leaves, entering the children farthest first: a segment runs from a
// surface towards a light, and an occluder found early ends it.
// Measured by a CPU count on the 1M-triangle frame's light segments:
// farthest-first visits 8 % fewer nodes per 64-segment wave than the
// children's slot order, nearest-first 1 % more.
template <int W, int LDS_DEPTH, bool Q = false>
__device__ __forceinline__ bool trace_occluded_steps(const DevScene &S, f3 o, f3 d, float tmax, Stack<LDS_DEPTH> &st) {
    for (uint32_t s = 0; s < S.nspheres; ++s) {
        double ts;
        if (sphere_hit(S.spheres + 4 * s, o, d, ts) && ts <= double(tmax)) return true;
    }
    if (S.ntris == 0) return false;
    RayBox rb = make_raybox(o, d);
    ShearRay sr = make_shear(o, d);
    float limit = tmax * 1.00000012f;
    st.reset();
    uint32_t cur = 0;
    int r;
    while ((r = occluded_iter<W, LDS_DEPTH, Q>(S, rb, sr, limit, tmax, st, cur)) < 0) {
    }
    return r == 1;
}

// trace_occluded_steps' walk over the triangles resumed from a state (the
// next ref `cur` and the stack), for at most `steps` iterations: 1 when a
// triangle blocks the segment, 0 when the walk ends without one, -1 when the
// steps ran out first. Pieces of the walk visit what the whole walk visits,
// in the same order. (The spheres are tested once, before the walk.)
template <int W, int LDS_DEPTH, bool Q>
__device__ __forceinline__ int occluded_walk(const DevScene &S, const RayBox &rb, const ShearRay &sr, float limit,
                                             float tmax, Stack<LDS_DEPTH> &st, uint32_t &cur, uint32_t steps) {
    for (uint32_t s = 0; s < steps; ++s) {
        const int r = occluded_iter<W, LDS_DEPTH, Q>(S, rb, sr, limit, tmax, st, cur);
        if (r >= 0) return r;
    }
    return -1;
}

// ------------------------------------------------ stepwise traversal --
// The any-hit traversal cut into steps (one inner node, or one leaf and the
// pop after it), so that a persistent kernel can refill lanes whose segment
// is done with new segments while the rest of the wave keeps traversing
// (Aila and Laine, "Understanding the Efficiency of Ray Traversal on GPUs",
// HPG 2009). Results are identical to trace_occluded.
struct TravState {
    RayBox rb;
    ShearRay sr;
    float limit;  // segment end (with the tri_hit margin); tmax
    float tmax;
    uint32_t cur;
    bool done;
    bool occluded;
};

template <int W, int LDS_DEPTH>
__device__ __forceinline__ void occluded_begin(const DevScene &S, f3 o, f3 d, float tmax, Stack<LDS_DEPTH> &st,
                                               TravState &ts) {
    ts.occluded = false;
    for (uint32_t s = 0; s < S.nspheres; ++s) {
        double tsph;
        if (sphere_hit(S.spheres + 4 * s, o, d, tsph) && tsph <= double(tmax)) ts.occluded = true;
    }
    ts.done = ts.occluded || S.ntris == 0;
    ts.rb = make_raybox(o, d);
    ts.sr = make_shear(o, d);
    ts.tmax = tmax;
    ts.limit = tmax * 1.00000012f;
    ts.cur = 0;
    st.reset();
}

template <int W, int LDS_DEPTH, bool Q>
__device__ __forceinline__ void occluded_step(const DevScene &S, Stack<LDS_DEPTH> &st, TravState &ts) {
    uint32_t cur = ts.cur;
    if (!(cur & kLeafBit)) {
        NodeView<W> nv;
        fetch_node<W, Q>(S.nodes, cur, nv);
        uint32_t next = kEmpty;
#pragma unroll
        for (int c = 0; c < W; ++c) {
            float tn = box_enter(ts.rb, nv.lo[0][c], nv.hi[0][c], nv.lo[1][c], nv.hi[1][c], nv.lo[2][c], nv.hi[2][c],
                                 ts.limit);
            if (nv.ref[c] != kEmpty && tn != __builtin_inff()) {
                if (next == kEmpty)
                    next = nv.ref[c];
                else
                    st.push(nv.ref[c], tn);
            }
        }
        if (next != kEmpty) {
            ts.cur = next;
            return;
        }
    } else {
        uint32_t first = cur & kLeafFirstMask;
        uint32_t cnt = ((cur >> kLeafCountShift) & 0xFu) + 1u;
        for (uint32_t k = 0; k < cnt; ++k) {
            const float4 *tp = S.tris + 3 * size_t(first + k);
            float4 a = tp[0], b = tp[1], c = tp[2];
            float tt, b1, b2;
            if (tri_hit(ts.sr, f3{a.x, a.y, a.z}, f3{b.x, b.y, b.z}, f3{c.x, c.y, c.z}, ts.limit, tt, b1, b2) &&
                tt <= ts.tmax) {
                ts.occluded = true;
                ts.done = true;
                return;
            }
        }
    }
    if (st.empty()) {
        ts.done = true;
        return;
    }
    float tn;
    st.pop(ts.cur, tn);
}

// ------------------------------------------------------------- shading --
// H::fromEular (src/util/hemisphere.hpp:66-69)
__device__ __forceinline__ f3 from_euler(float sinth, float costh, float phi) {
    float sp, cp;
    sincosf(phi, &sp, &cp);
    return {sinth * sp, sinth * cp, costh};
}

// H::sampleCos (src/util/hemisphere.hpp:76-95)
__device__ __forceinline__ f3 sample_cos(float r0, float r1) {
    float ux = 2.f * r0 - 1.f, uy = 2.f * r1 - 1.f;
    if (ux == 0.f && uy == 0.f) return {0.f, 0.f, 1.f};
    float r, phi;
    if (ux * ux > uy * uy) {
        r = ux;
        phi = (kPi / 4.f) * (uy / ux);
    } else {
        r = uy;
        phi = (kPi / 2.f) - ((kPi / 4.f) * (ux / uy));
    }
    float sp, cp;
    sincosf(phi, &sp, &cp);
    return {r * cp, r * sp, sqrtf(smax(0.f, 1.f - r * r))};
}

// BeckmannDistribution::sample / f (src/ext/distribution/beckmann.hpp:19-38)
__device__ __forceinline__ f3 beckmann_sample(float u0, float u1, float alpha2) {
    float tg2th = -alpha2 * logf(u0);
    float phi = u1 * 2.f * kPi;
    float costh = 1.f / sqrtf(1.f + tg2th);
    float sinth = sqrtf(smax(0.f, 1.f - costh * costh));
    return from_euler(sinth, costh, phi);
}
__device__ __forceinline__ float beckmann_f(f3 w, float alpha2) {
    float cos2 = w.z * w.z;
    float tg2 = smax(0.f, 1.f - cos2) / cos2;
    return expf(-tg2 / alpha2) / (kPi * alpha2 * (cos2 * cos2)) * fabsf(w.z);
}
__device__ __forceinline__ float cos_f(f3 w) { return w.z >= 0.f ? kInvPi : 0.f; }

// Fresnel::compute (src/ext/scala/fresnel.hpp:15-30): dielectric
// reflectance of wo for a surface of index ior.
__device__ __forceinline__ float fresnel(float ior, f3 wo) {
    float cosi = wo.z;
    float e = cosi > 0.f ? 1.f / ior : ior;
    cosi = fabsf(cosi);
    float sini = sqrtf(smax(0.f, 1.f - cosi * cosi));
    float sinr = e * sini;
    if (sinr >= 1.f) return 1.f;
    float cosr = sqrtf(smax(0.f, 1.f - sinr * sinr));
    float cosior = cosi / cosr;
    float rparl = (cosior - e) / (cosior + e);
    float rperp = (e * cosior - 1.f) / (e * cosior + 1.f);
    return .5f * (rparl * rparl + rperp * rperp);
}

// refract (src/vec/vmath.hpp:495-503): Snell through the microfacet normal N.
__device__ __forceinline__ bool refract_dir(f3 I, f3 N, float eta, f3 &out) {
    float costh = dot(N, I);
    float sinth2 = smax(0.f, 1.f - costh * costh);
    float sinphi2 = sinth2 * eta * eta;
    if (sinphi2 >= 1.f) return false;
    float cosphi = sqrtf(1.f - sinphi2);
    out = normalize((-I) * eta + N * (eta * costh - cosphi));
    return true;
}

__device__ __forceinline__ float dist_f(const ShaderRec &sh, f3 w) {
    return sh.dist == KR_DIST_COSINE ? cos_f(w) : beckmann_f(w, sh.alpha2);
}
__device__ __forceinline__ f3 dist_sample(const ShaderRec &sh, float u0, float u1) {
    return sh.dist == KR_DIST_COSINE ? sample_cos(u0, u1) : beckmann_sample(u0, u1, sh.alpha2);
}

// Shader::execute(compute_f_by_wi_wo) of a leaf shader: Lambert / Emission
// (lambert.hpp:31-33, emission.hpp:22-27), GlossyBSDF (glossy.hpp:28-30),
// GlassBSDF (glass.hpp:27-34, 48-51; uc is its Fresnel pick), and f = 0 for
// RefractionBSDF, Transparent, Blackhole.
__device__ __forceinline__ f3 shader_eval(const ShaderRec &sh, f3 wo, f3 wi, float uc = 0.f) {
    f3 c{sh.cr[0], sh.cr[1], sh.cr[2]};
    if (sh.type == KR_SHADER_GLOSSY) return c * dist_f(sh, normalize(wo + wi));
    if (sh.type == KR_SHADER_GLASS) {
        if (uc < fresnel(sh.ior, wo)) return c * dist_f(sh, normalize(wo + wi));
        return c * dist_f(sh, normalize(wo + wi * sh.ior));
    }
    if (sh.type == KR_SHADER_LAMBERT || sh.type == KR_SHADER_EMISSION) return c * cos_f(wi);
    return f3{0.f, 0.f, 0.f};
}

// Refraction branch of RefractionBSDF / GlassBSDF (refraction.hpp:24-34,
// glass.hpp:37-46); wi stays as it was when the refraction fails.
__device__ __forceinline__ void refraction_sample(const ShaderRec &sh, f3 wo, float u0, float u1, f3 &wi, f3 &f) {
    f3 h = dist_sample(sh, u0, u1);
    float e = wo.z > 0.f ? 1.f / sh.ior : sh.ior;
    if (wo.z < 0.f) h = -h;
    f3 r;
    if (refract_dir(wo, h, e, r)) {
        wi = r;
        float den = e < 1.f ? e * e : 1.f;
        f = f3{sh.cr[0] / den, sh.cr[1] / den, sh.cr[2] / den};
    } else {
        f = f3{0.f, 0.f, 0.f};
    }
}

// Shader::execute(sample_wi_f_by_wo) of a leaf shader (lambert.hpp:26-30,
// glossy.hpp:24-27, glass.hpp:22-46, refraction.hpp:24-34,
// transparent.hpp:20-23, blackhole.hpp:14-21). wi is in/out: a failed
// refraction keeps the caller's value, as the reference's Varyings does.
__device__ __forceinline__ void shader_sample(const ShaderRec &sh, f3 wo, float u0, float u1, float uc, f3 &wi,
                                              f3 &f) {
    f3 c{sh.cr[0], sh.cr[1], sh.cr[2]};
    if (sh.type == KR_SHADER_LAMBERT || sh.type == KR_SHADER_EMISSION) {
        wi = sample_cos(u0, u1);
        f = c * cos_f(wi) / kInvPi;
    } else if (sh.type == KR_SHADER_GLOSSY || (sh.type == KR_SHADER_GLASS && uc < fresnel(sh.ior, wo))) {
        f3 h = dist_sample(sh, u0, u1);
        wi = h * (dot(h, wo) * 2.f) - wo;  // reflect (src/vec/vmath.hpp:470-473)
        f = wo.z * wi.z > 0.f ? c : f3{0.f, 0.f, 0.f};
    } else if (sh.type == KR_SHADER_GLASS || sh.type == KR_SHADER_REFRACTION) {
        refraction_sample(sh, wo, u0, u1, wi, f);
    } else if (sh.type == KR_SHADER_TRANSPARENT) {
        wi = -wo;
        f = c;
    } else {  // Blackhole
        wi = f3{0.f, 0.f, 1.f};
        f = f3{0.f, 0.f, 0.f};
    }
}

// Mixed::execute (src/ext/shader/mixed.hpp:22-27) down to a leaf: one draw
// per Mixed level (b when u < fac), then GlassBSDF's Fresnel pick, all from
// one Philox block of the choice stream (at most 3 levels + 1 pick).
constexpr uint32_t kStreamChoice = 1u;
__device__ __forceinline__ float pick4(const Philox4 &r, int k) {
    return u01(k == 0 ? r.v[0] : (k == 1 ? r.v[1] : (k == 2 ? r.v[2] : r.v[3])));
}
__device__ __forceinline__ uint32_t resolve_leaf(const DevScene &S, uint32_t sid, f3 wo, uint32_t pixel,
                                                 uint32_t sample, uint32_t block, uint32_t seed_lo, uint32_t seed_hi,
                                                 float &uc) {
    uc = 0.f;
    uint32_t t = S.shaders[sid].type;
    if (t != KR_SHADER_MIXED && t != KR_SHADER_GLASS) return sid;
    Philox4 r = philox(pixel, sample, block, kStreamChoice, seed_lo, seed_hi);
    int k = 0;
    for (int level = 0; level < 3 && t == KR_SHADER_MIXED; ++level) {
        const ShaderRec &m = S.shaders[sid];
        float fc = m.fac_type == KR_SCALA_FRESNEL ? fresnel(m.fac_value, wo) : m.fac_value;
        float u = pick4(r, k++);
        sid = u < fc ? m.child1 : m.child0;
        t = S.shaders[sid].type;
    }
    if (t == KR_SHADER_GLASS) uc = pick4(r, k);
    return sid;
}

}  // namespace kr
