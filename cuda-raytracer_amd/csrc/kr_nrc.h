// kr_nrc.h -- neural radiance cache on gfx950 MFMA: a 96-64-64-64-64-64-16
// ReLU MLP with BF16 weights and activations and FP32 accumulation, trained
// online with Adam (FP32 master weights).
//
// Every layer is evaluated transposed, Y^T = W^T X^T, on a wave of 64
// samples (4 tiles of 16). In that form the 16x16 output tile a lane holds
// (4 consecutive output features of one sample) is exactly the B-operand
// fragment the next layer needs, so activations stay in registers from the
// input encoding to the output.
//
// Fragment layout of v_mfma_f32_16x16x16_bf16 (wave64), lane l = 16 q + s:
//   A[m][k]: lane holds A[s][4q + r], r = 0..3
//   B[k][n]: lane holds B[4q + r][s]
//   D[m][n]: lane holds D[4q + r][s]
// The forward pass uses gfx950's v_mfma_f32_16x16x32_bf16 (twice the rate):
// its A and B fragments hold 8 k values per lane, and a lane's 8 slots are
// contracted against the same 8 slots of the other operand. Filling slots
// 0-3 from 16-wide k tile 2p and slots 4-7 from tile 2p + 1, for A and B
// alike, assigns every k of the 32 to exactly one (lane group, slot) pair in
// both operands, so the sum is the same contraction without moving data
// between lanes.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace kr {
namespace nrc {

constexpr int kIn = 96, kInFeatures = 80, kWidth = 64, kOut = 16, kLayers = 6, kFreqs = 12;
constexpr int kRaw = 16;  // vertex record: pos(3) wo(3) normal(3) roughness(1) albedo(3) pad(3)
constexpr int kDims[kLayers + 1] = {kIn, kWidth, kWidth, kWidth, kWidth, kWidth, kOut};
// flat FP32 parameter order: W0 (in x out, row major), b0, W1, b1, ...
constexpr int w_off(int l) {
    int o = 0;
    for (int i = 0; i < l; ++i) o += kDims[i] * kDims[i + 1] + kDims[i + 1];
    return o;
}
constexpr int b_off(int l) { return w_off(l) + kDims[l] * kDims[l + 1]; }
constexpr int kNumParams = w_off(kLayers);
// BF16 weight copies (no biases): layer l starts at bw_off(l)
constexpr int bw_off(int l) {
    int o = 0;
    for (int i = 0; i < l; ++i) o += kDims[i] * kDims[i + 1];
    return o;
}
constexpr int kNumWeights = bw_off(kLayers);
// biases alone, layer after layer: layer l starts at bb_off(l)
constexpr int bb_off(int l) {
    int o = 0;
    for (int i = 0; i < l; ++i) o += kDims[i + 1];
    return o;
}
constexpr int kNumBiases = bb_off(kLayers);

typedef short v4s __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// One 16x16x32 step over k tiles (2p, 2p + 1): acc += A B with the pairing
// described above.
__device__ __forceinline__ v4f mfma_pair(v4s a0, v4s a1, v4s b0, v4s b1, v4f acc) {
    v8s a = __builtin_shufflevector(a0, a1, 0, 1, 2, 3, 4, 5, 6, 7);
    v8s b = __builtin_shufflevector(b0, b1, 0, 1, 2, 3, 4, 5, 6, 7);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), acc, 0,
                                                   0, 0);
}

struct Weights {
    const uint16_t *wt;  // per layer W^T [out][in], bf16
    const uint16_t *w;   // per layer W   [in][out], bf16
    const float *master; // FP32 parameters (biases read from here)
    float box_min[3], box_scale[3];
};

typedef __bf16 v4bf __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (gfx950's v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint16_t to_bf16(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

__device__ __forceinline__ v4s pack4(float a, float b, float c, float d) {
    v4bf v = {(__bf16)a, (__bf16)b, (__bf16)c, (__bf16)d};
    return __builtin_bit_cast(v4s, v);
}

__device__ __forceinline__ void angles(float x, float y, float z, float &th, float &ph) {
    th = acosf(fminf(fmaxf(z, -1.f), 1.f)) * (1.0f / 3.14159265358979f);
    ph = atan2f(y, x) * (0.5f / 3.14159265358979f) + 0.5f;
}

// A vertex record's fields in registers.
struct RawRec {
    float px, py, pz, wx, wy, wz, nx, ny, nz, rough, ar, ag, ab;
};

__device__ __forceinline__ RawRec load_raw(const float *raw) {
    const float4 *r4 = reinterpret_cast<const float4 *>(raw);
    float4 a = r4[0], b = r4[1], c = r4[2], d = r4[3];
    return RawRec{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w, d.x};
}

// Position feature idx < 72 of the normalised position (px, py, pz):
// sin / cos(2^k pi p_c) with k = idx / 6, j = idx % 6, c = j % 3, cos for
// j >= 3. 2^k pi p = 2 pi (2^(k-1) p), and both the revolution count
// 2^(k-1) p and its fractional part are exact in FP32, so the hardware's
// v_sin / v_cos (which take revolutions) see an exactly reduced argument.
__device__ __forceinline__ float pos_feature(int idx, float px, float py, float pz) {
    int k = idx / 6, j = idx - 6 * k;
    int c = j >= 3 ? j - 3 : j;
    float p = c == 0 ? px : (c == 1 ? py : pz);
    float rev = __builtin_amdgcn_fractf(ldexpf(p, k - 1));
    return j >= 3 ? __builtin_amdgcn_cosf(rev) : __builtin_amdgcn_sinf(rev);
}

// Feature `idx` of a vertex record: the position features above, the
// outgoing direction and normal as (theta/pi, phi/2pi + 1/2), roughness,
// albedo, zero padding.
__device__ __forceinline__ float feature(int idx, const RawRec &r, const Weights &W) {
    if (idx < 72)
        return pos_feature(idx, (r.px - W.box_min[0]) * W.box_scale[0], (r.py - W.box_min[1]) * W.box_scale[1],
                           (r.pz - W.box_min[2]) * W.box_scale[2]);
    if (idx < 76) {
        bool nrm = idx >= 74;
        float x = nrm ? r.nx : r.wx, y = nrm ? r.ny : r.wy, z = nrm ? r.nz : r.wz;
        float th, ph;
        angles(x, y, z, th, ph);
        return (idx & 1) ? ph : th;
    }
    if (idx == 76) return r.rough;
    if (idx == 77) return r.ar;
    if (idx == 78) return r.ag;
    if (idx == 79) return r.ab;
    return 0.f;
}

// Input fragments of a wave's 64 samples: in[kt][nt] for kt < 6, nt < 4.
// raw_of(n) gives the record of sample n of the wave (nullptr: padding,
// encoded as zeros). Lane (s, q) holds features 16 kt + 4 q .. + 3: for
// kt < 4 and for kt = 4, q < 2 these are position features; kt = 4, q = 2
// the two direction pairs; kt = 4, q = 3 roughness and albedo; kt = 5 is
// padding. Each lane evaluates only its own features.
template <class RawOf>
__device__ __forceinline__ void encode_inputs(v4s in[6][4], RawOf raw_of, const Weights &W) {
    int lane = __lane_id(), s = lane & 15, q = lane >> 4;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const float *raw = raw_of(nt * 16 + s);
        float f[20];
#pragma unroll
        for (int i = 0; i < 20; ++i) f[i] = 0.f;
        if (raw) {
            RawRec r = load_raw(raw);
            float px = (r.px - W.box_min[0]) * W.box_scale[0], py = (r.py - W.box_min[1]) * W.box_scale[1],
                  pz = (r.pz - W.box_min[2]) * W.box_scale[2];
#pragma unroll
            for (int kt = 0; kt < 4; ++kt)
#pragma unroll
                for (int i = 0; i < 4; ++i) f[4 * kt + i] = pos_feature(kt * 16 + 4 * q + i, px, py, pz);
            if (q < 2) {
#pragma unroll
                for (int i = 0; i < 4; ++i) f[16 + i] = pos_feature(64 + 4 * q + i, px, py, pz);
            } else if (q == 2) {
                angles(r.wx, r.wy, r.wz, f[16], f[17]);
                angles(r.nx, r.ny, r.nz, f[18], f[19]);
            } else {
                f[16] = r.rough;
                f[17] = r.ar;
                f[18] = r.ag;
                f[19] = r.ab;
            }
        }
#pragma unroll
        for (int kt = 0; kt < 5; ++kt) in[kt][nt] = pack4(f[4 * kt], f[4 * kt + 1], f[4 * kt + 2], f[4 * kt + 3]);
        in[5][nt] = v4s{0, 0, 0, 0};
    }
}

// Hidden layer: Y(mt, nt) = ReLU(sum_kt W^T(mt, kt) X^T(kt, nt) + b), packed
// to BF16 fragments for the next layer. KT = in/16 (even), MT = out/16.
template <int KT, int MT>
__device__ __forceinline__ void hidden_layer(const v4s x[KT][4], v4s y[MT][4], const uint16_t *wt, const float *bias) {
    static_assert(KT % 2 == 0, "k tiles are contracted in pairs");
    int lane = __lane_id(), s = lane & 15, q = lane >> 4;
    constexpr int IN = KT * 16;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        v4s a[KT];
#pragma unroll
        for (int kt = 0; kt < KT; ++kt)
            a[kt] = *reinterpret_cast<const v4s *>(wt + (mt * 16 + s) * IN + kt * 16 + 4 * q);
        float4 bb = *reinterpret_cast<const float4 *>(bias + mt * 16 + 4 * q);
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int p = 0; p < KT / 2; ++p) acc = mfma_pair(a[2 * p], a[2 * p + 1], x[2 * p][nt], x[2 * p + 1][nt], acc);
            y[mt][nt] = pack4(fmaxf(acc[0] + bb.x, 0.f), fmaxf(acc[1] + bb.y, 0.f), fmaxf(acc[2] + bb.z, 0.f),
                              fmaxf(acc[3] + bb.w, 0.f));
        }
    }
}

// Output layer (64 -> 16, linear): out[nt] = outputs 4q..4q+3 of sample nt*16 + s.
__device__ __forceinline__ void output_layer(const v4s x[4][4], v4f out[4], const uint16_t *wt, const float *bias) {
    int lane = __lane_id(), s = lane & 15, q = lane >> 4;
    v4s a[4];
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) a[kt] = *reinterpret_cast<const v4s *>(wt + s * kWidth + kt * 16 + 4 * q);
    float4 bb = *reinterpret_cast<const float4 *>(bias + 4 * q);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int p = 0; p < 2; ++p) acc = mfma_pair(a[2 * p], a[2 * p + 1], x[2 * p][nt], x[2 * p + 1][nt], acc);
        acc[0] += bb.x;
        acc[1] += bb.y;
        acc[2] += bb.z;
        acc[3] += bb.w;
        out[nt] = acc;
    }
}

// Forward of a wave's 64 samples; out[nt] in lanes q = 0 holds the RGB
// prediction of sample nt*16 + s. wt: the BF16 W^T of every layer
// (bw_off); bias: the FP32 biases, layer l at b_off(l) of the parameter
// vector (PACKED = false) or at bb_off(l) of a biases-only copy (PACKED =
// true, the LDS copy of k_nrc_infer).
template <bool PACKED>
__device__ __forceinline__ void forward_from(const v4s in[6][4], v4f out[4], const uint16_t *wt, const float *bias) {
    // The empty asm makes the offsets opaque on every call, so the compiler
    // cannot hoist all 48 KB of weight fragments out of a caller's sample
    // loop into registers (which leaves one wave per SIMD, or spills); they
    // are reloaded per layer instead (from L1/L2, or from LDS). The pointers
    // themselves stay visible, so a __shared__ copy is still read with LDS
    // instructions.
    uint32_t z = 0;
    asm volatile("" : "+s"(z));
    wt += z;
    bias += z;
    auto bo = [](int l) { return PACKED ? bb_off(l) : b_off(l); };
    v4s h[4][4], g[4][4];
    hidden_layer<6, 4>(in, h, wt + bw_off(0), bias + bo(0));
    hidden_layer<4, 4>(h, g, wt + bw_off(1), bias + bo(1));
    hidden_layer<4, 4>(g, h, wt + bw_off(2), bias + bo(2));
    hidden_layer<4, 4>(h, g, wt + bw_off(3), bias + bo(3));
    hidden_layer<4, 4>(g, h, wt + bw_off(4), bias + bo(4));
    output_layer(h, out, wt + bw_off(5), bias + bo(5));
}

__device__ __forceinline__ void forward(const v4s in[6][4], v4f out[4], const Weights &W) {
    forward_from<false>(in, out, W.wt, W.master);
}

}  // namespace nrc
}  // namespace kr
