// kr_lib.hip -- C ABI (include/kr_api.h) and the wavefront Radiance kernels
// for gfx950.
//
// Per chunk of P camera paths (one slot per path), each bounce b runs:
//   k_trace   closest hit of every queued ray; hits go to per-material
//             buckets (Lambert / GlossyBSDF / Emission) by wave-aggregated
//             appends, misses end their path
//   k_shade   one pass over the buckets: frame, area-light pick and sample,
//             BSDF eval (NEE), emission, BSDF sample, throughput, Russian
//             roulette; appends shadow segments and continuation rays
//   k_shadow  any-hit test of the shadow segments; adds unoccluded NEE
// and after the last bounce k_accumulate adds each path's radiance to its
// pixel in FP64 (NaN/Inf samples rejected, src/core/tracer/cpuMulticore.cu:58-64).
// Every kernel is a grid-stride loop over a queue length held in device
// memory, so no bounce needs a host round trip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "bvh_build.h"
#include "kr_api.h"
#include "kr_device.h"
#include "kr_layout.h"
#include "kr_nrc.h"

using namespace kr;

// ------------------------------------------------------------- errors --
namespace {
thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define KR_HIP(call)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail(e_ == hipErrorOutOfMemory ? KR_E_OOM : KR_E_HIP,                       \
                        std::string(#call) + ": " + hipGetErrorString(e_));                    \
    } while (0)

constexpr int kBlock = 256;
// Traversal stack entries per lane kept in LDS, and the waves per SIMD the
// traversal kernels and the shading kernel ask the register allocator for
// (1: its own choice). At 8 waves per SIMD the traversal kernels fit 64
// VGPRs with a few spilled; with a 10-entry LDS stack (20 KB per block, 8
// blocks = the CU's 160 KB) that measured 3 % faster on the 100k-triangle
// frame and 3 % on the 1M-triangle frame than 6 waves (70 VGPRs, no spill);
// 7 waves measured 1 % slower on the four-stream 1M-triangle frame.
#ifndef KR_LDS_DEPTH
#define KR_LDS_DEPTH 10
#endif
#ifndef KR_TRACE_WAVES
#define KR_TRACE_WAVES 8
#endif
#ifndef KR_SHADE_OCTANT_SORT
#define KR_SHADE_OCTANT_SORT 1
#endif

#ifndef KR_SHADOW_STEPS
#define KR_SHADOW_STEPS 1
#endif
// The plain single-render shading kernel asks for 6 waves per SIMD (80
// VGPRs, no spill; its own choice was 90 VGPRs, 5 waves): the four-stream
// 1M-triangle frame 71.5 -> 71.0 ms. The cache, profiler and batch
// instantiations need more registers and keep the compiler's choice.
#ifndef KR_SHADE_WAVES
#define KR_SHADE_WAVES 6
#endif
constexpr int kLdsDepth = KR_LDS_DEPTH;
// The wave hint applies to the BVH2 / BVH4 traversals of single renders
// (full or compressed nodes); the BVH8 and job-mixing batch variants need
// more registers and would spill.
constexpr int trace_waves(int w, bool batch, bool) { return (w <= 4 && !batch) ? KR_TRACE_WAVES : 1; }
constexpr int kLdsDepthMin = kLdsDepth;
// traversal stack entries per lane beyond the LDS part, in HBM (108 in all:
// BVH8 trees to depth 14, BVH4 to depth 34)
constexpr int kSpillDepth = 108 - kLdsDepthMin;

// Host Philox4x32-10 for the light tables (same function as the device one).
void philox_host(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = uint64_t(0xD2511F53u) * c0;
        uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
        uint32_t n0 = uint32_t(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = uint32_t(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = uint32_t(p1);
        c2 = n2;
        c3 = uint32_t(p0);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}
inline float u01h(uint32_t x) { return float(x >> 8) * (1.0f / 16777216.0f); }
constexpr uint32_t kStreamLight = 0x4C495445u;

struct H3 {
    float x, y, z;
};
inline H3 hsub(H3 a, H3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline H3 hadd(H3 a, H3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline H3 hmul(H3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float hdot(H3 a, H3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline H3 hcross(H3 a, H3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline H3 hnorm(H3 a) { return hmul(a, 1.0f / std::sqrt(hdot(a, a))); }
inline H3 hload(const float *p) { return {p[0], p[1], p[2]}; }

}  // namespace

// ---------------------------------------------------------- handle types --
// Device memory cache: every buffer of the library's scenes, renders,
// batches and radiance caches comes from its device's cache. hipFree waits
// for the whole device and unmaps the pages (40 to 300 ms for the ~60
// buffers of a 1080p render), so freed buffers are kept, keyed by size, and
// handed to the next request they fit (at most 1/8 larger). kr_device_trim
// and kr_device_close give them back to HIP, and a request that HIP cannot
// serve first releases the idle buffers and then tries once more. Callers
// free a buffer only after the work that uses it is done (every *_free
// synchronises its streams first).
struct MemCache {
    std::mutex mu;
    std::multimap<size_t, void *> idle;           // bytes -> idle buffer
    std::unordered_map<void *, size_t> bytes_of;  // every buffer made, idle or in use
    size_t idle_bytes = 0;
};

struct kr_device {
    int ordinal = 0;
    hipStream_t stream = nullptr;
    int num_cu = 0;
    std::string name;
    // device address of this device's copy of `scratch` (kr_trace_closest_noalloc)
    uint2 *scratch = nullptr;
    mutable MemCache mem;
};

namespace {
// requests round up to 256 B, and from 64 KiB on to 64 KiB
size_t cache_class(size_t bytes) {
    if (bytes == 0) bytes = 1;
    const size_t g = bytes >= (size_t(64) << 10) ? (size_t(64) << 10) : 256;
    return (bytes + g - 1) / g * g;
}

void cache_release(const kr_device *dev) {
    std::lock_guard<std::mutex> lk(dev->mem.mu);
    for (auto &kv : dev->mem.idle) {
        dev->mem.bytes_of.erase(kv.second);
        (void)hipFree(kv.second);
    }
    dev->mem.idle.clear();
    dev->mem.idle_bytes = 0;
}

int cache_alloc(const kr_device *dev, size_t bytes, void **p) {
    *p = nullptr;
    const size_t want = cache_class(bytes);
    {
        std::lock_guard<std::mutex> lk(dev->mem.mu);
        auto it = dev->mem.idle.lower_bound(want);
        if (it != dev->mem.idle.end() && it->first <= want + want / 8) {
            *p = it->second;
            dev->mem.idle_bytes -= it->first;
            dev->mem.idle.erase(it);
            return KR_OK;
        }
    }
    hipError_t e = hipMalloc(p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        cache_release(dev);
        e = hipMalloc(p, want);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(KR_E_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    std::lock_guard<std::mutex> lk(dev->mem.mu);
    dev->mem.bytes_of[*p] = want;
    return KR_OK;
}

template <class T>
int dalloc(const kr_device *dev, T **p, size_t n) {
    void *v = nullptr;
    int rc = cache_alloc(dev, (n ? n : 1) * sizeof(T), &v);
    *p = static_cast<T *>(v);
    return rc;
}

// returns a buffer of dalloc to the device's cache
void dfree(const kr_device *dev, void *p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(dev->mem.mu);
    auto it = dev->mem.bytes_of.find(p);
    if (it == dev->mem.bytes_of.end()) return;
    dev->mem.idle.emplace(it->second, p);
    dev->mem.idle_bytes += it->second;
}
}  // namespace

// Scratch memory of kr_trace_closest_noalloc: the traversal spill stacks of
// the lanes of one launch, (ref, t) pairs, entry k of lane g at k * lanes + g
// (the layout make_stack expects). A file-scope device array, so that the
// entry point allocates nothing: every GPU the library runs on holds its own
// copy, whose device address kr_device_open looks up once per device with
// hipGetSymbolAddress. Host code uses that address, never the array's name
// (on the host the name is only a stand-in symbol). All calls on one GPU
// share the array, so calls that overlap on one GPU race.
constexpr size_t kScratchBytes = size_t(32) << 20;
static __device__ __align__(256) uint2 scratch[kScratchBytes / sizeof(uint2)];

struct kr_scene {
    kr_device *dev = nullptr;
    DevScene view{};
    uint32_t width = 4;
    bool qnodes = false;  // compressed nodes (qnode_bytes) or full FP32 boxes
    std::vector<void *> allocs;
    // host copies needed to answer queries
    std::vector<kr_prim_desc> prims;
    std::vector<uint32_t> tri_prim;
    std::vector<uint32_t> sphere_prim;
    std::vector<LightRec> lights;
    std::vector<float> light_table;
    kr_scene_info info{};
    // spill stack for the ray-query entry points
    uint32_t *spill = nullptr;
    size_t spill_lanes = 0;
};

struct RayQ {
    uint32_t *slot;
    float *ox, *oy, *oz, *dx, *dy, *dz, *bx, *by, *bz;
    uint32_t *flags;  // bit 0: training path; bits 1-2: diffuse/rough vertices so far
};

// Radiance-cache state of a render (all null / 0 when no cache is attached).
struct NrcRender {
    int enabled;
    uint32_t train_log2;   // one path in every group of 2^train_log2 consecutive paths trains
    uint32_t train_mask;   // 2^train_log2 - 1
    float rough_threshold; // GlossyBSDF roughness counted as rough
    uint32_t end_at;       // a path ends at its end_at-th diffuse or rough vertex
    int indirect;          // the cache holds what a vertex reflects from its
                           // continuation, without its own light sample and
                           // emission (KR_NRC_INDIRECT)
    uint32_t *tslot_of;    // slot -> training-path index or kMissCode
    float *rec;            // training vertex records [train_cap][nv][kRec]
    uint32_t *tcount;      // vertices recorded per training path
    uint32_t train_cap, nv;
    uint32_t *qslot;       // cache queries: slot, throughput, vertex record
    float *qbeta, *qraw;
};
struct ShadowQ {
    uint32_t *slot;
    float *ox, *oy, *oz, *dx, *dy, *dz, *tmax, *cx, *cy, *cz;
};
struct HitQ {
    float *t, *u, *v;
    uint32_t *code;
};
// Rays or segments whose traversal used up its step budget in a launch, with
// what the next launch needs to resume it (SoA, `cap` entries): the ray's
// index in its queue, the next node or leaf ref, the stack depth (and, for a
// closest hit, the material bucket of the hit so far in bits 16-17; the hit
// itself waits in the hit queue at the ray's index) and the stack entries,
// entry k of slot j at stk[k * cap + j].
struct TailQ {
    uint32_t *idx, *cur, *info;
    uint2 *stk;
    uint32_t cap;      // slots
    uint32_t stk_cap;  // stack entries per slot: a walk with a deeper stack finishes in its launch
};
struct LensDev {
    float n[3], u[3], v[3], du[3], dv[3], o[3];
    uint32_t lens, pw, w, h;
};
// Order of a chunk's camera rays in the first ray queue. A chunk of whole
// samples (g0 and P multiples of the pixel count) is queued tile by tile:
// queue position pos holds path g0 + j with
//   pos = ((group * ntiles + tile) * sg + k) * tw * th + p,
// pixel p of a tw x th pixel tile and sample k of a group of sg samples, so a
// wave of 64 positions traces tw * th neighbouring pixels at sg sub-pixel
// strata each: nearly identical primary rays, which visit the same BVH nodes.
// The path index, its RNG streams and its pixel are those of row-major order;
// only the queue order differs, so images are bit-identical. Disabled
// (row-major positions) when tw == 0.
struct TileOrder {
    uint32_t tw, th, sg;   // tile width, height (pixels), samples per group
    uint32_t tiles_x, ntiles;
};
// Profiler slices (core::Profiler over RadianceKernel::Slice): the paths of
// the pixels in [x0, x1) x [y0, y1) record their first `depth` vertices.
constexpr int kProfFloats = 14;  // o, d, f, L, theta, shader id (uint32 bits)
struct ProfView {
    int enabled;
    uint32_t x0, y0, x1, y1, w, npix, spp, depth;
    float *rec;  // [slice][depth][kProfFloats]
    float *fin;  // [slice][4]: path radiance, loop counter (uint32 bits)
};

// One job of a batch, as the kernels see it. A batch's path space is the
// concatenation of its jobs' sample spaces: job k owns global path indices
// [begin, begin + npix * spp), local index = s * npix + pixel as in a
// single render, so a job's paths, RNG streams and pixels are those of the
// same render done alone.
struct JobRec {
    DevScene S;
    LensDev lens;
    uint64_t begin;
    uint32_t npix, max_bounce, seed_lo, seed_hi;
};
struct JobTable {
    const JobRec *rec;
    const uint64_t *begin;  // begin of every job, ascending
    uint32_t n;
};

// counters layout
enum : int {
    C_RAY0 = 0,
    C_RAY1 = 1,
    C_BUCKET = 2,  // 3 entries
    C_SHADOW = 5,
    C_NRCQ = 6,   // cache queries
    C_TRAIN = 7,  // training paths
    C_FETCH = 8,  // persistent closest-hit traversal: next ray to fetch
    C_FETCH_SH = 9,  // persistent shadow traversal: next segment to fetch
    C_TAIL = 10,     // closest-hit walks handed to tail round r: C_TAIL + r (kMaxTailRounds + 1 entries)
    C_TAIL_SH = 15,  // the same for shadow segments
    C_NUM = 20
};
constexpr int kMaxTailRounds = 4;
static_assert(C_TAIL + kMaxTailRounds < C_TAIL_SH && C_TAIL_SH + kMaxTailRounds < C_NUM, "tail counters");
// Each counter sits on its own 128-byte line, so that the wave-aggregated
// atomics of different queues (fetch, buckets, shadow, next rays) do not
// serialise on one L2 channel.
constexpr int kCtrStride = 32;
#define CTR(i) (kCtrStride * (i))
enum : int { S_CLOSEST = 0, S_SHADOW = 1, S_INVALID = 2, S_PATHS = 3, S_NODES = 4, S_TRIS = 5, S_NRCQ = 6, S_NUM = 8 };
enum : int { K_TRACE = 0, K_SHADE = 1, K_SHADOW = 2, K_OTHER = 3, K_KINDS = 4 };

struct kr_nrc {
    kr_device *dev = nullptr;
    kr_nrc_config cfg{};
    float *master = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    uint16_t *wt = nullptr, *w = nullptr;
    // what renders query: the bias-corrected moving average of the trained
    // parameters (FP32) and its BF16 W^T; null when renders query the
    // trained parameters themselves
    float *ema = nullptr;
    uint16_t *wt_ema = nullptr;
    float *partial = nullptr;  // [batch / 64][kNumParams]
    float *loss = nullptr;     // device scalar
    uint32_t rows = 0;         // batch / 64
    uint32_t step = 0;         // Adam steps taken
    uint64_t samples = 0;
    float last_loss = 0.f;
    // training pool filled by renders (ring buffer)
    float *pool_raw = nullptr, *pool_tgt = nullptr;
    uint32_t *pool_total = nullptr;
    uint32_t pool_cap = 0;
    std::vector<void *> allocs;
};

// One set of chunk buffers and the stream its kernels run on. A render has
// KR_RENDER_PIPES pipes when its frames are large: consecutive chunks cycle
// through them, so the short late bounces of one chunk run beside the long
// early bounces of the others instead of leaving the GPU half idle; only the
// accumulations into the shared image are ordered between the pipes. Four
// pipes, one per hardware queue HIP gives a process by default, measured
// 1.16x (100k triangles) and 1.10x (1M) the frame rate of two; six and eight
// share those queues and were slower than four.
#ifndef KR_RENDER_PIPES
#define KR_RENDER_PIPES 4
#endif
constexpr int kMaxPipes = 4;
static_assert(KR_RENDER_PIPES >= 1 && KR_RENDER_PIPES <= kMaxPipes, "pipes per render");
struct Pipe {
    RayQ q[2]{};
    HitQ hit{};
    uint32_t *bucket[3]{};
    ShadowQ sq{};
    float *L = nullptr;  // 3 * P
    uint32_t *counters = nullptr;
    uint32_t *spill = nullptr;
    size_t spill_lanes = 0;
    hipStream_t stream = nullptr;  // pipe 0: the device's stream
    bool own_stream = false;
    hipEvent_t acc_done = nullptr;  // recorded after each of the pipe's accumulations
    TailQ tail[2]{};                // walks handed between tail rounds (cap 0: no tail rounds)
    NrcRender nr{};                 // radiance-cache buffers of the pipe's chunks
    double *acc = nullptr;          // the image the pipe accumulates into: the render's
    uint32_t *valid = nullptr;      // (shared, ordered by acc_done) or, in a batch, its own
};

struct kr_render {
    kr_scene *scene = nullptr;
    kr_camera cam{};
    kr_params prm{};
    LensDev lens{};
    uint32_t P = 0;  // paths per chunk
    uint32_t npix = 0;
    Pipe pipe[kMaxPipes];
    int npipes = 1;
    hipEvent_t fork = nullptr;  // pipes 1.. start after what pipe 0 has queued
    unsigned long long *stats = nullptr;
    double *acc = nullptr;
    uint32_t *valid = nullptr;
    int grid = 0;
    // tail rounds (k_trace_tail, k_shadow_tail): iteration budget of a launch,
    // follow-up launches per queue (the last one finishes every walk), first
    // bounce whose closest-hit walks use them (shadow segments: every bounce)
    uint32_t tail_steps = 0, tail_rounds = 0, tail_from = 1;
    std::vector<void *> allocs;
    double last_seconds = 0;
    uint64_t paths = 0;
    // profiling (KR_PROFILE_*): pending event pairs per kernel kind
    uint32_t profile = 0;
    std::vector<hipEvent_t> event_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
    double kind_ms[K_KINDS] = {0, 0, 0, 0};
    double kind_busy_ms[K_KINDS] = {0, 0, 0, 0};
    uint64_t kind_launches[K_KINDS] = {0, 0, 0, 0};
    // attached radiance cache
    kr_nrc *nrc = nullptr;
    NrcRender nr{};
    std::vector<void *> nrc_allocs;
    // profiler slices
    ProfView pv{};
    uint64_t prof_slices = 0;
    std::vector<void *> prof_allocs;
};

namespace {
void resolve_events(kr_render *r);
void nrc_after_chunk(kr_render *r, Pipe &pp, kr_nrc *n, const NrcRender &nr);
int nrc_after_run(kr_render *r, kr_nrc *n);
}  // namespace

// ------------------------------------------------------------- kernels --
namespace {

__device__ __forceinline__ uint32_t wave_append(uint32_t *counter, bool pred) {
    unsigned long long mask = __ballot(pred);
    uint32_t lane = __lane_id();
    uint32_t base = 0;
    if (mask) {
        int leader = __ffsll((unsigned long long)mask) - 1;
        if (int(lane) == leader) base = atomicAdd(counter, uint32_t(__popcll(mask)));
        base = __shfl(base, leader);
    }
    return base + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
}

// Reserves n consecutive entries of *counter for every lane of the wave with
// one atomic; returns the lane's first entry (lanes in order).
__device__ __forceinline__ uint32_t wave_reserve(uint32_t *counter, uint32_t n) {
    uint32_t lane = __lane_id();
    uint32_t incl = n;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(incl, o);
        if (lane >= uint32_t(o)) incl += y;
    }
    uint32_t total = __shfl(incl, 63);
    uint32_t base = 0;
    if (lane == 63 && total) base = atomicAdd(counter, total);
    base = __shfl(base, 63);
    return base + incl - n;
}

// Job owning global path g: the last k with begin[k] <= g.
__device__ __forceinline__ uint32_t find_job(const JobTable &T, uint64_t g) {
    uint32_t lo = 0, hi = T.n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (T.begin[mid] <= g)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Slice of global path g of a render (Profiler::at, kernel.hpp:62-68:
// ((y - y0) * (x1 - x0) + (x - x0)) * spp + k), or -1 outside the area.
__device__ __forceinline__ int64_t prof_slice(const ProfView &pv, uint64_t g) {
    uint32_t k = uint32_t(g / pv.npix), pixel = uint32_t(g % pv.npix);
    uint32_t x = pixel % pv.w, y = pixel / pv.w;
    if (x < pv.x0 || x >= pv.x1 || y < pv.y0 || y >= pv.y1) return -1;
    return (int64_t(y - pv.y0) * (pv.x1 - pv.x0) + (x - pv.x0)) * pv.spp + k;
}

// An unoccluded light sample's contribution, added to the L of the
// profiler record of its vertex (the reference adds it before recording).
__device__ __forceinline__ void prof_add_light(const ProfView &pv, uint64_t g, uint32_t bounce, float cx, float cy,
                                               float cz) {
    int64_t sl = prof_slice(pv, g);
    if (sl < 0 || bounce >= pv.depth) return;
    float *pr = pv.rec + (size_t(sl) * pv.depth + bounce) * kProfFloats;
    pr[9] += cx;
    pr[10] += cy;
    pr[11] += cz;
}

// The traversal stack of a lane of a kBlock-thread block over the block's
// __shared__ uint32_t[2 * LDS * kBlock] (refs, then entry distances).
template <int LDS>
__device__ __forceinline__ Stack<LDS, kBlock> make_stack(uint32_t *lds, uint32_t *spill, size_t spill_lanes) {
    Stack<LDS, kBlock> st;
    st.lds = reinterpret_cast<char *>(lds);
    st.off = threadIdx.x * 4u;
    st.gstride = uint32_t(spill_lanes);
    st.spill = reinterpret_cast<uint2 *>(spill) + size_t(blockIdx.x) * kBlock;
    return st;
}

// Ray-query entry point: closest hit of n rays (parity hook).
template <int W, bool Q>
__global__ __launch_bounds__(kBlock) void k_query_closest(DevScene S, const float *o, const float *d, uint32_t n,
                                                          float *t, uint32_t *prim, uint32_t *tri, float *u, float *v,
                                                          uint32_t *spill, size_t spill_lanes, uint64_t *visits) {
    __shared__ uint32_t stk[2 * kLdsDepth * kBlock];
    auto st = make_stack<kLdsDepth>(stk, spill, spill_lanes);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        f3 ro{o[3 * i], o[3 * i + 1], o[3 * i + 2]};
        f3 rd{d[3 * i], d[3 * i + 1], d[3 * i + 2]};
        uint64_t cnt[2] = {0, 0};
        HitRec h = trace_closest<W, kLdsDepth, Q>(S, ro, rd, st, visits ? cnt : nullptr);
        if (visits) {
            visits[2 * i] = cnt[0];
            visits[2 * i + 1] = cnt[1];
        }
        t[i] = h.t;
        if (h.code == kMissCode) {
            prim[i] = KR_MISS;
            tri[i] = 0;
            u[i] = v[i] = 0.f;
        } else if (h.code & kSphereBit) {
            prim[i] = S.sphere_prim[h.code & ~kSphereBit];
            tri[i] = 0;
            u[i] = v[i] = 0.f;
        } else {
            uint32_t p = S.tri_prim[h.code];
            prim[i] = p;
            tri[i] = h.code - S.prims[p].tri_begin;
            u[i] = h.u;
            v[i] = h.v;
        }
    }
}

template <int W, bool Q>
__global__ __launch_bounds__(kBlock) void k_query_shadow(DevScene S, const float *o, const float *d, const float *tmax,
                                                         uint32_t n, uint8_t *occ, uint32_t *spill,
                                                         size_t spill_lanes) {
    __shared__ uint32_t stk[2 * kLdsDepth * kBlock];
    auto st = make_stack<kLdsDepth>(stk, spill, spill_lanes);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        f3 ro{o[3 * i], o[3 * i + 1], o[3 * i + 2]};
        f3 rd{d[3 * i], d[3 * i + 1], d[3 * i + 2]};
        occ[i] = trace_occluded<W, kLdsDepth, Q>(S, ro, rd, tmax[i], st) ? 1 : 0;
    }
}

// Plain HBM copy (the bandwidth / counter calibration probe).
__global__ __launch_bounds__(kBlock) void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

// Clears the counters whose bit is set in `mask`.
__global__ void k_reset(uint32_t *counters, uint32_t mask) {
    int i = threadIdx.x;
    if (i < C_NUM && ((mask >> i) & 1u)) counters[CTR(i)] = 0;
}

// With a cache attached, the global path indices fall into groups of
// 2^train_log2 consecutive paths (consecutive pixels of one sample); in each
// group the member picked by a Philox hash of the group index is a training
// path, and its record slot is the group's index within the chunk, so no
// atomics are needed to hand out slots.
constexpr uint32_t kStreamTrainPick = 0x5452414Eu;
constexpr int kRecFloats = 28;  // vertex record: raw 16 | beta 3 | c 3 | beta*E 3 | pad 3

// PinholeLens / OrthographicLens::sample (src/core/misc/lens.hpp:55-67,
// 40-53) for sample s of pixel px.
__device__ __forceinline__ void camera_ray(const LensDev &lens, uint32_t s, uint32_t px, f3 &o, f3 &d) {
    uint32_t x = px % lens.w, y = px / lens.w;
    float fx = float(x) - float(lens.w) * 0.5f, fy = float(y) - float(lens.h) * 0.5f;
    float kx = float(s % lens.pw), ky = float(s / lens.pw);
    f3 n{lens.n[0], lens.n[1], lens.n[2]}, u{lens.u[0], lens.u[1], lens.u[2]};
    f3 v{lens.v[0], lens.v[1], lens.v[2]};
    f3 du{lens.du[0], lens.du[1], lens.du[2]}, dv{lens.dv[0], lens.dv[1], lens.dv[2]};
    o = f3{lens.o[0], lens.o[1], lens.o[2]};
    if (lens.lens == KR_LENS_ORTHOGRAPHIC) {
        f3 dd = u * fx + v * fy;
        dd = dd + (du * kx + dv * ky);
        o = o + dd * 4.f / float(lens.w);
        d = normalize(n);
    } else {
        f3 t = n + u * fx + v * fy;
        d = normalize(t + du * kx + dv * ky);
    }
}

// Camera paths of a chunk: slot j is global path g0 + j. BATCH: the path's
// job (and its lens and local index) comes from the job table. With a
// profiler only the paths of its area are traced (CPUMulticoreTracer traces
// only the profiled area, src/core/tracer/cpuMulticore.cu:17): the queue is
// then compacted and the other slots keep zero radiance.
template <bool BATCH>
__global__ __launch_bounds__(kBlock) void k_camera(LensDev lens, JobTable T, uint64_t g0, uint32_t P, uint32_t npix,
                                                   RayQ q, float *L, uint32_t *counters, int qi, NrcRender nr,
                                                   uint32_t seed_lo, uint32_t seed_hi, ProfView pv, TileOrder to) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < P; base += stride) {
        uint32_t lin = base + threadIdx.x;
        bool active = lin < P;
        // the path this queue position holds (TileOrder)
        uint32_t j = lin;
        if (!BATCH && to.tw && active) {
            uint32_t tp = to.tw * to.th;
            uint32_t p = lin % tp, r = lin / tp;
            uint32_t k = r % to.sg;
            r /= to.sg;
            uint32_t tile = r % to.ntiles, grp = r / to.ntiles;
            uint32_t x = (tile % to.tiles_x) * to.tw + p % to.tw, y = (tile / to.tiles_x) * to.th + p / to.tw;
            j = (grp * to.sg + k) * npix + y * lens.w + x;
        }
        if (active) {
            // every slot's radiance starts at zero (a bijection of positions)
            L[3 * lin] = 0.f;
            L[3 * lin + 1] = 0.f;
            L[3 * lin + 2] = 0.f;
        }
        bool train = false;
        f3 o{0.f, 0.f, 0.f}, d{0.f, 0.f, 0.f};
        if (active) {
            uint64_t g = g0 + j;
            uint32_t s, px;
            if (BATCH) {
                const JobRec &J = T.rec[find_job(T, g)];
                uint64_t l = g - J.begin;
                s = uint32_t(l / J.npix);
                px = uint32_t(l % J.npix);
                camera_ray(J.lens, s, px, o, d);
            } else {
                s = uint32_t(g / npix);
                px = uint32_t(g % npix);
                camera_ray(lens, s, px, o, d);
            }
        }
        uint32_t flags = 0;
        if (!BATCH && nr.enabled && active) {
            uint64_t g = g0 + j;
            uint64_t grp = g >> nr.train_log2;
            uint32_t ts = uint32_t(grp - (g0 >> nr.train_log2));
            uint32_t member = uint32_t(g) & nr.train_mask;
            Philox4 h = philox(uint32_t(grp), uint32_t(grp >> 32), 0xFFFFFFFFu, kStreamTrainPick, seed_lo, seed_hi);
            train = member == (h.v[0] & nr.train_mask);
            // the group's first slot in the chunk clears its record count, so
            // a group whose training member lies in another chunk records nothing
            if (train || member == 0 || j == 0) nr.tcount[ts] = 0;
            if (train) flags = 1u;
            nr.tslot_of[j] = train ? ts : kMissCode;
        }
        bool emit = active;
        uint32_t pos = lin;
        if (!BATCH && pv.enabled) {
            emit = active && prof_slice(pv, g0 + j) >= 0;
            pos = wave_append(&counters[CTR(qi)], emit);
        }
        if (emit) {
            q.slot[pos] = j;
            q.ox[pos] = o.x;
            q.oy[pos] = o.y;
            q.oz[pos] = o.z;
            q.dx[pos] = d.x;
            q.dy[pos] = d.y;
            q.dz[pos] = d.z;
            q.bx[pos] = 1.f;
            q.by[pos] = 1.f;
            q.bz[pos] = 1.f;
            q.flags[pos] = flags;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (BATCH || !pv.enabled) counters[CTR(qi)] = P;
        if (!BATCH && nr.enabled)
            counters[CTR(C_TRAIN)] = uint32_t(((g0 + P - 1) >> nr.train_log2) - (g0 >> nr.train_log2) + 1);
    }
}

template <int W, bool COUNT, bool BATCH, bool Q>
__global__ __launch_bounds__(kBlock, trace_waves(W, BATCH, Q)) void k_trace(DevScene S, JobTable T, uint64_t g0, RayQ q, int qi, HitQ hit,
                                                  uint32_t *b0, uint32_t *b1, uint32_t *b2, uint32_t *counters,
                                                  unsigned long long *stats, int sort, uint32_t *spill,
                                                  size_t spill_lanes) {
    __shared__ uint32_t stk[2 * kLdsDepth * kBlock];
    auto st = make_stack<kLdsDepth>(stk, spill, spill_lanes);
    uint32_t n = counters[CTR(qi)];
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&stats[S_CLOSEST], (unsigned long long)n);
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        uint32_t i = base + threadIdx.x;
        bool active = i < n;
        HitRec h;
        h.code = kMissCode;
        DevScene Sx = S;
        if (BATCH && active) Sx = T.rec[find_job(T, g0 + q.slot[i])].S;
        if (active) {
            f3 o{q.ox[i], q.oy[i], q.oz[i]}, d{q.dx[i], q.dy[i], q.dz[i]};
            if (COUNT) {
                uint64_t cnt[2] = {0, 0};
                h = trace_closest<W, kLdsDepth, Q>(Sx, o, d, st, cnt);
                atomicAdd(&stats[S_NODES], (unsigned long long)cnt[0]);
                atomicAdd(&stats[S_TRIS], (unsigned long long)cnt[1]);
            } else {
                h = trace_closest<W, kLdsDepth, Q>(Sx, o, d, st);
            }
            hit.t[i] = h.t;
            hit.u[i] = h.u;
            hit.v[i] = h.v;
            hit.code[i] = h.code;
        }
        bool hitp = active && h.code != kMissCode;
        uint32_t mat = 0;
        if (hitp && sort) {
            // Lambert, Glossy, everything else: stored with the triangle, looked up for spheres
            mat = (h.code & kSphereBit) ? min(Sx.shaders[Sx.prims[Sx.sphere_prim[h.code & ~kSphereBit]].shader].type, 2u)
                                        : h.mat;
        }
        uint32_t p0 = wave_append(&counters[CTR(C_BUCKET + 0)], hitp && mat == 0);
        if (hitp && mat == 0) b0[p0] = i;
        if (sort) {
            uint32_t p1 = wave_append(&counters[CTR(C_BUCKET + 1)], hitp && mat == 1);
            if (hitp && mat == 1) b1[p1] = i;
            uint32_t p2 = wave_append(&counters[CTR(C_BUCKET + 2)], hitp && mat == 2);
            if (hitp && mat == 2) b2[p2] = i;
        }
    }
}

// Closest-hit traversal in rounds ("tail rounds"). A wave of incoherent rays
// runs until its slowest ray is done: on the 1M-triangle frame a bounce ray
// takes 17 steps (one node or one leaf each) at the median and the slowest of
// 64 about 48, so most lanes idle through the end of every wave. Here a walk
// that has run `steps` iterations of closest_iter (a node and the leaf it
// enters, or a leaf) leaves its state in the tail queue `tout` and
// the lane moves on; the next launch (FIRST = false) resumes the queued walks
// packed into full waves from `tin`. A walk whose stack is deeper than the
// queue keeps, or that finds the queue full, finishes in its launch, and so
// does every walk of a launch with steps = ~0u. Walks are resumed exactly
// where they stopped, so hits are those of k_trace; only the order of the
// bucket entries differs.
template <int W, bool COUNT, bool Q, bool FIRST>
__global__ __launch_bounds__(kBlock, trace_waves(W, false, Q)) void k_trace_tail(
    DevScene S, RayQ q, int qi, HitQ hit, uint32_t *b0, uint32_t *b1, uint32_t *b2, uint32_t *counters,
    unsigned long long *stats, int sort, uint32_t *spill, size_t spill_lanes, TailQ tin, int cin, TailQ tout,
    int cout, uint32_t steps) {
    __shared__ uint32_t stk[2 * kLdsDepth * kBlock];
    auto st = make_stack<kLdsDepth>(stk, spill, spill_lanes);
    const uint32_t n = FIRST ? counters[CTR(qi)] : min(counters[CTR(cin)], tin.cap);
    if (FIRST && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&stats[S_CLOSEST], (unsigned long long)n);
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t j = base + threadIdx.x;
        const bool active = j < n;
        uint32_t i = 0;
        HitRec h;
        h.code = kMissCode;
        h.mat = 0;
        bool done = false;
        if (active) {
            i = FIRST ? j : tin.idx[j];
            f3 o{q.ox[i], q.oy[i], q.oz[i]}, d{q.dx[i], q.dy[i], q.dz[i]};
            RayBox rb = make_raybox(o, d);
            ShearRay sr = make_shear(o, d);
            uint32_t cur = 0;
            st.reset();
            if (FIRST) {
                h = closest_spheres(S, o, d);
                done = S.ntris == 0;
            } else {
                h.t = hit.t[i];
                h.u = hit.u[i];
                h.v = hit.v[i];
                h.code = hit.code[i];
                cur = tin.cur[j];
                const uint32_t info = tin.info[j];
                h.mat = info >> 16;
                const uint32_t dep = info & 0xFFFFu;
                for (uint32_t k = 0; k < dep; ++k) {
                    uint2 e = tin.stk[size_t(k) * tin.cap + j];
                    st.push(e.x, __uint_as_float(e.y));
                }
            }
            uint32_t nvis = 0, ntri = 0, budget = steps;
            while (!done) {
                done = closest_walk<W, kLdsDepth, Q>(S, rb, sr, st, cur, h, budget, nvis, ntri);
                if (done) break;
                // out of steps: queue the walk for the next round if its stack fits
                const uint32_t dep = st.depth();
                const bool fits = dep <= tout.stk_cap;
                const uint32_t pos = wave_append(&counters[CTR(cout)], fits);
                if (fits && pos < tout.cap) {
                    tout.idx[pos] = i;
                    tout.cur[pos] = cur;
                    tout.info[pos] = dep | (h.mat << 16);
                    for (uint32_t k = 0; k < dep; ++k) tout.stk[size_t(k) * tout.cap + pos] = st.entry(k);
                    break;
                }
                budget = ~0u;
            }
            if (COUNT) {
                atomicAdd(&stats[S_NODES], (unsigned long long)nvis);
                atomicAdd(&stats[S_TRIS], (unsigned long long)ntri);
            }
            // the hit so far (a queued walk resumes from it) or the final one
            hit.t[i] = h.t;
            hit.u[i] = h.u;
            hit.v[i] = h.v;
            hit.code[i] = h.code;
        }
        const bool hitp = active && done && h.code != kMissCode;
        uint32_t mat = 0;
        if (hitp && sort)
            mat = (h.code & kSphereBit) ? min(S.shaders[S.prims[S.sphere_prim[h.code & ~kSphereBit]].shader].type, 2u)
                                        : h.mat;
        uint32_t p0 = wave_append(&counters[CTR(C_BUCKET + 0)], hitp && mat == 0);
        if (hitp && mat == 0) b0[p0] = i;
        if (sort) {
            uint32_t p1 = wave_append(&counters[CTR(C_BUCKET + 1)], hitp && mat == 1);
            if (hitp && mat == 1) b1[p1] = i;
            uint32_t p2 = wave_append(&counters[CTR(C_BUCKET + 2)], hitp && mat == 2);
            if (hitp && mat == 2) b2[p2] = i;
        }
    }
}

// Closest-hit traversal with lane refill (Aila and Laine, "Understanding the
// Efficiency of Ray Traversal on GPUs", HPG 2009): a wave walks its rays one
// closest_iter per iteration and, once fewer than kRefillBelow of
// its lanes still walk, hands the finished rays to their material buckets and
// gives every idle lane the next ray of the queue (one atomic per wave for
// each). With
// incoherent rays a wave otherwise runs until its slowest ray is done while
// most lanes idle. The walk is trace_closest's, so hits are identical; only
// the order of the bucket entries differs.
#ifndef KR_REFILL_BELOW
#define KR_REFILL_BELOW 48
#endif
constexpr int kRefillBelow = KR_REFILL_BELOW;
static_assert(kRefillBelow >= 1 && kRefillBelow <= 64, "refill threshold in lanes of a wave");

template <int W, bool COUNT, bool Q>
__global__ __launch_bounds__(kBlock, trace_waves(W, false, Q)) void k_trace_refill(
    DevScene S, RayQ q, int qi, HitQ hit, uint32_t *b0, uint32_t *b1, uint32_t *b2, uint32_t *counters,
    unsigned long long *stats, int sort, uint32_t *spill, size_t spill_lanes) {
    __shared__ uint32_t stk[2 * kLdsDepth * kBlock];
    auto st = make_stack<kLdsDepth>(stk, spill, spill_lanes);
    const uint32_t n = counters[CTR(qi)];
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&stats[S_CLOSEST], (unsigned long long)n);
    uint32_t i = kMissCode;  // the lane's ray (queue position), kMissCode when idle
    bool walking = false;    // false with i set: done, waiting for its bucket
    bool exhausted = false;  // the queue has no more rays (wave-uniform)
    HitRec h;
    h.code = kMissCode;
    h.mat = 0;
    RayBox rb;
    ShearRay sr;
    uint32_t cur = 0;
    uint64_t nvis = 0, ntri = 0;
    while (true) {
        if (__popcll(__ballot(walking)) < kRefillBelow) {
            bool fin = !walking && i != kMissCode;
            bool hitp = fin && h.code != kMissCode;
            uint32_t mat = 0;
            if (hitp && sort)
                mat = (h.code & kSphereBit) ? min(S.shaders[S.prims[S.sphere_prim[h.code & ~kSphereBit]].shader].type, 2u)
                                            : h.mat;
            uint32_t p0 = wave_append(&counters[CTR(C_BUCKET + 0)], hitp && mat == 0);
            if (hitp && mat == 0) b0[p0] = i;
            if (sort) {
                uint32_t p1 = wave_append(&counters[CTR(C_BUCKET + 1)], hitp && mat == 1);
                if (hitp && mat == 1) b1[p1] = i;
                uint32_t p2 = wave_append(&counters[CTR(C_BUCKET + 2)], hitp && mat == 2);
                if (hitp && mat == 2) b2[p2] = i;
            }
            if (fin) i = kMissCode;
            if (!exhausted) {
                bool want = i == kMissCode;
                uint32_t k = wave_append(&counters[CTR(C_FETCH)], want);
                exhausted = __ballot(want && k >= n) != 0ull;
                if (want && k < n) {
                    i = k;
                    f3 o{q.ox[i], q.oy[i], q.oz[i]}, d{q.dx[i], q.dy[i], q.dz[i]};
                    h = closest_spheres(S, o, d);
                    rb = make_raybox(o, d);
                    sr = make_shear(o, d);
                    cur = 0;
                    st.reset();
                    walking = S.ntris != 0;
                    if (!walking) {
                        hit.t[i] = h.t;
                        hit.u[i] = h.u;
                        hit.v[i] = h.v;
                        hit.code[i] = h.code;
                    }
                }
            }
            if (__ballot(i != kMissCode) == 0ull) break;
        }
        if (walking) {
            // one iteration of trace_closest's walk (closest_iter)
            uint32_t nv = 0, nt = 0;
            const bool over = closest_iter<W, kLdsDepth, Q>(S, rb, sr, st, cur, h, nv, nt);
            if (COUNT) {
                nvis += nv;
                ntri += nt;
            }
            if (over) {
                walking = false;
                hit.t[i] = h.t;
                hit.u[i] = h.u;
                hit.v[i] = h.v;
                hit.code[i] = h.code;
            }
        }
    }
    if (COUNT) {
        atomicAdd(&stats[S_NODES], (unsigned long long)nvis);
        atomicAdd(&stats[S_TRIS], (unsigned long long)ntri);
    }
}

// Scene::intersect's shading normal and primitive of a hit
// (src/core/meta/scene.hpp:48-53): sphere normal in FP64
// (src/core/primitive/sphere.hpp:52), smooth mesh normal
// (src/core/primitive/mesh.hpp:85-91).
__device__ __forceinline__ void hit_surface(const DevScene &S, f3 o, f3 d, float t, uint32_t code, float bu,
                                            float bv, f3 &nrm, uint32_t &prim) {
    if (code & kSphereBit) {
        uint32_t s = code & ~kSphereBit;
        const double *sp = S.spheres + 4 * s;
        prim = S.sphere_prim[s];
        double hx = double(o.x) + double(t) * double(d.x) - sp[0];
        double hy = double(o.y) + double(t) * double(d.y) - sp[1];
        double hz = double(o.z) + double(t) * double(d.z) - sp[2];
        double inv = 1.0 / sqrt(hx * hx + hy * hy + hz * hz);
        nrm = f3{float(hx * inv), float(hy * inv), float(hz * inv)};
    } else {
        prim = S.tri_prim[code];
        const float *nn = S.tri_nrm + 9 * size_t(code);
        float b0w = 1.f - bu - bv;
        f3 n0{nn[0], nn[1], nn[2]}, n1{nn[3], nn[4], nn[5]}, n2{nn[6], nn[7], nn[8]};
        nrm = normalize(n0 * b0w + n1 * bu + n2 * bv);
    }
}

// Light::sample's point and pdf for shading point p: a point light
// (pointLight.hpp:20-27; pdf unused), a mesh light's table entry
// (mesh.hpp:92-98), a sphere's theta-uniform point (sphere.hpp:22-30);
// area pdfs become solid-angle pdfs (areaLight.hpp:26).
__device__ __forceinline__ void light_point(const DevScene &S, const LightRec &lr, f3 p, float u0, float u1, f3 &lp,
                                            float &pdf) {
    f3 ln{0.f, 0.f, 1.f};
    pdf = 1.f;
    if (lr.type == kLightPoint) {
        lp = f3{lr.center[0], lr.center[1], lr.center[2]};
        return;
    }
    if (lr.type == KR_PRIM_MESH) {
        uint32_t e = min(uint32_t(floorf(u0 * 1024.f)), 1023u);
        const float *te = S.light_table + 6 * (size_t(lr.table) + e);
        lp = f3{te[0], te[1], te[2]};
        ln = f3{te[3], te[4], te[5]};
        pdf = lr.pdf;
    } else {
        float st0, ct0;
        sincosf(u0 * kPi, &st0, &ct0);
        f3 w = from_euler(st0, ct0, u1 * kPi * 2.f);
        lp = f3{lr.center[0], lr.center[1], lr.center[2]} + w * sqrtf(lr.center[3]);
        ln = w;
        pdf = 1.f / (4.f * kPi * lr.center[3]);
    }
    f3 wl = p - lp;
    float d2 = dot(wl, wl);
    pdf *= d2 / fabsf(dot(ln, normalize(wl)));
}

struct ShadeArgs {
    DevScene S;
    RayQ q, qn;
    int qi, qni;
    HitQ hit;
    uint32_t *b0, *b1, *b2;
    ShadowQ sq;
    float *L;
    uint32_t *counters;
    unsigned long long *stats;
    uint64_t g0;
    uint32_t npix, bounce, max_bounce, seed_lo, seed_hi;
    NrcRender nr;
    JobTable T;  // batch renders: per-path scene, pixel count, depth and seed
    ProfView pv;
};

// RadianceKernel::execute loop body (src/core/kernel/radiance.cu:26-71) for
// one path vertex. NRC: a radiance cache may be attached (a.nr); PROF: a
// profiler may record vertices (a.pv). The plain instantiation carries
// neither's code or registers.
template <bool BATCH, bool NRC, bool PROF>
__global__ __launch_bounds__(kBlock, (!BATCH && !NRC && !PROF) ? KR_SHADE_WAVES : 1) void k_shade(ShadeArgs a) {
    const bool nr_on = NRC && a.nr.enabled, pv_on = PROF && a.pv.enabled;
    uint32_t n0 = a.counters[CTR(C_BUCKET + 0)], n1 = a.counters[CTR(C_BUCKET + 1)], n2 = a.counters[CTR(C_BUCKET + 2)];
    uint32_t total = n0 + n1 + n2;
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += stride) {
        uint32_t k = base + threadIdx.x;
        bool active = k < total;
        bool push_shadow = false, push_ray = false, push_query = false;
        uint32_t slot = 0, nflags = 0;
        f3 so{0, 0, 0}, sd{0, 0, 0}, contrib{0, 0, 0}, no{0, 0, 0}, nd{0, 0, 0}, nb{0, 0, 0};
        float stmax = 0.f;
        float qrec[13];
        if (active) {
            uint32_t i = k < n0 ? a.b0[k] : (k < n0 + n1 ? a.b1[k - n0] : a.b2[k - n0 - n1]);
            slot = a.q.slot[i];
            f3 o{a.q.ox[i], a.q.oy[i], a.q.oz[i]}, d{a.q.dx[i], a.q.dy[i], a.q.dz[i]};
            f3 beta{a.q.bx[i], a.q.by[i], a.q.bz[i]};
            float t = a.hit.t[i];
            uint32_t code = a.hit.code[i];
            uint64_t g = a.g0 + slot;
            DevScene S = a.S;
            uint32_t npix = a.npix, max_bounce = a.max_bounce, seed_lo = a.seed_lo, seed_hi = a.seed_hi;
            if (BATCH) {
                const JobRec &J = a.T.rec[find_job(a.T, g)];
                S = J.S;
                g -= J.begin;
                npix = J.npix;
                max_bounce = J.max_bounce;
                seed_lo = J.seed_lo;
                seed_hi = J.seed_hi;
            }
            uint32_t sample = uint32_t(g / npix), pixel = uint32_t(g % npix);
            f3 nrm;
            uint32_t prim;
            hit_surface(S, o, d, t, code, a.hit.u[i], a.hit.v[i], nrm, prim);
            f3 fu = normalize(cross(nrm, f3{0.f, 1.f, 0.f}));
            f3 fv = normalize(cross(nrm, fu));
            f3 p = o + d * t;
            f3 md = -d;
            f3 wo = normalize(f3{dot(md, fu), dot(md, fv), dot(md, nrm)});
            const uint32_t sid = S.prims[prim].shader;
            const ShaderRec sh = S.shaders[sid];  // the primitive's shader (a Mixed is resolved per call)
            f3 L{0.f, 0.f, 0.f};
            bool touch_L = false;
            // radiance cache: training paths record this vertex; other paths
            // end at their second diffuse or rough vertex with a cache query
            // (KR_NRC_INDIRECT: at their first, after its light sample)
            uint32_t flags = a.q.flags[i];
            bool train = (flags & 1u) != 0;
            uint32_t ndr = (flags >> 1) & 3u;
            bool terminate = false;
            float *rec = nullptr;
            if (nr_on) {
                bool dr = sh.type == KR_SHADER_LAMBERT || sh.type == KR_SHADER_EMISSION ||
                          sh.type == KR_SHADER_MIXED ||
                          ((sh.type == KR_SHADER_GLOSSY || sh.type == KR_SHADER_GLASS ||
                            sh.type == KR_SHADER_REFRACTION) &&
                           sh.dist == KR_DIST_BECKMANN && sh.roughness >= a.nr.rough_threshold);
                if (!train && dr && ++ndr >= a.nr.end_at) terminate = true;
                if (train) {
                    uint32_t ts = a.nr.tslot_of[slot];
                    rec = a.nr.rec + (size_t(ts) * a.nr.nv + a.bounce) * kRecFloats;
                    a.nr.tcount[ts] = a.bounce + 1;
                }
                if (terminate || rec) {
                    float v[13] = {p.x, p.y, p.z, md.x, md.y, md.z, nrm.x, nrm.y, nrm.z, sh.roughness,
                                   sh.cr[0], sh.cr[1], sh.cr[2]};
                    for (int c = 0; c < 13; ++c) qrec[c] = v[c];
                    if (rec) {
                        for (int c = 0; c < 13; ++c) rec[c] = v[c];
                        rec[16] = beta.x;
                        rec[17] = beta.y;
                        rec[18] = beta.z;
                        for (int c = 19; c < 25; ++c) rec[c] = 0.f;
                    }
                }
            }
            if (terminate && !a.nr.indirect) {
                // emission as usual; the cache stands for everything scattered here
                if (sh.type == KR_SHADER_EMISSION && S.nlights) {
                    f3 e = beta * f3{sh.emission[0], sh.emission[1], sh.emission[2]};
                    a.L[3 * slot] += e.x;
                    a.L[3 * slot + 1] += e.y;
                    a.L[3 * slot + 2] += e.z;
                }
                push_query = true;
                nb = beta;
            } else {
            Philox4 r0 = philox(pixel, sample, 2 * a.bounce, 0u, seed_lo, seed_hi);
            Philox4 r1 = philox(pixel, sample, 2 * a.bounce + 1, 0u, seed_lo, seed_hi);
            // wi of the reference's Varyings: the light direction after NEE, else
            // zero; a failed refraction leaves it in place
            f3 wi_v{0.f, 0.f, 0.f};
            f3 f_nee{0.f, 0.f, 0.f};  // Varyings::f after the light sample (zero without lights)
            uint32_t eval_leaf = sid;
            if (S.nlights) {
                // light pick (radiance.cu:32-34) and AreaLight::sample (areaLight.hpp:19-31)
                float xi = u01(r0.v[0]);
                uint32_t li_idx = min(uint32_t(floorf(xi * float(S.nlights))), S.nlights - 1u);
                float lpdf = 1.f / float(S.nlights);
                const LightRec lr = S.lights[li_idx];
                float u0 = u01(r0.v[1]), u1 = u01(r0.v[2]);
                f3 lp;
                float pdf;
                light_point(S, lr, p, u0, u1, lp, pdf);
                // LocalVaryings::emitSeg (varyings.hpp:27-37)
                f3 sdv = lp - p;
                f3 sdir = normalize(sdv);
                f3 diff = normalize(sdir) * kSpawnEps;
                f3 sorig = p + diff;
                float slen = length(sdv - diff * 2.f);
                f3 wi = normalize(lp - p);
                wi = normalize(f3{dot(wi, fu), dot(wi, fv), dot(wi, nrm)});
                wi_v = wi;
                float uce;
                eval_leaf = resolve_leaf(S, sid, wo, pixel, sample, 2 * a.bounce, seed_lo, seed_hi, uce);
                f3 f = shader_eval(S.shaders[eval_leaf], wo, wi, uce);
                f_nee = f;
                f3 E{lr.emission[0], lr.emission[1], lr.emission[2]};
                f3 li = lr.type == kLightPoint ? E : E / pdf;
                f3 c = beta * f * li * fabsf(wi.z) / lpdf;
                if (!finite3(c)) {
                    // The sample is rejected whether or not the segment is
                    // occluded (an occluded li is 0/pdf), so no query is needed.
                    L = L + c;
                    touch_L = true;
                    if (rec) {
                        rec[19] += c.x;
                        rec[20] += c.y;
                        rec[21] += c.z;
                    }
                } else if (c.x != 0.f || c.y != 0.f || c.z != 0.f) {
                    push_shadow = true;
                    so = sorig;
                    sd = sdir;
                    stmax = slen;
                    contrib = c;
                }
            }
            if (S.nlights && S.shaders[eval_leaf].type == KR_SHADER_EMISSION) {
                const ShaderRec &es = S.shaders[eval_leaf];
                f3 e = beta * f3{es.emission[0], es.emission[1], es.emission[2]};
                L = L + e;
                touch_L = true;
                if (rec) {
                    rec[19] += e.x;
                    rec[20] += e.y;
                    rec[21] += e.z;
                    rec[22] = e.x;
                    rec[23] = e.y;
                    rec[24] = e.z;
                }
            }
            if (touch_L) {
                a.L[3 * slot] += L.x;
                a.L[3 * slot + 1] += L.y;
                a.L[3 * slot + 2] += L.z;
            }
            // profiler record of this vertex (radiance.cu:38-50); k_shadow adds
            // the light sample's contribution to its L once it is unoccluded
            int64_t psl = pv_on ? prof_slice(a.pv, a.g0 + slot) : -1;
            if (psl >= 0 && a.bounce < a.pv.depth) {
                float *pr = a.pv.rec + (size_t(psl) * a.pv.depth + a.bounce) * kProfFloats;
                pr[0] = o.x;
                pr[1] = o.y;
                pr[2] = o.z;
                pr[3] = d.x;
                pr[4] = d.y;
                pr[5] = d.z;
                pr[6] = f_nee.x;
                pr[7] = f_nee.y;
                pr[8] = f_nee.z;
                pr[9] = a.L[3 * slot];
                pr[10] = a.L[3 * slot + 1];
                pr[11] = a.L[3 * slot + 2];
                // acos(wo . reflect(wi, n)), reflect(wi, n) = 2 (n . wi) n - wi, n = +z
                pr[12] = acosf(dot(wo, f3{-wi_v.x, -wi_v.y, wi_v.z}));
                pr[13] = __uint_as_float(sid);
            }
            if (terminate) {
                // KR_NRC_INDIRECT: light sample and emission done above; the
                // cache stands for what the BSDF-sampled continuation brings
                push_query = true;
                nb = beta;
            } else {
            // BSDF sample (radiance.cu:56-60)
            float ucs;
            uint32_t sample_leaf = resolve_leaf(S, sid, wo, pixel, sample, 2 * a.bounce + 1, seed_lo, seed_hi, ucs);
            f3 wi = wi_v, f;
            shader_sample(S.shaders[sample_leaf], wo, u01(r0.v[3]), u01(r1.v[0]), ucs, wi, f);
            beta = beta * (f * fabsf(wi.z));
            f3 gd = normalize(fu * wi.x + fv * wi.y + nrm * wi.z);
            f3 go = p + gd * kSpawnEps;
            // Russian roulette (radiance.cu:62-68)
            bool alive = a.bounce < max_bounce, rr_kill = false;
            float rr = smax(beta.x, smax(beta.y, beta.z));
            if (rr < 1.f && a.bounce > 3) {
                float qq = smax(0.05f, 1.f - rr);
                if (u01(r1.v[2]) < qq) {
                    alive = false;
                    rr_kill = true;
                } else {
                    beta = beta / (1.f - qq);
                }
            }
            // the profiler's loop counter: a path ended by Russian roulette
            // breaks before ++bounce (radiance.cu:66), any other goes on to
            // the next intersect, which ends the loop
            if (psl >= 0) a.pv.fin[4 * size_t(psl) + 3] = __uint_as_float(rr_kill ? a.bounce : a.bounce + 1);
            // a ray with a NaN direction or origin (NaN frame, failed refraction
            // with no light sampled) hits nothing, in the reference's traversal
            // as in this one: end the path here instead of tracing it
            if (alive && !(finite3(gd) && finite3(go))) alive = false;
            if (alive) {
                push_ray = true;
                no = go;
                nd = gd;
                nb = beta;
                nflags = (train ? 1u : 0u) | (min(ndr, 3u) << 1);
            }
            }  // continued
            }  // not terminated before the light sample
        }
        if (nr_on) {
            uint32_t pq = wave_append(&a.counters[CTR(C_NRCQ)], push_query);
            if (push_query) {
                a.nr.qslot[pq] = slot;
                a.nr.qbeta[3 * size_t(pq)] = nb.x;
                a.nr.qbeta[3 * size_t(pq) + 1] = nb.y;
                a.nr.qbeta[3 * size_t(pq) + 2] = nb.z;
                float4 *dst = reinterpret_cast<float4 *>(a.nr.qraw + size_t(pq) * nrc::kRaw);
                dst[0] = make_float4(qrec[0], qrec[1], qrec[2], qrec[3]);
                dst[1] = make_float4(qrec[4], qrec[5], qrec[6], qrec[7]);
                dst[2] = make_float4(qrec[8], qrec[9], qrec[10], qrec[11]);
                dst[3] = make_float4(qrec[12], 0.f, 0.f, 0.f);
            }
        }
#if KR_SHADE_OCTANT_SORT
        // shadow segments grouped by direction octant within the block, as
        // the next rays below
        __shared__ uint32_t s_soct[8];
        __shared__ uint32_t s_sbase;
        if (threadIdx.x < 8) s_soct[threadIdx.x] = 0;
        __syncthreads();
        uint32_t soct = (sd.x < 0.f ? 1u : 0u) | (sd.y < 0.f ? 2u : 0u) | (sd.z < 0.f ? 4u : 0u);
        uint32_t srank = push_shadow ? atomicAdd(&s_soct[soct], 1u) : 0u;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t tot = 0;
            for (int o = 0; o < 8; ++o) {
                uint32_t c = s_soct[o];
                s_soct[o] = tot;
                tot += c;
            }
            s_sbase = tot ? atomicAdd(&a.counters[CTR(C_SHADOW)], tot) : 0u;
        }
        __syncthreads();
        uint32_t ps = s_sbase + s_soct[soct] + srank;
#else
        uint32_t ps = wave_append(&a.counters[CTR(C_SHADOW)], push_shadow);
#endif
        if (push_shadow) {
            a.sq.slot[ps] = slot;
            a.sq.ox[ps] = so.x;
            a.sq.oy[ps] = so.y;
            a.sq.oz[ps] = so.z;
            a.sq.dx[ps] = sd.x;
            a.sq.dy[ps] = sd.y;
            a.sq.dz[ps] = sd.z;
            a.sq.tmax[ps] = stmax;
            a.sq.cx[ps] = contrib.x;
            a.sq.cy[ps] = contrib.y;
            a.sq.cz[ps] = contrib.z;
        }
#if KR_SHADE_OCTANT_SORT
        // the block's next rays enter the queue grouped by direction octant,
        // so that a traversal wave holds rays of one or two octants, which
        // order their nodes alike (one queue atomic per block)
        __shared__ uint32_t s_oct[8];
        __shared__ uint32_t s_base;
        if (threadIdx.x < 8) s_oct[threadIdx.x] = 0;
        __syncthreads();
        uint32_t oct = (nd.x < 0.f ? 1u : 0u) | (nd.y < 0.f ? 2u : 0u) | (nd.z < 0.f ? 4u : 0u);
        uint32_t rank = push_ray ? atomicAdd(&s_oct[oct], 1u) : 0u;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t tot = 0;
            for (int o = 0; o < 8; ++o) {
                uint32_t c = s_oct[o];
                s_oct[o] = tot;
                tot += c;
            }
            s_base = tot ? atomicAdd(&a.counters[CTR(a.qni)], tot) : 0u;
        }
        __syncthreads();
        uint32_t pr = s_base + s_oct[oct] + rank;
        __syncthreads();  // s_oct and s_base are rewritten by the next iteration
#else
        uint32_t pr = wave_append(&a.counters[CTR(a.qni)], push_ray);
#endif
        if (push_ray) {
            a.qn.slot[pr] = slot;
            a.qn.ox[pr] = no.x;
            a.qn.oy[pr] = no.y;
            a.qn.oz[pr] = no.z;
            a.qn.dx[pr] = nd.x;
            a.qn.dy[pr] = nd.y;
            a.qn.dz[pr] = nd.z;
            a.qn.bx[pr] = nb.x;
            a.qn.by[pr] = nb.y;
            a.qn.bz[pr] = nb.z;
            a.qn.flags[pr] = nflags;
        }
    }
}

// An unoccluded light sample: its contribution goes into the path's L (and
// into the vertex record of a training path and the profiler's record).
__device__ __forceinline__ void shadow_unoccluded(const ShadowQ &sq, uint32_t i, uint32_t slot, float *L,
                                                  const NrcRender &nr, uint32_t bounce, const ProfView &pv,
                                                  uint64_t g0) {
    L[3 * slot] += sq.cx[i];
    L[3 * slot + 1] += sq.cy[i];
    L[3 * slot + 2] += sq.cz[i];
    if (nr.enabled) {
        uint32_t ts = nr.tslot_of[slot];
        if (ts != kMissCode) {
            float *rec = nr.rec + (size_t(ts) * nr.nv + bounce) * kRecFloats;
            rec[19] += sq.cx[i];
            rec[20] += sq.cy[i];
            rec[21] += sq.cz[i];
        }
    }
    if (pv.enabled) prof_add_light(pv, g0 + slot, bounce, sq.cx[i], sq.cy[i], sq.cz[i]);
}

template <int W, bool BATCH, bool Q>
__global__ __launch_bounds__(kBlock, trace_waves(W, BATCH, Q)) void k_shadow(DevScene S, JobTable T, uint64_t g0, ShadowQ sq, uint32_t *counters,
                                                   float *L, unsigned long long *stats, uint32_t *spill,
                                                   size_t spill_lanes, NrcRender nr, uint32_t bounce,
                                                   ProfView pv) {
    __shared__ uint32_t stk[2 * kLdsDepth * kBlock];
    auto st = make_stack<kLdsDepth>(stk, spill, spill_lanes);
    uint32_t n = counters[CTR(C_SHADOW)];
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&stats[S_SHADOW], (unsigned long long)n);
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        uint32_t i = base + threadIdx.x;
        bool active = i < n;
        if (active) {
            f3 o{sq.ox[i], sq.oy[i], sq.oz[i]}, d{sq.dx[i], sq.dy[i], sq.dz[i]};
            uint32_t slot = sq.slot[i];
            DevScene Sx = S;
            if (BATCH) Sx = T.rec[find_job(T, g0 + slot)].S;
            bool occ = KR_SHADOW_STEPS ? trace_occluded_steps<W, kLdsDepth, Q>(Sx, o, d, sq.tmax[i], st)
                                       : trace_occluded<W, kLdsDepth, Q>(Sx, o, d, sq.tmax[i], st);
            if (!occ) shadow_unoccluded(sq, i, slot, L, nr, bounce, pv, g0);
        }
    }
}

// Shadow segments in tail rounds, as k_trace_tail does for closest hits: a
// walk that has run `steps` iterations is queued in `tout` with its state, and
// the next launch (FIRST = false) resumes the queued walks packed into full
// waves. Single renders and one-job batch chunks only.
template <int W, bool Q, bool FIRST>
__global__ __launch_bounds__(kBlock, trace_waves(W, false, Q)) void k_shadow_tail(
    DevScene S, uint64_t g0, ShadowQ sq, uint32_t *counters, float *L, unsigned long long *stats, uint32_t *spill,
    size_t spill_lanes, NrcRender nr, uint32_t bounce, ProfView pv, TailQ tin, int cin, TailQ tout, int cout,
    uint32_t steps) {
    __shared__ uint32_t stk[2 * kLdsDepth * kBlock];
    auto st = make_stack<kLdsDepth>(stk, spill, spill_lanes);
    const uint32_t n = FIRST ? counters[CTR(C_SHADOW)] : min(counters[CTR(cin)], tin.cap);
    if (FIRST && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&stats[S_SHADOW], (unsigned long long)n);
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t j = base + threadIdx.x;
        if (j >= n) continue;
        const uint32_t i = FIRST ? j : tin.idx[j];
        f3 o{sq.ox[i], sq.oy[i], sq.oz[i]}, d{sq.dx[i], sq.dy[i], sq.dz[i]};
        const float tmax = sq.tmax[i];
        RayBox rb = make_raybox(o, d);
        ShearRay sr = make_shear(o, d);
        // tri_hit accepts t < limit; the segment includes its end point
        const float limit = tmax * 1.00000012f;
        uint32_t cur = 0;
        int occ = -1;
        st.reset();
        if (FIRST) {
            for (uint32_t s = 0; s < S.nspheres && occ < 0; ++s) {
                double ts;
                if (sphere_hit(S.spheres + 4 * s, o, d, ts) && ts <= double(tmax)) occ = 1;
            }
            if (occ < 0 && S.ntris == 0) occ = 0;
        } else {
            cur = tin.cur[j];
            const uint32_t dep = tin.info[j];
            for (uint32_t k = 0; k < dep; ++k) {
                uint2 e = tin.stk[size_t(k) * tin.cap + j];
                st.push(e.x, __uint_as_float(e.y));
            }
        }
        uint32_t budget = steps;
        while (occ < 0) {
            occ = occluded_walk<W, kLdsDepth, Q>(S, rb, sr, limit, tmax, st, cur, budget);
            if (occ >= 0) break;
            const uint32_t dep = st.depth();
            const bool fits = dep <= tout.stk_cap;
            const uint32_t pos = wave_append(&counters[CTR(cout)], fits);
            if (fits && pos < tout.cap) {
                tout.idx[pos] = i;
                tout.cur[pos] = cur;
                tout.info[pos] = dep;
                for (uint32_t k = 0; k < dep; ++k) tout.stk[size_t(k) * tout.cap + pos] = st.entry(k);
                break;
            }
            budget = ~0u;
        }
        if (occ == 0) shadow_unoccluded(sq, i, sq.slot[i], L, nr, bounce, pv, g0);
    }
}

// NormalKernel / VisibilityKernel / HalfVectorKernel at the first hit
// (src/core/kernel/normal.hpp:17-28, visibility.hpp:17-35,
// halfVector.hpp:17-37). Visibility queues a shadow segment to light 0 with
// contribution 1 (k_shadow adds it when unoccluded); the others write L.
template <bool BATCH>
__global__ __launch_bounds__(kBlock) void k_shade_debug(ShadeArgs a, uint32_t kernel) {
    uint32_t n0 = a.counters[CTR(C_BUCKET + 0)], n1 = a.counters[CTR(C_BUCKET + 1)],
             n2 = a.counters[CTR(C_BUCKET + 2)];
    uint32_t total = n0 + n1 + n2;
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += stride) {
        uint32_t k = base + threadIdx.x;
        bool active = k < total, push_shadow = false;
        uint32_t slot = 0;
        f3 so{0, 0, 0}, sd{0, 0, 0};
        float stmax = 0.f;
        if (active) {
            uint32_t i = k < n0 ? a.b0[k] : (k < n0 + n1 ? a.b1[k - n0] : a.b2[k - n0 - n1]);
            slot = a.q.slot[i];
            f3 o{a.q.ox[i], a.q.oy[i], a.q.oz[i]}, d{a.q.dx[i], a.q.dy[i], a.q.dz[i]};
            float t = a.hit.t[i];
            uint32_t code = a.hit.code[i];
            uint64_t g = a.g0 + slot;
            DevScene S = a.S;
            uint32_t npix = a.npix, seed_lo = a.seed_lo, seed_hi = a.seed_hi;
            if (BATCH) {
                const JobRec &J = a.T.rec[find_job(a.T, g)];
                S = J.S;
                g -= J.begin;
                npix = J.npix;
                seed_lo = J.seed_lo;
                seed_hi = J.seed_hi;
            }
            uint32_t sample = uint32_t(g / npix), pixel = uint32_t(g % npix);
            f3 nrm;
            uint32_t prim;
            hit_surface(S, o, d, t, code, a.hit.u[i], a.hit.v[i], nrm, prim);
            f3 L{0.f, 0.f, 0.f};
            if (kernel == KR_KERNEL_NORMAL) {
                L = nrm * .5f + f3{.5f, .5f, .5f};
            } else if (S.nlights) {
                f3 fu = normalize(cross(nrm, f3{0.f, 1.f, 0.f}));
                f3 fv = normalize(cross(nrm, fu));
                f3 p = o + d * t;
                f3 md = -d;
                f3 wo = normalize(f3{dot(md, fu), dot(md, fv), dot(md, nrm)});
                Philox4 r0 = philox(pixel, sample, 0u, 0u, seed_lo, seed_hi);
                const LightRec lr = S.lights[0];
                f3 lp;
                float pdf;
                light_point(S, lr, p, u01(r0.v[1]), u01(r0.v[2]), lp, pdf);
                f3 wi = normalize(lp - p);
                wi = normalize(f3{dot(wi, fu), dot(wi, fv), dot(wi, nrm)});
                if (kernel == KR_KERNEL_VISIBILITY) {
                    f3 E{lr.emission[0], lr.emission[1], lr.emission[2]};
                    f3 li = lr.type == kLightPoint ? E : E / pdf;
                    if (!(li.x == 0.f && li.y == 0.f && li.z == 0.f)) {
                        // LocalVaryings::emitSeg (varyings.hpp:27-37)
                        f3 sdv = lp - p;
                        sd = normalize(sdv);
                        f3 diff = normalize(sd) * kSpawnEps;
                        so = p + diff;
                        stmax = length(sdv - diff * 2.f);
                        push_shadow = true;
                    }
                } else {
                    float v = acosf(normalize(wi + wo).z) * 2.f / kPi;
                    L = f3{v, v, v};
                }
            }
            a.L[3 * slot] = L.x;
            a.L[3 * slot + 1] = L.y;
            a.L[3 * slot + 2] = L.z;
        }
#if KR_SHADE_OCTANT_SORT
        // shadow segments grouped by direction octant within the block, as
        // the next rays below
        __shared__ uint32_t s_soct[8];
        __shared__ uint32_t s_sbase;
        if (threadIdx.x < 8) s_soct[threadIdx.x] = 0;
        __syncthreads();
        uint32_t soct = (sd.x < 0.f ? 1u : 0u) | (sd.y < 0.f ? 2u : 0u) | (sd.z < 0.f ? 4u : 0u);
        uint32_t srank = push_shadow ? atomicAdd(&s_soct[soct], 1u) : 0u;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t tot = 0;
            for (int o = 0; o < 8; ++o) {
                uint32_t c = s_soct[o];
                s_soct[o] = tot;
                tot += c;
            }
            s_sbase = tot ? atomicAdd(&a.counters[CTR(C_SHADOW)], tot) : 0u;
        }
        __syncthreads();
        uint32_t ps = s_sbase + s_soct[soct] + srank;
#else
        uint32_t ps = wave_append(&a.counters[CTR(C_SHADOW)], push_shadow);
#endif
        if (push_shadow) {
            a.sq.slot[ps] = slot;
            a.sq.ox[ps] = so.x;
            a.sq.oy[ps] = so.y;
            a.sq.oz[ps] = so.z;
            a.sq.dx[ps] = sd.x;
            a.sq.dy[ps] = sd.y;
            a.sq.dz[ps] = sd.z;
            a.sq.tmax[ps] = stmax;
            a.sq.cx[ps] = 1.f;
            a.sq.cy[ps] = 1.f;
            a.sq.cz[ps] = 1.f;
        }
    }
}

// Persistent shadow traversal: every lane keeps a segment until it is done
// and takes a new one from the queue (one atomic per wave) once fewer than
// kMinActive lanes of its wave still walk (Aila and Laine, HPG 2009). Same
// additions to L and to the training records as k_shadow.
constexpr int kMinActive = 32;
template <int W, bool BATCH, bool Q>
__global__ __launch_bounds__(kBlock) void k_shadow_dyn(DevScene S, JobTable T, uint64_t g0, ShadowQ sq,
                                                       uint32_t *counters, float *L, unsigned long long *stats,
                                                       uint32_t *spill, size_t spill_lanes, NrcRender nr,
                                                       uint32_t bounce, ProfView pv) {
    __shared__ uint32_t stk[2 * kLdsDepth * kBlock];
    auto st = make_stack<kLdsDepth>(stk, spill, spill_lanes);
    const uint32_t n = counters[CTR(C_SHADOW)];
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&stats[S_SHADOW], (unsigned long long)n);
    uint32_t i = 0;
    bool has = false, exhausted = false;
    TravState ts;
    DevScene Sx = S;
    while (true) {
        bool want = !has && !exhausted;
        uint32_t k = wave_append(&counters[CTR(C_FETCH_SH)], want);
        if (want) {
            if (k < n) {
                i = k;
                has = true;
                if (BATCH) Sx = T.rec[find_job(T, g0 + sq.slot[i])].S;
                f3 o{sq.ox[i], sq.oy[i], sq.oz[i]}, d{sq.dx[i], sq.dy[i], sq.dz[i]};
                occluded_begin<W, kLdsDepth>(Sx, o, d, sq.tmax[i], st, ts);
            } else {
                exhausted = true;
            }
        }
        if (__ballot(has) == 0ull) break;
        while (true) {
            if (has && !ts.done) occluded_step<W, kLdsDepth, Q>(Sx, st, ts);
            if (__popcll(__ballot(has && !ts.done)) < kMinActive) break;
        }
        bool fin = has && ts.done;
        if (fin && !ts.occluded) {
            uint32_t slot = sq.slot[i];
            L[3 * slot] += sq.cx[i];
            L[3 * slot + 1] += sq.cy[i];
            L[3 * slot + 2] += sq.cz[i];
            if (nr.enabled) {
                uint32_t t = nr.tslot_of[slot];
                if (t != kMissCode) {
                    float *rec = nr.rec + (size_t(t) * nr.nv + bounce) * kRecFloats;
                    rec[19] += sq.cx[i];
                    rec[20] += sq.cy[i];
                    rec[21] += sq.cz[i];
                }
            }
            if (pv.enabled) prof_add_light(pv, g0 + slot, bounce, sq.cx[i], sq.cy[i], sq.cz[i]);
        }
        if (fin) has = false;
    }
}

// Final path radiance of the chunk's profiled paths (Slice::L, radiance.cu:74-78).
__global__ __launch_bounds__(kBlock) void k_prof_final(ProfView pv, const float *L, uint64_t g0, uint32_t P) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < P; j += gridDim.x * blockDim.x) {
        int64_t sl = prof_slice(pv, g0 + j);
        if (sl < 0) continue;
        pv.fin[4 * size_t(sl)] = L[3 * size_t(j)];
        pv.fin[4 * size_t(sl) + 1] = L[3 * size_t(j) + 1];
        pv.fin[4 * size_t(sl) + 2] = L[3 * size_t(j) + 2];
    }
}

// Per-pixel FP64 accumulation of the chunk's finished paths; no atomics:
// one lane owns a pixel and visits its slots j = j0, j0 + npix, ... < P.
__global__ __launch_bounds__(kBlock) void k_accumulate(const float *L, uint64_t g0, uint32_t P, uint32_t npix,
                                                       double *acc, uint32_t *valid, unsigned long long *stats) {
    uint32_t off = uint32_t(g0 % npix);
    for (uint32_t px = blockIdx.x * blockDim.x + threadIdx.x; px < npix; px += gridDim.x * blockDim.x) {
        uint32_t j0 = (px + npix - off) % npix;
        double sx = 0, sy = 0, sz = 0;
        uint32_t nv = 0, bad = 0;
        for (uint32_t j = j0; j < P; j += npix) {
            float lx = L[3 * j], ly = L[3 * j + 1], lz = L[3 * j + 2];
            if (isfinite(lx) && isfinite(ly) && isfinite(lz)) {
                sx += lx;
                sy += ly;
                sz += lz;
                ++nv;
            } else {
                ++bad;
            }
        }
        if (j0 < P) {
            acc[3 * size_t(px)] += sx;
            acc[3 * size_t(px) + 1] += sy;
            acc[3 * size_t(px) + 2] += sz;
            valid[px] += nv;
            if (bad) atomicAdd(&stats[S_INVALID], (unsigned long long)bad);
        }
    }
}

}  // namespace

// ------------------------------------------------------- radiance cache --
namespace {
constexpr uint32_t kStreamTrain = 0x5452414Eu;  // training-path selection
constexpr int kRec = 28;  // vertex record: raw 16 | beta 3 | c 3 | beta*E 3 | pad 3

// Sample k of a training batch: explicit arrays (pool_total == nullptr:
// index k) or the render's ring buffer (window of the most recent samples).
struct TrainSource {
    const float *raw;      // kRaw per sample
    const float *target;   // 4 per sample
    const uint32_t *pool_total;
    uint32_t n, cap, step, steps;
};

__device__ __forceinline__ void train_window(const TrainSource &src, uint32_t batch, uint32_t &start,
                                             uint32_t &count, uint32_t &mod) {
    if (!src.pool_total) {
        start = 0;
        count = src.n;
        mod = src.n ? src.n : 1;
        return;
    }
    uint32_t total = *src.pool_total;
    if (total == 0) {
        start = 0;
        count = 0;
        mod = 1;
        return;
    }
    if (total <= src.cap) {
        // ring not yet wrapped: samples live in [0, total); cycle through them
        mod = total;
        count = batch < total ? batch : total;
        start = uint32_t(uint64_t(src.step) * count % total);
    } else {
        // most recent steps * count samples, oldest window first
        mod = src.cap;
        count = batch < src.cap ? batch : src.cap;
        uint64_t back = uint64_t(src.steps - src.step) * count % src.cap;
        start = uint32_t((uint64_t(total) % src.cap + src.cap - back) % src.cap);
    }
}

// Forward + backward of one wave (64 samples) of a training batch; writes
// the wave's partial gradient row and adds its loss to *loss.
__global__ __launch_bounds__(64) void k_nrc_fwd_bwd(nrc::Weights W, TrainSource src, uint32_t batch, float *partial,
                                                    float *loss) {
    using namespace nrc;
    __shared__ v4s Hs[24 + 5 * 16][64];  // input fragments of every layer
    __shared__ v4f Ds[16][64];           // dZ fragments of the current layer
    int lane = threadIdx.x, s = lane & 15, q = lane >> 4;
    uint32_t start, count, mod;
    train_window(src, batch, start, count, mod);
    uint32_t base = blockIdx.x * 64;
    auto index_of = [&](int n) -> int64_t {
        uint32_t k = base + uint32_t(n);
        if (k >= count) return -1;
        return int64_t((uint64_t(start) + k) % mod);
    };
    v4s in[6][4];
    encode_inputs(in, [&](int n) -> const float * {
        int64_t i = index_of(n);
        return i < 0 ? nullptr : src.raw + size_t(i) * kRaw;
    }, W);
#pragma unroll
    for (int kt = 0; kt < 6; ++kt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) Hs[kt * 4 + nt][lane] = in[kt][nt];
    // forward, keeping each layer's input fragments in LDS
    v4s h[4][4], g[4][4];
    hidden_layer<6, 4>(in, h, W.wt + bw_off(0), W.master + b_off(0));
#pragma unroll
    for (int l = 1; l <= 4; ++l) {
#pragma unroll
        for (int kt = 0; kt < 4; ++kt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) Hs[24 + (l - 1) * 16 + kt * 4 + nt][lane] = h[kt][nt];
        hidden_layer<4, 4>(h, g, W.wt + bw_off(l), W.master + b_off(l));
#pragma unroll
        for (int kt = 0; kt < 4; ++kt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) h[kt][nt] = g[kt][nt];
    }
#pragma unroll
    for (int kt = 0; kt < 4; ++kt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) Hs[24 + 4 * 16 + kt * 4 + nt][lane] = h[kt][nt];
    v4f zo[4];
    output_layer(h, zo, W.wt + bw_off(5), W.master + b_off(5));
    // relative L2 loss: (y - t)^2 / (y^2 + 0.01), mean over samples and RGB
    float inv = count ? 1.f / (3.f * float(count)) : 0.f;
    float lsum = 0.f;
    v4f dz[4][4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        v4f d = {0.f, 0.f, 0.f, 0.f};
        int64_t i = index_of(nt * 16 + s);
        if (q == 0 && i >= 0) {
            const float *t = src.target + size_t(i) * 4;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                float y = zo[nt][c];
                float den = y * y + 0.01f;
                float diff = y - t[c];
                lsum += diff * diff / den;
                d[c] = 2.f * diff / den * inv;
            }
        }
        dz[0][nt] = d;
    }
    // wave reduction of the loss
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
    if (lane == 0 && loss) atomicAdd(loss, lsum * inv);
    float *grow = partial + size_t(blockIdx.x) * kNumParams;
    // backward, layer 5 .. 0
#pragma unroll
    for (int l = kLayers - 1; l >= 0; --l) {
        const int IN = kDims[l], OUT = kDims[l + 1];
        const int MT = OUT / 16, KT = IN / 16;
        const int hbase = l == 0 ? 0 : 24 + (l - 1) * 16;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
                if (mt < MT) Ds[mt * 4 + nt][lane] = dz[mt][nt];
        __syncthreads();
        // dW[i][o] = sum_n H^T[i][n] dZ[n][o] on MFMA, the 64 samples as two
        // k = 32 steps (k tiles 2p, 2p + 1 paired as in mfma_pair). A = H^T
        // (exactly BF16: the activations are), B = dZ split into BF16 high
        // and low parts, dZ = hi + lo + O(2^-17 dZ), so the FP32 sums keep
        // FP32-level accuracy. Lane (s, q) needs H^T[16 it + s][n] and
        // dZ[n][16 ot + s] for n = 32 p + 16 h + 4 q + r, which the forward
        // layouts of Hs and Ds hold at lane 16 (s >> 2) + 4 q + r, element
        // s & 3, of tile (2 p + h).
        {
            const int src_lane = 16 * (s >> 2) + 4 * q, el = s & 3;
            v4s bhi[4][2][2], blo[4][2][2];
#pragma unroll
            for (int ot = 0; ot < 4; ++ot)
#pragma unroll
                for (int p = 0; p < 2; ++p)
#pragma unroll
                    for (int h = 0; h < 2; ++h)
                        if (ot < MT) {
                            float f[4], lo[4];
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                f[r] = Ds[ot * 4 + 2 * p + h][src_lane + r][el];
                                lo[r] = f[r] - from_bf16(to_bf16(f[r]));
                            }
                            bhi[ot][p][h] = pack4(f[0], f[1], f[2], f[3]);
                            blo[ot][p][h] = pack4(lo[0], lo[1], lo[2], lo[3]);
                        }
#pragma unroll
            for (int it = 0; it < 6; ++it) {
                if (it >= KT) break;
                v4s a[2][2];
#pragma unroll
                for (int p = 0; p < 2; ++p)
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const v4s *src = &Hs[hbase + it * 4 + 2 * p + h][src_lane];
                        a[p][h] = v4s{src[0][el], src[1][el], src[2][el], src[3][el]};
                    }
#pragma unroll
                for (int ot = 0; ot < 4; ++ot) {
                    if (ot >= MT) break;
                    v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int p = 0; p < 2; ++p) {
                        acc = mfma_pair(a[p][0], a[p][1], bhi[ot][p][0], bhi[ot][p][1], acc);
                        acc = mfma_pair(a[p][0], a[p][1], blo[ot][p][0], blo[ot][p][1], acc);
                    }
                    // lane (s, q) holds dW[16 it + 4 q + r][16 ot + s]
#pragma unroll
                    for (int r = 0; r < 4; ++r) grow[w_off(l) + (16 * it + 4 * q + r) * OUT + 16 * ot + s] = acc[r];
                }
            }
        }
        if (lane < OUT) {
            int dmt = lane >> 4, dq = (lane & 15) >> 2, dr = lane & 3;
            float acc = 0.f;
            for (int n = 0; n < 64; ++n) acc += Ds[dmt * 4 + (n >> 4)][(n & 15) + 16 * dq][dr];
            grow[b_off(l) + lane] = acc;
        }
        __syncthreads();
        if (l == 0) break;
        // dH^T (in x samples) = W (in x out) dZ^T; dZ_{l-1} = dH * (H_l > 0)
        v4s dzb[4][4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
                if (mt < MT) dzb[mt][nt] = pack4(dz[mt][nt][0], dz[mt][nt][1], dz[mt][nt][2], dz[mt][nt][3]);
        const uint16_t *w = W.w + bw_off(l);
#pragma unroll
        for (int mi = 0; mi < 4; ++mi) {  // IN = 64 for l >= 1
            v4s a[4];
#pragma unroll
            for (int kt = 0; kt < 4; ++kt)
                if (kt < MT) a[kt] = *reinterpret_cast<const v4s *>(w + (mi * 16 + s) * OUT + kt * 16 + 4 * q);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kt = 0; kt < 4; ++kt)
                    if (kt < MT) acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[kt], dzb[kt][nt], acc, 0, 0, 0);
                v4s hv = Hs[hbase + mi * 4 + nt][lane];
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (!(from_bf16(uint16_t(hv[r])) > 0.f)) acc[r] = 0.f;
                dz[mi][nt] = acc;
            }
        }
        (void)KT;
    }
}

// Sums the partial rows, takes one Adam step on the FP32 masters and
// refreshes the BF16 copies (W^T for the forward, W for the backward). With
// ema non-null it also moves the average towards the new masters by
// ema_a = (1 - d) / (1 - d^t) at step t, the bias-corrected exponential
// moving average of decay d (ema_a = 1 at the first step), and refreshes
// its BF16 W^T.
// Blocks of kAdamParams x kAdamGroups threads: the kAdamGroups threads of a
// parameter sum every kAdamGroups-th partial row (coalesced across the
// block's consecutive parameters), the block adds the group sums in LDS in a
// fixed order, and the first group's thread takes the step.
constexpr int kAdamParams = 32, kAdamGroups = 8;
__global__ __launch_bounds__(kAdamParams *kAdamGroups) void k_nrc_adam(
    float *master, uint16_t *wt, uint16_t *w, float *m, float *v, const float *partial, uint32_t rows, float lr,
    float b1, float b2, float eps, float c1, float c2, const uint32_t *pool_total, float *ema, uint16_t *wt_ema,
    float ema_a) {
    using namespace nrc;
    if (pool_total && *pool_total == 0) return;
    __shared__ float gs[kAdamGroups][kAdamParams];
    const int x = threadIdx.x % kAdamParams, y = threadIdx.x / kAdamParams;
    auto adam_update = [&](int p, int x) {
        float g = 0.f;
#pragma unroll
        for (int k = 0; k < kAdamGroups; ++k) g += gs[k][x];
        float mm = b1 * m[p] + (1.f - b1) * g;
        float vv = b2 * v[p] + (1.f - b2) * g * g;
        m[p] = mm;
        v[p] = vv;
        float val = master[p] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
        master[p] = val;
        float av = 0.f;
        if (ema) {
            av = ema[p] + (val - ema[p]) * ema_a;
            ema[p] = av;
        }
        for (int l = 0; l < kLayers; ++l) {
            int in = kDims[l], out = kDims[l + 1];
            if (p >= w_off(l) && p < b_off(l)) {
                int e = p - w_off(l), i = e / out, o = e % out;
                uint16_t hb = to_bf16(val);
                w[bw_off(l) + i * out + o] = hb;
                wt[bw_off(l) + o * in + i] = hb;
                if (ema) wt_ema[bw_off(l) + o * in + i] = to_bf16(av);
            }
        }
    };
    for (int pb = blockIdx.x * kAdamParams; pb < kNumParams; pb += gridDim.x * kAdamParams) {
        const int p = pb + x;
        float part = 0.f;
        if (p < kNumParams)
            for (uint32_t r = uint32_t(y); r < rows; r += kAdamGroups) part += partial[size_t(r) * kNumParams + p];
        gs[y][x] = part;
        __syncthreads();
        // every thread reaches both barriers of an iteration; only the first
        // group's threads of real parameters take the step
        if (y == 0 && p < kNumParams) adam_update(p, x);
        __syncthreads();  // gs is rewritten by the next parameter block
    }
}

__global__ void k_nrc_sum_rows(const float *partial, uint32_t rows, float *out) {
    using namespace nrc;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < kNumParams; p += gridDim.x * blockDim.x) {
        float g = 0.f;
        for (uint32_t r = 0; r < rows; ++r) g += partial[size_t(r) * kNumParams + p];
        out[p] = g;
    }
}

__global__ void k_nrc_encode(nrc::Weights W, const float *raw, uint32_t n, float *features) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        nrc::RawRec r = nrc::load_raw(raw + size_t(i) * nrc::kRaw);
        for (int f = 0; f < nrc::kIn; ++f) features[size_t(i) * nrc::kIn + f] = nrc::feature(f, r, W);
    }
}

// Prediction for explicit records; one wave per 64 records.
__global__ __launch_bounds__(256) void k_nrc_predict(nrc::Weights W, const float *raw, uint32_t n, float *rgb) {
    using namespace nrc;
    int lane = __lane_id(), s = lane & 15, q = lane >> 4;
    uint32_t waves = gridDim.x * (blockDim.x / 64);
    for (uint32_t base = (blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64) * 64; base < n; base += waves * 64) {
        v4s in[6][4];
        encode_inputs(in, [&](int k) -> const float * {
            return base + k < n ? raw + size_t(base + k) * kRaw : nullptr;
        }, W);
        v4f out[4];
        forward(in, out, W);
        if (q == 0)
            for (int nt = 0; nt < 4; ++nt) {
                uint32_t i = base + nt * 16 + s;
                if (i < n) {
                    rgb[3 * size_t(i)] = out[nt][0];
                    rgb[3 * size_t(i) + 1] = out[nt][1];
                    rgb[3 * size_t(i) + 2] = out[nt][2];
                }
            }
    }
}
}  // namespace

// ------------------------------------------------------------ C ABI --
#define KR_EXPORT extern "C" __attribute__((visibility("default")))


KR_EXPORT const char *kr_version(void) { return "koishi-mi355x kr 1 (gfx950 wavefront)"; }

KR_EXPORT int kr_last_error(char *buf, size_t len) {
    if (!buf || len == 0) return KR_E_INVALID;
    std::snprintf(buf, len, "%s", g_err.c_str());
    return KR_OK;
}

KR_EXPORT int kr_device_count(int *count) {
    if (!count) return fail(KR_E_INVALID, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return KR_OK;
}

KR_EXPORT int kr_device_open(int ordinal, kr_device **out) {
    if (!out) return fail(KR_E_INVALID, "out is NULL");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(KR_E_NO_DEVICE, "no HIP device visible");
    if (ordinal < 0 || ordinal >= n) return fail(KR_E_INVALID, "device ordinal out of range");
    KR_HIP(hipSetDevice(ordinal));
    hipDeviceProp_t prop;
    KR_HIP(hipGetDeviceProperties(&prop, ordinal));
    auto *d = new kr_device();
    d->ordinal = ordinal;
    d->num_cu = prop.multiProcessorCount;
    d->name = prop.gcnArchName;
    void *sp = nullptr;
    hipError_t e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !sp) {
        delete d;
        return fail(KR_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    d->scratch = static_cast<uint2 *>(sp);
    e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete d;
        return fail(KR_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    *out = d;
    return KR_OK;
}

KR_EXPORT int kr_device_close(kr_device *dev) {
    if (!dev) return KR_OK;
    (void)hipSetDevice(dev->ordinal);
    (void)hipStreamSynchronize(dev->stream);
    (void)hipStreamDestroy(dev->stream);
    cache_release(dev);
    delete dev;
    return KR_OK;
}

KR_EXPORT int kr_device_trim(kr_device *dev, size_t *released) {
    if (!dev) return fail(KR_E_INVALID, "NULL device");
    (void)hipSetDevice(dev->ordinal);
    size_t idle;
    {
        std::lock_guard<std::mutex> lk(dev->mem.mu);
        idle = dev->mem.idle_bytes;
    }
    cache_release(dev);
    if (released) *released = idle;
    return KR_OK;
}

KR_EXPORT int kr_device_bandwidth(kr_device *dev, size_t bytes, uint32_t iters, double *gbps) {
    if (!dev || !gbps || bytes < 16 || iters == 0) return fail(KR_E_INVALID, "bad argument");
    (void)hipSetDevice(dev->ordinal);
    size_t n = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    int rc = dalloc(dev, &a, n);
    if (rc) return rc;
    rc = dalloc(dev, &b, n);
    if (rc) {
        dfree(dev, a);
        return rc;
    }
    hipStream_t st = dev->stream;
    int grid = int(std::min<size_t>((n + kBlock - 1) / kBlock, size_t(std::max(dev->num_cu, 1)) * 32));
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipMemsetAsync(a, 0, n * 16, st);
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(kBlock), 0, st, a, b, n);  // warm-up
    (void)hipEventRecord(e0, st);
    for (uint32_t i = 0; i < iters; ++i) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(kBlock), 0, st, a, b, n);
    (void)hipEventRecord(e1, st);
    hipError_t e = hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dfree(dev, a);
    dfree(dev, b);
    if (e != hipSuccess) return fail(KR_E_HIP, std::string("k_copy: ") + hipGetErrorString(e));
    *gbps = 2.0 * double(n * 16) * iters / (double(ms) * 1e-3) / 1e9;
    return KR_OK;
}

KR_EXPORT int kr_scene_free(kr_scene *s) {
    if (!s) return KR_OK;
    (void)hipSetDevice(s->dev->ordinal);
    (void)hipStreamSynchronize(s->dev->stream);
    for (void *p : s->allocs) dfree(s->dev, p);
    delete s;
    return KR_OK;
}

namespace {
template <class T>
int upload(kr_scene *s, const T *host, size_t n, const T **dst) {
    T *p = nullptr;
    int rc = dalloc(s->dev, &p, n);
    if (rc) return rc;
    s->allocs.push_back(p);
    if (n) KR_HIP(hipMemcpyAsync(p, host, n * sizeof(T), hipMemcpyHostToDevice, s->dev->stream));
    *dst = p;
    return KR_OK;
}

// Blocks per CU of a render's grid-stride launches (batches choose their
// own, see kr_batch_tuning). The traversal kernels fit 6 blocks per CU
// (registers), so most blocks queue behind resident ones: a block that
// finishes early is replaced, which balances the uneven cost of rays. 64
// blocks per CU was measured 21 % faster than 8 on the 100k-triangle frame
// and 13 % on the 1M-triangle frame; the price is the spill stack of every
// lane of the grid, 8 B per entry the scene's tree depth can need beyond the
// LDS part (32 entries, 1.1 GB on 256 CUs, for a BVH4 of depth 14).
#ifndef KR_BLOCKS_PER_CU
#define KR_BLOCKS_PER_CU 64
#endif
size_t grid_lanes(const kr_device *dev) { return size_t(std::max(dev->num_cu, 1)) * KR_BLOCKS_PER_CU * kBlock; }

// Renders of at least this many paths get KR_RENDER_PIPES pipes (struct Pipe).
constexpr uint64_t kPipeMinPaths = 8u << 20;

// Spill entries a traversal of a scene with this stack capacity can need.
uint32_t spill_depth(uint32_t stack_cap) {
    return stack_cap > uint32_t(kLdsDepthMin) ? stack_cap - kLdsDepthMin : 1u;
}

// Spill stacks: `depth` (ref, t) entries per lane beyond the LDS part,
// entry k of lane g at k * lanes + g; depth 0 = the largest a scene may need.
int ensure_spill(uint32_t *&spill, size_t &lanes, std::vector<void *> &allocs, const kr_device *dev,
                 size_t want_lanes = 0, uint32_t depth = 0) {
    if (spill) return KR_OK;
    lanes = want_lanes ? want_lanes : grid_lanes(dev);
    if (depth == 0 || depth > uint32_t(kSpillDepth)) depth = kSpillDepth;
    int rc = dalloc(dev, &spill, lanes * depth * 2);
    if (rc) return rc;
    allocs.push_back(spill);
    return KR_OK;
}
}  // namespace

KR_EXPORT int kr_scene_upload(kr_device *dev, const kr_scene_desc *d, kr_scene **out) {
    if (!dev || !d || !out) return fail(KR_E_INVALID, "NULL argument");
    *out = nullptr;
    auto t0 = std::chrono::steady_clock::now();
    if (d->num_prims == 0) return fail(KR_E_INVALID, "scene has no primitives");
    if (d->num_triangles && (!d->positions || !d->normals || !d->indices))
        return fail(KR_E_INVALID, "mesh arrays missing");
    uint32_t width = d->bvh_width ? d->bvh_width : 4;
    if (width != 2 && width != 4 && width != 8) return fail(KR_E_INVALID, "bvh_width must be 2, 4 or 8");
    for (uint32_t t = 0; t < 3 * d->num_triangles; ++t)
        if (d->indices[t] >= d->num_vertices) return fail(KR_E_INVALID, "triangle index out of range");
    std::vector<uint32_t> tri_prim(d->num_triangles, KR_MISS);
    std::vector<PrimRec> prec(d->num_prims);
    std::vector<double> sph;
    std::vector<uint32_t> sph_prim;
    for (uint32_t i = 0; i < d->num_prims; ++i) {
        const kr_prim_desc &p = d->prims[i];
        if (p.shader >= d->num_shaders) return fail(KR_E_INVALID, "primitive shader index out of range");
        prec[i] = {p.type, p.shader, p.tri_begin, 0};
        if (p.type == KR_PRIM_MESH) {
            if (p.tri_count == 0 || uint64_t(p.tri_begin) + p.tri_count > d->num_triangles)
                return fail(KR_E_INVALID, "mesh triangle range out of bounds");
            for (uint32_t t = p.tri_begin; t < p.tri_begin + p.tri_count; ++t) {
                if (tri_prim[t] != KR_MISS) return fail(KR_E_INVALID, "triangle owned by two primitives");
                tri_prim[t] = i;
            }
        } else if (p.type == KR_PRIM_SPHERE) {
            prec[i].sphere = uint32_t(sph_prim.size());
            sph_prim.push_back(i);
            sph.push_back(p.center[0]);
            sph.push_back(p.center[1]);
            sph.push_back(p.center[2]);
            sph.push_back(double(p.radius) * double(p.radius));
        } else {
            return fail(KR_E_INVALID, "unknown primitive type");
        }
    }
    for (uint32_t t = 0; t < d->num_triangles; ++t)
        if (tri_prim[t] == KR_MISS) return fail(KR_E_INVALID, "triangle not owned by any primitive");
    std::vector<ShaderRec> shd(d->num_shaders);
    for (uint32_t i = 0; i < d->num_shaders; ++i) {
        const kr_shader_desc &s = d->shaders[i];
        ShaderRec r{};
        r.type = s.type;
        r.dist = s.distribution;
        r.alpha2 = s.roughness * s.roughness;
        bool micro = s.type == KR_SHADER_GLOSSY || s.type == KR_SHADER_GLASS || s.type == KR_SHADER_REFRACTION;
        r.roughness = micro ? s.roughness : 0.f;
        r.ior = s.ior;
        if (s.type == KR_SHADER_LAMBERT) {
            for (int c = 0; c < 3; ++c) r.cr[c] = s.color[c] * s.R[c];
        } else if (s.type == KR_SHADER_EMISSION) {
            for (int c = 0; c < 3; ++c) {
                r.cr[c] = 1.f;
                r.emission[c] = s.color[c] * s.strength;
            }
        } else if (micro || s.type == KR_SHADER_TRANSPARENT || s.type == KR_SHADER_BLACKHOLE) {
            for (int c = 0; c < 3; ++c) r.cr[c] = s.color[c];
        } else if (s.type == KR_SHADER_MIXED) {
            if (s.child[0] >= d->num_shaders || s.child[1] >= d->num_shaders)
                return fail(KR_E_INVALID, "Mixed shader child index out of range");
            if (s.fac_type != KR_SCALA_CONSTANT && s.fac_type != KR_SCALA_FRESNEL)
                return fail(KR_E_INVALID, "unknown Mixed fac type");
            r.fac_type = s.fac_type;
            r.fac_value = s.fac_value;
            r.child0 = s.child[0];
            r.child1 = s.child[1];
        } else {
            return fail(KR_E_INVALID, "unknown shader type");
        }
        if (micro && s.distribution != KR_DIST_BECKMANN && s.distribution != KR_DIST_COSINE)
            return fail(KR_E_INVALID, "unknown distribution type");
        if ((s.type == KR_SHADER_GLASS || s.type == KR_SHADER_REFRACTION) && !(s.ior > 0.f))
            return fail(KR_E_INVALID, "GlassBSDF / RefractionBSDF need ior > 0");
        shd[i] = r;
    }
    // Mixed trees: at most 3 Mixed levels below any shader (one Philox block
    // of choices per execute call), which also rules out cycles
    for (uint32_t i = 0; i < d->num_shaders; ++i) {
        std::vector<std::pair<uint32_t, int>> todo{{i, 0}};
        while (!todo.empty()) {
            auto [k, depth] = todo.back();
            todo.pop_back();
            if (d->shaders[k].type != KR_SHADER_MIXED) continue;
            if (depth >= 3) return fail(KR_E_UNSUPPORTED, "Mixed shaders nest deeper than 3 levels");
            todo.push_back({d->shaders[k].child[0], depth + 1});
            todo.push_back({d->shaders[k].child[1], depth + 1});
        }
    }
    auto *s = new kr_scene();
    s->dev = dev;
    s->width = width;
    s->prims.assign(d->prims, d->prims + d->num_prims);
    s->tri_prim = tri_prim;
    s->sphere_prim = sph_prim;
    // point lights first, in file order (scene.cu:158-168)
    if (d->num_point_lights && !d->point_lights) {
        delete s;
        return fail(KR_E_INVALID, "point light array missing");
    }
    for (uint32_t i = 0; i < d->num_point_lights; ++i) {
        LightRec lr{};
        lr.prim = KR_MISS;
        lr.type = kLightPoint;
        for (int c = 0; c < 3; ++c) {
            lr.center[c] = d->point_lights[i].position[c];
            lr.emission[c] = d->point_lights[i].emission[c];
        }
        s->lights.push_back(lr);
    }
    // then every Emission primitive as an area light, in primitive order
    // (scene.cu:200-216)
    for (uint32_t i = 0; i < d->num_prims && !(d->flags & KR_SCENE_NO_AREA_LIGHTS); ++i) {
        const kr_prim_desc &p = d->prims[i];
        const kr_shader_desc &sh = d->shaders[p.shader];
        if (sh.type != KR_SHADER_EMISSION) continue;
        LightRec lr{};
        lr.prim = i;
        lr.type = p.type;
        for (int c = 0; c < 3; ++c) lr.emission[c] = sh.color[c] * sh.strength;
        if (p.type == KR_PRIM_SPHERE) {
            for (int c = 0; c < 3; ++c) lr.center[c] = p.center[c];
            lr.center[3] = p.radius * p.radius;
        } else {
            // Mesh::generateSamples (mesh.hpp:61-85) over the face list padded to
            // a multiple of 4 with copies of the last triangle (mesh.cu:135-143).
            uint32_t light = uint32_t(s->lights.size());
            lr.table = uint32_t(s->light_table.size() / 6);
            uint32_t n = p.tri_count, F = (n + 3) / 4 * 4;
            auto face = [&](uint32_t j, H3 &o, H3 &d1, H3 &d2, H3 &n0, H3 &n1, H3 &n2) {
                uint32_t t = p.tri_begin + (j < n ? j : n - 1);
                const uint32_t *ix = d->indices + 3 * size_t(t);
                o = hload(d->positions + 3 * size_t(ix[0]));
                d1 = hsub(hload(d->positions + 3 * size_t(ix[1])), o);
                d2 = hsub(hload(d->positions + 3 * size_t(ix[2])), o);
                n0 = hload(d->normals + 3 * size_t(ix[0]));
                n1 = hload(d->normals + 3 * size_t(ix[1]));
                n2 = hload(d->normals + 3 * size_t(ix[2]));
            };
            uint32_t key[2] = {uint32_t(d->light_seed), uint32_t(d->light_seed >> 32)};
            for (uint32_t e = 0; e < kLightTable; ++e) {
                uint32_t ctr[4] = {light, e, 0, kStreamLight}, r[4];
                philox_host(ctr, key, r);
                float xi = u01h(r[0]);
                uint32_t idx = std::min(uint32_t(std::floor(xi * float(F))), F - 1);
                float a = u01h(r[1]), b = u01h(r[2]);
                float ux = a < b ? a : b, uy = a > b ? a : b;
                H3 o, d1, d2, n0, n1, n2;
                face(idx, o, d1, d2, n0, n1, n2);
                H3 pp = hadd(hadd(o, hmul(d1, ux)), hmul(d2, uy));
                H3 nn = hnorm(hadd(hadd(hmul(n0, 1.f - ux - uy), hmul(n1, ux)), hmul(n2, uy)));
                float ent[6] = {pp.x, pp.y, pp.z, nn.x, nn.y, nn.z};
                s->light_table.insert(s->light_table.end(), ent, ent + 6);
            }
            float area = 0.f;
            for (uint32_t j = 0; j < F; ++j) {
                H3 o, d1, d2, n0, n1, n2;
                face(j, o, d1, d2, n0, n1, n2);
                H3 c = hcross(d1, d2);
                area = float(double(area) + double(std::sqrt(hdot(c, c))) * .5);
            }
            lr.pdf = 1.f / area;
        }
        s->lights.push_back(lr);
    }
    // BVH over all triangles
    WideBvh bvh;
    std::vector<float> tv(9 * size_t(d->num_triangles));
    for (uint32_t t = 0; t < d->num_triangles; ++t)
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 3; ++c) tv[9 * size_t(t) + 3 * k + c] = d->positions[3 * size_t(d->indices[3 * t + k]) + c];
    if (d->num_triangles) {
        std::string err;
        if (!build_wide_bvh(tv.data(), d->num_triangles, width, d->leaf_size, bvh, err)) {
            delete s;
            return fail(KR_E_INVALID, "bvh: " + err);
        }
    }
    std::vector<float4> tris(3 * size_t(d->num_triangles));
    std::vector<float> nrm(9 * size_t(d->num_triangles));
    for (uint32_t k = 0; k < d->num_triangles; ++k) {
        uint32_t t = bvh.tri_order[k];
        const float *v = &tv[9 * size_t(t)];
        float w0, w1, w2;
        uint32_t pb = tri_prim[t];
        // the material bucket of the triangle's shader (Lambert, GlossyBSDF,
        // everything else), so that a hit is bucketed without gathers
        uint32_t mat = std::min(d->shaders[d->prims[pb].shader].type, 2u);
        std::memcpy(&w0, &pb, 4);
        std::memcpy(&w1, &t, 4);
        std::memcpy(&w2, &mat, 4);
        tris[3 * size_t(k) + 0] = make_float4(v[0], v[1], v[2], w0);
        tris[3 * size_t(k) + 1] = make_float4(v[3], v[4], v[5], w1);
        tris[3 * size_t(k) + 2] = make_float4(v[6], v[7], v[8], w2);
    }
    for (uint32_t t = 0; t < d->num_triangles; ++t)
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 3; ++c) nrm[9 * size_t(t) + 3 * k + c] = d->normals[3 * size_t(d->indices[3 * t + k]) + c];
    (void)hipSetDevice(dev->ordinal);
    DevScene &v = s->view;
    int rc = KR_OK;
    s->qnodes = (d->flags & KR_SCENE_FULL_NODES) == 0;
    // the kernels address nodes with 32-bit byte offsets
    const uint32_t node_bytes = s->qnodes ? qnode_bytes(width) : node_floats(width) * 4;
    if (uint64_t(bvh.num_nodes) * node_bytes >= (uint64_t(1) << 32)) {
        delete s;
        return fail(KR_E_INVALID, "BVH of " + std::to_string(bvh.num_nodes) + " nodes exceeds 4 GB");
    }
    if (s->qnodes) {
        std::vector<uint32_t> qn;
        compress_wide_nodes(bvh, qn);
        // every decoded child box must contain the exact one
        const uint32_t NF = node_floats(width), NW = qnode_bytes(width) / 4;
        for (uint32_t n = 0; n < bvh.num_nodes && !rc; ++n) {
            const float *f = bvh.nodes.data() + size_t(n) * NF;
            for (uint32_t c = 0; c < width; ++c) {
                if (qn[size_t(n) * NW + 4 + c] == kEmpty) continue;
                float lo[3], hi[3];
                decode_qnode_box(qn.data() + size_t(n) * NW, width, c, lo, hi);
                for (int a = 0; a < 3; ++a)
                    if (!(lo[a] <= f[(2 * a) * width + c] && hi[a] >= f[(2 * a + 1) * width + c]))
                        rc = fail(KR_E_OVERFLOW, "compressed BVH node does not contain its child box");
            }
        }
        // inner-node refs become offsets in 16-byte units (see node_ref_units)
        for (uint32_t n = 0; n < bvh.num_nodes; ++n)
            for (uint32_t c = 0; c < width; ++c) {
                uint32_t &ref = qn[size_t(n) * NW + 4 + c];
                if (ref != kEmpty && !(ref & kLeafBit)) ref *= node_bytes / 16;
            }
        const uint32_t *nodes_q = nullptr;
        rc = rc ? rc : upload(s, qn.data(), qn.size(), &nodes_q);
        v.nodes = reinterpret_cast<const float4 *>(nodes_q);
    } else {
        const uint32_t NF = node_floats(width);
        for (uint32_t n = 0; n < bvh.num_nodes; ++n)
            for (uint32_t c = 0; c < width; ++c) {
                float &f = bvh.nodes[size_t(n) * NF + 6 * width + c];
                uint32_t ref;
                std::memcpy(&ref, &f, 4);
                if (ref != kEmpty && !(ref & kLeafBit)) ref *= node_bytes / 16;
                std::memcpy(&f, &ref, 4);
            }
        const float *nodes_f = nullptr;
        rc = rc ? rc : upload(s, bvh.nodes.data(), bvh.nodes.size(), &nodes_f);
        v.nodes = reinterpret_cast<const float4 *>(nodes_f);
    }
    rc = rc ? rc : upload(s, tris.data(), tris.size(), &v.tris);
    rc = rc ? rc : upload(s, nrm.data(), nrm.size(), &v.tri_nrm);
    rc = rc ? rc : upload(s, tri_prim.data(), tri_prim.size(), &v.tri_prim);
    rc = rc ? rc : upload(s, sph.data(), sph.size(), &v.spheres);
    rc = rc ? rc : upload(s, sph_prim.data(), sph_prim.size(), &v.sphere_prim);
    rc = rc ? rc : upload(s, prec.data(), prec.size(), &v.prims);
    rc = rc ? rc : upload(s, shd.data(), shd.size(), &v.shaders);
    rc = rc ? rc : upload(s, s->lights.data(), s->lights.size(), &v.lights);
    rc = rc ? rc : upload(s, s->light_table.data(), s->light_table.size(), &v.light_table);
    if (rc) {
        kr_scene_free(s);
        return rc;
    }
    v.nspheres = uint32_t(sph_prim.size());
    v.nlights = uint32_t(s->lights.size());
    v.ntris = d->num_triangles;
    uint32_t cap = (width - 1) * (bvh.depth + 1) + 1;
    v.stack_cap = cap;
    if (cap > uint32_t(kLdsDepthMin + kSpillDepth)) {
        kr_scene_free(s);
        return fail(KR_E_OVERFLOW, "BVH too deep for the traversal stack");
    }
    hipError_t e = hipStreamSynchronize(dev->stream);
    if (e != hipSuccess) {
        kr_scene_free(s);
        return fail(KR_E_HIP, std::string("upload: ") + hipGetErrorString(e));
    }
    kr_scene_info &in = s->info;
    in.num_prims = d->num_prims;
    in.num_triangles = d->num_triangles;
    in.num_spheres = v.nspheres;
    in.num_lights = v.nlights;
    in.bvh_width = width;
    in.bvh_nodes = bvh.num_nodes;
    in.bvh_leaves = bvh.num_leaves;
    in.bvh_depth = bvh.depth;
    in.node_bytes = s->qnodes ? qnode_bytes(width) : node_floats(width) * 4;
    in.tri_bytes = 48;
    in.device_bytes = size_t(bvh.num_nodes) * in.node_bytes + tris.size() * 16 + nrm.size() * 4 + tri_prim.size() * 4 +
                      sph.size() * 8 + s->light_table.size() * 4;
    in.build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    *out = s;
    return KR_OK;
}

KR_EXPORT int kr_scene_get_info(const kr_scene *s, kr_scene_info *info) {
    if (!s || !info) return fail(KR_E_INVALID, "NULL argument");
    *info = s->info;
    return KR_OK;
}

KR_EXPORT int kr_scene_light_samples(const kr_scene *s, uint32_t light, float *p_n, float *pdf) {
    if (!s || !p_n || !pdf) return fail(KR_E_INVALID, "NULL argument");
    if (light >= s->lights.size()) return fail(KR_E_INVALID, "light index out of range");
    const LightRec &lr = s->lights[light];
    if (lr.type != KR_PRIM_MESH) return fail(KR_E_INVALID, "light is not a mesh light");
    std::memcpy(p_n, &s->light_table[6 * size_t(lr.table)], sizeof(float) * 6 * kLightTable);
    *pdf = lr.pdf;
    return KR_OK;
}

namespace {
// Node format of a traversal instantiation: width W, compressed (Q) or not.
template <int W_, bool Q_>
struct Fmt {
    static constexpr int W = W_;
    static constexpr bool Q = Q_;
};
template <class F>
int dispatch_fmt(uint32_t w, bool q, F f) {
    if (w == 2) return q ? f(Fmt<2, true>{}) : f(Fmt<2, false>{});
    if (w == 4) return q ? f(Fmt<4, true>{}) : f(Fmt<4, false>{});
    if (w == 8) return q ? f(Fmt<8, true>{}) : f(Fmt<8, false>{});
    return fail(KR_E_INVALID, "bad bvh width");
}


struct DevBufs {
    const kr_device *dev;
    std::vector<void *> ptrs;
    explicit DevBufs(const kr_device *d) : dev(d) {}
    ~DevBufs() {
        for (void *p : ptrs) dfree(dev, p);
    }
    template <class T>
    int get(T **p, size_t n) {
        int rc = dalloc(dev, p, n);
        if (!rc) ptrs.push_back(*p);
        return rc;
    }
};

int query_grid(const kr_device *dev, uint32_t n) {
    size_t blocks = (size_t(n) + kBlock - 1) / kBlock;
    size_t cap = grid_lanes(dev) / kBlock;
    return int(std::max<size_t>(1, std::min(blocks, cap)));
}
}  // namespace

KR_EXPORT int kr_trace_closest(kr_scene *s, const kr_rays *rays, uint32_t n, kr_hits *hits, uint32_t mem) {
    if (!s || !rays || !hits || !rays->origin || !rays->direction || !hits->t || !hits->prim || !hits->tri ||
        !hits->u || !hits->v)
        return fail(KR_E_INVALID, "NULL argument");
    if (n == 0) return KR_OK;
    (void)hipSetDevice(s->dev->ordinal);
    hipStream_t stream = s->dev->stream;
    int rc = ensure_spill(s->spill, s->spill_lanes, s->allocs, s->dev, 0, spill_depth(s->view.stack_cap));
    if (rc) return rc;
    DevBufs b(s->dev);
    const float *o = rays->origin, *d = rays->direction;
    float *t = hits->t, *u = hits->u, *v = hits->v;
    uint32_t *prim = hits->prim, *tri = hits->tri;
    if (mem == KR_MEM_HOST) {
        float *o_, *d_;
        if ((rc = b.get(&o_, 3 * size_t(n))) || (rc = b.get(&d_, 3 * size_t(n))) || (rc = b.get(&t, n)) ||
            (rc = b.get(&u, n)) || (rc = b.get(&v, n)) || (rc = b.get(&prim, n)) || (rc = b.get(&tri, n)))
            return rc;
        KR_HIP(hipMemcpyAsync(o_, rays->origin, 12 * size_t(n), hipMemcpyHostToDevice, stream));
        KR_HIP(hipMemcpyAsync(d_, rays->direction, 12 * size_t(n), hipMemcpyHostToDevice, stream));
        o = o_;
        d = d_;
    } else if (mem != KR_MEM_DEVICE) {
        return fail(KR_E_INVALID, "mem must be KR_MEM_HOST or KR_MEM_DEVICE");
    }
    int grid = query_grid(s->dev, n);
    auto launch = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, stream, s->view, o, d, n, t, prim, tri, u, v, s->spill,
                           s->spill_lanes, (uint64_t *)nullptr);
        return KR_OK;
    };
    rc = dispatch_fmt(s->width, s->qnodes, [&](auto fm) {
        using FM = decltype(fm);
        return launch(k_query_closest<FM::W, FM::Q>);
    });
    if (rc) return rc;
    KR_HIP(hipGetLastError());
    if (mem == KR_MEM_HOST) {
        KR_HIP(hipMemcpyAsync(hits->t, t, 4 * size_t(n), hipMemcpyDeviceToHost, stream));
        KR_HIP(hipMemcpyAsync(hits->u, u, 4 * size_t(n), hipMemcpyDeviceToHost, stream));
        KR_HIP(hipMemcpyAsync(hits->v, v, 4 * size_t(n), hipMemcpyDeviceToHost, stream));
        KR_HIP(hipMemcpyAsync(hits->prim, prim, 4 * size_t(n), hipMemcpyDeviceToHost, stream));
        KR_HIP(hipMemcpyAsync(hits->tri, tri, 4 * size_t(n), hipMemcpyDeviceToHost, stream));
    }
    KR_HIP(hipStreamSynchronize(stream));
    return KR_OK;
}

KR_EXPORT int kr_trace_closest_noalloc(kr_scene *s, const kr_rays *rays, uint32_t n, kr_hits *hits) {
    if (!s || !rays || !hits || !rays->origin || !rays->direction || !hits->t || !hits->prim || !hits->tri ||
        !hits->u || !hits->v)
        return fail(KR_E_INVALID, "NULL argument");
    if (n == 0) return KR_OK;
    kr_device *dev = s->dev;
    if (!dev->scratch) return fail(KR_E_HIP, "device scratch address unknown");
    // The scratch holds `depth` spill entries for every lane in flight: the
    // launch is cut to as many lanes as fit, and its grid-stride loop covers
    // any n. A scene whose stack needs more than one block's worth is refused.
    const uint32_t depth = spill_depth(s->view.stack_cap);
    size_t lanes = kScratchBytes / (sizeof(uint2) * size_t(depth)) / kBlock * kBlock;
    if (lanes < size_t(kBlock))
        return fail(KR_E_OVERFLOW, "kr_trace_closest_noalloc: the scene's traversal stack needs " +
                                       std::to_string(size_t(depth) * kBlock * sizeof(uint2)) +
                                       " bytes of scratch per block, more than the fixed " +
                                       std::to_string(kScratchBytes) + " bytes");
    lanes = std::min(lanes, grid_lanes(dev));
    size_t blocks = (size_t(n) + kBlock - 1) / kBlock;
    int grid = int(std::max<size_t>(1, std::min(blocks, lanes / kBlock)));
    KR_HIP(hipSetDevice(dev->ordinal));
    hipStream_t stream = dev->stream;
    uint32_t *spill = reinterpret_cast<uint32_t *>(dev->scratch);
    int rc = dispatch_fmt(s->width, s->qnodes, [&](auto fm) {
        using FM = decltype(fm);
        auto kern = k_query_closest<FM::W, FM::Q>;
        hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, stream, s->view,
                           rays->origin, rays->direction, n, hits->t, hits->prim, hits->tri, hits->u, hits->v, spill,
                           lanes, (uint64_t *)nullptr);
        return KR_OK;
    });
    if (rc) return rc;
    KR_HIP(hipGetLastError());
    KR_HIP(hipStreamSynchronize(stream));
    return KR_OK;
}

KR_EXPORT int kr_trace_shadow(kr_scene *s, const kr_rays *rays, uint32_t n, uint8_t *occluded, uint32_t mem) {
    if (!s || !rays || !occluded || !rays->origin || !rays->direction || !rays->tmax)
        return fail(KR_E_INVALID, "NULL argument");
    if (n == 0) return KR_OK;
    (void)hipSetDevice(s->dev->ordinal);
    hipStream_t stream = s->dev->stream;
    int rc = ensure_spill(s->spill, s->spill_lanes, s->allocs, s->dev, 0, spill_depth(s->view.stack_cap));
    if (rc) return rc;
    DevBufs b(s->dev);
    const float *o = rays->origin, *d = rays->direction, *tm = rays->tmax;
    uint8_t *occ = occluded;
    if (mem == KR_MEM_HOST) {
        float *o_, *d_, *t_;
        if ((rc = b.get(&o_, 3 * size_t(n))) || (rc = b.get(&d_, 3 * size_t(n))) || (rc = b.get(&t_, n)) ||
            (rc = b.get(&occ, n)))
            return rc;
        KR_HIP(hipMemcpyAsync(o_, rays->origin, 12 * size_t(n), hipMemcpyHostToDevice, stream));
        KR_HIP(hipMemcpyAsync(d_, rays->direction, 12 * size_t(n), hipMemcpyHostToDevice, stream));
        KR_HIP(hipMemcpyAsync(t_, rays->tmax, 4 * size_t(n), hipMemcpyHostToDevice, stream));
        o = o_;
        d = d_;
        tm = t_;
    } else if (mem != KR_MEM_DEVICE) {
        return fail(KR_E_INVALID, "mem must be KR_MEM_HOST or KR_MEM_DEVICE");
    }
    int grid = query_grid(s->dev, n);
    auto launch = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, stream, s->view, o, d, tm, n, occ, s->spill,
                           s->spill_lanes);
        return KR_OK;
    };
    rc = dispatch_fmt(s->width, s->qnodes, [&](auto fm) {
        using FM = decltype(fm);
        return launch(k_query_shadow<FM::W, FM::Q>);
    });
    if (rc) return rc;
    KR_HIP(hipGetLastError());
    if (mem == KR_MEM_HOST) KR_HIP(hipMemcpyAsync(occluded, occ, n, hipMemcpyDeviceToHost, stream));
    KR_HIP(hipStreamSynchronize(stream));
    return KR_OK;
}

// ----------------------------------------------------------------- render --
KR_EXPORT int kr_render_free(kr_render *r) {
    if (!r) return KR_OK;
    (void)hipSetDevice(r->scene->dev->ordinal);
    (void)hipStreamSynchronize(r->scene->dev->stream);
    for (Pipe &pp : r->pipe)
        if (pp.own_stream) (void)hipStreamSynchronize(pp.stream);
    for (void *p : r->allocs) dfree(r->scene->dev, p);
    for (auto &pe : r->pending) {
        (void)hipEventDestroy(pe.second.first);
        (void)hipEventDestroy(pe.second.second);
    }
    for (hipEvent_t e : r->event_pool) (void)hipEventDestroy(e);
    for (void *p : r->nrc_allocs) dfree(r->scene->dev, p);
    for (void *p : r->prof_allocs) dfree(r->scene->dev, p);
    for (Pipe &pp : r->pipe) {
        if (pp.acc_done) (void)hipEventDestroy(pp.acc_done);
        if (pp.own_stream) (void)hipStreamDestroy(pp.stream);
    }
    if (r->fork) (void)hipEventDestroy(r->fork);
    delete r;
    return KR_OK;
}

namespace {
// Lens ctor (src/core/misc/lens.hpp:15-27), FP32 as in the reference.
LensDev make_lens(const kr_camera *cam, const kr_params *prm) {
    LensDev L{};
    L.lens = cam->lens;
    L.w = prm->width;
    L.h = prm->height;
    L.pw = uint32_t(std::ceil(std::sqrt(double(prm->spp))));
    H3 tgt = hnorm(hload(cam->target));
    float den = float(2 * std::tan(3.14159265358979323846 * (double(cam->fovx) * .5) / 180.));
    H3 n = hmul(tgt, float(prm->width));
    n = {n.x / den, n.y / den, n.z / den};
    H3 u = hnorm(hcross(n, hload(cam->upaxis)));
    H3 v = hnorm(hcross(n, u));
    float pw = float(L.pw);
    H3 du{u.x / pw, u.y / pw, u.z / pw}, dv{v.x / pw, v.y / pw, v.z / pw};
    auto st3 = [](float *dst, H3 a) {
        dst[0] = a.x;
        dst[1] = a.y;
        dst[2] = a.z;
    };
    st3(L.n, n);
    st3(L.u, u);
    st3(L.v, v);
    st3(L.du, du);
    st3(L.dv, dv);
    st3(L.o, hload(cam->position));
    return L;
}

int check_view(const kr_camera *cam, const kr_params *prm) {
    if (prm->width == 0 || prm->height == 0 || prm->spp == 0) return fail(KR_E_INVALID, "empty render");
    if (uint64_t(prm->width) * prm->height > (1ull << 31)) return fail(KR_E_INVALID, "resolution too large");
    if (cam->lens != KR_LENS_PINHOLE && cam->lens != KR_LENS_ORTHOGRAPHIC) return fail(KR_E_INVALID, "unknown lens");
    if (prm->kernel > KR_KERNEL_HALFVECTOR) return fail(KR_E_INVALID, "unknown kernel");
    return KR_OK;
}

// Tail rounds (k_trace_tail): a launch's iteration budget, the follow-up launches
// per queue and the stack entries a queued walk keeps; KR_TAIL_STEPS,
// KR_TAIL_ROUNDS, KR_TAIL_STACK and KR_TAIL_FROM override them.
uint32_t env_u32(const char *name, uint32_t dflt) {
    const char *e = std::getenv(name);
    return e && *e ? uint32_t(std::strtoul(e, nullptr, 10)) : dflt;
}
#ifndef KR_TAIL_STEPS
#define KR_TAIL_STEPS 28
#endif
#ifndef KR_TAIL_ROUNDS
#define KR_TAIL_ROUNDS 2
#endif
constexpr uint32_t kTailMinPaths = 2u << 20;
uint32_t tail_stack_entries() { return std::max(1u, std::min(64u, env_u32("KR_TAIL_STACK", 12))); }
void tail_config(kr_render *r) {
    r->tail_steps = (r->prm.flags & (KR_RENDER_NO_TAIL_ROUNDS | KR_RENDER_PERSISTENT_TRAVERSAL))
                        ? 0u
                        : env_u32("KR_TAIL_STEPS", KR_TAIL_STEPS);
    r->tail_rounds = std::min<uint32_t>(env_u32("KR_TAIL_ROUNDS", KR_TAIL_ROUNDS), kMaxTailRounds);
    r->tail_from = env_u32("KR_TAIL_FROM", 1);
    if (r->tail_rounds == 0) r->tail_steps = 0;
}

// Queues for r->P paths per pipe (r->npipes of them), accumulators for
// r->npix pixels, counters, the traversal spill stack for `lanes` lanes per
// pipe (spill_entries per lane, 0 = the most any scene may need); r->grid =
// lanes / kBlock. Pipe 0 runs on the device's stream, the others on streams
// of their own (default priority).
int render_buffers(kr_render *r, size_t lanes, uint32_t spill_entries) {
    size_t P = r->P;
    auto get = [&](auto **p, size_t cnt) {
        int rc = dalloc(r->scene->dev, p, cnt);
        if (!rc) r->allocs.push_back(*p);
        return rc;
    };
    int rc = KR_OK;
    for (int pi = 0; pi < r->npipes && !rc; ++pi) {
        Pipe &pp = r->pipe[pi];
        for (int k = 0; k < 2 && !rc; ++k) {
            RayQ &q = pp.q[k];
            rc = get(&q.slot, P);
            float **fs[9] = {&q.ox, &q.oy, &q.oz, &q.dx, &q.dy, &q.dz, &q.bx, &q.by, &q.bz};
            for (auto f : fs)
                if (!rc) rc = get(f, P);
            if (!rc) rc = get(&q.flags, P);
        }
        if (!rc) rc = get(&pp.hit.t, P);
        if (!rc) rc = get(&pp.hit.u, P);
        if (!rc) rc = get(&pp.hit.v, P);
        if (!rc) rc = get(&pp.hit.code, P);
        for (int m = 0; m < 3 && !rc; ++m) rc = get(&pp.bucket[m], P);
        if (!rc) rc = get(&pp.sq.slot, P);
        {
            float **fs[10] = {&pp.sq.ox, &pp.sq.oy, &pp.sq.oz, &pp.sq.dx, &pp.sq.dy,
                              &pp.sq.dz, &pp.sq.tmax, &pp.sq.cx, &pp.sq.cy, &pp.sq.cz};
            for (auto f : fs)
                if (!rc) rc = get(f, P);
        }
        if (!rc) rc = get(&pp.L, 3 * P);
        if (!rc) rc = get(&pp.counters, size_t(C_NUM) * kCtrStride);
        if (!rc) rc = ensure_spill(pp.spill, pp.spill_lanes, r->allocs, r->scene->dev, lanes, spill_entries);
        if (r->tail_steps) {
            // room for half the queue per round; walks beyond it finish in place
            const uint32_t cap = uint32_t((P + 1) / 2), scap = tail_stack_entries();
            for (int k = 0; k < 2 && !rc; ++k) {
                TailQ &t = pp.tail[k];
                t.cap = cap;
                t.stk_cap = scap;
                rc = get(&t.idx, cap);
                if (!rc) rc = get(&t.cur, cap);
                if (!rc) rc = get(&t.info, cap);
                if (!rc) rc = get(&t.stk, size_t(cap) * scap);
            }
        }
        if (!rc) {
            if (pi == 0) {
                pp.stream = r->scene->dev->stream;
            } else {
                hipError_t e = hipStreamCreateWithFlags(&pp.stream, hipStreamNonBlocking);
                if (e != hipSuccess) return fail(KR_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
                pp.own_stream = true;
            }
            hipError_t e = hipEventCreateWithFlags(&pp.acc_done, hipEventDisableTiming);
            if (e != hipSuccess) return fail(KR_E_HIP, std::string("hipEventCreate: ") + hipGetErrorString(e));
        }
    }
    if (!rc && !r->fork) {
        hipError_t e = hipEventCreateWithFlags(&r->fork, hipEventDisableTiming);
        if (e != hipSuccess) return fail(KR_E_HIP, std::string("hipEventCreate: ") + hipGetErrorString(e));
    }
    if (!rc) rc = get(&r->stats, S_NUM);
    if (!rc) rc = get(&r->acc, 3 * size_t(r->npix));
    if (!rc) rc = get(&r->valid, r->npix);
    for (int pi = 0; pi < r->npipes && !rc; ++pi) {
        r->pipe[pi].acc = r->acc;
        r->pipe[pi].valid = r->valid;
    }
    if (!rc) r->grid = int(r->pipe[0].spill_lanes / kBlock);
    return rc;
}
}  // namespace

KR_EXPORT int kr_render_create(kr_scene *s, const kr_camera *cam, const kr_params *prm, kr_render **out) {
    if (!s || !cam || !prm || !out) return fail(KR_E_INVALID, "NULL argument");
    *out = nullptr;
    int rc = check_view(cam, prm);
    if (rc) return rc;
    (void)hipSetDevice(s->dev->ordinal);
    auto *r = new kr_render();
    r->scene = s;
    r->cam = *cam;
    r->prm = *prm;
    r->npix = prm->width * prm->height;
    r->lens = make_lens(cam, prm);
    uint64_t total = uint64_t(r->npix) * prm->spp;
    // Default chunk: up to 32M paths in flight (~5.5 GB of queues), so that
    // each bounce's kernels see tens of millions of rays and their tails and
    // launch gaps are amortised.
    uint64_t P = prm->chunk_paths ? prm->chunk_paths : (32u << 20);
    P = std::min<uint64_t>(P, total);
    P = std::max<uint64_t>(P, 1);
    r->P = uint32_t(P);
    // more pipes once a frame is large enough to split into chunks that each
    // still fill the GPU
    r->npipes = total >= kPipeMinPaths ? KR_RENDER_PIPES : 1;
    if (r->npipes > 1 && !prm->chunk_paths) {
        // a multiple of npipes equal chunks of at most P paths
        uint64_t per = (total + r->npipes - 1) / r->npipes;
        uint64_t m = (per + P - 1) / P;
        r->P = uint32_t((per + m - 1) / m);
    }
    tail_config(r);
    rc = render_buffers(r, 0, spill_depth(s->view.stack_cap));
    if (rc) {
        kr_render_free(r);
        return rc;
    }
    rc = kr_render_reset(r);
    if (rc) {
        kr_render_free(r);
        return rc;
    }
    *out = r;
    return KR_OK;
}

KR_EXPORT int kr_render_reset(kr_render *r) {
    if (!r) return fail(KR_E_INVALID, "NULL render");
    hipStream_t st = r->scene->dev->stream;
    KR_HIP(hipMemsetAsync(r->acc, 0, sizeof(double) * 3 * size_t(r->npix), st));
    KR_HIP(hipMemsetAsync(r->valid, 0, sizeof(uint32_t) * size_t(r->npix), st));
    KR_HIP(hipMemsetAsync(r->stats, 0, sizeof(unsigned long long) * S_NUM, st));
    for (int pi = 0; pi < r->npipes; ++pi) {
        const Pipe &pp = r->pipe[pi];
        KR_HIP(hipMemsetAsync(pp.counters, 0, sizeof(uint32_t) * C_NUM * kCtrStride, st));
        if (pp.acc != r->acc) {
            KR_HIP(hipMemsetAsync(pp.acc, 0, sizeof(double) * 3 * size_t(r->npix), st));
            KR_HIP(hipMemsetAsync(pp.valid, 0, sizeof(uint32_t) * size_t(r->npix), st));
        }
    }
    if (r->pv.enabled) {
        KR_HIP(hipMemsetAsync(r->pv.rec, 0, sizeof(float) * kProfFloats * r->prof_slices * r->pv.depth, st));
        KR_HIP(hipMemsetAsync(r->pv.fin, 0, sizeof(float) * 4 * r->prof_slices, st));
    }
    KR_HIP(hipStreamSynchronize(st));
    resolve_events(r);
    for (int k = 0; k < K_KINDS; ++k) {
        r->kind_ms[k] = 0;
        r->kind_busy_ms[k] = 0;
        r->kind_launches[k] = 0;
    }
    r->paths = 0;
    return KR_OK;
}

namespace {
// Brackets one launch with HIP events on the render's stream when
// KR_PROFILE_EVENTS is on; the pairs are resolved at the next sync.
struct Timed {
    kr_render *r;
    int kind;
    hipStream_t st;
    hipEvent_t a = nullptr, b = nullptr;
    Timed(kr_render *r_, int k, hipStream_t s = nullptr) : r(r_), kind(k), st(s ? s : r_->scene->dev->stream) {
        if (!(r->profile & KR_PROFILE_EVENTS) && !((r->profile & KR_PROFILE_TRACE_EVENTS) && kind == K_TRACE)) return;
        a = take();
        b = take();
        if (a && b) (void)hipEventRecord(a, st);
    }
    hipEvent_t take() {
        hipEvent_t e = nullptr;
        if (!r->event_pool.empty()) {
            e = r->event_pool.back();
            r->event_pool.pop_back();
        } else if (hipEventCreate(&e) != hipSuccess) {
            e = nullptr;
        }
        return e;
    }
    ~Timed() {
        if (!a || !b) return;
        (void)hipEventRecord(b, st);
        r->pending.push_back({kind, {a, b}});
    }
};

// Adds the resolved launches to the per-kind sums. With several pipes,
// launches of one kind overlap in time, so besides the summed durations the union of
// their intervals (measured from the first pending event) is kept: the time
// during which at least one launch of the kind ran. Every pending event has
// completed when this runs (after a synchronisation of every pipe), and later
// events are later in time, so unions of successive calls add up.
void resolve_events(kr_render *r) {
    std::vector<std::pair<float, float>> iv[K_KINDS];
    hipEvent_t ref = r->pending.empty() ? nullptr : r->pending.front().second.first;
    for (auto &pe : r->pending) {
        float ms = 0.f, t0 = 0.f, t1 = 0.f;
        if (hipEventElapsedTime(&ms, pe.second.first, pe.second.second) == hipSuccess) {
            r->kind_ms[pe.first] += ms;
            r->kind_launches[pe.first] += 1;
            if (hipEventElapsedTime(&t0, ref, pe.second.first) == hipSuccess &&
                hipEventElapsedTime(&t1, ref, pe.second.second) == hipSuccess)
                iv[pe.first].push_back({t0, t1});
        }
    }
    for (int k = 0; k < K_KINDS; ++k) {
        std::sort(iv[k].begin(), iv[k].end());
        double busy = 0.0, lo = 0.0, hi = -1.0;
        for (auto &x : iv[k]) {
            if (x.first > hi) {
                if (hi > lo) busy += hi - lo;
                lo = x.first;
                hi = x.second;
            } else {
                hi = std::max<double>(hi, x.second);
            }
        }
        if (hi > lo) busy += hi - lo;
        r->kind_busy_ms[k] += busy;
    }
    for (auto &pe : r->pending) {
        r->event_pool.push_back(pe.second.first);
        r->event_pool.push_back(pe.second.second);
    }
    r->pending.clear();
}

// A batch's jobs as the host sees them: path ranges and accumulator slices.
struct HostJob {
    uint64_t begin, paths;
    uint32_t npix, acc_off;
};
struct BatchView {
    JobTable T;
    const std::vector<HostJob> *jobs;
};

// A batch chunk that lies inside one job, seen as a chunk of that job's own
// render: its scene, lens, pixel count, depth, seed and accumulator slice.
struct JobChunk {
    DevScene S;
    LensDev lens;
    uint32_t npix, max_bounce;
    uint64_t seed;
    double *acc;
    uint32_t *valid;
    kr_nrc *nrc;   // the job's radiance cache, or null
    NrcRender nr;  // its view over the pipe's cache buffers (enabled = 0 without one)
};

// Tile order of a chunk's camera rays (TileOrder): 2 x 2 pixels with 16
// samples each fill a wave when the chunk's sample count allows, else 4 x 4
// pixels with 4 samples, else 8 x 8 pixels with one (on the 1M-triangle
// 1080p frame, against row-major order: 108.0 / 109.3 / 111.3 against
// 117.4 ms per frame); used when the chunk holds whole samples of the frame
// and the tiles divide it evenly, else the queue stays in row-major order.
// KR_CAMERA_TILE="WxHxS" (e.g. "8x8x1") overrides the shape; "0" turns the
// tile order off.
TileOrder tile_order(uint32_t w, uint32_t h, uint64_t g0, uint32_t P) {
    TileOrder to{};
    const uint32_t npix = w * h;
    if (g0 % npix || P % npix) return to;
    const uint32_t sc = P / npix;  // samples in the chunk
    uint32_t tw, th, sg;
    if (const char *e = std::getenv("KR_CAMERA_TILE")) {
        unsigned a = 0, b = 0, c = 0;
        if (std::sscanf(e, "%ux%ux%u", &a, &b, &c) != 3) return to;
        tw = a, th = b, sg = c;
    } else if (sc % 16 == 0 && w % 2 == 0 && h % 2 == 0) {
        tw = th = 2, sg = 16;
    } else if (sc % 4 == 0 && w % 4 == 0 && h % 4 == 0) {
        tw = th = 4, sg = 4;
    } else {
        tw = th = 8, sg = 1;
    }
    if (tw == 0 || th == 0 || sg == 0 || w % tw || h % th || sc % sg) return to;
    to.tw = tw;
    to.th = th;
    to.sg = sg;
    to.tiles_x = w / tw;
    to.ntiles = (w / tw) * (h / th);
    return to;
}

// Enqueues one chunk of P paths starting at global path g0 on pipe pp:
// camera, then maxBounce + 1 rounds of trace / shade / shadow, then the FP64
// accumulation, which first waits for `acc_after` (the other pipe's last
// accumulation, or nullptr) and then records pp.acc_done. Only kernel
// launches and events on the pipe's stream, so the sequence can be captured
// into a HIP graph. bv != nullptr: a batch chunk that mixes jobs (BATCH =
// true); jc != nullptr: a batch chunk inside one job, run by the single-render
// kernels with g0 the job's own path index.
template <int W, bool BATCH, bool Q>
int run_chunk(kr_render *r, Pipe &pp, uint64_t g0, uint32_t P, const BatchView *bv, hipEvent_t acc_after,
              const JobChunk *jc = nullptr) {
    kr_scene *s = r->scene;
    hipStream_t st = pp.stream;
    const DevScene &S = jc ? jc->S : s->view;
    const LensDev &lens = jc ? jc->lens : r->lens;
    const uint32_t npix = jc ? jc->npix : r->npix, max_bounce = jc ? jc->max_bounce : r->prm.max_bounce;
    const uint64_t seed = jc ? jc->seed : r->prm.seed;
    double *acc = jc ? jc->acc : pp.acc;
    uint32_t *valid = jc ? jc->valid : pp.valid;
    JobTable T = bv ? bv->T : JobTable{nullptr, nullptr, 0};
    // the profiler (single renders only); the one-hit kernels record no vertices
    ProfView pv = (BATCH || jc) ? ProfView{} : r->pv;
    if (r->prm.kernel != KR_KERNEL_RADIANCE) pv.depth = 0;
    // the radiance cache of the chunk: the render's, or a batch job's
    const NrcRender nrv = jc ? jc->nr : (BATCH ? NrcRender{} : pp.nr);
    kr_nrc *const nrc = jc ? jc->nrc : r->nrc;
    // no more lanes than the chunk has paths: a small chunk's launches would
    // otherwise start blocks that find no work
    int grid = int(std::min<size_t>((P + kBlock - 1) / kBlock, size_t(r->grid)));
    int cam_grid = grid;
    int cur = 0;
    bool count = (r->profile & KR_PROFILE_COUNTERS) != 0;
    {
        Timed tm(r, K_OTHER, st);
        hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, st, pp.counters, 0xFFFFFFFFu);
        TileOrder to{};
        if (!BATCH && !pv.enabled)
            to = jc ? tile_order(jc->lens.w, jc->lens.h, g0, P) : tile_order(r->prm.width, r->prm.height, g0, P);
        hipLaunchKernelGGL((k_camera<BATCH>), dim3(cam_grid), dim3(kBlock), 0, st, lens, T, g0, P, npix,
                           pp.q[0], pp.L, pp.counters, C_RAY0, nrv, uint32_t(seed), uint32_t(seed >> 32), pv, to);
    }
    int sort = (r->prm.flags & KR_RENDER_NO_MATERIAL_SORT) ? 0 : 1;
    const bool dyn = (r->prm.flags & KR_RENDER_PERSISTENT_TRAVERSAL) != 0;
    // tail rounds pay on large queues; on a small chunk their extra launches
    // cost more than the waves they pack (256 x 256 x 16-spp sphere frame:
    // 0.50 -> 0.56 ms with them)
    const bool tails = r->tail_steps != 0 && pp.tail[0].cap != 0 && P >= kTailMinPaths;
    const uint32_t kern = r->prm.kernel;
    const uint32_t rounds = kern == KR_KERNEL_RADIANCE ? max_bounce + 1 : 1;
    for (uint32_t b = 0; b < rounds; ++b) {
        int qi = cur == 0 ? C_RAY0 : C_RAY1;
        int qni = cur == 0 ? C_RAY1 : C_RAY0;
        {
            // clear buckets, shadow count and the next ray queue's count
            Timed tm(r, K_OTHER, st);
            uint32_t mask = (0x7u << C_BUCKET) | (1u << C_SHADOW) | (1u << qni) | (1u << C_FETCH) | (1u << C_FETCH_SH) |
                            (0x1Fu << C_TAIL) | (0x1Fu << C_TAIL_SH);
            hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, st, pp.counters, mask);
        }
        {
            Timed tm(r, K_TRACE, st);
            auto trace = [&](auto kern) {
                hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, st, S, T, g0, pp.q[cur], qi, pp.hit,
                                   pp.bucket[0], pp.bucket[1], pp.bucket[2], pp.counters, r->stats, sort, pp.spill,
                                   pp.spill_lanes);
            };
            auto trace_refill = [&](auto kern) {
                hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, st, S, pp.q[cur], qi, pp.hit, pp.bucket[0],
                                   pp.bucket[1], pp.bucket[2], pp.counters, r->stats, sort, pp.spill, pp.spill_lanes);
            };
            // tail rounds: the first launch walks every ray for up to tail_steps
            // iterations, each follow-up resumes the walks queued by the
            // previous one, the last without a limit
            auto trace_tail = [&](auto kern, int round) {
                const TailQ &tin = pp.tail[(round + 1) & 1], &tout = pp.tail[round & 1];
                uint32_t steps = round < int(r->tail_rounds) ? r->tail_steps : ~0u;
                hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, st, S, pp.q[cur], qi, pp.hit, pp.bucket[0],
                                   pp.bucket[1], pp.bucket[2], pp.counters, r->stats, sort, pp.spill,
                                   pp.spill_lanes, tin, C_TAIL + std::max(round - 1, 0), tout, C_TAIL + round,
                                   steps);
            };
            if (!BATCH && dyn) {
                count ? trace_refill(k_trace_refill<W, true, Q>) : trace_refill(k_trace_refill<W, false, Q>);
            } else if (!BATCH && tails && b >= r->tail_from) {
                count ? trace_tail(k_trace_tail<W, true, Q, true>, 0) : trace_tail(k_trace_tail<W, false, Q, true>, 0);
                for (int rd = 1; rd <= int(r->tail_rounds); ++rd)
                    count ? trace_tail(k_trace_tail<W, true, Q, false>, rd)
                          : trace_tail(k_trace_tail<W, false, Q, false>, rd);
            } else {
                count ? trace(k_trace<W, true, BATCH, Q>) : trace(k_trace<W, false, BATCH, Q>);
            }
        }
        ShadeArgs a;
        a.S = S;
        a.q = pp.q[cur];
        a.qn = pp.q[1 - cur];
        a.qi = qi;
        a.qni = qni;
        a.hit = pp.hit;
        a.b0 = pp.bucket[0];
        a.b1 = pp.bucket[1];
        a.b2 = pp.bucket[2];
        a.sq = pp.sq;
        a.L = pp.L;
        a.counters = pp.counters;
        a.stats = r->stats;
        a.g0 = g0;
        a.npix = npix;
        a.bounce = b;
        a.max_bounce = max_bounce;
        a.seed_lo = uint32_t(seed);
        a.seed_hi = uint32_t(seed >> 32);
        a.nr = nrv;
        a.T = T;
        a.pv = pv;
        {
            Timed tm(r, K_SHADE, st);
            auto shade = [&](auto kfn) { hipLaunchKernelGGL(kfn, dim3(grid), dim3(kBlock), 0, st, a); };
            if (kern == KR_KERNEL_RADIANCE) {
                // the profiler records single renders only
                if (nrv.enabled)
                    (!BATCH && pv.enabled) ? shade(k_shade<false, true, true>) : shade(k_shade<BATCH, true, false>);
                else
                    (!BATCH && pv.enabled) ? shade(k_shade<false, false, true>) : shade(k_shade<BATCH, false, false>);
            } else
                hipLaunchKernelGGL((k_shade_debug<BATCH>), dim3(grid), dim3(kBlock), 0, st, a, kern);
        }
        {
            Timed tm(r, K_SHADOW, st);
            auto shadow_tail = [&](auto kern, int round) {
                const TailQ &tin = pp.tail[(round + 1) & 1], &tout = pp.tail[round & 1];
                uint32_t steps = round < int(r->tail_rounds) ? r->tail_steps : ~0u;
                hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, st, S, g0, pp.sq, pp.counters, pp.L, r->stats,
                                   pp.spill, pp.spill_lanes, nrv, b, pv, tin, C_TAIL_SH + std::max(round - 1, 0),
                                   tout, C_TAIL_SH + round, steps);
            };
            if (dyn) {
                hipLaunchKernelGGL((k_shadow_dyn<W, BATCH, Q>), dim3(grid), dim3(kBlock), 0, st, S, T, g0, pp.sq,
                                   pp.counters, pp.L, r->stats, pp.spill, pp.spill_lanes, nrv, b, pv);
            } else if (!BATCH && tails) {
                shadow_tail(k_shadow_tail<W, Q, true>, 0);
                for (int rd = 1; rd <= int(r->tail_rounds); ++rd) shadow_tail(k_shadow_tail<W, Q, false>, rd);
            } else
                hipLaunchKernelGGL((k_shadow<W, BATCH, Q>), dim3(grid), dim3(kBlock), 0, st, S, T, g0, pp.sq,
                                   pp.counters, pp.L, r->stats, pp.spill, pp.spill_lanes, nrv, b, pv);
        }
        cur = 1 - cur;
    }
    if (!BATCH && nrv.enabled && nrc && kern == KR_KERNEL_RADIANCE) nrc_after_chunk(r, pp, nrc, nrv);
    if (pv.enabled) {
        Timed tm(r, K_OTHER, st);
        hipLaunchKernelGGL(k_prof_final, dim3(cam_grid), dim3(kBlock), 0, st, pv, pp.L, g0, P);
    }
    if (acc_after) KR_HIP(hipStreamWaitEvent(st, acc_after, 0));
    {
        Timed tm(r, K_OTHER, st);
        if (!BATCH) {
            int acc_grid = int(std::min<size_t>((npix + kBlock - 1) / kBlock, size_t(grid)));
            hipLaunchKernelGGL(k_accumulate, dim3(acc_grid), dim3(kBlock), 0, st, pp.L, g0, P, npix, acc, valid,
                               r->stats);
        } else {
            // one accumulation per job overlapping the chunk, on its slice
            for (const HostJob &J : *bv->jobs) {
                uint64_t lo = std::max(g0, J.begin), hi = std::min(g0 + P, J.begin + J.paths);
                if (lo >= hi) continue;
                int acc_grid = int(std::min<size_t>((J.npix + kBlock - 1) / kBlock, size_t(grid)));
                hipLaunchKernelGGL(k_accumulate, dim3(acc_grid), dim3(kBlock), 0, st, pp.L + 3 * (lo - g0),
                                   lo - J.begin, uint32_t(hi - lo), J.npix, pp.acc + 3 * size_t(J.acc_off),
                                   pp.valid + J.acc_off, r->stats);
            }
        }
    }
    KR_HIP(hipEventRecord(pp.acc_done, st));
    KR_HIP(hipGetLastError());
    return KR_OK;
}

// Runs f(pipe, acc_after) for consecutive chunks, cycling through the
// render's pipes: pipes 1.. first wait for everything queued on pipe 0
// (resets, earlier runs), each accumulation waits for the previous one, and
// pipe 0 finally waits for every other pipe, so that work queued afterwards
// on the device's stream sees the whole run.
template <class F>
int pipelined(kr_render *r, int npipes, uint64_t nchunks, F &&f) {
    npipes = int(std::min<uint64_t>(uint64_t(npipes), nchunks));
    if (npipes > 1) {
        KR_HIP(hipEventRecord(r->fork, r->pipe[0].stream));
        for (int pi = 1; pi < npipes; ++pi) KR_HIP(hipStreamWaitEvent(r->pipe[pi].stream, r->fork, 0));
    } else {
        npipes = 1;
    }
    for (uint64_t k = 0; k < nchunks; ++k) {
        Pipe &pp = r->pipe[k % npipes];
        hipEvent_t after = k > 0 && npipes > 1 ? r->pipe[(k - 1) % npipes].acc_done : nullptr;
        int rc = f(pp, after, k);
        if (rc) return rc;
    }
    for (int pi = 1; pi < npipes; ++pi) KR_HIP(hipStreamWaitEvent(r->pipe[0].stream, r->pipe[pi].acc_done, 0));
    return KR_OK;
}
}  // namespace

KR_EXPORT int kr_render_run(kr_render *r, uint32_t spp_begin, uint32_t spp_count) {
    if (!r) return fail(KR_E_INVALID, "NULL render");
    if (spp_count == 0) return KR_OK;
    if (uint64_t(spp_begin) + spp_count > 0xFFFFFFFFull) return fail(KR_E_INVALID, "sample range overflows");
    if (uint64_t(spp_begin) + spp_count > r->prm.spp)
        return fail(KR_E_INVALID, "sample range exceeds the render's spp (the lens strata depend on it)");
    (void)hipSetDevice(r->scene->dev->ordinal);
    auto t0 = std::chrono::steady_clock::now();
    const uint64_t g0 = uint64_t(spp_begin) * r->npix, gend = uint64_t(spp_begin + spp_count) * r->npix;
    const uint64_t nchunks = (gend - g0 + r->P - 1) / r->P;
    const int np = (r->prm.flags & KR_RENDER_ONE_STREAM) ? 1 : r->npipes;
    int rc = pipelined(r, np, nchunks, [&](Pipe &pp, hipEvent_t after, uint64_t k) {
        uint64_t g = g0 + k * r->P;
        uint32_t P = uint32_t(std::min<uint64_t>(r->P, gend - g));
        int rc2 = dispatch_fmt(r->scene->width, r->scene->qnodes, [&](auto fm) {
            using FM = decltype(fm);
            return run_chunk<FM::W, false, FM::Q>(r, pp, g, P, nullptr, after);
        });
        // keep the number of outstanding event pairs bounded
        if (!rc2 && r->pending.size() > 4096) {
            for (int pi = 0; pi < r->npipes; ++pi) KR_HIP(hipStreamSynchronize(r->pipe[pi].stream));
            resolve_events(r);
        }
        return rc2;
    });
    if (rc) return rc;
    r->paths += gend - g0;
    if (r->nr.enabled) {
        int rc = nrc_after_run(r, r->nrc);
        if (rc) return rc;
    }
    r->last_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return KR_OK;
}

KR_EXPORT int kr_render_sync(kr_render *r) {
    if (!r) return fail(KR_E_INVALID, "NULL render");
    auto t0 = std::chrono::steady_clock::now();
    KR_HIP(hipStreamSynchronize(r->scene->dev->stream));
    r->last_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    resolve_events(r);
    return KR_OK;
}

KR_EXPORT int kr_render_read(kr_render *r, double *rgb, uint32_t *valid) {
    if (!r) return fail(KR_E_INVALID, "NULL render");
    hipStream_t st = r->scene->dev->stream;
    if (rgb) KR_HIP(hipMemcpyAsync(rgb, r->acc, sizeof(double) * 3 * size_t(r->npix), hipMemcpyDeviceToHost, st));
    if (valid) KR_HIP(hipMemcpyAsync(valid, r->valid, sizeof(uint32_t) * r->npix, hipMemcpyDeviceToHost, st));
    KR_HIP(hipStreamSynchronize(st));
    resolve_events(r);
    return KR_OK;
}

static_assert(sizeof(kr_slice_record) == kProfFloats * sizeof(float), "slice record layout");

KR_EXPORT int kr_render_set_profiler(kr_render *r, const kr_profiler_config *cfg) {
    if (!r) return fail(KR_E_INVALID, "NULL render");
    if (cfg && r->nrc) return fail(KR_E_UNSUPPORTED, "a render with a radiance cache takes no profiler");
    (void)hipSetDevice(r->scene->dev->ordinal);
    KR_HIP(hipStreamSynchronize(r->scene->dev->stream));
    for (void *p : r->prof_allocs) dfree(r->scene->dev, p);
    r->prof_allocs.clear();
    r->pv = ProfView{};
    r->prof_slices = 0;
    if (!cfg) return KR_OK;
    // Profiler::Profiler (kernel.cu:34-35): area = max(0, min(area, (w, h, w, h)))
    uint32_t w = r->prm.width, h = r->prm.height;
    uint32_t x0 = std::min(cfg->area[0], w), y0 = std::min(cfg->area[1], h);
    uint32_t x1 = std::min(cfg->area[2], w), y1 = std::min(cfg->area[3], h);
    if (x1 < x0 || y1 < y0) return fail(KR_E_INVALID, "profiler area has x1 < x0 or y1 < y0");
    uint64_t slices = uint64_t(x1 - x0) * (y1 - y0) * r->prm.spp;
    ProfView pv{};
    pv.enabled = 1;
    pv.x0 = x0;
    pv.y0 = y0;
    pv.x1 = x1;
    pv.y1 = y1;
    pv.w = w;
    pv.npix = r->npix;
    pv.spp = r->prm.spp;
    pv.depth = cfg->depth;
    size_t nrec = std::max<size_t>(1, size_t(slices) * cfg->depth * kProfFloats);
    size_t nfin = std::max<size_t>(1, size_t(slices) * 4);
    int rc = dalloc(r->scene->dev, &pv.rec, nrec);
    if (!rc) r->prof_allocs.push_back(pv.rec);
    if (!rc) rc = dalloc(r->scene->dev, &pv.fin, nfin);
    if (!rc) r->prof_allocs.push_back(pv.fin);
    if (rc) {
        for (void *p : r->prof_allocs) dfree(r->scene->dev, p);
        r->prof_allocs.clear();
        return rc;
    }
    hipStream_t st = r->scene->dev->stream;
    KR_HIP(hipMemsetAsync(pv.rec, 0, sizeof(float) * nrec, st));
    KR_HIP(hipMemsetAsync(pv.fin, 0, sizeof(float) * nfin, st));
    KR_HIP(hipStreamSynchronize(st));
    r->pv = pv;
    r->prof_slices = slices;
    return KR_OK;
}

KR_EXPORT int kr_render_profiler_size(kr_render *r, uint64_t *num_slices, uint32_t *depth) {
    if (!r) return fail(KR_E_INVALID, "NULL render");
    if (num_slices) *num_slices = r->prof_slices;
    if (depth) *depth = r->pv.enabled ? r->pv.depth : 0;
    return KR_OK;
}

KR_EXPORT int kr_render_read_slices(kr_render *r, uint64_t first, uint64_t count, float *L, uint32_t *bounce,
                                    kr_slice_record *recs) {
    if (!r) return fail(KR_E_INVALID, "NULL render");
    if (!r->pv.enabled) return fail(KR_E_INVALID, "render has no profiler");
    if (first > r->prof_slices || count > r->prof_slices - first)
        return fail(KR_E_INVALID, "slice range exceeds the profiler's slices");
    if (count == 0) return KR_OK;
    hipStream_t st = r->scene->dev->stream;
    if (L || bounce) {
        std::vector<float> fin(4 * size_t(count));
        KR_HIP(hipMemcpyAsync(fin.data(), r->pv.fin + 4 * size_t(first), sizeof(float) * fin.size(),
                              hipMemcpyDeviceToHost, st));
        KR_HIP(hipStreamSynchronize(st));
        for (uint64_t i = 0; i < count; ++i) {
            if (L)
                for (int c = 0; c < 3; ++c) L[3 * i + c] = fin[4 * i + c];
            if (bounce) std::memcpy(&bounce[i], &fin[4 * i + 3], 4);
        }
    }
    if (recs && r->pv.depth) {
        size_t per = size_t(r->pv.depth) * kProfFloats;
        KR_HIP(hipMemcpyAsync(recs, r->pv.rec + per * first, sizeof(float) * per * count, hipMemcpyDeviceToHost, st));
    }
    KR_HIP(hipStreamSynchronize(st));
    resolve_events(r);
    return KR_OK;
}

KR_EXPORT int kr_render_get_stats(kr_render *r, kr_render_stats *out) {
    if (!r || !out) return fail(KR_E_INVALID, "NULL argument");
    unsigned long long h[S_NUM];
    KR_HIP(hipMemcpyAsync(h, r->stats, sizeof(h), hipMemcpyDeviceToHost, r->scene->dev->stream));
    KR_HIP(hipStreamSynchronize(r->scene->dev->stream));
    resolve_events(r);
    std::memset(out, 0, sizeof(*out));
    out->closest_rays = h[S_CLOSEST];
    out->shadow_rays = h[S_SHADOW];
    out->invalid_samples = h[S_INVALID];
    out->node_visits = h[S_NODES];
    out->tri_tests = h[S_TRIS];
    out->cache_queries = h[S_NRCQ];
    out->paths = r->paths;
    out->seconds = r->last_seconds;
    out->trace_ms = r->kind_ms[K_TRACE];
    out->shade_ms = r->kind_ms[K_SHADE];
    out->shadow_ms = r->kind_ms[K_SHADOW];
    out->other_ms = r->kind_ms[K_OTHER];
    out->trace_launches = r->kind_launches[K_TRACE];
    out->shade_launches = r->kind_launches[K_SHADE];
    out->shadow_launches = r->kind_launches[K_SHADOW];
    out->trace_busy_ms = r->kind_busy_ms[K_TRACE];
    out->shade_busy_ms = r->kind_busy_ms[K_SHADE];
    out->shadow_busy_ms = r->kind_busy_ms[K_SHADOW];
    return KR_OK;
}

KR_EXPORT int kr_render_set_profiling(kr_render *r, uint32_t flags) {
    if (!r) return fail(KR_E_INVALID, "NULL render");
    r->profile = flags;
    return KR_OK;
}

KR_EXPORT int kr_render_radiance(kr_scene *s, const kr_camera *cam, const kr_params *prm, double *rgb, uint32_t *valid) {
    kr_render *r = nullptr;
    int rc = kr_render_create(s, cam, prm, &r);
    if (rc) return rc;
    rc = kr_render_run(r, 0, prm->spp);
    if (!rc) rc = kr_render_sync(r);
    if (!rc) rc = kr_render_read(r, rgb, valid);
    kr_render_free(r);
    return rc;
}


// ------------------------------------------------------- radiance cache --

namespace {
nrc::Weights nrc_view(const kr_nrc *n) {
    nrc::Weights W;
    W.wt = n->wt;
    W.w = n->w;
    W.master = n->master;
    for (int c = 0; c < 3; ++c) {
        W.box_min[c] = n->cfg.box_min[c];
        float ext = n->cfg.box_max[c] - n->cfg.box_min[c];
        W.box_scale[c] = ext > 0.f ? 1.f / ext : 1.f;
    }
    return W;
}

// The weights renders query (k_nrc_infer).
nrc::Weights nrc_render_view(const kr_nrc *n) {
    nrc::Weights W = nrc_view(n);
    if (n->ema) {
        W.wt = n->wt_ema;
        W.master = n->ema;
    }
    return W;
}

// BF16 copies of the master weights, made on the host (set_params / create).
void nrc_host_bf16(const float *master, std::vector<uint16_t> &w, std::vector<uint16_t> &wt) {
    using namespace nrc;
    w.assign(kNumWeights, 0);
    wt.assign(kNumWeights, 0);
    for (int l = 0; l < kLayers; ++l) {
        int in = kDims[l], out = kDims[l + 1];
        for (int i = 0; i < in; ++i)
            for (int o = 0; o < out; ++o) {
                float v = master[w_off(l) + i * out + o];
                uint32_t u;
                std::memcpy(&u, &v, 4);
                u += 0x7FFFu + ((u >> 16) & 1u);
                uint16_t h = uint16_t(u >> 16);
                w[bw_off(l) + i * out + o] = h;
                wt[bw_off(l) + o * in + i] = h;
            }
    }
}

int nrc_upload_params(kr_nrc *n, const float *host) {
    std::vector<uint16_t> w, wt;
    nrc_host_bf16(host, w, wt);
    hipStream_t st = n->dev->stream;
    KR_HIP(hipMemcpyAsync(n->master, host, sizeof(float) * nrc::kNumParams, hipMemcpyHostToDevice, st));
    KR_HIP(hipMemcpyAsync(n->w, w.data(), 2 * w.size(), hipMemcpyHostToDevice, st));
    KR_HIP(hipMemcpyAsync(n->wt, wt.data(), 2 * wt.size(), hipMemcpyHostToDevice, st));
    if (n->ema) {
        KR_HIP(hipMemcpyAsync(n->ema, host, sizeof(float) * nrc::kNumParams, hipMemcpyHostToDevice, st));
        KR_HIP(hipMemcpyAsync(n->wt_ema, wt.data(), 2 * wt.size(), hipMemcpyHostToDevice, st));
    }
    KR_HIP(hipMemsetAsync(n->adam_m, 0, sizeof(float) * nrc::kNumParams, st));
    KR_HIP(hipMemsetAsync(n->adam_v, 0, sizeof(float) * nrc::kNumParams, st));
    KR_HIP(hipStreamSynchronize(st));
    n->step = 0;
    return KR_OK;
}

// One training step on a source; the caller syncs.
int nrc_step(kr_nrc *n, const TrainSource &src, uint32_t batch, uint32_t rows) {
    hipStream_t st = n->dev->stream;
    KR_HIP(hipMemsetAsync(n->loss, 0, sizeof(float), st));
    hipLaunchKernelGGL(k_nrc_fwd_bwd, dim3(rows), dim3(64), 0, st, nrc_view(n), src, batch, n->partial, n->loss);
    n->step += 1;
    const float b1 = 0.9f, b2 = 0.99f, eps = 1e-8f;
    float c1 = float(1.0 - std::pow(double(b1), double(n->step)));
    float c2 = float(1.0 - std::pow(double(b2), double(n->step)));
    float lr = n->cfg.learning_rate > 0.f ? n->cfg.learning_rate : 1e-2f;
    const double d = n->cfg.ema_decay;
    float ema_a = n->ema ? float((1.0 - d) / (1.0 - std::pow(d, double(n->step)))) : 0.f;
    hipLaunchKernelGGL(k_nrc_adam, dim3((nrc::kNumParams + kAdamParams - 1) / kAdamParams),
                       dim3(kAdamParams * kAdamGroups), 0, st, n->master, n->wt, n->w, n->adam_m, n->adam_v,
                       n->partial, rows, lr, b1, b2, eps, c1, c2, src.pool_total, n->ema, n->wt_ema, ema_a);
    KR_HIP(hipGetLastError());
    return KR_OK;
}

template <class T>
int nrc_alloc(kr_nrc *n, T **p, size_t cnt) {
    int rc = dalloc(n->dev, p, cnt);
    if (!rc) n->allocs.push_back(*p);
    return rc;
}
}  // namespace

KR_EXPORT int kr_nrc_free(kr_nrc *n) {
    if (!n) return KR_OK;
    (void)hipSetDevice(n->dev->ordinal);
    (void)hipStreamSynchronize(n->dev->stream);
    for (void *p : n->allocs) dfree(n->dev, p);
    delete n;
    return KR_OK;
}

KR_EXPORT int kr_nrc_create(kr_device *dev, const kr_nrc_config *cfg, kr_nrc **out) {
    using namespace nrc;
    if (!dev || !cfg || !out) return fail(KR_E_INVALID, "NULL argument");
    *out = nullptr;
    (void)hipSetDevice(dev->ordinal);
    auto *n = new kr_nrc();
    n->dev = dev;
    n->cfg = *cfg;
    if (!n->cfg.batch) n->cfg.batch = 16384;
    if (!n->cfg.steps_per_frame) n->cfg.steps_per_frame = 4;
    if (!n->cfg.train_log2)
        n->cfg.train_log2 = 4;
    else if (n->cfg.train_log2 == KR_NRC_TRAIN_ALL)
        n->cfg.train_log2 = 0;
    else if (n->cfg.train_log2 > 31) {
        delete n;
        return fail(KR_E_INVALID, "train_log2 must be at most 31 (or KR_NRC_TRAIN_ALL)");
    }
    if (n->cfg.rough_threshold <= 0.f) n->cfg.rough_threshold = 0.3f;
    if (n->cfg.ema_decay == 0.f) n->cfg.ema_decay = 0.99f;
    if (!(n->cfg.ema_decay < 1.f)) {
        delete n;
        return fail(KR_E_INVALID, "ema_decay must be below 1 (0 for the default, negative for none)");
    }
    n->cfg.batch = (n->cfg.batch + 63) / 64 * 64;
    n->rows = n->cfg.batch / 64;
    n->pool_cap = 1u << 18;
    int rc = KR_OK;
    if (!rc) rc = nrc_alloc(n, &n->master, kNumParams);
    if (!rc) rc = nrc_alloc(n, &n->adam_m, kNumParams);
    if (!rc) rc = nrc_alloc(n, &n->adam_v, kNumParams);
    if (!rc) rc = nrc_alloc(n, &n->wt, kNumWeights);
    if (!rc) rc = nrc_alloc(n, &n->w, kNumWeights);
    if (!rc && n->cfg.ema_decay > 0.f) rc = nrc_alloc(n, &n->ema, kNumParams);
    if (!rc && n->cfg.ema_decay > 0.f) rc = nrc_alloc(n, &n->wt_ema, kNumWeights);
    if (!rc) rc = nrc_alloc(n, &n->partial, size_t(n->rows) * kNumParams);
    if (!rc) rc = nrc_alloc(n, &n->loss, 1);
    if (!rc) rc = nrc_alloc(n, &n->pool_raw, size_t(n->pool_cap) * kRaw);
    if (!rc) rc = nrc_alloc(n, &n->pool_tgt, size_t(n->pool_cap) * 4);
    if (!rc) rc = nrc_alloc(n, &n->pool_total, 1);
    if (rc) {
        kr_nrc_free(n);
        return rc;
    }
    // Glorot-uniform weights from a SplitMix64 stream, zero biases
    std::vector<float> host(kNumParams, 0.f);
    uint64_t state = 0x9E3779B97F4A7C15ull * (uint64_t(cfg->seed) + 1);
    auto next = [&]() {
        state += 0x9E3779B97F4A7C15ull;
        uint64_t z = state;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return double(z >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int l = 0; l < kLayers; ++l) {
        double lim = std::sqrt(6.0 / (kDims[l] + kDims[l + 1]));
        for (int e = 0; e < kDims[l] * kDims[l + 1]; ++e) host[w_off(l) + e] = float((2 * next() - 1) * lim);
    }
    rc = nrc_upload_params(n, host.data());
    if (!rc) {
        hipError_t e = hipMemsetAsync(n->pool_total, 0, sizeof(uint32_t), dev->stream);
        if (e != hipSuccess) rc = fail(KR_E_HIP, "hipMemsetAsync");
    }
    if (rc) {
        kr_nrc_free(n);
        return rc;
    }
    *out = n;
    return KR_OK;
}

KR_EXPORT int kr_nrc_get_params(kr_nrc *n, float *params) {
    if (!n || !params) return fail(KR_E_INVALID, "NULL argument");
    KR_HIP(hipMemcpyAsync(params, n->master, sizeof(float) * nrc::kNumParams, hipMemcpyDeviceToHost, n->dev->stream));
    KR_HIP(hipStreamSynchronize(n->dev->stream));
    return KR_OK;
}

KR_EXPORT int kr_nrc_get_render_params(kr_nrc *n, float *params) {
    if (!n || !params) return fail(KR_E_INVALID, "NULL argument");
    KR_HIP(hipMemcpyAsync(params, n->ema ? n->ema : n->master, sizeof(float) * nrc::kNumParams,
                          hipMemcpyDeviceToHost, n->dev->stream));
    KR_HIP(hipStreamSynchronize(n->dev->stream));
    return KR_OK;
}

KR_EXPORT int kr_nrc_set_params(kr_nrc *n, const float *params) {
    if (!n || !params) return fail(KR_E_INVALID, "NULL argument");
    return nrc_upload_params(n, params);
}

KR_EXPORT int kr_nrc_encode(kr_nrc *n, const float *raw, uint32_t cnt, float *features) {
    if (!n || !raw || !features) return fail(KR_E_INVALID, "NULL argument");
    if (!cnt) return KR_OK;
    DevBufs b(n->dev);
    float *r, *f;
    int rc;
    if ((rc = b.get(&r, size_t(cnt) * nrc::kRaw)) || (rc = b.get(&f, size_t(cnt) * nrc::kIn))) return rc;
    hipStream_t st = n->dev->stream;
    KR_HIP(hipMemcpyAsync(r, raw, sizeof(float) * cnt * nrc::kRaw, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_nrc_encode, dim3((cnt + 255) / 256), dim3(256), 0, st, nrc_view(n), r, cnt, f);
    KR_HIP(hipGetLastError());
    KR_HIP(hipMemcpyAsync(features, f, sizeof(float) * cnt * nrc::kIn, hipMemcpyDeviceToHost, st));
    KR_HIP(hipStreamSynchronize(st));
    return KR_OK;
}

KR_EXPORT int kr_nrc_predict(kr_nrc *n, const float *raw, uint32_t cnt, float *rgb) {
    if (!n || !raw || !rgb) return fail(KR_E_INVALID, "NULL argument");
    if (!cnt) return KR_OK;
    DevBufs b(n->dev);
    float *r, *o;
    int rc;
    if ((rc = b.get(&r, size_t(cnt) * nrc::kRaw)) || (rc = b.get(&o, size_t(cnt) * 3))) return rc;
    hipStream_t st = n->dev->stream;
    KR_HIP(hipMemcpyAsync(r, raw, sizeof(float) * cnt * nrc::kRaw, hipMemcpyHostToDevice, st));
    uint32_t waves = (cnt + 63) / 64;
    hipLaunchKernelGGL(k_nrc_predict, dim3((waves + 3) / 4), dim3(256), 0, st, nrc_view(n), r, cnt, o);
    KR_HIP(hipGetLastError());
    KR_HIP(hipMemcpyAsync(rgb, o, sizeof(float) * cnt * 3, hipMemcpyDeviceToHost, st));
    KR_HIP(hipStreamSynchronize(st));
    return KR_OK;
}

namespace {
int nrc_explicit_source(kr_nrc *n, DevBufs &b, const float *raw, const float *target, uint32_t cnt, TrainSource &src,
                        float **partial, uint32_t *rows) {
    float *r, *t;
    int rc;
    if ((rc = b.get(&r, size_t(cnt) * nrc::kRaw)) || (rc = b.get(&t, size_t(cnt) * 4))) return rc;
    std::vector<float> t4(size_t(cnt) * 4, 0.f);
    for (uint32_t i = 0; i < cnt; ++i)
        for (int c = 0; c < 3; ++c) t4[4 * size_t(i) + c] = target[3 * size_t(i) + c];
    hipStream_t st = n->dev->stream;
    KR_HIP(hipMemcpyAsync(r, raw, sizeof(float) * cnt * nrc::kRaw, hipMemcpyHostToDevice, st));
    KR_HIP(hipMemcpyAsync(t, t4.data(), sizeof(float) * t4.size(), hipMemcpyHostToDevice, st));
    KR_HIP(hipStreamSynchronize(st));
    src.raw = r;
    src.target = t;
    src.pool_total = nullptr;
    src.n = cnt;
    src.cap = cnt;
    src.step = 0;
    src.steps = 1;
    *rows = (cnt + 63) / 64;
    if ((rc = b.get(partial, size_t(*rows) * nrc::kNumParams))) return rc;
    return KR_OK;
}
}  // namespace

KR_EXPORT int kr_nrc_gradient(kr_nrc *n, const float *raw, const float *target, uint32_t cnt, float *grad,
                              float *loss) {
    if (!n || !raw || !target || !grad || !cnt) return fail(KR_E_INVALID, "bad argument");
    DevBufs b(n->dev);
    TrainSource src;
    float *part;
    uint32_t rows;
    int rc = nrc_explicit_source(n, b, raw, target, cnt, src, &part, &rows);
    if (rc) return rc;
    float *g;
    if ((rc = b.get(&g, nrc::kNumParams))) return rc;
    hipStream_t st = n->dev->stream;
    KR_HIP(hipMemsetAsync(n->loss, 0, sizeof(float), st));
    hipLaunchKernelGGL(k_nrc_fwd_bwd, dim3(rows), dim3(64), 0, st, nrc_view(n), src, cnt, part, n->loss);
    hipLaunchKernelGGL(k_nrc_sum_rows, dim3(96), dim3(256), 0, st, part, rows, g);
    KR_HIP(hipGetLastError());
    KR_HIP(hipMemcpyAsync(grad, g, sizeof(float) * nrc::kNumParams, hipMemcpyDeviceToHost, st));
    if (loss) KR_HIP(hipMemcpyAsync(loss, n->loss, sizeof(float), hipMemcpyDeviceToHost, st));
    KR_HIP(hipStreamSynchronize(st));
    return KR_OK;
}

KR_EXPORT int kr_nrc_train(kr_nrc *n, const float *raw, const float *target, uint32_t cnt, uint32_t steps, float *loss) {
    if (!n || !raw || !target || !cnt) return fail(KR_E_INVALID, "bad argument");
    DevBufs b(n->dev);
    TrainSource src;
    float *part;
    uint32_t rows;
    int rc = nrc_explicit_source(n, b, raw, target, cnt, src, &part, &rows);
    if (rc) return rc;
    float *saved = n->partial;
    n->partial = part;
    for (uint32_t s = 0; s < steps && !rc; ++s) rc = nrc_step(n, src, cnt, rows);
    n->partial = saved;
    if (rc) return rc;
    float l = 0.f;
    KR_HIP(hipMemcpyAsync(&l, n->loss, sizeof(float), hipMemcpyDeviceToHost, n->dev->stream));
    KR_HIP(hipStreamSynchronize(n->dev->stream));
    n->samples += uint64_t(cnt) * steps;
    n->last_loss = l;
    if (loss) *loss = l;
    return KR_OK;
}

KR_EXPORT int kr_nrc_read_pool(kr_nrc *n, uint32_t max, float *raw, float *target, uint32_t *count) {
    if (!n || !count || (max && (!raw || !target))) return fail(KR_E_INVALID, "bad argument");
    hipStream_t st = n->dev->stream;
    uint32_t total = 0;
    KR_HIP(hipMemcpyAsync(&total, n->pool_total, 4, hipMemcpyDeviceToHost, st));
    KR_HIP(hipStreamSynchronize(st));
    uint32_t avail = std::min(total, n->pool_cap);
    uint32_t cnt = std::min(avail, max);
    if (cnt) {
        std::vector<float> t4(size_t(cnt) * 4);
        KR_HIP(hipMemcpyAsync(raw, n->pool_raw, sizeof(float) * nrc::kRaw * cnt, hipMemcpyDeviceToHost, st));
        KR_HIP(hipMemcpyAsync(t4.data(), n->pool_tgt, sizeof(float) * 4 * cnt, hipMemcpyDeviceToHost, st));
        KR_HIP(hipStreamSynchronize(st));
        for (uint32_t i = 0; i < cnt; ++i)
            for (int c = 0; c < 3; ++c) target[3 * size_t(i) + c] = t4[4 * size_t(i) + c];
    }
    *count = cnt;
    return KR_OK;
}

KR_EXPORT int kr_nrc_get_stats(kr_nrc *n, uint64_t *samples_trained, float *last_loss) {
    if (!n) return fail(KR_E_INVALID, "NULL argument");
    if (samples_trained) *samples_trained = n->samples;
    if (last_loss) {
        // the loss of the last step, whether kr_nrc_train or a render ran it
        KR_HIP(hipMemcpyAsync(&n->last_loss, n->loss, sizeof(float), hipMemcpyDeviceToHost, n->dev->stream));
        KR_HIP(hipStreamSynchronize(n->dev->stream));
        *last_loss = n->last_loss;
    }
    return KR_OK;
}

// Cache queries and training-sample extraction after a chunk's bounces.
// Left alone, the compiler keeps all 48 KB of BF16 weights in registers
// across the query loop (316 VGPRs, one wave per SIMD); asking for two waves
// per SIMD makes it reload them from L1/L2 instead and hides the latency of
// the records and of L: 6 % faster per cached frame on the closed room.
#ifndef KR_NRC_INFER_WAVES
#define KR_NRC_INFER_WAVES 2
#endif
__global__ __launch_bounds__(256, KR_NRC_INFER_WAVES) void k_nrc_infer(nrc::Weights W, NrcRender nr, const uint32_t *counters, float *L,
                                                                         unsigned long long *stats) {
    using namespace nrc;
    // The block's copy of the weights and biases: every layer's fragments
    // are read from LDS instead of from L1/L2 per 64 queries.
    __shared__ __align__(16) uint16_t s_wt[kNumWeights];
    __shared__ float s_bias[kNumBiases];
    static_assert(kNumWeights % 8 == 0, "weights copied in 16-byte pieces");
    for (int i = threadIdx.x; i < kNumWeights / 8; i += blockDim.x)
        reinterpret_cast<uint4 *>(s_wt)[i] = reinterpret_cast<const uint4 *>(W.wt)[i];
    for (int l = 0; l < kLayers; ++l)
        for (int i = threadIdx.x; i < kDims[l + 1]; i += blockDim.x) s_bias[bb_off(l) + i] = W.master[b_off(l) + i];
    __syncthreads();
    int lane = __lane_id(), s = lane & 15, q = lane >> 4;
    uint32_t n = counters[CTR(C_NRCQ)];
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&stats[S_NRCQ], (unsigned long long)n);
    uint32_t waves = gridDim.x * (blockDim.x / 64);
    for (uint32_t base = (blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64) * 64; base < n; base += waves * 64) {
        v4s in[6][4];
        encode_inputs(in, [&](int k) -> const float * {
            return base + k < n ? nr.qraw + size_t(base + k) * kRaw : nullptr;
        }, W);
        v4f out[4];
        forward_from<true>(in, out, s_wt, s_bias);
        if (q == 0)
            for (int nt = 0; nt < 4; ++nt) {
                uint32_t i = base + nt * 16 + s;
                if (i < n) {
                    uint32_t slot = nr.qslot[i];
                    // a radiance prediction below zero is clamped to zero
                    L[3 * size_t(slot)] += nr.qbeta[3 * size_t(i)] * fmaxf(out[nt][0], 0.f);
                    L[3 * size_t(slot) + 1] += nr.qbeta[3 * size_t(i) + 1] * fmaxf(out[nt][1], 0.f);
                    L[3 * size_t(slot) + 2] += nr.qbeta[3 * size_t(i) + 2] * fmaxf(out[nt][2], 0.f);
                }
            }
    }
}

// Training targets of the chunk's training paths: for vertex j the radiance
// the rest of the path carried back to it, (sum_{k>=j} c_k - beta_j E_j) / beta_j,
// where c_k is everything vertex k added to the path's L; with nr.indirect
// what the vertices after j carried back, (sum_{k>j} c_k) / beta_j.
__global__ __launch_bounds__(256) void k_nrc_extract(NrcRender nr, const uint32_t *counters, float *pool_raw,
                                                     float *pool_tgt, uint32_t *pool_total, uint32_t cap,
                                                     float clamp) {
    uint32_t ntrain = min(counters[CTR(C_TRAIN)], nr.train_cap);
    uint32_t stride = gridDim.x * blockDim.x;
    // whole waves iterate together so that the ring slots can be reserved
    // with one atomic per wave
    for (uint32_t base = blockIdx.x * blockDim.x; base < ntrain; base += stride) {
        uint32_t ts = base + threadIdx.x;
        uint32_t nv = ts < ntrain ? nr.tcount[ts] : 0u;
        // pass 1: how many of this path's vertices give a usable target
        uint32_t keep = 0;
        {
            float sx = 0.f, sy = 0.f, sz = 0.f;
            for (int j = int(nv) - 1; j >= 0; --j) {
                const float *rec = nr.rec + (size_t(ts) * nr.nv + uint32_t(j)) * kRecFloats;
                sx += rec[19];
                sy += rec[20];
                sz += rec[21];
                float bx = rec[16], by = rec[17], bz = rec[18];
                if (!(bx > 0.f && by > 0.f && bz > 0.f)) continue;
                const int own = nr.indirect ? 19 : 22;
                float tx = (sx - rec[own]) / bx, ty = (sy - rec[own + 1]) / by, tz = (sz - rec[own + 2]) / bz;
                if (isfinite(tx) && isfinite(ty) && isfinite(tz)) ++keep;
            }
        }
        uint32_t slot = wave_reserve(pool_total, keep);
        // pass 2: write them
        float sx = 0.f, sy = 0.f, sz = 0.f;
        for (int j = int(nv) - 1; j >= 0; --j) {
            const float *rec = nr.rec + (size_t(ts) * nr.nv + uint32_t(j)) * kRecFloats;
            sx += rec[19];
            sy += rec[20];
            sz += rec[21];
            float bx = rec[16], by = rec[17], bz = rec[18];
            if (!(bx > 0.f && by > 0.f && bz > 0.f)) continue;
            const int own = nr.indirect ? 19 : 22;
            float tx = (sx - rec[own]) / bx, ty = (sy - rec[own + 1]) / by, tz = (sz - rec[own + 2]) / bz;
            if (!(isfinite(tx) && isfinite(ty) && isfinite(tz))) continue;
            if (clamp > 0.f) {
                tx = fminf(tx, clamp);
                ty = fminf(ty, clamp);
                tz = fminf(tz, clamp);
            }
            uint32_t r = slot++ % cap;
            float4 *dst = reinterpret_cast<float4 *>(pool_raw + size_t(r) * nrc::kRaw);
            dst[0] = make_float4(rec[0], rec[1], rec[2], rec[3]);
            dst[1] = make_float4(rec[4], rec[5], rec[6], rec[7]);
            dst[2] = make_float4(rec[8], rec[9], rec[10], rec[11]);
            dst[3] = make_float4(rec[12], 0.f, 0.f, 0.f);
            reinterpret_cast<float4 *>(pool_tgt)[r] = make_float4(tx, ty, tz, 0.f);
        }
    }
}

namespace {
void nrc_after_chunk(kr_render *r, Pipe &pp, kr_nrc *n, const NrcRender &nr) {
    hipStream_t st = pp.stream;
    {
        Timed tm(r, K_OTHER, st);
        // three resident blocks per CU (48 KB of LDS weights each, 158 VGPRs), looping
        int grid = int(std::min<size_t>((size_t(r->P) + 255) / 256, size_t(r->scene->dev->num_cu) * 3));
        hipLaunchKernelGGL(k_nrc_infer, dim3(grid), dim3(256), 0, st, nrc_render_view(n), nr, pp.counters, pp.L,
                           r->stats);
        int egrid = int(std::min<size_t>((size_t(nr.train_cap) + 255) / 256, size_t(r->scene->dev->num_cu) * 8));
        hipLaunchKernelGGL(k_nrc_extract, dim3(egrid), dim3(256), 0, st, nr, pp.counters, n->pool_raw, n->pool_tgt,
                           n->pool_total, n->pool_cap, n->cfg.target_clamp);
    }
}

int nrc_after_run(kr_render *r, kr_nrc *n) {
    if (n->cfg.flags & KR_NRC_FROZEN) return KR_OK;
    Timed tm(r, K_OTHER);
    for (uint32_t i = 0; i < n->cfg.steps_per_frame; ++i) {
        TrainSource src;
        src.raw = n->pool_raw;
        src.target = n->pool_tgt;
        src.pool_total = n->pool_total;
        src.n = 0;
        src.cap = n->pool_cap;
        src.step = i;
        src.steps = n->cfg.steps_per_frame;
        int rc = nrc_step(n, src, n->cfg.batch, n->rows);
        if (rc) return rc;
        n->samples += n->cfg.batch;
    }
    return KR_OK;
}
}  // namespace

KR_EXPORT int kr_render_attach_nrc(kr_render *r, kr_nrc *n) {
    if (!r) return fail(KR_E_INVALID, "NULL render");
    if (n && r->pv.enabled) return fail(KR_E_UNSUPPORTED, "a render with a profiler takes no radiance cache");
    (void)hipSetDevice(r->scene->dev->ordinal);
    (void)hipStreamSynchronize(r->scene->dev->stream);
    for (int pi = 0; pi < r->npipes; ++pi) (void)hipStreamSynchronize(r->pipe[pi].stream);
    for (void *p : r->nrc_allocs) dfree(r->scene->dev, p);
    r->nrc_allocs.clear();
    r->nr = NrcRender{};
    for (Pipe &pp : r->pipe) pp.nr = NrcRender{};
    r->nrc = nullptr;
    if (!n) return KR_OK;
    if (n->dev != r->scene->dev) return fail(KR_E_INVALID, "cache and render live on different devices");
    NrcRender &nr = r->nr;
    uint32_t log2 = n->cfg.train_log2;
    nr.train_log2 = log2;               // <= 31 (kr_nrc_create)
    nr.train_mask = (1u << log2) - 1u;
    nr.rough_threshold = n->cfg.rough_threshold;
    nr.indirect = (n->cfg.flags & KR_NRC_INDIRECT) ? 1 : 0;
    nr.end_at = nr.indirect ? 1u : 2u;
    nr.nv = r->prm.max_bounce + 1;
    // groups a chunk of P paths can touch: P / 2^log2 + 2 partial ones
    nr.train_cap = uint32_t(std::min<uint64_t>(r->P, (uint64_t(r->P) >> log2) + 2));
    auto get = [&](auto **p, size_t cnt) {
        int rc = dalloc(r->scene->dev, p, cnt);
        if (!rc) r->nrc_allocs.push_back(*p);
        return rc;
    };
    int rc = KR_OK;
    nr.enabled = 1;
    // each pipe records and queries through buffers of its own
    for (int pi = 0; pi < r->npipes && !rc; ++pi) {
        NrcRender &pn = r->pipe[pi].nr;
        pn = nr;
        if (!rc) rc = get(&pn.tslot_of, r->P);
        if (!rc) rc = get(&pn.rec, size_t(nr.train_cap) * nr.nv * kRecFloats);
        if (!rc) rc = get(&pn.tcount, nr.train_cap);
        if (!rc) rc = get(&pn.qslot, r->P);
        if (!rc) rc = get(&pn.qbeta, 3 * size_t(r->P));
        if (!rc) rc = get(&pn.qraw, size_t(r->P) * nrc::kRaw);
    }
    if (rc) {
        for (void *p : r->nrc_allocs) dfree(r->scene->dev, p);
        r->nrc_allocs.clear();
        r->nr = NrcRender{};
        for (Pipe &pp : r->pipe) pp.nr = NrcRender{};
        return rc;
    }
    r->nrc = n;
    return KR_OK;
}

// ------------------------------------------------------------- batches --
namespace {
// Owned copy of a scene description (the batch rebuilds BVHs at other
// widths while autotuning).
struct SceneCopy {
    std::vector<float> pos, nrm;
    std::vector<uint32_t> idx;
    std::vector<kr_prim_desc> prims;
    std::vector<kr_shader_desc> shaders;
    kr_scene_desc desc{};
};
}  // namespace

struct kr_batch {
    kr_device *dev = nullptr;
    std::vector<SceneCopy> scenes;
    std::vector<kr_scene *> up;  // uploaded at tuning.bvh_width
    std::vector<kr_batch_job> jobs;
    std::vector<HostJob> hjobs;  // full sample ranges
    std::vector<uint32_t> acc_off;
    uint64_t total_paths = 0, total_pix = 0;
    uint32_t max_bounce = 0;
    kr_batch_tuning tuning{};
    kr_render *r = nullptr;       // queues, counters and the batch's accumulators
    kr_scene anchor;              // r->scene: device and stream only (the jobs' scenes come from the table)
    JobRec *d_rec = nullptr;      // full job table
    uint64_t *d_begin = nullptr;
    JobRec *d_trec = nullptr;     // trial job table (autotuner)
    uint64_t *d_tbegin = nullptr;
    std::vector<void *> allocs;
    hipGraph_t graph[kMaxPipes] = {};  // one per pipe: its chunks, replayed on its stream
    hipGraphExec_t exec[kMaxPipes] = {};
    uint32_t graph_nodes = 0, chunks = 0;
    double capture_seconds = 0, seconds = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool timed = false;
    // radiance caches of the jobs (kr_batch_attach_nrc): per job a cache or
    // null; per pipe one set of cache buffers, sized for the batch's chunks
    // and shared by the pipe's cached chunks (a pipe's chunks run in order)
    std::vector<kr_nrc *> nrc;
    NrcRender pnr[kMaxPipes] = {};
    std::vector<void *> nrc_allocs;
};

namespace {
constexpr uint32_t kMaxBlocksPerCu = 64;
constexpr uint32_t kBatchMaxChunk = 32u << 20;

void batch_drop_graph(kr_batch *b) {
    for (int pi = 0; pi < kMaxPipes; ++pi) {
        if (b->exec[pi]) (void)hipGraphExecDestroy(b->exec[pi]);
        if (b->graph[pi]) (void)hipGraphDestroy(b->graph[pi]);
        b->exec[pi] = nullptr;
        b->graph[pi] = nullptr;
    }
}

// Job tables for the first spp_limit samples of every job (0: all).
// spp_limit: every job's first min(spp_limit, spp) samples (0: all); frac
// (with spp_limit 0): the first max(1, spp * frac) samples of every job, a
// scaled copy of the batch's mix of job sizes.
void batch_tables(const kr_batch *b, uint32_t spp_limit, std::vector<JobRec> &rec, std::vector<uint64_t> &begin,
                  std::vector<HostJob> &hj, double frac = 1.0) {
    rec.resize(b->jobs.size());
    begin.resize(b->jobs.size());
    hj.resize(b->jobs.size());
    uint64_t g = 0;
    for (size_t k = 0; k < b->jobs.size(); ++k) {
        const kr_batch_job &J = b->jobs[k];
        uint32_t npix = J.params.width * J.params.height;
        uint32_t spp = spp_limit ? std::min(spp_limit, J.params.spp) : J.params.spp;
        if (!spp_limit && frac < 1.0)
            spp = std::max<uint32_t>(1, std::min<uint32_t>(J.params.spp, uint32_t(double(J.params.spp) * frac + 0.5)));
        JobRec &R = rec[k];
        R.S = b->up[J.scene]->view;
        R.lens = make_lens(&J.camera, &J.params);
        R.begin = g;
        R.npix = npix;
        R.max_bounce = J.params.max_bounce;
        R.seed_lo = uint32_t(J.params.seed);
        R.seed_hi = uint32_t(J.params.seed >> 32);
        begin[k] = g;
        hj[k] = HostJob{g, uint64_t(npix) * spp, npix, b->acc_off[k]};
        g += uint64_t(npix) * spp;
    }
}

int batch_upload_tables(kr_batch *b, JobRec *d_rec, uint64_t *d_begin, uint32_t spp_limit,
                        std::vector<HostJob> &hj, double frac = 1.0) {
    std::vector<JobRec> rec;
    std::vector<uint64_t> begin;
    batch_tables(b, spp_limit, rec, begin, hj, frac);
    hipStream_t st = b->dev->stream;
    KR_HIP(hipMemcpyAsync(d_rec, rec.data(), sizeof(JobRec) * rec.size(), hipMemcpyHostToDevice, st));
    KR_HIP(hipMemcpyAsync(d_begin, begin.data(), sizeof(uint64_t) * begin.size(), hipMemcpyHostToDevice, st));
    KR_HIP(hipStreamSynchronize(st));
    return KR_OK;
}

// (Re)uploads every scene at `width` and rebuilds the full job table.
int batch_set_width(kr_batch *b, uint32_t width) {
    if (!b->up.empty() && b->up[0] && b->up[0]->width == width) return KR_OK;
    batch_drop_graph(b);
    (void)hipStreamSynchronize(b->dev->stream);
    for (kr_scene *s : b->up) kr_scene_free(s);
    b->up.assign(b->scenes.size(), nullptr);
    for (size_t i = 0; i < b->scenes.size(); ++i) {
        kr_scene_desc d = b->scenes[i].desc;
        d.bvh_width = width;
        // one node format for all scenes: the kernels are compiled per format
        const uint32_t node_fmt = KR_SCENE_COMPRESSED_NODES | KR_SCENE_FULL_NODES;
        d.flags = (d.flags & ~node_fmt) | (b->scenes[0].desc.flags & node_fmt);
        int rc = kr_scene_upload(b->dev, &d, &b->up[i]);
        if (rc) {
            // leave no half-uploaded set behind: the next call re-uploads all
            for (kr_scene *s : b->up) kr_scene_free(s);
            b->up.clear();
            return rc;
        }
    }
    b->tuning.bvh_width = width;
    return batch_upload_tables(b, b->d_rec, b->d_begin, 0, b->hjobs);
}

// The chunks of a batch's path space of `total` paths, at most P paths each:
// by default every job's range is cut on its own, so no chunk mixes jobs and
// every chunk runs the single-render kernels; with KR_BATCH_MIX_JOBS the
// space is cut into equal chunks regardless of job boundaries, and the
// chunks that straddle one run the batch kernels (each path looks up its job).
std::vector<std::pair<uint64_t, uint32_t>> batch_chunks(const kr_batch *b, const BatchView &bv, uint64_t total,
                                                        uint32_t P) {
    std::vector<std::pair<uint64_t, uint32_t>> c;
    if (b->tuning.flags & KR_BATCH_MIX_JOBS) {
        for (uint64_t g = 0; g < total; g += P) c.push_back({g, uint32_t(std::min<uint64_t>(P, total - g))});
    } else {
        // a job's chunks hold whole samples of its frame, as many as fit in P,
        // spread evenly over the fewest chunks (so that their camera rays
        // enter the queue in tile order, as a single render's)
        for (const HostJob &J : *bv.jobs) {
            uint64_t step = P;
            if (J.npix <= P && J.paths % J.npix == 0) {
                const uint64_t spp = J.paths / J.npix, fit = P / J.npix;
                const uint64_t m = (spp + fit - 1) / fit;
                step = (spp + m - 1) / m * J.npix;
            }
            for (uint64_t g = J.begin; g < J.begin + J.paths; g += step)
                c.push_back({g, uint32_t(std::min<uint64_t>(step, J.begin + J.paths - g))});
        }
    }
    return c;
}

// A cached job's view over pipe pi's cache buffers: the buffers, and the
// job's cache's path selection and ending rule (as kr_render_attach_nrc).
NrcRender batch_nrc_view(const kr_batch *b, int pi, const kr_nrc *n) {
    NrcRender nr = b->pnr[pi];
    nr.enabled = 1;
    nr.train_log2 = n->cfg.train_log2;
    nr.train_mask = (1u << n->cfg.train_log2) - 1u;
    nr.rough_threshold = n->cfg.rough_threshold;
    nr.indirect = (n->cfg.flags & KR_NRC_INDIRECT) ? 1 : 0;
    nr.end_at = nr.indirect ? 1u : 2u;
    nr.train_cap = uint32_t(std::min<uint64_t>(b->r->P, (uint64_t(b->r->P) >> n->cfg.train_log2) + 2));
    return nr;
}

// Enqueues on pipe pi its share of the batch's chunk list. A batch's pipes
// accumulate into images of their own (read back summed), so the pipes never
// wait for each other and each pipe's chunks form a graph of their own.
int batch_enqueue_pipe(kr_batch *b, const BatchView &bv, uint64_t total, uint32_t P, int pi) {
    kr_render *r = b->r;
    const std::vector<HostJob> &hj = *bv.jobs;
    Pipe &pp = r->pipe[pi];
    auto chunks = batch_chunks(b, bv, total, P);
    // each chunk in list order goes to the pipe with the fewest paths so far,
    // so that the pipes, which run concurrently, end together
    std::vector<uint64_t> load(size_t(r->npipes), 0);
    for (size_t k = 0; k < chunks.size(); ++k) {
        const int owner = int(std::min_element(load.begin(), load.end()) - load.begin());
        load[size_t(owner)] += chunks[k].second;
        if (owner != pi) continue;
        uint64_t g = chunks[k].first;
        uint32_t n = chunks[k].second;
        // a chunk inside one job runs the single-render kernels (scene views
        // as kernel arguments, fewer registers); one that mixes jobs the
        // batch kernels
        uint32_t j = 0;
        while (j + 1 < hj.size() && hj[j + 1].begin <= g) ++j;
        bool single = g >= hj[j].begin && g + n <= hj[j].begin + hj[j].paths;
        int rc = dispatch_fmt(b->tuning.bvh_width, b->up[0]->qnodes, [&](auto fm) {
            using FM = decltype(fm);
            if (!single) return run_chunk<FM::W, true, FM::Q>(r, pp, g, n, &bv, nullptr);
            const kr_batch_job &J = b->jobs[j];
            JobChunk jc{b->up[J.scene]->view, make_lens(&J.camera, &J.params), hj[j].npix, J.params.max_bounce,
                        J.params.seed, pp.acc + 3 * size_t(hj[j].acc_off), pp.valid + hj[j].acc_off};
            if (j < b->nrc.size() && b->nrc[j]) {
                jc.nrc = b->nrc[j];
                jc.nr = batch_nrc_view(b, pi, b->nrc[j]);
            }
            return run_chunk<FM::W, false, FM::Q>(r, pp, g - hj[j].begin, n, nullptr, nullptr, &jc);
        });
        if (rc) return rc;
    }
    return KR_OK;
}

// Every pipe's chunks, pipes 1.. forked from and joined back into the
// device's stream.
int batch_enqueue(kr_batch *b, const BatchView &bv, uint64_t total, uint32_t P) {
    kr_render *r = b->r;
    if (r->npipes > 1) {
        KR_HIP(hipEventRecord(r->fork, r->pipe[0].stream));
        for (int pi = 1; pi < r->npipes; ++pi) KR_HIP(hipStreamWaitEvent(r->pipe[pi].stream, r->fork, 0));
    }
    for (int pi = 0; pi < r->npipes; ++pi) {
        int rc = batch_enqueue_pipe(b, bv, total, P, pi);
        if (rc) return rc;
    }
    for (int pi = 1; pi < r->npipes; ++pi) KR_HIP(hipStreamWaitEvent(r->pipe[0].stream, r->pipe[pi].acc_done, 0));
    return KR_OK;
}

void batch_apply_launch(kr_batch *b) {
    b->r->grid = int(std::min<size_t>(size_t(std::max(b->dev->num_cu, 1)) * b->tuning.blocks_per_cu,
                                      b->r->pipe[0].spill_lanes / kBlock));
    b->r->prm.flags = b->tuning.flags & ~KR_BATCH_MIX_JOBS;
}

int batch_capture(kr_batch *b) {
    batch_drop_graph(b);
    batch_apply_launch(b);
    auto t0 = std::chrono::steady_clock::now();
    BatchView bv{JobTable{b->d_rec, b->d_begin, uint32_t(b->jobs.size())}, &b->hjobs};
    b->graph_nodes = 0;
    for (int pi = 0; pi < b->r->npipes; ++pi) {
        hipStream_t st = b->r->pipe[pi].stream;
        KR_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = batch_enqueue_pipe(b, bv, b->total_paths, b->tuning.chunk_paths, pi);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(st, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(KR_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        b->graph[pi] = g;
        size_t nodes = 0;
        KR_HIP(hipGraphGetNodes(g, nullptr, &nodes));
        b->graph_nodes += uint32_t(nodes);
        KR_HIP(hipGraphInstantiate(&b->exec[pi], g, nullptr, nullptr, 0));
    }
    b->chunks = uint32_t(batch_chunks(b, bv, b->total_paths, b->tuning.chunk_paths).size());
    b->capture_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return KR_OK;
}
}  // namespace

KR_EXPORT int kr_batch_free(kr_batch *b) {
    if (!b) return KR_OK;
    (void)hipSetDevice(b->dev->ordinal);
    (void)hipStreamSynchronize(b->dev->stream);
    batch_drop_graph(b);
    if (b->e0) (void)hipEventDestroy(b->e0);
    if (b->e1) (void)hipEventDestroy(b->e1);
    for (void *p : b->nrc_allocs) dfree(b->dev, p);
    if (b->r) kr_render_free(b->r);
    for (kr_scene *s : b->up) kr_scene_free(s);
    for (void *p : b->allocs) dfree(b->dev, p);
    delete b;
    return KR_OK;
}

KR_EXPORT int kr_batch_create(kr_device *dev, const kr_scene_desc *scenes, uint32_t num_scenes,
                              const kr_batch_job *jobs, uint32_t num_jobs, const kr_batch_tuning *tuning,
                              kr_batch **out) {
    if (!dev || !scenes || !jobs || !out) return fail(KR_E_INVALID, "NULL argument");
    *out = nullptr;
    if (!num_scenes || !num_jobs) return fail(KR_E_INVALID, "empty batch");
    kr_batch_tuning tu{4, kBatchMaxChunk, KR_BLOCKS_PER_CU, 0, 0.0};
    if (tuning) tu = *tuning;
    if (tu.bvh_width != 2 && tu.bvh_width != 4 && tu.bvh_width != 8)
        return fail(KR_E_INVALID, "bvh_width must be 2, 4 or 8");
    if (tu.blocks_per_cu < 1 || tu.blocks_per_cu > kMaxBlocksPerCu)
        return fail(KR_E_INVALID, "blocks_per_cu must be 1..64");
    if (tu.chunk_paths == 0) tu.chunk_paths = kBatchMaxChunk;
    uint64_t total = 0, pix = 0;
    uint32_t mb = 0;
    for (uint32_t k = 0; k < num_jobs; ++k) {
        const kr_batch_job &J = jobs[k];
        if (J.scene >= num_scenes) return fail(KR_E_INVALID, "job scene index out of range");
        int rc = check_view(&J.camera, &J.params);
        if (rc) return rc;
        if (J.params.kernel != KR_KERNEL_RADIANCE)
            return fail(KR_E_UNSUPPORTED, "batched jobs render the Radiance kernel only");
        total += uint64_t(J.params.width) * J.params.height * J.params.spp;
        pix += uint64_t(J.params.width) * J.params.height;
        mb = std::max(mb, J.params.max_bounce);
    }
    if (pix > 0xFFFFFFFFull) return fail(KR_E_INVALID, "batch has too many pixels");
    (void)hipSetDevice(dev->ordinal);
    auto *b = new kr_batch();
    b->dev = dev;
    b->jobs.assign(jobs, jobs + num_jobs);
    b->total_paths = total;
    b->total_pix = pix;
    b->max_bounce = mb;
    b->tuning = tu;
    b->tuning.bvh_width = 0;  // set by batch_set_width
    uint32_t off = 0;
    for (uint32_t k = 0; k < num_jobs; ++k) {
        b->acc_off.push_back(off);
        off += jobs[k].params.width * jobs[k].params.height;
    }
    b->scenes.resize(num_scenes);
    for (uint32_t i = 0; i < num_scenes; ++i) {
        const kr_scene_desc &d = scenes[i];
        SceneCopy &c = b->scenes[i];
        if (d.num_triangles && (!d.positions || !d.normals || !d.indices)) {
            delete b;
            return fail(KR_E_INVALID, "mesh arrays missing");
        }
        if (d.num_prims && !d.prims) {
            delete b;
            return fail(KR_E_INVALID, "primitive array missing");
        }
        if (d.num_shaders && !d.shaders) {
            delete b;
            return fail(KR_E_INVALID, "shader array missing");
        }
        if (d.num_triangles) {
            c.pos.assign(d.positions, d.positions + 3 * size_t(d.num_vertices));
            c.nrm.assign(d.normals, d.normals + 3 * size_t(d.num_vertices));
            c.idx.assign(d.indices, d.indices + 3 * size_t(d.num_triangles));
        }
        c.prims.assign(d.prims, d.prims + d.num_prims);
        c.shaders.assign(d.shaders, d.shaders + d.num_shaders);
        c.desc = d;
        c.desc.positions = c.pos.data();
        c.desc.normals = c.nrm.data();
        c.desc.indices = c.idx.data();
        c.desc.prims = c.prims.data();
        c.desc.shaders = c.shaders.data();
    }
    auto get = [&](auto **p, size_t cnt) {
        int rc = dalloc(b->dev, p, cnt);
        if (!rc) b->allocs.push_back(*p);
        return rc;
    };
    int rc = get(&b->d_rec, num_jobs);
    if (!rc) rc = get(&b->d_begin, num_jobs);
    if (!rc) rc = get(&b->d_trec, num_jobs);
    if (!rc) rc = get(&b->d_tbegin, num_jobs);
    if (!rc) rc = batch_set_width(b, tu.bvh_width);
    if (!rc) {
        auto *r = new kr_render();
        b->r = r;
        b->anchor.dev = dev;
        r->scene = &b->anchor;
        r->prm.max_bounce = mb;
        r->prm.flags = tu.flags;
        r->npix = uint32_t(pix);
        r->P = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(std::max(tu.chunk_paths, kBatchMaxChunk), total)));
        r->npipes = total >= kPipeMinPaths ? KR_RENDER_PIPES : 1;
        tail_config(r);
        // the autotuner may re-upload the scenes with deeper trees: room for any
        rc = render_buffers(r, size_t(std::max(dev->num_cu, 1)) * kMaxBlocksPerCu * kBlock, 0);
        // pipes 1.. accumulate into images of their own (kr_batch_read sums them)
        for (int pi = 1; pi < r->npipes && !rc; ++pi) {
            Pipe &pp = r->pipe[pi];
            rc = dalloc(b->dev, &pp.acc, 3 * size_t(r->npix));
            if (!rc) r->allocs.push_back(pp.acc);
            if (!rc) rc = dalloc(b->dev, &pp.valid, size_t(r->npix));
            if (!rc) r->allocs.push_back(pp.valid);
        }
        if (!rc) rc = kr_render_reset(r);
    }
    if (!rc && (hipEventCreate(&b->e0) != hipSuccess || hipEventCreate(&b->e1) != hipSuccess))
        rc = fail(KR_E_HIP, "hipEventCreate");
    if (rc) {
        kr_batch_free(b);
        return rc;
    }
    b->tuning.chunk_paths = std::min(tu.chunk_paths, b->r->P);
    *out = b;
    return KR_OK;
}

KR_EXPORT int kr_batch_attach_nrc(kr_batch *b, uint32_t job, kr_nrc *n) {
    if (!b) return fail(KR_E_INVALID, "NULL batch");
    if (job >= b->jobs.size()) return fail(KR_E_INVALID, "job index out of range");
    if (n && n->dev != b->dev) return fail(KR_E_INVALID, "cache and batch live on different devices");
    if (n && (b->tuning.flags & KR_BATCH_MIX_JOBS))
        return fail(KR_E_UNSUPPORTED, "a batch whose chunks mix jobs takes no radiance cache");
    (void)hipSetDevice(b->dev->ordinal);
    (void)hipStreamSynchronize(b->dev->stream);
    for (int pi = 0; pi < b->r->npipes; ++pi) (void)hipStreamSynchronize(b->r->pipe[pi].stream);
    b->nrc.resize(b->jobs.size(), nullptr);
    b->nrc[job] = n;
    // buffers for the smallest training group any attached cache uses
    for (void *p : b->nrc_allocs) dfree(b->dev, p);
    b->nrc_allocs.clear();
    for (NrcRender &x : b->pnr) x = NrcRender{};
    batch_drop_graph(b);  // the captured chunks carry the cache views
    uint32_t log2 = 32;
    for (kr_nrc *c : b->nrc)
        if (c) log2 = std::min(log2, c->cfg.train_log2);
    if (log2 == 32) return KR_OK;
    const uint32_t P = b->r->P;
    const uint32_t nv = b->max_bounce + 1;
    const uint32_t cap = uint32_t(std::min<uint64_t>(P, (uint64_t(P) >> log2) + 2));
    auto get = [&](auto **p, size_t cnt) {
        int rc = dalloc(b->dev, p, cnt);
        if (!rc) b->nrc_allocs.push_back(*p);
        return rc;
    };
    int rc = KR_OK;
    for (int pi = 0; pi < b->r->npipes && !rc; ++pi) {
        NrcRender &pn = b->pnr[pi];
        pn.nv = nv;
        if (!rc) rc = get(&pn.tslot_of, P);
        if (!rc) rc = get(&pn.rec, size_t(cap) * nv * kRecFloats);
        if (!rc) rc = get(&pn.tcount, cap);
        if (!rc) rc = get(&pn.qslot, P);
        if (!rc) rc = get(&pn.qbeta, 3 * size_t(P));
        if (!rc) rc = get(&pn.qraw, size_t(P) * nrc::kRaw);
    }
    if (rc) {
        for (void *p : b->nrc_allocs) dfree(b->dev, p);
        b->nrc_allocs.clear();
        for (NrcRender &x : b->pnr) x = NrcRender{};
        std::fill(b->nrc.begin(), b->nrc.end(), nullptr);
        return rc;
    }
    return KR_OK;
}

KR_EXPORT int kr_batch_run(kr_batch *b) {
    if (!b) return fail(KR_E_INVALID, "NULL batch");
    (void)hipSetDevice(b->dev->ordinal);
    if (!b->exec[0]) {
        int rc = batch_capture(b);
        if (rc) return rc;
    }
    kr_render *r = b->r;
    hipStream_t st = b->dev->stream;
    KR_HIP(hipEventRecord(b->e0, st));
    for (int pi = 1; pi < r->npipes; ++pi) KR_HIP(hipStreamWaitEvent(r->pipe[pi].stream, b->e0, 0));
    for (int pi = 0; pi < r->npipes; ++pi) KR_HIP(hipGraphLaunch(b->exec[pi], r->pipe[pi].stream));
    // join: a wait takes the event's latest record, so one event serves every pipe
    for (int pi = 1; pi < r->npipes; ++pi) {
        KR_HIP(hipEventRecord(r->fork, r->pipe[pi].stream));
        KR_HIP(hipStreamWaitEvent(st, r->fork, 0));
    }
    // every attached cache takes its Adam steps on what its jobs gathered
    std::vector<kr_nrc *> trained;
    for (kr_nrc *n : b->nrc)
        if (n && std::find(trained.begin(), trained.end(), n) == trained.end()) {
            trained.push_back(n);
            int rc = nrc_after_run(r, n);
            if (rc) return rc;
        }
    KR_HIP(hipEventRecord(b->e1, st));
    b->timed = true;
    b->r->paths += b->total_paths;
    return KR_OK;
}

KR_EXPORT int kr_batch_sync(kr_batch *b) {
    if (!b) return fail(KR_E_INVALID, "NULL batch");
    KR_HIP(hipStreamSynchronize(b->dev->stream));
    if (b->timed) {
        float ms = 0.f;
        KR_HIP(hipEventElapsedTime(&ms, b->e0, b->e1));
        b->seconds = ms * 1e-3;
        b->timed = false;
    }
    return KR_OK;
}

KR_EXPORT int kr_batch_reset(kr_batch *b) {
    if (!b) return fail(KR_E_INVALID, "NULL batch");
    return kr_render_reset(b->r);
}

KR_EXPORT int kr_batch_read(kr_batch *b, uint32_t job, double *rgb, uint32_t *valid) {
    if (!b) return fail(KR_E_INVALID, "NULL batch");
    if (job >= b->jobs.size()) return fail(KR_E_INVALID, "job index out of range");
    hipStream_t st = b->dev->stream;
    size_t off = b->acc_off[job], n = size_t(b->jobs[job].params.width) * b->jobs[job].params.height;
    if (rgb) KR_HIP(hipMemcpyAsync(rgb, b->r->acc + 3 * off, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, st));
    if (valid) KR_HIP(hipMemcpyAsync(valid, b->r->valid + off, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, st));
    KR_HIP(hipStreamSynchronize(st));
    // the other pipes' images: the job's sums are the images' sums, in pipe order
    std::vector<double> rgbp(rgb ? 3 * n : 0);
    std::vector<uint32_t> validp(valid ? n : 0);
    for (int pi = 1; pi < b->r->npipes; ++pi) {
        const Pipe &pp = b->r->pipe[pi];
        if (pp.acc == b->r->acc) continue;
        if (rgb)
            KR_HIP(hipMemcpyAsync(rgbp.data(), pp.acc + 3 * off, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, st));
        if (valid)
            KR_HIP(hipMemcpyAsync(validp.data(), pp.valid + off, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, st));
        KR_HIP(hipStreamSynchronize(st));
        if (rgb)
            for (size_t i = 0; i < 3 * n; ++i) rgb[i] += rgbp[i];
        if (valid)
            for (size_t i = 0; i < n; ++i) valid[i] += validp[i];
    }
    return KR_OK;
}

KR_EXPORT int kr_batch_get_info(kr_batch *b, kr_batch_info *info) {
    if (!b || !info) return fail(KR_E_INVALID, "NULL argument");
    info->total_paths = b->total_paths;
    info->total_pixels = b->total_pix;
    info->num_jobs = uint32_t(b->jobs.size());
    info->num_scenes = uint32_t(b->scenes.size());
    info->graph_nodes = b->graph_nodes;
    info->chunks = b->chunks;
    info->tuning = b->tuning;
    info->seconds = b->seconds;
    info->capture_seconds = b->capture_seconds;
    return KR_OK;
}

KR_EXPORT int kr_batch_autotune(kr_batch *b, uint32_t trial_spp, kr_batch_tuning *chosen) {
    if (!b) return fail(KR_E_INVALID, "NULL batch");
    if (std::any_of(b->nrc.begin(), b->nrc.end(), [](kr_nrc *n) { return n != nullptr; }))
        return fail(KR_E_UNSUPPORTED, "tune a batch before attaching radiance caches (the trials would train them)");
    (void)hipSetDevice(b->dev->ordinal);
    hipStream_t st = b->dev->stream;
    // trial_spp 0: the same fraction of every job's samples, about 256 M
    // paths in all, so that the trial has the batch's mix of job sizes (a
    // uniform sample count favoured small chunks that the full batch does not)
    double frac = 1.0;
    if (trial_spp == 0) frac = std::min(1.0, double(256ull << 20) / double(std::max<uint64_t>(b->total_paths, 1)));
    const uint32_t widths[3] = {2, 4, 8};
    const uint32_t chunks[2] = {16u << 20, 32u << 20};
    const uint32_t bpcs[2] = {32, 64};
    kr_batch_tuning best = b->tuning;
    best.paths_per_second = -1.0;
    uint32_t flags = b->tuning.flags;
    for (uint32_t w : widths) {
        int rc = batch_set_width(b, w);
        if (rc == KR_E_OVERFLOW) continue;  // a BVH of this width is too deep for the stack
        if (rc) return rc;
        std::vector<HostJob> thj;
        rc = batch_upload_tables(b, b->d_trec, b->d_tbegin, trial_spp, thj, frac);
        if (rc) return rc;
        uint64_t ttotal = 0;
        for (const HostJob &h : thj) ttotal += h.paths;
        BatchView bv{JobTable{b->d_trec, b->d_tbegin, uint32_t(b->jobs.size())}, &thj};
        uint32_t last_P = 0;
        for (uint32_t Pc : chunks) {
            uint32_t P = std::min(Pc, b->r->P);  // the queues hold at most r->P paths
            if (P == last_P) continue;
            last_P = P;
            for (uint32_t bpc : bpcs) {
                b->tuning.chunk_paths = P;
                b->tuning.blocks_per_cu = bpc;
                batch_apply_launch(b);
                double best_s = 1e30;
                for (int rep = 0; rep < 2; ++rep) {
                    KR_HIP(hipEventRecord(b->e0, st));
                    rc = batch_enqueue(b, bv, ttotal, P);
                    if (rc) return rc;
                    KR_HIP(hipEventRecord(b->e1, st));
                    KR_HIP(hipEventSynchronize(b->e1));
                    float ms = 0.f;
                    KR_HIP(hipEventElapsedTime(&ms, b->e0, b->e1));
                    best_s = std::min(best_s, double(ms) * 1e-3);
                }
                double rate = double(ttotal) / best_s;
                if (rate > best.paths_per_second) {
                    best = b->tuning;
                    best.bvh_width = w;
                    best.chunk_paths = P;
                    best.blocks_per_cu = bpc;
                    best.flags = flags;
                    best.paths_per_second = rate;
                }
            }
        }
    }
    int rc = batch_set_width(b, best.bvh_width);
    if (rc) return rc;
    b->tuning = best;
    batch_drop_graph(b);
    rc = kr_render_reset(b->r);
    if (rc) return rc;
    if (chosen) *chosen = best;
    return KR_OK;
}
