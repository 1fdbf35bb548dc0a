/*
 * kr_api.h -- C ABI of the MI355X (gfx950) Koishi Radiance path tracer.
 *
 * This is the drop-in boundary for the reference's hot path. The reference
 * moves a polymorphic scene to the GPU with poly::object/poly::vector deep
 * copies (src/poly/object.hpp, src/poly/vector.hpp) and launches a megakernel
 * through poly::kernel (src/poly/kernel.hpp:329-333) from
 * CudaSingleGPUTracer::execute (src/core/tracer/cudaSingleGPU.hpp:52-93).
 * Here the scene crosses as flat arrays, the integrator is a wavefront of
 * hand-written HIP kernels, and every call returns an int status.
 *
 * Conventions
 *  - Every function returns KR_OK (0) or a KR_E_* code. No C++ exception
 *    crosses this boundary. kr_last_error() returns the message of the last
 *    failure on the calling thread (replaces KTHROW, src/util/debug.hpp:68-75).
 *  - Host arrays passed in are copied before the call returns; the caller
 *    keeps ownership. Output arrays are caller-allocated.
 *  - A kr_device owns one HIP stream. Calls on one device (and on the scenes
 *    and renders made from it) must be serialised by the caller; separate
 *    devices are independent.
 *  - Coordinates follow the reference: Z-up camera basis by default
 *    (upaxis (0,0,1), src/util/config.hpp:32).
 */
#ifndef KR_API_H
#define KR_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KR_API_VERSION 4

enum kr_status {
    KR_OK = 0,
    KR_E_INVALID = 1,     /* bad argument or malformed scene            */
    KR_E_HIP = 2,         /* HIP runtime failure                         */
    KR_E_OOM = 3,         /* device or host allocation failed            */
    KR_E_OVERFLOW = 4,    /* a queue or stack bound was exceeded         */
    KR_E_UNSUPPORTED = 5, /* feature not provided by this build          */
    KR_E_NO_DEVICE = 6    /* no HIP device / extension not usable        */
};

typedef struct kr_device kr_device;
typedef struct kr_scene kr_scene;
typedef struct kr_render kr_render;

/* ---- scene description (flat, host memory) ---------------------------- */

/* Shader type tags. Replace the Factory<Shader> names
 * (src/ext/shader/shader.impl.cu:15-24): "Lambert", "GlossyBSDF",
 * "Emission", "GlassBSDF", "RefractionBSDF", "Transparent", "Blackhole",
 * "Mixed". */
enum kr_shader_type {
    KR_SHADER_LAMBERT = 0,
    KR_SHADER_GLOSSY = 1,
    KR_SHADER_EMISSION = 2,
    KR_SHADER_GLASS = 3,
    KR_SHADER_REFRACTION = 4,
    KR_SHADER_TRANSPARENT = 5,
    KR_SHADER_BLACKHOLE = 6,
    KR_SHADER_MIXED = 7
};

/* Scala<float> tags for Mixed's "fac" (src/ext/scala/scala.impl.cu:9-16). */
enum kr_scala_type {
    KR_SCALA_CONSTANT = 0, /* "Constant": value                              */
    KR_SCALA_FRESNEL = 1   /* "Fresnel": dielectric reflectance of wo, ior   */
};

/* Distribution tags for GlossyBSDF's "distribution" entry. Replace
 * Factory<SphericalDistribution> (src/ext/distribution/distribution.impl.cu:8-11). */
enum kr_distribution_type {
    KR_DIST_BECKMANN = 0,
    KR_DIST_COSINE = 1
};

typedef struct kr_shader_desc {
    uint32_t type;         /* kr_shader_type                                   */
    uint32_t distribution; /* kr_distribution_type (GlossyBSDF only)           */
    float color[3];        /* Lambert/GlossyBSDF/Emission "color", default 1   */
    float R[3];            /* Lambert "R", default 1                            */
    float roughness;       /* Beckmann "roughness", default 0.5 (alpha2 = r^2) */
    float strength;        /* Emission "strength", default 2                    */
    float ior;             /* GlassBSDF / RefractionBSDF "ior"                  */
    uint32_t fac_type;     /* Mixed: kr_scala_type of "fac"                     */
    float fac_value;       /* Mixed: Constant "value" or Fresnel "ior"          */
    uint32_t child[2];     /* Mixed: "shaders" [a, b] as indices into the
                              shader array; b is taken when u < fac
                              (src/ext/shader/mixed.hpp:22-27). Nesting depth
                              at most 3 Mixed levels.                         */
} kr_shader_desc;

enum kr_prim_type {
    KR_PRIM_MESH = 0,
    KR_PRIM_SPHERE = 1
};

/* One reference Primitive (src/core/primitive/primitive.hpp:10-37): a Mesh
 * (one per imported object/material group) or a Sphere. Order matters: it is
 * the order of Scene::primitives, which fixes the order of area lights. */
typedef struct kr_prim_desc {
    uint32_t type;      /* kr_prim_type                                     */
    uint32_t shader;    /* index into kr_scene_desc.shaders                 */
    uint32_t tri_begin; /* mesh: first triangle in kr_scene_desc.indices/3  */
    uint32_t tri_count; /* mesh: number of triangles                        */
    float center[3];    /* sphere: centre "o"                               */
    float radius;       /* sphere: "r"                                      */
} kr_prim_desc;

typedef struct kr_scene_desc {
    const float *positions;   /* num_vertices * 3                              */
    const float *normals;     /* num_vertices * 3 (per-vertex shading normals) */
    uint32_t num_vertices;
    const uint32_t *indices;  /* num_triangles * 3, into positions/normals     */
    uint32_t num_triangles;
    const kr_prim_desc *prims;
    uint32_t num_prims;
    const kr_shader_desc *shaders;
    uint32_t num_shaders;
    uint64_t light_seed;      /* seeds the 1024-entry area-light sample tables */
    uint32_t bvh_width;       /* 2, 4 or 8; 0 = library default (4)            */
    uint32_t leaf_size;       /* max triangles per leaf; 0 = default (1)       */
    uint32_t flags;           /* KR_SCENE_* bits                               */
    const struct kr_point_light *point_lights; /* "pointLight" assets, in file order */
    uint32_t num_point_lights;
} kr_scene_desc;

/* A "pointLight" asset (src/core/light/pointLight.hpp:10-36): li is the
 * emission color * strength with no distance falloff, occluded by the scene.
 * Point lights come first in the light list, in file order, then the area
 * lights (src/core/meta/scene.cu:158-168, 200-216). */
typedef struct kr_point_light {
    float position[3];
    float emission[3]; /* color * strength */
} kr_point_light;

/* No area lights are derived: the reference skips light promotion when any
 * shader name has no configuration (src/core/meta/scene.cu:193-216). */
#define KR_SCENE_NO_AREA_LIGHTS 0x1u
/* BVH child boxes quantised to 8 bits per plane (16 + 10W bytes per node),
 * decoded conservatively by traversal: same hits as full boxes, fewer bytes
 * per node visit. The default; the flag is accepted and changes nothing. */
#define KR_SCENE_COMPRESSED_NODES 0x2u
/* Full FP32 child boxes instead (28W bytes per node, 16-byte padded). */
#define KR_SCENE_FULL_NODES 0x4u

typedef struct kr_scene_info {
    uint32_t num_prims, num_triangles, num_spheres, num_lights;
    uint32_t bvh_width, bvh_nodes, bvh_leaves, bvh_depth;
    uint64_t device_bytes;     /* geometry + BVH bytes resident in HBM         */
    uint32_t node_bytes;       /* bytes per BVH node                           */
    uint32_t tri_bytes;        /* bytes per triangle read by traversal         */
    double build_seconds;
} kr_scene_info;

/* ---- camera and render parameters ------------------------------------- */

enum kr_lens_type { KR_LENS_PINHOLE = 0, KR_LENS_ORTHOGRAPHIC = 1 };

/* Mirrors CameraConfig (src/util/config.hpp:12-35). "target" is a
 * direction, not a point (src/core/misc/lens.hpp:21). */
typedef struct kr_camera {
    uint32_t lens;
    float fovx;           /* degrees */
    float position[3];
    float target[3];
    float upaxis[3];
} kr_camera;

typedef struct kr_params {
    uint32_t width, height;
    uint32_t spp;        /* total samples per pixel; fixes the lens strata
                            pw = ceil(sqrt(spp)) (src/core/misc/lens.hpp:19) */
    uint32_t max_bounce; /* RadianceKernel "maxBounce" (src/core/kernel/radiance.hpp:14) */
    uint64_t seed;       /* counter-based RNG key                               */
    uint32_t flags;      /* KR_RENDER_* bits                                    */
    uint32_t chunk_paths;/* paths in flight per wavefront chunk; 0 = default   */
    uint32_t kernel;     /* kr_kernel_type: "Radiance" (default) or one of the
                            one-hit debug kernels                             */
} kr_params;

/* Factory<Kernel> names (src/core/kernel/kernel.impl.cu:10-15). */
enum kr_kernel_type {
    KR_KERNEL_RADIANCE = 0,   /* RadianceKernel (src/core/kernel/radiance.cu)       */
    KR_KERNEL_NORMAL = 1,     /* n/2 + 1/2 at the first hit (normal.hpp:17-28)       */
    KR_KERNEL_VISIBILITY = 2, /* 1 if light 0 is visible from the first hit
                                 (visibility.hpp:17-35)                              */
    KR_KERNEL_HALFVECTOR = 3  /* acos(cos theta of the half vector to light 0) 2/pi
                                 (halfVector.hpp:17-37)                              */
};

#define KR_RENDER_NO_MATERIAL_SORT 0x1u /* shade with one mixed-material queue */
#define KR_RENDER_PERSISTENT_TRAVERSAL 0x2u /* traversal kernels refill lanes whose ray
                                               is done from the queue (persistent
                                               threads) instead of taking a fixed
                                               share of 64 rays per pass (same
                                               image; slower on the measured
                                               scenes) */
#define KR_RENDER_ONE_STREAM 0x4u /* run every chunk on the device's stream; by
                                     default a render of >= 8 M paths cycles its
                                     chunks through four streams (same image)   */
#define KR_RENDER_NO_TAIL_ROUNDS 0x8u /* traverse each bounce's queue in one launch;
                                         by default walks that outlast a step
                                         budget are resumed, packed into full
                                         waves, by follow-up launches (same
                                         image)                                 */

/* ---- ray query buffers -------------------------------------------------- */

#define KR_MEM_HOST 0u
#define KR_MEM_DEVICE 1u /* all pointers of the call are HIP device pointers */

typedef struct kr_rays {
    const float *origin;    /* n * 3 */
    const float *direction; /* n * 3, unit length */
    const float *tmax;      /* n; segment length for shadow queries, else NULL */
} kr_rays;

#define KR_MISS 0xFFFFFFFFu

typedef struct kr_hits {
    float *t;       /* INFINITY on miss                                   */
    uint32_t *prim; /* primitive index, KR_MISS on miss                   */
    uint32_t *tri;  /* triangle index within the primitive (input order)  */
    float *u, *v;   /* barycentric weights of vertex 1 and vertex 2       */
} kr_hits;

typedef struct kr_render_stats {
    uint64_t paths;          /* camera paths traced                        */
    uint64_t closest_rays;   /* closest-hit queries                        */
    uint64_t shadow_rays;    /* shadow (segment) queries                   */
    uint64_t invalid_samples;/* samples rejected as NaN/Inf                */
    double seconds;          /* wall time of the last kr_render_run + sync */
    /* filled when KR_PROFILE_COUNTERS is on: closest-hit traversal work */
    uint64_t node_visits;    /* wide-BVH nodes loaded                      */
    uint64_t tri_tests;      /* ray/triangle tests                         */
    /* filled when KR_PROFILE_EVENTS is on: HIP-event time of the render's
       launches, summed over launches since the last reset                 */
    double trace_ms, shade_ms, shadow_ms, other_ms;
    uint64_t trace_launches, shade_launches, shadow_launches;
    /* the same launches' busy time: the union of their intervals. A render
       of at least 8 M paths cycles its chunks through four streams, so
       launches of one kind overlap and the sums above exceed this        */
    double trace_busy_ms, shade_busy_ms, shadow_busy_ms;
    uint64_t cache_queries;  /* radiance-cache queries (paths ended by it)  */
} kr_render_stats;

#define KR_PROFILE_EVENTS 0x1u   /* time every kernel with HIP events        */
#define KR_PROFILE_COUNTERS 0x2u /* count node visits / triangle tests      */
#define KR_PROFILE_TRACE_EVENTS 0x4u /* time only the closest-hit traversal
                                        launches (fewer events in the stream)  */

/* ---- library and device ------------------------------------------------- */

/* Library version string. */
const char *kr_version(void);

/* Copies the calling thread's last error message into buf. */
int kr_last_error(char *buf, size_t len);

/* Number of visible HIP devices. */
int kr_device_count(int *count);

/* Opens device `ordinal` and creates its stream. Closing it releases the
 * device memory its scenes, renders, batches and caches left cached. */
int kr_device_open(int ordinal, kr_device **out);
int kr_device_close(kr_device *dev);

/* The library keeps the device buffers of freed scenes, renders, batches and
 * caches for later ones of a size they fit (hipFree waits for the whole
 * device and unmaps pages: tens to hundreds of ms per 1080p render).
 * kr_device_trim gives the idle ones back to HIP and reports their bytes.
 * No reference counterpart: the reference allocates and frees device memory
 * with each launch's arguments (cudaMallocManaged / cudaFree,
 * src/poly/kernel.hpp:83-116). */
int kr_device_trim(kr_device *dev, size_t *released);

/* Streams a plain device-to-device copy of `bytes` bytes (16-byte vector
 * loads and stores, kernel k_copy) `iters` times on the device's stream and
 * reports the achieved bandwidth (read + write bytes / time). No reference
 * counterpart: the HBM probe the roofline and the traffic counters are
 * calibrated against. */
int kr_device_bandwidth(kr_device *dev, size_t bytes, uint32_t iters, double *gbps);

/* ---- scene -------------------------------------------------------------
 * Replaces the transfer of core::Scene (src/core/meta/scene.hpp:15-71,
 * built by src/core/meta/scene.cu:18-219) through poly::vector/object.
 * Builds the wide BVH on the host, derives the area lights (every primitive
 * whose shader is Emission, in primitive order, src/core/meta/scene.cu:200-216),
 * builds each light's 1024-entry sample table
 * (src/core/primitive/mesh.hpp:61-98) and uploads everything to HBM. */
int kr_scene_upload(kr_device *dev, const kr_scene_desc *desc, kr_scene **out);
int kr_scene_free(kr_scene *scene);
int kr_scene_get_info(const kr_scene *scene, kr_scene_info *info);

/* Copies light `light`'s sample table: p_n receives 1024 * 6 floats
 * (position xyz, normal xyz), pdf the area pdf (1 / total area). */
int kr_scene_light_samples(const kr_scene *scene, uint32_t light, float *p_n, float *pdf);

/* ---- ray queries --------------------------------------------------------
 * kr_trace_closest replaces Scene::intersect(Ray) (src/core/meta/scene.hpp:35-55)
 * over Mesh::intersect(Ray) (src/core/primitive/mesh.cu:200-267) and
 * Sphere::intersect (src/core/primitive/sphere.hpp:31-54).
 * kr_trace_shadow replaces Scene::intersect(Seg) (src/core/meta/scene.hpp:57-67);
 * occluded[i] = 1 when segment i (origin, direction, tmax) is blocked. */
int kr_trace_closest(kr_scene *scene, const kr_rays *rays, uint32_t n, kr_hits *hits,
                     uint32_t mem);
int kr_trace_shadow(kr_scene *scene, const kr_rays *rays, uint32_t n, uint8_t *occluded,
                    uint32_t mem);

/* kr_trace_closest with every pointer of rays and hits a HIP device pointer
 * (KR_MEM_DEVICE), calling no hipMalloc, hipFree or allocating library
 * routine. No counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation). The traversal spill stacks live in a fixed
 * 32 MiB device array per GPU, held until the process ends; the launch is
 * cut to the lanes whose stacks fit in it, so any n is accepted, and a scene
 * whose stack depth needs more than 256 lanes' worth is refused with
 * KR_E_OVERFLOW. Every call on one GPU shares that array: calls that overlap
 * on one GPU (from two threads, or two kr_device handles of one ordinal)
 * race. Synchronous on the device's stream, like kr_trace_closest. */
int kr_trace_closest_noalloc(kr_scene *scene, const kr_rays *rays, uint32_t n, kr_hits *hits);

/* ---- Radiance integrator -------------------------------------------------
 * Replaces Tracer::execute (src/core/tracer/tracer.hpp:20-21) running
 * RadianceKernel::execute (src/core/kernel/radiance.cu:14-81) per sample and
 * the per-pixel averaging of CPUMulticoreTracer
 * (src/core/tracer/cpuMulticore.cu:47-64).
 * A render accumulates in FP64 on the device: rgb[3*(y*w+x)+c] is the sum
 * of the valid (finite) samples of pixel (x, y), valid[y*w+x] their count;
 * the reference's pixel value is rgb / valid. */
int kr_render_create(kr_scene *scene, const kr_camera *camera, const kr_params *params,
                     kr_render **out);
/* Traces samples [spp_begin, spp_begin + spp_count) of every pixel and adds
 * them to the accumulators. Asynchronous on the device stream. */
int kr_render_run(kr_render *render, uint32_t spp_begin, uint32_t spp_count);
int kr_render_sync(kr_render *render);
/* Clears the accumulators. */
int kr_render_reset(kr_render *render);
/* Copies the accumulators to host memory (either pointer may be NULL). */
int kr_render_read(kr_render *render, double *rgb, uint32_t *valid);
int kr_render_get_stats(kr_render *render, kr_render_stats *stats);
/* KR_PROFILE_* bits; takes effect from the next kr_render_run. */
int kr_render_set_profiling(kr_render *render, uint32_t flags);
int kr_render_free(kr_render *render);

/* One-shot render of all params->spp samples into caller buffers. */
int kr_render_radiance(kr_scene *scene, const kr_camera *camera, const kr_params *params,
                       double *rgb, uint32_t *valid);

/* ---- profiler (bounce dump) -----------------------------------------------
 * Replaces core::Profiler with RadianceKernel::Slice
 * (src/core/kernel/kernel.hpp:35-81, src/core/kernel/kernel.cu:26-73,
 * src/core/kernel/radiance.hpp:17-55, radiance.cu:38-50, 74-78). While a
 * profiler is set, a render traces only the pixels x0 <= x < x1,
 * y0 <= y < y1 of its area (CPUMulticoreTracer traces only the profiled area,
 * src/core/tracer/cpuMulticore.cu:17); the other pixels keep zero radiance.
 * Every path (x, y, k) of the area is a slice holding its first `depth`
 * vertices. Slices are numbered as the reference writes them:
 * ((y - y0) * (x1 - x0) + (x - x0)) * spp + k. */
typedef struct kr_profiler_config {
    uint32_t area[4]; /* x0, y0, x1, y1; clamped to [0, w] x [0, h] as Profiler's ctor does */
    uint32_t depth;   /* vertices recorded per path (the slice's "depth" property, 8 in the reference) */
} kr_profiler_config;

/* RadianceKernel::Slice::BounceRecord, one per vertex below depth. */
typedef struct kr_slice_record {
    float o[3], d[3]; /* the ray that reached the vertex */
    float f[3];       /* BSDF value of the light sample (zero when the scene has no lights) */
    float L[3];       /* path radiance after the vertex's light sample and emission */
    float theta;      /* acos(wo . reflect(wi, n)) in the vertex's local frame */
    uint32_t shader;  /* shader id of the hit primitive */
} kr_slice_record;

/* Sets (config non-NULL) or removes the profiler of a Radiance / Normal /
 * Visibility / HalfVector render; clears the slices. Not combined with a
 * radiance cache (KR_E_UNSUPPORTED). The debug kernels record no vertices. */
int kr_render_set_profiler(kr_render *render, const kr_profiler_config *config);
/* Slice count (area pixels x spp) and depth of the profiler in force. */
int kr_render_profiler_size(kr_render *render, uint64_t *num_slices, uint32_t *depth);
/* Copies slices [first, first + count): per slice the path radiance (3
 * floats, L), the loop counter of RadianceKernel::execute (bounce: the
 * vertices processed, not counting one ended by Russian roulette) and `depth`
 * records (recs, zero beyond the recorded ones). Any pointer may be NULL.
 * Synchronises. */
int kr_render_read_slices(kr_render *render, uint64_t first, uint64_t count, float *L, uint32_t *bounce,
                          kr_slice_record *recs);

/* ---- neural radiance cache ----------------------------------------------
 * No counterpart in the reference: an online-trained radiance cache
 * (Mueller et al., SIGGRAPH 2021) that ends paths early. The MLP is
 * 96-64-64-64-64-64-16 (ReLU, 3 outputs used), BF16 weights and activations
 * on MFMA with FP32 accumulation, Adam on FP32 master weights.
 *
 * Input of one query, a "vertex record" of KR_NRC_RAW floats:
 *   position xyz, outgoing direction xyz (towards the previous vertex),
 *   shading normal xyz, roughness (0 for Lambert / Emission), albedo rgb, 0 0 0.
 * Parameters are KR_NRC_PARAMS floats: W0 (96 x 64, row major, in x out),
 * b0 (64), W1 (64 x 64), b1, ..., W5 (64 x 16), b5 (16). */
#define KR_NRC_RAW 16
#define KR_NRC_FEATURES 96
#define KR_NRC_PARAMS 23888

typedef struct kr_nrc kr_nrc;

#define KR_NRC_TRAIN_ALL 0xFFFFFFFFu /* kr_nrc_config.train_log2: every path trains */

typedef struct kr_nrc_config {
    float box_min[3], box_max[3]; /* scene box for the position encoding      */
    float learning_rate;          /* Adam step size; 0 = 1e-2                  */
    uint32_t seed;                /* weight initialisation                     */
    uint32_t batch;               /* training samples per step; 0 = 16384      */
    uint32_t steps_per_frame;     /* Adam steps after each render run; 0 = 4   */
    uint32_t train_log2;          /* 1 / 2^train_log2 of paths train (at most
                                     31); 0 = 4; KR_NRC_TRAIN_ALL: every path  */
    float rough_threshold;        /* GlossyBSDF roughness counted as rough;
                                     0 = 0.3                                   */
    uint32_t flags;               /* KR_NRC_* bits                             */
    float target_clamp;           /* > 0: training targets are clamped to this
                                     radiance per channel (e.g. the brightest
                                     emitter's), 0: unclamped                   */
    float ema_decay;              /* renders query the bias-corrected moving
                                     average of the trained parameters with this
                                     decay per Adam step (0 < d < 1); 0 = 0.99;
                                     < 0: they query the trained parameters     */
} kr_nrc_config;

#define KR_NRC_FROZEN 0x1u /* renders query the cache but never train it */
/* The cache holds the radiance a vertex reflects from its BSDF-sampled
 * continuation (everything the later vertices of a path carry back), not its
 * own light sample and emission; paths end at their first diffuse or rough
 * vertex, after its light sample, with a query. Without the flag the cache
 * holds everything a vertex scatters except its emission, and paths end at
 * their second diffuse or rough vertex without a light sample there. */
#define KR_NRC_INDIRECT 0x2u

int kr_nrc_create(kr_device *dev, const kr_nrc_config *config, kr_nrc **out);
int kr_nrc_free(kr_nrc *nrc);
int kr_nrc_get_params(kr_nrc *nrc, float *params);
int kr_nrc_set_params(kr_nrc *nrc, const float *params); /* also resets Adam
                                                          and the average */
/* The parameters renders query: the moving average of the trained ones (the
 * trained ones themselves when ema_decay < 0). kr_nrc_predict, _gradient and
 * _train work on the trained ones. */
int kr_nrc_get_render_params(kr_nrc *nrc, float *params);
/* Feature vectors (n x KR_NRC_FEATURES floats) of n vertex records. */
int kr_nrc_encode(kr_nrc *nrc, const float *raw, uint32_t n, float *features);
/* Cache prediction (n x 3 floats) for n vertex records. */
int kr_nrc_predict(kr_nrc *nrc, const float *raw, uint32_t n, float *rgb);
/* Relative-L2 loss and its gradient (KR_NRC_PARAMS floats) over n records
 * with targets (n x 3); no update. */
int kr_nrc_gradient(kr_nrc *nrc, const float *raw, const float *target, uint32_t n, float *grad,
                    float *loss);
/* `steps` Adam steps, each on all n records; loss of the last step. */
int kr_nrc_train(kr_nrc *nrc, const float *raw, const float *target, uint32_t n, uint32_t steps,
                 float *loss);
/* Copies up to `max` of the training samples renders have gathered (vertex
 * records, max x KR_NRC_RAW floats; targets, max x 3) in ring order; *count
 * receives how many were copied. */
int kr_nrc_read_pool(kr_nrc *nrc, uint32_t max, float *raw, float *target, uint32_t *count);
/* Samples trained on so far, and the mean loss of the last training step. */
int kr_nrc_get_stats(kr_nrc *nrc, uint64_t *samples_trained, float *last_loss);

/* Attaches a cache to a render (NULL detaches). While attached, one path in
 * 2^train_log2 is traced unbiased to max_bounce and its vertices become
 * training samples; every other path ends at its second diffuse or rough
 * vertex, where the cache's prediction replaces the rest of the path. After
 * each kr_render_run the cache takes steps_per_frame Adam steps on the most
 * recent samples. */
int kr_render_attach_nrc(kr_render *render, kr_nrc *nrc);

/* ---- batched renders ------------------------------------------------------
 * No counterpart in the reference, which renders one scene per `cr` process
 * (src/cr.cu:53-60). A batch holds many jobs (scene, camera, params) of
 * uneven size and renders all their samples as ONE wavefront: every chunk of
 * paths may mix jobs, each path finds its job by its index, and the whole
 * sequence of kernels for all chunks is captured once into a HIP graph and
 * replayed by kr_batch_run with a single graph launch. A job's image is
 * bit-identical to kr_render_radiance of the same scene, camera and params.
 * Registered on the host side as the tracer "HIPBatched". */
typedef struct kr_batch kr_batch;

typedef struct kr_batch_job {
    uint32_t scene;    /* index into the batch's scene list                  */
    kr_camera camera;
    kr_params params;  /* width, height, spp, max_bounce, seed; chunk_paths
                          and flags are batch-wide (kr_batch_tuning)         */
} kr_batch_job;

typedef struct kr_batch_tuning {
    uint32_t bvh_width;     /* 2, 4 or 8                                     */
    uint32_t chunk_paths;   /* paths in flight per wavefront chunk           */
    uint32_t blocks_per_cu; /* 256-thread blocks per CU of every launch (1..64) */
    uint32_t flags;         /* KR_RENDER_* bits, KR_BATCH_MIX_JOBS           */
    double paths_per_second;/* kr_batch_autotune: the winner's trial rate    */
} kr_batch_tuning;

/* kr_batch_tuning.flags: cut the batch's path space into equal chunks that
 * may straddle jobs (their paths look up their job on the device) instead of
 * cutting every job's range on its own. */
#define KR_BATCH_MIX_JOBS 0x100u

typedef struct kr_batch_info {
    uint64_t total_paths;   /* sum of width * height * spp over the jobs     */
    uint64_t total_pixels;
    uint32_t num_jobs, num_scenes;
    uint32_t graph_nodes;   /* kernels captured in the batch's HIP graph     */
    uint32_t chunks;
    kr_batch_tuning tuning; /* in force                                      */
    double seconds;         /* device time of the last kr_batch_run (events) */
    double capture_seconds; /* host time of the last graph capture + instantiate */
} kr_batch_info;

/* Copies the scene descriptions (the batch rebuilds their BVHs when the
 * tuning changes the width) and uploads them. tuning NULL: width 4, 32 M
 * paths per chunk, 64 blocks per CU. */
int kr_batch_create(kr_device *dev, const kr_scene_desc *scenes, uint32_t num_scenes, const kr_batch_job *jobs,
                    uint32_t num_jobs, const kr_batch_tuning *tuning, kr_batch **out);
/* Autotuner: times trial renders of the first trial_spp samples of every
 * job (0: the same fraction of every job's samples, about 256 M paths in
 * all) for every BVH width in
 * {2, 4, 8}, chunk size in {16, 32} M paths and blocks per CU in {32, 64},
 * applies the fastest and returns it. */
int kr_batch_autotune(kr_batch *batch, uint32_t trial_spp, kr_batch_tuning *chosen);
/* Renders every sample of every job (adds to the accumulators) with one
 * launch of the captured graph; asynchronous. */
int kr_batch_run(kr_batch *batch);
int kr_batch_sync(kr_batch *batch);
int kr_batch_reset(kr_batch *batch);
/* Accumulators of one job, as kr_render_read. */
int kr_batch_read(kr_batch *batch, uint32_t job, double *rgb, uint32_t *valid);
int kr_batch_get_info(kr_batch *batch, kr_batch_info *info);
int kr_batch_free(kr_batch *batch);
/* Attaches a radiance cache to job `job` of the batch (NULL detaches), as
 * kr_render_attach_nrc does for a render: in every kr_batch_run the job's
 * paths are selected for training or end early with a cache query by the
 * cache's configuration, and after the run every attached cache takes its
 * steps_per_frame Adam steps on the samples its jobs gathered. A cache is
 * specific to its scene; several jobs of one scene may share one. The next
 * kr_batch_run re-captures the batch's graphs. KR_E_UNSUPPORTED for batches
 * whose chunks mix jobs (KR_BATCH_MIX_JOBS); kr_batch_autotune refuses a
 * batch with caches attached (its trials would train them). No counterpart
 * in the reference. */
int kr_batch_attach_nrc(kr_batch *batch, uint32_t job, kr_nrc *nrc);

#ifdef __cplusplus
}
#endif

#endif /* KR_API_H */
