"""ctypes binding of the C ABI in include/kr_api.h (libkr.so, gfx950).

The Python side is a thin caller: scenes, rays and images cross as numpy
arrays into the library, which does all work in HIP kernels. Importing this
module never falls back to anything else: if libkr.so is missing or has no
usable device, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_ROOT, "libkr.so")

KR_OK = 0
KR_MISS = 0xFFFFFFFF
KR_MEM_HOST = 0
KR_MEM_DEVICE = 1
KR_RENDER_NO_MATERIAL_SORT = 0x1
KR_RENDER_PERSISTENT_TRAVERSAL = 0x2
KR_RENDER_ONE_STREAM = 0x4
KR_RENDER_NO_TAIL_ROUNDS = 0x8

SHADER_LAMBERT, SHADER_GLOSSY, SHADER_EMISSION = 0, 1, 2
SHADER_GLASS, SHADER_REFRACTION, SHADER_TRANSPARENT, SHADER_BLACKHOLE, SHADER_MIXED = 3, 4, 5, 6, 7
SCALA_CONSTANT, SCALA_FRESNEL = 0, 1
KERNEL_RADIANCE, KERNEL_NORMAL, KERNEL_VISIBILITY, KERNEL_HALFVECTOR = 0, 1, 2, 3
DIST_BECKMANN, DIST_COSINE = 0, 1
PRIM_MESH, PRIM_SPHERE = 0, 1
LENS_PINHOLE, LENS_ORTHOGRAPHIC = 0, 1

u32, f32, u64 = C.c_uint32, C.c_float, C.c_uint64


class ShaderDesc(C.Structure):
    _fields_ = [("type", u32), ("distribution", u32), ("color", f32 * 3), ("R", f32 * 3),
                ("roughness", f32), ("strength", f32), ("ior", f32), ("fac_type", u32), ("fac_value", f32),
                ("child", u32 * 2)]


class PrimDesc(C.Structure):
    _fields_ = [("type", u32), ("shader", u32), ("tri_begin", u32), ("tri_count", u32),
                ("center", f32 * 3), ("radius", f32)]


class SceneDesc(C.Structure):
    _fields_ = [("positions", C.POINTER(f32)), ("normals", C.POINTER(f32)), ("num_vertices", u32),
                ("indices", C.POINTER(u32)), ("num_triangles", u32),
                ("prims", C.POINTER(PrimDesc)), ("num_prims", u32),
                ("shaders", C.POINTER(ShaderDesc)), ("num_shaders", u32),
                ("light_seed", u64), ("bvh_width", u32), ("leaf_size", u32), ("flags", u32),
                ("point_lights", C.c_void_p), ("num_point_lights", u32)]


class PointLight(C.Structure):
    _fields_ = [("position", f32 * 3), ("emission", f32 * 3)]


KR_SCENE_NO_AREA_LIGHTS = 0x1
KR_SCENE_COMPRESSED_NODES = 0x2  # the default node format
KR_SCENE_FULL_NODES = 0x4


class SceneInfo(C.Structure):
    _fields_ = [("num_prims", u32), ("num_triangles", u32), ("num_spheres", u32), ("num_lights", u32),
                ("bvh_width", u32), ("bvh_nodes", u32), ("bvh_leaves", u32), ("bvh_depth", u32),
                ("device_bytes", u64), ("node_bytes", u32), ("tri_bytes", u32), ("build_seconds", C.c_double)]


class Camera(C.Structure):
    _fields_ = [("lens", u32), ("fovx", f32), ("position", f32 * 3), ("target", f32 * 3), ("upaxis", f32 * 3)]


class Params(C.Structure):
    _fields_ = [("width", u32), ("height", u32), ("spp", u32), ("max_bounce", u32), ("seed", u64),
                ("flags", u32), ("chunk_paths", u32), ("kernel", u32)]


class Rays(C.Structure):
    _fields_ = [("origin", C.c_void_p), ("direction", C.c_void_p), ("tmax", C.c_void_p)]


class Hits(C.Structure):
    _fields_ = [("t", C.c_void_p), ("prim", C.c_void_p), ("tri", C.c_void_p), ("u", C.c_void_p),
                ("v", C.c_void_p)]


class ProfilerConfig(C.Structure):
    _fields_ = [("area", u32 * 4), ("depth", u32)]


# kr_slice_record
SLICE_RECORD = np.dtype([("o", "<f4", 3), ("d", "<f4", 3), ("f", "<f4", 3), ("L", "<f4", 3), ("theta", "<f4"),
                         ("shader", "<u4")])


class RenderStats(C.Structure):
    _fields_ = [("paths", u64), ("closest_rays", u64), ("shadow_rays", u64), ("invalid_samples", u64),
                ("seconds", C.c_double), ("node_visits", u64), ("tri_tests", u64), ("trace_ms", C.c_double),
                ("shade_ms", C.c_double), ("shadow_ms", C.c_double), ("other_ms", C.c_double),
                ("trace_launches", u64), ("shade_launches", u64), ("shadow_launches", u64),
                ("trace_busy_ms", C.c_double), ("shade_busy_ms", C.c_double), ("shadow_busy_ms", C.c_double),
                ("cache_queries", u64)]


KR_PROFILE_EVENTS = 0x1
KR_BATCH_MIX_JOBS = 0x100  # BatchTuning(flags=...): chunks may straddle jobs
KR_PROFILE_COUNTERS = 0x2
KR_PROFILE_TRACE_EVENTS = 0x4  # time only the closest-hit traversal launches

NRC_RAW, NRC_FEATURES, NRC_PARAMS = 16, 96, 23888
NRC_TRAIN_ALL = 0xFFFFFFFF  # Nrc(train_log2=...): every path trains
NRC_FROZEN = 0x1  # Nrc(flags=...): renders query the cache but never train it
NRC_INDIRECT = 0x2  # Nrc(flags=...): the cache holds a vertex's reflected continuation; paths end at the first


class BatchJob(C.Structure):
    _fields_ = [("scene", u32), ("camera", Camera), ("params", Params)]


class BatchTuning(C.Structure):
    _fields_ = [("bvh_width", u32), ("chunk_paths", u32), ("blocks_per_cu", u32), ("flags", u32),
                ("paths_per_second", C.c_double)]

    def as_dict(self):
        return {"bvh_width": self.bvh_width, "chunk_paths": self.chunk_paths, "blocks_per_cu": self.blocks_per_cu,
                "flags": self.flags, "trial_paths_per_second": self.paths_per_second}


class BatchInfo(C.Structure):
    _fields_ = [("total_paths", u64), ("total_pixels", u64), ("num_jobs", u32), ("num_scenes", u32),
                ("graph_nodes", u32), ("chunks", u32), ("tuning", BatchTuning), ("seconds", C.c_double),
                ("capture_seconds", C.c_double)]


class NrcConfig(C.Structure):
    _fields_ = [("box_min", f32 * 3), ("box_max", f32 * 3), ("learning_rate", f32), ("seed", u32),
                ("batch", u32), ("steps_per_frame", u32), ("train_log2", u32), ("rough_threshold", f32),
                ("flags", u32), ("target_clamp", f32), ("ema_decay", f32)]


EXPORTS = ["kr_version", "kr_last_error", "kr_device_count", "kr_device_open", "kr_device_close",
           "kr_scene_upload", "kr_scene_free", "kr_scene_get_info", "kr_scene_light_samples",
           "kr_trace_closest", "kr_trace_shadow", "kr_render_create", "kr_render_run", "kr_render_sync",
           "kr_render_reset", "kr_render_read", "kr_render_get_stats", "kr_render_free",
           "kr_render_radiance", "kr_render_set_profiling", "kr_device_bandwidth", "kr_nrc_create", "kr_nrc_free",
           "kr_nrc_get_params", "kr_nrc_set_params", "kr_nrc_get_render_params", "kr_nrc_encode", "kr_nrc_predict", "kr_nrc_gradient",
           "kr_nrc_train", "kr_nrc_get_stats", "kr_render_attach_nrc", "kr_nrc_read_pool",
           "kr_batch_create", "kr_batch_autotune", "kr_batch_run", "kr_batch_sync", "kr_batch_reset", "kr_batch_read",
           "kr_batch_get_info", "kr_batch_free", "kr_trace_closest_noalloc", "kr_batch_attach_nrc",
           "kr_device_trim"]

_lib = None


def load(path: str = LIB_PATH):
    """Loads libkr.so (raises OSError if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"{path} is not built: run __graft_entry__.build() or make -C cuda-raytracer_amd")
    lib = C.CDLL(path)
    vp = C.c_void_p
    lib.kr_version.restype = C.c_char_p
    lib.kr_last_error.argtypes = [C.c_char_p, C.c_size_t]
    lib.kr_device_count.argtypes = [C.POINTER(C.c_int)]
    lib.kr_device_open.argtypes = [C.c_int, C.POINTER(vp)]
    lib.kr_device_close.argtypes = [vp]
    lib.kr_scene_upload.argtypes = [vp, C.POINTER(SceneDesc), C.POINTER(vp)]
    lib.kr_scene_free.argtypes = [vp]
    lib.kr_scene_get_info.argtypes = [vp, C.POINTER(SceneInfo)]
    lib.kr_scene_light_samples.argtypes = [vp, u32, vp, C.POINTER(f32)]
    lib.kr_trace_closest.argtypes = [vp, C.POINTER(Rays), u32, C.POINTER(Hits), u32]
    lib.kr_trace_shadow.argtypes = [vp, C.POINTER(Rays), u32, vp, u32]
    lib.kr_trace_closest_noalloc.argtypes = [vp, C.POINTER(Rays), u32, C.POINTER(Hits)]
    lib.kr_render_create.argtypes = [vp, C.POINTER(Camera), C.POINTER(Params), C.POINTER(vp)]
    lib.kr_render_run.argtypes = [vp, u32, u32]
    lib.kr_render_sync.argtypes = [vp]
    lib.kr_render_reset.argtypes = [vp]
    lib.kr_render_read.argtypes = [vp, vp, vp]
    lib.kr_render_get_stats.argtypes = [vp, C.POINTER(RenderStats)]
    lib.kr_render_free.argtypes = [vp]
    lib.kr_render_radiance.argtypes = [vp, C.POINTER(Camera), C.POINTER(Params), vp, vp]
    lib.kr_render_set_profiling.argtypes = [vp, u32]
    lib.kr_render_set_profiler.argtypes = [vp, C.POINTER(ProfilerConfig)]
    lib.kr_render_profiler_size.argtypes = [vp, C.POINTER(u64), C.POINTER(u32)]
    lib.kr_render_read_slices.argtypes = [vp, u64, u64, vp, vp, vp]
    lib.kr_device_bandwidth.argtypes = [vp, C.c_size_t, u32, C.POINTER(C.c_double)]
    lib.kr_device_trim.argtypes = [vp, C.POINTER(C.c_size_t)]
    lib.kr_nrc_create.argtypes = [vp, C.POINTER(NrcConfig), C.POINTER(vp)]
    lib.kr_nrc_free.argtypes = [vp]
    lib.kr_nrc_get_params.argtypes = [vp, vp]
    lib.kr_nrc_get_render_params.argtypes = [vp, vp]
    lib.kr_nrc_set_params.argtypes = [vp, vp]
    lib.kr_nrc_encode.argtypes = [vp, vp, u32, vp]
    lib.kr_nrc_predict.argtypes = [vp, vp, u32, vp]
    lib.kr_nrc_gradient.argtypes = [vp, vp, vp, u32, vp, C.POINTER(f32)]
    lib.kr_nrc_train.argtypes = [vp, vp, vp, u32, u32, C.POINTER(f32)]
    lib.kr_nrc_get_stats.argtypes = [vp, C.POINTER(u64), C.POINTER(f32)]
    lib.kr_render_attach_nrc.argtypes = [vp, vp]
    lib.kr_nrc_read_pool.argtypes = [vp, u32, vp, vp, C.POINTER(u32)]
    lib.kr_batch_create.argtypes = [vp, C.POINTER(SceneDesc), u32, C.POINTER(BatchJob), u32,
                                    C.POINTER(BatchTuning), C.POINTER(vp)]
    lib.kr_batch_autotune.argtypes = [vp, u32, C.POINTER(BatchTuning)]
    lib.kr_batch_run.argtypes = [vp]
    lib.kr_batch_sync.argtypes = [vp]
    lib.kr_batch_reset.argtypes = [vp]
    lib.kr_batch_read.argtypes = [vp, u32, vp, vp]
    lib.kr_batch_get_info.argtypes = [vp, C.POINTER(BatchInfo)]
    lib.kr_batch_free.argtypes = [vp]
    lib.kr_batch_attach_nrc.argtypes = [vp, u32, vp]
    for name in EXPORTS:
        if name not in ("kr_version",):
            getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib


class KrError(RuntimeError):
    def __init__(self, code, where):
        buf = C.create_string_buffer(1024)
        if _lib is not None:
            _lib.kr_last_error(buf, 1024)
        super().__init__(f"{where} failed with status {code}: {buf.value.decode(errors='replace')}")
        self.code = code


def check(rc, where):
    if rc != KR_OK:
        raise KrError(rc, where)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Device:
    def __init__(self, ordinal: int = 0):
        lib = load()
        self.lib = lib
        h = C.c_void_p()
        check(lib.kr_device_open(ordinal, C.byref(h)), "kr_device_open")
        self.h = h

    def bandwidth(self, nbytes: int, iters: int = 10) -> float:
        """GB/s of the plain copy kernel k_copy over nbytes (read + write)."""
        g = C.c_double()
        check(self.lib.kr_device_bandwidth(self.h, nbytes, iters, C.byref(g)), "kr_device_bandwidth")
        return g.value

    def trim(self) -> int:
        """Gives the device buffers that freed scenes, renders, batches and
        caches left cached back to HIP; returns their bytes."""
        n = C.c_size_t()
        check(self.lib.kr_device_trim(self.h, C.byref(n)), "kr_device_trim")
        return n.value

    def close(self):
        if self.h:
            self.lib.kr_device_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class GpuScene:
    """A scene uploaded through kr_scene_upload."""

    def __init__(self, dev: Device, desc_holder):
        self.dev = dev
        self.lib = dev.lib
        h = C.c_void_p()
        check(self.lib.kr_scene_upload(dev.h, C.byref(desc_holder.desc), C.byref(h)), "kr_scene_upload")
        self.h = h

    def info(self) -> SceneInfo:
        inf = SceneInfo()
        check(self.lib.kr_scene_get_info(self.h, C.byref(inf)), "kr_scene_get_info")
        return inf

    def light_samples(self, light: int):
        pn = np.zeros((1024, 6), np.float32)
        pdf = f32()
        check(self.lib.kr_scene_light_samples(self.h, light, _ptr(pn), C.byref(pdf)), "kr_scene_light_samples")
        return pn, pdf.value

    def trace_closest(self, origin, direction):
        o = np.ascontiguousarray(origin, np.float32)
        d = np.ascontiguousarray(direction, np.float32)
        n = o.shape[0]
        t = np.empty(n, np.float32)
        prim = np.empty(n, np.uint32)
        tri = np.empty(n, np.uint32)
        u = np.empty(n, np.float32)
        v = np.empty(n, np.float32)
        rays = Rays(_ptr(o), _ptr(d), None)
        hits = Hits(_ptr(t), _ptr(prim), _ptr(tri), _ptr(u), _ptr(v))
        check(self.lib.kr_trace_closest(self.h, C.byref(rays), n, C.byref(hits), KR_MEM_HOST), "kr_trace_closest")
        return t, prim, tri, u, v

    def trace_closest_device(self, origin, direction, n, t, prim, tri, u, v, noalloc=False):
        """Closest hit of n rays whose buffers are HIP device addresses
        (ints): kr_trace_closest with KR_MEM_DEVICE, or
        kr_trace_closest_noalloc."""
        rays = Rays(C.c_void_p(origin), C.c_void_p(direction), None)
        hits = Hits(*(C.c_void_p(x) for x in (t, prim, tri, u, v)))
        if noalloc:
            check(self.lib.kr_trace_closest_noalloc(self.h, C.byref(rays), n, C.byref(hits)),
                  "kr_trace_closest_noalloc")
        else:
            check(self.lib.kr_trace_closest(self.h, C.byref(rays), n, C.byref(hits), KR_MEM_DEVICE),
                  "kr_trace_closest")

    def trace_shadow(self, origin, direction, tmax):
        o = np.ascontiguousarray(origin, np.float32)
        d = np.ascontiguousarray(direction, np.float32)
        tm = np.ascontiguousarray(tmax, np.float32)
        n = o.shape[0]
        occ = np.empty(n, np.uint8)
        rays = Rays(_ptr(o), _ptr(d), _ptr(tm))
        check(self.lib.kr_trace_shadow(self.h, C.byref(rays), n, _ptr(occ), KR_MEM_HOST), "kr_trace_shadow")
        return occ

    def free(self):
        if self.h:
            self.lib.kr_scene_free(self.h)
            self.h = None


class Render:
    """kr_render handle: device FP64 accumulators for one camera/params."""

    def __init__(self, scene: GpuScene, camera: Camera, params: Params):
        self.scene = scene
        self.lib = scene.lib
        self.params = params
        h = C.c_void_p()
        check(self.lib.kr_render_create(scene.h, C.byref(camera), C.byref(params), C.byref(h)), "kr_render_create")
        self.h = h
        self.npix = params.width * params.height

    def run(self, spp_begin, spp_count):
        check(self.lib.kr_render_run(self.h, spp_begin, spp_count), "kr_render_run")

    def sync(self):
        check(self.lib.kr_render_sync(self.h), "kr_render_sync")

    def reset(self):
        check(self.lib.kr_render_reset(self.h), "kr_render_reset")

    def read(self):
        rgb = np.zeros((self.npix, 3), np.float64)
        valid = np.zeros(self.npix, np.uint32)
        check(self.lib.kr_render_read(self.h, _ptr(rgb), _ptr(valid)), "kr_render_read")
        return rgb, valid

    def attach_nrc(self, nrc):
        check(self.lib.kr_render_attach_nrc(self.h, nrc.h if nrc is not None else None), "kr_render_attach_nrc")

    def set_profiling(self, flags: int):
        check(self.lib.kr_render_set_profiling(self.h, flags), "kr_render_set_profiling")

    def set_profiler(self, area=None, depth=8):
        """kr_render_set_profiler: area (x0, y0, x1, y1) or None to remove."""
        if area is None:
            check(self.lib.kr_render_set_profiler(self.h, None), "kr_render_set_profiler")
            return
        cfg = ProfilerConfig((u32 * 4)(*[min(int(a), 0xFFFFFFFF) for a in area]), depth)
        check(self.lib.kr_render_set_profiler(self.h, C.byref(cfg)), "kr_render_set_profiler")

    def profiler_size(self):
        n, d = u64(0), u32(0)
        check(self.lib.kr_render_profiler_size(self.h, C.byref(n), C.byref(d)), "kr_render_profiler_size")
        return n.value, d.value

    def read_slices(self, first=0, count=None):
        """Slices [first, first + count): (L (n, 3) float32, bounce (n,) uint32,
        records (n, depth) of SLICE_RECORD)."""
        n, depth = self.profiler_size()
        if count is None:
            count = n - first
        L = np.zeros((count, 3), np.float32)
        bounce = np.zeros(count, np.uint32)
        recs = np.zeros((count, depth), SLICE_RECORD)
        check(self.lib.kr_render_read_slices(self.h, first, count, _ptr(L), _ptr(bounce), _ptr(recs)),
              "kr_render_read_slices")
        return L, bounce, recs

    def stats(self) -> RenderStats:
        s = RenderStats()
        check(self.lib.kr_render_get_stats(self.h, C.byref(s)), "kr_render_get_stats")
        return s

    def free(self):
        if self.h:
            self.lib.kr_render_free(self.h)
            self.h = None


def device_count() -> int:
    lib = load()
    n = C.c_int(0)
    check(lib.kr_device_count(C.byref(n)), "kr_device_count")
    return n.value


class Nrc:
    """kr_nrc handle: the neural radiance cache on one device."""

    def __init__(self, dev: Device, box_min, box_max, seed=5, learning_rate=0.0, batch=0, steps_per_frame=0,
                 train_log2=0, rough_threshold=0.0, flags=0, target_clamp=0.0, ema_decay=0.0):
        self.lib = dev.lib
        cfg = NrcConfig((f32 * 3)(*box_min), (f32 * 3)(*box_max), learning_rate, seed, batch, steps_per_frame,
                        train_log2, rough_threshold, flags, target_clamp, ema_decay)
        h = C.c_void_p()
        check(self.lib.kr_nrc_create(dev.h, C.byref(cfg), C.byref(h)), "kr_nrc_create")
        self.h = h

    def get_params(self):
        p = np.zeros(NRC_PARAMS, np.float32)
        check(self.lib.kr_nrc_get_params(self.h, _ptr(p)), "kr_nrc_get_params")
        return p

    def get_render_params(self):
        """The parameters renders query (the moving average of the trained ones)."""
        p = np.zeros(NRC_PARAMS, np.float32)
        check(self.lib.kr_nrc_get_render_params(self.h, _ptr(p)), "kr_nrc_get_render_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        check(self.lib.kr_nrc_set_params(self.h, _ptr(p)), "kr_nrc_set_params")

    def encode(self, raw):
        raw = np.ascontiguousarray(raw, np.float32)
        out = np.zeros((raw.shape[0], NRC_FEATURES), np.float32)
        check(self.lib.kr_nrc_encode(self.h, _ptr(raw), raw.shape[0], _ptr(out)), "kr_nrc_encode")
        return out

    def predict(self, raw):
        raw = np.ascontiguousarray(raw, np.float32)
        out = np.zeros((raw.shape[0], 3), np.float32)
        check(self.lib.kr_nrc_predict(self.h, _ptr(raw), raw.shape[0], _ptr(out)), "kr_nrc_predict")
        return out

    def gradient(self, raw, target):
        raw = np.ascontiguousarray(raw, np.float32)
        tgt = np.ascontiguousarray(target, np.float32)
        g = np.zeros(NRC_PARAMS, np.float32)
        loss = f32()
        check(self.lib.kr_nrc_gradient(self.h, _ptr(raw), _ptr(tgt), raw.shape[0], _ptr(g), C.byref(loss)),
              "kr_nrc_gradient")
        return g, loss.value

    def train(self, raw, target, steps=1):
        raw = np.ascontiguousarray(raw, np.float32)
        tgt = np.ascontiguousarray(target, np.float32)
        loss = f32()
        check(self.lib.kr_nrc_train(self.h, _ptr(raw), _ptr(tgt), raw.shape[0], steps, C.byref(loss)),
              "kr_nrc_train")
        return loss.value

    def read_pool(self, max_samples):
        raw = np.zeros((max_samples, NRC_RAW), np.float32)
        tgt = np.zeros((max_samples, 3), np.float32)
        cnt = u32()
        check(self.lib.kr_nrc_read_pool(self.h, max_samples, _ptr(raw), _ptr(tgt), C.byref(cnt)), "kr_nrc_read_pool")
        return raw[:cnt.value], tgt[:cnt.value]

    def stats(self):
        n, l = u64(), f32()
        check(self.lib.kr_nrc_get_stats(self.h, C.byref(n), C.byref(l)), "kr_nrc_get_stats")
        return n.value, l.value

    def free(self):
        if self.h:
            self.lib.kr_nrc_free(self.h)
            self.h = None


class Batch:
    """kr_batch handle: many (scene, camera, params) jobs rendered as one
    wavefront, replayed from one captured HIP graph.

    scenes: objects with .to_desc() (SceneData, host.LoadedScene);
    jobs: (scene index, Camera, Params) triples."""

    def __init__(self, dev: Device, scenes, jobs, tuning: BatchTuning | None = None):
        self.lib = dev.lib
        holders = [sc.to_desc() for sc in scenes]
        descs = (SceneDesc * len(holders))(*[hd.desc for hd in holders])
        js = (BatchJob * len(jobs))(*[BatchJob(k, cam, prm) for k, cam, prm in jobs])
        self.sizes = [(prm.width * prm.height) for _, _, prm in jobs]
        h = C.c_void_p()
        check(self.lib.kr_batch_create(dev.h, descs, len(holders), js, len(jobs),
                                       C.byref(tuning) if tuning is not None else None, C.byref(h)),
              "kr_batch_create")
        self.h = h

    def autotune(self, trial_spp: int = 0) -> BatchTuning:
        t = BatchTuning()
        check(self.lib.kr_batch_autotune(self.h, trial_spp, C.byref(t)), "kr_batch_autotune")
        return t

    def attach_nrc(self, job: int, nrc):
        """kr_batch_attach_nrc: the job's paths use `nrc` (None detaches)."""
        check(self.lib.kr_batch_attach_nrc(self.h, job, nrc.h if nrc is not None else None), "kr_batch_attach_nrc")

    def run(self):
        check(self.lib.kr_batch_run(self.h), "kr_batch_run")

    def sync(self):
        check(self.lib.kr_batch_sync(self.h), "kr_batch_sync")

    def reset(self):
        check(self.lib.kr_batch_reset(self.h), "kr_batch_reset")

    def read(self, job: int):
        n = self.sizes[job] if 0 <= job < len(self.sizes) else 0  # out of range: the library reports it
        rgb = np.zeros((n, 3), np.float64)
        valid = np.zeros(n, np.uint32)
        check(self.lib.kr_batch_read(self.h, job, _ptr(rgb), _ptr(valid)), "kr_batch_read")
        return rgb, valid

    def info(self) -> BatchInfo:
        i = BatchInfo()
        check(self.lib.kr_batch_get_info(self.h, C.byref(i)), "kr_batch_get_info")
        return i

    def free(self):
        if self.h:
            self.lib.kr_batch_free(self.h)
            self.h = None
