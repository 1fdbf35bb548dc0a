"""ctypes access to the CPU oracle (oracle/liboracle.so) and to the
reference-header shim (oracle/_ref/libkoishi_ref.so). Test infrastructure:
used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline only."""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "cuda-raytracer_amd"))
from koishi_amd import kr  # noqa: E402

ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libkoishi_ref.so")

_orc = None
_ref = None
D = C.c_double
PD = C.POINTER(C.c_double)
VP = C.c_void_p


def build_oracle():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def oracle():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            build_oracle()
        lib = C.CDLL(ORACLE_SO)
        for sfx in ("_f64", "_f32"):
            g = lambda n: getattr(lib, n + sfx)  # noqa: E731
            g("oracle_scene_create").argtypes = [C.POINTER(kr.SceneDesc), C.POINTER(VP)]
            g("oracle_scene_free").argtypes = [VP]
            g("oracle_scene_set_traversal").argtypes = [VP, C.c_int]
            g("oracle_scene_counts").argtypes = [VP, C.POINTER(C.c_uint64)]
            g("oracle_trace_closest").argtypes = [VP, VP, VP, C.c_uint32, VP, VP, VP, VP, VP, VP]
            g("oracle_trace_shadow").argtypes = [VP, VP, VP, VP, C.c_uint32, VP]
            g("oracle_render").argtypes = [VP, C.POINTER(kr.Camera), C.POINTER(kr.Params), C.c_uint32, C.c_uint32,
                                           C.c_int, VP, C.c_uint32, VP, VP, VP]
            g("oracle_render").restype = D
            g("oracle_path_radiance").argtypes = [VP, PD, PD, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32, PD]
            g("oracle_profile_path").argtypes = [VP, C.POINTER(kr.Camera), C.POINTER(kr.Params), C.c_uint32,
                                                 C.c_uint32, C.c_uint32, C.c_uint32, VP, VP, C.POINTER(C.c_uint32)]
            g("oracle_light_samples").argtypes = [VP, C.c_uint32, VP, PD]
            g("oracle_lens_ray").argtypes = [C.POINTER(kr.Camera), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                             C.c_uint32, C.c_uint32, PD, PD]
            g("oracle_ray_bbox").argtypes = [PD, PD, PD, PD]
            g("oracle_ray_triangle").argtypes = [PD, PD, PD, PD, PD, PD]
            g("oracle_seg_bbox").argtypes = [PD, PD, D, PD, PD]
            g("oracle_seg_triangle").argtypes = [PD, PD, D, PD, PD, PD]
            g("oracle_sphere_intersect").argtypes = [PD, D, PD, PD, PD, PD]
            g("oracle_sphere_sample").argtypes = [PD, D, D, D, PD, PD, PD]
            g("oracle_sample_cos").argtypes = [D, D, PD]
            g("oracle_beckmann_sample").argtypes = [D, D, D, PD]
            g("oracle_beckmann_f").argtypes = [PD, D]
            g("oracle_beckmann_f").restype = D
            g("oracle_reflect").argtypes = [PD, PD, PD]
            g("oracle_shader_sample").argtypes = [C.POINTER(kr.ShaderDesc), PD, D, D, D, D, PD, PD, PD]
            g("oracle_shader_eval").argtypes = [C.POINTER(kr.ShaderDesc), PD, PD, D, PD, PD]
            g("oracle_fresnel").argtypes = [D, PD]
            g("oracle_fresnel").restype = D
            g("oracle_refract").argtypes = [PD, PD, D, PD]
            g("oracle_refract").restype = C.c_int
            g("oracle_frame").argtypes = [PD, PD, PD, PD, PD, PD]
            g("oracle_emit_seg").argtypes = [PD, PD, PD, PD, PD]
        lib.oracle_philox.argtypes = [VP, VP, VP]
        lib.oracle_u01.argtypes = [C.c_uint32]
        lib.oracle_u01.restype = C.c_float
        _orc = lib
    return _orc


def ref():
    """The reference-header shim, or None when it was not built (no reference sources)."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        lib = C.CDLL(REF_SO)
        lib.ref_ray_bbox.argtypes = [PD, PD, PD, PD]
        lib.ref_ray_triangle.argtypes = [PD, PD, PD, PD, PD, PD]
        lib.ref_seg_bbox.argtypes = [PD, PD, D, PD, PD]
        lib.ref_seg_triangle.argtypes = [PD, PD, D, PD, PD, PD]
        lib.ref_sphere_intersect.argtypes = [PD, D, PD, PD, PD, PD]
        lib.ref_sphere_sample.argtypes = [PD, D, D, D, PD, PD, PD]
        lib.ref_sample_cos.argtypes = [D, D, PD]
        lib.ref_beckmann_sample.argtypes = [D, D, D, PD, PD]
        lib.ref_beckmann_f.argtypes = [PD, D]
        lib.ref_beckmann_f.restype = D
        lib.ref_cos_f.argtypes = [PD]
        lib.ref_cos_f.restype = D
        lib.ref_reflect.argtypes = [PD, PD, PD]
        lib.ref_lens_ray.argtypes = [C.c_int, D, PD, PD, PD, C.c_uint, C.c_uint, C.c_uint, C.c_uint, C.c_uint,
                                     C.c_uint, PD, PD]
        lib.ref_frame.argtypes = [PD, PD, PD, PD, PD, PD]
        lib.ref_emit_seg.argtypes = [PD, PD, PD, PD, PD]
        lib.ref_emit_ray.argtypes = [PD, PD, PD, PD]
        _ref = lib
    return _ref


def d3(v=(0.0, 0.0, 0.0)):
    return (D * 3)(*[float(x) for x in v])


def fn(name, real="f64"):
    return getattr(oracle(), f"{name}_{real}")


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class OracleScene:
    def __init__(self, scene, real: str = "f64", traversal: int = 0):
        self.real = real
        self.holder = scene.to_desc()
        h = C.c_void_p()
        rc = fn("oracle_scene_create", real)(C.byref(self.holder.desc), C.byref(h))
        if rc != 0:
            raise RuntimeError(f"oracle_scene_create failed: {rc}")
        self.h = h
        self.scene = scene
        fn("oracle_scene_set_traversal", real)(h, traversal)

    def counts(self):
        c = (C.c_uint64 * 4)()
        fn("oracle_scene_counts", self.real)(self.h, c)
        return dict(lights=c[0], prims=c[1], bvh_nodes=c[2], faces=c[3])

    def trace_closest(self, o, d, stats=False):
        o = np.ascontiguousarray(o, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        n = o.shape[0]
        t = np.empty(n)
        prim = np.empty(n, np.uint32)
        tri = np.empty(n, np.uint32)
        u = np.empty(n)
        v = np.empty(n)
        st = np.zeros(2, np.uint64)
        fn("oracle_trace_closest", self.real)(self.h, _p(o), _p(d), n, _p(t), _p(prim), _p(tri), _p(u), _p(v),
                                              _p(st) if stats else None)
        if stats:
            return t, prim, tri, u, v, st
        return t, prim, tri, u, v

    def trace_shadow(self, o, d, tmax):
        o = np.ascontiguousarray(o, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        tm = np.ascontiguousarray(tmax, np.float32)
        occ = np.empty(o.shape[0], np.uint8)
        fn("oracle_trace_shadow", self.real)(self.h, _p(o), _p(d), _p(tm), o.shape[0], _p(occ))
        return occ

    def render(self, params, spp_begin=0, spp_count=None, nthreads=0, pixels=None):
        cam = self.scene.camera_struct()
        spp_count = params.spp if spp_count is None else spp_count
        if pixels is None:
            npix = params.width * params.height
            pix = None
        else:
            pix = np.ascontiguousarray(pixels, np.uint32)
            npix = pix.shape[0]
        rgb = np.zeros((npix, 3))
        valid = np.zeros(npix, np.uint32)
        cnt = np.zeros(4, np.uint64)
        sec = fn("oracle_render", self.real)(self.h, C.byref(cam), C.byref(params), spp_begin, spp_count, nthreads,
                                             _p(pix) if pix is not None else None, npix, _p(rgb), _p(valid), _p(cnt))
        return rgb, valid, dict(seconds=sec, closest=int(cnt[0]), shadow=int(cnt[1]), nodes=int(cnt[2]),
                                tris=int(cnt[3]))

    def light_samples(self, light):
        pn = np.zeros((1024, 6))
        pdf = C.c_double()
        rc = fn("oracle_light_samples", self.real)(self.h, light, _p(pn), C.byref(pdf))
        if rc != 0:
            raise RuntimeError("not a mesh light")
        return pn, pdf.value

    def profile_path(self, params, x, y, k, depth=8):
        """The profiler slice of path (x, y, k): records (depth, 14) as
        (o, d, f, L, theta, shader), the path radiance and the loop counter."""
        cam = self.scene.camera_struct()
        rec = np.zeros((depth, 14))
        L = np.zeros(3)
        b = C.c_uint32(0)
        fn("oracle_profile_path", self.real)(self.h, C.byref(cam), C.byref(params), x, y, k, depth, _p(rec), _p(L),
                                             C.byref(b))
        return rec, L, b.value

    def free(self):
        if self.h:
            fn("oracle_scene_free", self.real)(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def lens_rays(scene, params, real="f64", spp_index=0, pixels=None):
    """Camera rays of the oracle's lens for every pixel (or `pixels`) at sample spp_index."""
    cam = scene.camera_struct()
    w, h = params.width, params.height
    pix = np.arange(w * h) if pixels is None else np.asarray(pixels)
    o = np.zeros((pix.shape[0], 3))
    d = np.zeros((pix.shape[0], 3))
    f = fn("oracle_lens_ray", real)
    oo, dd = d3(), d3()
    for k, p in enumerate(pix):
        f(C.byref(cam), w, h, params.spp, int(p % w), int(p // w), spp_index, oo, dd)
        o[k] = list(oo)
        d[k] = list(dd)
    return o, d


def philox(ctr, key):
    out = (C.c_uint32 * 4)()
    oracle().oracle_philox((C.c_uint32 * 4)(*ctr), (C.c_uint32 * 2)(*key), out)
    return list(out)
