"""ctypes binding of the host library (include/koishi.h, libkoishi.so): the
reference's scene JSON + OBJ loader and the `cr` render flow."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import kr

LIB_PATH = os.path.join(kr.PKG_ROOT, "libkoishi.so")
CR_PATH = os.path.join(kr.PKG_ROOT, "cr")

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is not built: run __graft_entry__.build()")
        lib = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        lib.koishi_last_error.argtypes = [C.c_char_p, C.c_size_t]
        lib.koishi_set_logging.argtypes = [C.c_int]
        lib.koishi_scene_load.argtypes = [C.c_char_p, C.POINTER(vp)]
        lib.koishi_scene_free.argtypes = [vp]
        lib.koishi_scene_desc.argtypes = [vp, C.POINTER(kr.SceneDesc)]
        lib.koishi_scene_summary.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_char_p,
                                             C.c_size_t, C.c_char_p, C.c_size_t]
        lib.koishi_scene_camera.argtypes = [vp, C.c_uint32, C.POINTER(kr.Camera)]
        lib.koishi_scene_shader_name.argtypes = [vp, C.c_uint32, C.c_char_p, C.c_size_t]
        lib.koishi_write_png.argtypes = [C.c_char_p, C.c_uint32, C.c_uint32, vp, vp]
        _lib = lib
    return _lib


class HostError(RuntimeError):
    pass


def _err(lib):
    buf = C.create_string_buffer(2048)
    lib.koishi_last_error(buf, 2048)
    return buf.value.decode(errors="replace")


class LoadedScene:
    """A scene file parsed by libkoishi.so; `holder` is a kr_scene_desc for
    kr.GpuScene / the oracle (valid while this object lives)."""

    def __init__(self, path: str, log: bool = False):
        lib = load()
        lib.koishi_set_logging(1 if log else 0)
        self.lib = lib
        h = C.c_void_p()
        if lib.koishi_scene_load(path.encode(), C.byref(h)) != 0:
            raise HostError(_err(lib))
        self.h = h
        self.desc = kr.SceneDesc()
        lib.koishi_scene_desc(h, C.byref(self.desc))
        ncam, mb = C.c_uint32(), C.c_uint32()
        tb, kb = C.create_string_buffer(128), C.create_string_buffer(128)
        lib.koishi_scene_summary(h, C.byref(ncam), C.byref(mb), tb, 128, kb, 128)
        self.num_cameras, self.max_bounce = ncam.value, mb.value
        self.tracer, self.kernel = tb.value.decode(), kb.value.decode()
        self.holder = kr_holder(self.desc)

    def to_desc(self):
        return self.holder

    def camera_struct(self):
        return self.camera(0)

    def camera(self, i=0) -> kr.Camera:
        c = kr.Camera()
        if self.lib.koishi_scene_camera(self.h, i, C.byref(c)) != 0:
            raise HostError("no camera %d" % i)
        return c

    def shader_names(self):
        out = []
        buf = C.create_string_buffer(256)
        for i in range(self.desc.num_shaders):
            self.lib.koishi_scene_shader_name(self.h, i, buf, 256)
            out.append(buf.value.decode())
        return out

    def arrays(self):
        d = self.desc
        pos = np.ctypeslib.as_array(d.positions, (d.num_vertices, 3)).copy() if d.num_vertices else np.zeros((0, 3))
        nrm = np.ctypeslib.as_array(d.normals, (d.num_vertices, 3)).copy() if d.num_vertices else np.zeros((0, 3))
        idx = np.ctypeslib.as_array(d.indices, (d.num_triangles, 3)).copy() if d.num_triangles else \
            np.zeros((0, 3), np.uint32)
        prims = [d.prims[i] for i in range(d.num_prims)]
        shaders = [d.shaders[i] for i in range(d.num_shaders)]
        return pos, nrm, idx, prims, shaders

    def free(self):
        if self.h:
            self.lib.koishi_scene_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class kr_holder:
    """Same shape as scenegen.DescHolder; the arrays belong to the LoadedScene."""

    def __init__(self, desc):
        self.desc = desc
        self.keep = []


def write_png(path, w, h, rgb_sum, valid):
    lib = load()
    rgb = np.ascontiguousarray(rgb_sum, np.float64)
    val = np.ascontiguousarray(valid, np.uint32)
    if lib.koishi_write_png(path.encode(), w, h, rgb.ctypes.data_as(C.c_void_p), val.ctypes.data_as(C.c_void_p)):
        raise HostError(_err(lib))
