"""The wide BVH builder and the compressed node encoding on the host (no
GPU): built from the library's own sources with g++ into a checker that
walks the tree (every triangle in exactly one leaf, inside every box on its
path) and decodes every compressed child box exactly as the kernels do (it
must contain the FP32 box). Geometry: the mesh workload, a cloud of tiny
triangles far from the origin (boxes much thinner than one FP32 ulp of the
coordinates would allow an 8-bit grid to resolve), and degenerate flat and
collinear triangles."""
import os
import struct
import subprocess

import numpy as np
import pytest

from koishi_amd import scenegen as sg

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "cuda-raytracer_amd", "csrc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("bvh") / "bvh_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "bvh_check.cpp"),
                    os.path.join(CSRC, "bvh_build.cpp"), "-o", exe, "-lpthread"], check=True, timeout=300)
    return exe


def _run(checker, tmp_path, tv, width):
    path = str(tmp_path / f"tris_{width}.bin")
    tv = np.ascontiguousarray(tv, np.float32).reshape(-1, 9)
    with open(path, "wb") as f:
        f.write(struct.pack("<I", tv.shape[0]) + tv.tobytes())
    out = subprocess.run([checker, path, str(width)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and out.stdout.startswith("ok"), out.stdout + out.stderr
    return out.stdout


def _mesh_tris():
    s = sg.mesh_scene(freq=30)
    p = np.asarray(s.positions, np.float32)
    i = np.asarray(s.indices, np.int64).reshape(-1, 3)
    return p[i].reshape(-1, 9)


def _far_tiny_tris(seed=3, n=20000):
    rng = np.random.default_rng(seed)
    c = np.array([1.0e5, -3.0e4, 7.5e3]) + rng.uniform(-2.0, 2.0, (n, 1, 3))
    return (c + rng.normal(0, 1e-3, (n, 3, 3))).astype(np.float32).reshape(-1, 9)


def _degenerate_tris(seed=4, n=3000):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 1, (n, 3))
    b = rng.uniform(-1, 1, (n, 3))
    flat = np.stack([a, b, a], axis=1)  # repeated vertex
    line = np.stack([a, (a + b) / 2, b], axis=1)  # collinear
    plane = np.stack([a, b, rng.uniform(-1, 1, (n, 3))], axis=1)
    plane[:, :, 2] = 0.25  # all in one axis-aligned plane: zero-extent boxes
    return np.concatenate([flat, line, plane]).astype(np.float32).reshape(-1, 9)


@pytest.mark.parametrize("width", [2, 4, 8])
@pytest.mark.parametrize("geometry", ["mesh", "far_tiny", "degenerate"])
def test_bvh_and_compressed_nodes(checker, tmp_path, width, geometry):
    tv = {"mesh": _mesh_tris, "far_tiny": _far_tiny_tris, "degenerate": _degenerate_tris}[geometry]()
    _run(checker, tmp_path, tv, width)
