"""Sanitizer runs of the CPU code, built here with g++ (no GPU code):

- AddressSanitizer + UndefinedBehaviorSanitizer on the host parsers
  (host/json.cpp, host/scene_loader.cpp: the scene JSON, OBJ and PLY
  readers), fed valid scene files and seeded malformed variants of them
  (truncated at random points, random bytes changed);
- ThreadSanitizer on the CPU oracle's multithreaded render
  (oracle/oracle_capi.cpp with the FP64 instantiation), 4 threads.

A sanitizer report makes the harness exit nonzero (-fno-sanitize-recover)."""
import os
import subprocess

import numpy as np
import pytest

from koishi_amd import scenegen as sg

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HOST = os.path.join(ROOT, "cuda-raytracer_amd", "host")
INC = os.path.join(ROOT, "include")


def _build(tmp_path, name, sources, flags, defines=()):
    exe = str(tmp_path / name)
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", *flags, "-I", INC, "-I", HOST,
           "-I", os.path.join(ROOT, "oracle"), *[f"-D{d}" for d in defines], *sources, "-o", exe, "-lpthread"]
    subprocess.run(cmd, check=True, timeout=600, capture_output=True)
    return exe


def _mutants(data: bytes, rng, n):
    out = []
    for k in range(n):
        if k % 2 == 0:
            out.append(data[:int(rng.integers(0, len(data)))])
        else:
            b = bytearray(data)
            for _ in range(int(rng.integers(1, 8))):
                b[int(rng.integers(0, len(b)))] = int(rng.integers(0, 256))
            out.append(bytes(b))
    return out


def test_host_parsers_under_asan_ubsan(tmp_path):
    exe = _build(tmp_path, "host_parsers", [os.path.join(HERE, "sanitize", "host_parsers.cpp"),
                                            os.path.join(HOST, "json.cpp"), os.path.join(HOST, "scene_loader.cpp")],
                 ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"])
    rng = np.random.default_rng(5)
    files = []
    # valid scenes (scene JSON + OBJ), and the same meshes as PLY
    for k, s in enumerate([sg.sphere_scene(), sg.mesh_scene(freq=6), sg.materials_scene()]):
        d = tmp_path / f"s{k}"
        d.mkdir()
        path = s.write(str(d))
        files.append(path)
        text = open(path, "rb").read()
        obj = [f for f in os.listdir(d) if f.endswith(".obj")]
        for j, m in enumerate(_mutants(text, rng, 24)):
            p = d / f"bad{j}.json"
            p.write_bytes(m)
            files.append(str(p))
        for o in obj:
            raw = (d / o).read_bytes()
            for j, m in enumerate(_mutants(raw, rng, 12)):
                dm = tmp_path / f"s{k}_obj{j}"
                dm.mkdir()
                (dm / o).write_bytes(m)
                (dm / os.path.basename(path)).write_bytes(text)
                files.append(str(dm / os.path.basename(path)))
    from test_host_loader import _ply_bytes
    pos = rng.uniform(-1, 1, (40, 3)).astype(np.float32)
    faces = [list(rng.choice(40, int(rng.integers(3, 6)), replace=False)) for _ in range(30)]
    for fmt in ("ascii", "binary_little_endian", "binary_big_endian"):
        data = _ply_bytes(fmt, pos, faces, extra=True)
        for j, m in enumerate([data] + _mutants(data, rng, 16)):
            dm = tmp_path / f"ply_{fmt}_{j}"
            dm.mkdir()
            (dm / "m.ply").write_bytes(m)
            doc = ('{"scene": {"camera": [{"position": [0, -5, 1], "target": [0, 1, 0]}], '
                   '"assets": [["import", {"path": "m.ply"}]], "shaders": {"DefaultMaterial": ["Lambert", {}]}}, '
                   '"tracer": ["CPUMulticore", {}]}')
            (dm / "s.json").write_text(doc)
            files.append(str(dm / "s.json"))
    # strings that need escaping survive dump -> parse
    (tmp_path / "esc.json").write_text('{"a\\"b": ["x\\\\y\\n\\t\\u0001", 1.5e-3, true, null, {}]}')
    files.append(str(tmp_path / "esc.json"))
    out = subprocess.run([exe, *files], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-4000:]
    assert out.stdout.startswith("ok"), out.stdout
    parsed = int(out.stdout.split()[2])
    loaded = int(out.stdout.split()[4])
    assert parsed >= 4 and loaded >= 4, out.stdout


def test_oracle_threads_under_tsan(tmp_path):
    exe = _build(tmp_path, "oracle_threads", [os.path.join(HERE, "sanitize", "oracle_threads.cpp"),
                                              os.path.join(ROOT, "oracle", "oracle_capi.cpp"),
                                              os.path.join(HOST, "json.cpp"), os.path.join(HOST, "scene_loader.cpp")],
                 ["-fsanitize=thread"], ["ORACLE_REAL=double", "ORACLE_SUFFIX=_f64", "ORACLE_WITH_PHILOX"])
    s = sg.mesh_scene(freq=8, width=40, height=24, spp=4)
    path = s.write(str(tmp_path))
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    out = subprocess.run([exe, path], capture_output=True, text=True, timeout=600, env=env)
    assert out.returncode == 0, (out.returncode, out.stderr[-4000:])
    assert out.stdout.startswith("ok"), out.stdout
