"""The drop-in `cr` end to end on the GPU: a scene JSON + OBJ written by the
generators, rendered with -t HIPWavefront, -t HIPBatched, -t CPUMulticore
(the reference's tracer name) and the file's own tracer, must give the same raw FP64 image as
kr_render_radiance on the loaded scene, and a PNG of the reference's 8-bit
conversion."""
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

from koishi_amd import host, kr, scenegen as sg

pytestmark = pytest.mark.gpu


def _read_raw(path):
    with open(path, "rb") as f:
        b = f.read()
    assert b[:6] == b"KRRAW1"
    w, h = struct.unpack("<II", b[8:16])
    mean = np.frombuffer(b[16:16 + 24 * w * h], np.float64).reshape(w * h, 3)
    valid = np.frombuffer(b[16 + 24 * w * h:], np.uint32)
    return w, h, mean, valid


def _read_png(path):
    """Decodes the 8-bit RGB PNG `cr` writes (filter type 0 rows)."""
    with open(path, "rb") as f:
        b = f.read()
    assert b[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w, h = 8, b"", 0, 0
    while pos < len(b):
        n, = struct.unpack(">I", b[pos:pos + 4])
        kind = b[pos + 4:pos + 8]
        data = b[pos + 8:pos + 8 + n]
        if kind == b"IHDR":
            w, h = struct.unpack(">II", data[:8])
        elif kind == b"IDAT":
            idat += data
        pos += 12 + n
    raw = zlib.decompress(idat)
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 3 * w)
    assert (rows[:, 0] == 0).all()
    return rows[:, 1:].reshape(h * w, 3)


def test_cr_tracers_agree_with_the_c_abi(tmp_path, device):
    s = sg.mesh_scene(freq=16)
    path = s.write(str(tmp_path), "mesh")
    w, h, spp = 80, 48, 8
    raws = {}
    for tracer in ("HIPWavefront", "HIPBatched", "CPUMulticore", None):
        out_png = str(tmp_path / f"{tracer}.png")
        out_raw = str(tmp_path / f"{tracer}.krraw")
        cmd = [host.CR_PATH, path, "-o", out_png, "-s", str(spp), f"--resolution={w}x{h}", "--seed", "5",
               "--dump-raw", out_raw]
        if tracer:
            cmd += ["-t", tracer]
        res = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
        assert res.returncode == 0, res.stdout + res.stderr
        assert "Render finished successfully" in res.stdout
        raws[tracer] = _read_raw(out_raw)
        png = _read_png(out_png)
        mean = raws[tracer][2]
        # the reference's conversion: clamp to [0, 1], scale by 255, truncate
        expect = (np.clip(mean, 0.0, 1.0) * 255.0).astype(np.float32).astype(np.uint8)
        assert np.abs(png.astype(int) - expect.astype(int)).max() <= 1
    for t in ("HIPBatched", "CPUMulticore", None):
        assert np.array_equal(raws[t][2], raws["HIPWavefront"][2]) and np.array_equal(raws[t][3],
                                                                                     raws["HIPWavefront"][3])
    # the same render through the C ABI on the scene as cr loaded it
    ls = host.LoadedScene(path)
    gs = kr.GpuScene(device, ls.to_desc())
    r = kr.Render(gs, ls.camera(0), kr.Params(w, h, spp, ls.max_bounce, 5, 0, 0))
    r.run(0, spp)
    rgb, val = r.read()
    r.free()
    gs.free()
    ls.free()
    assert np.array_equal(val, raws["HIPWavefront"][3])
    assert np.array_equal(rgb / np.maximum(val, 1)[:, None], raws["HIPWavefront"][2])


def test_cr_writes_the_profile(tmp_path, device):
    """A scene file with a "profiler" key: `cr` traces only the area and
    writes the reference's profile text (Profiler::~Profiler with
    RadianceKernel::Slice::writeSlice) for every slice of the area."""
    import json

    s = sg.room_scene(width=24, height=18, spp=2, freq=8)
    s.shaders["light"] = ["Emission", {"color": [1.0, 1.0, 1.0], "strength": 150.0}]
    path = s.write(str(tmp_path), "room")
    doc = json.load(open(path))
    prof = str(tmp_path / "room.prof")
    area, depth = [2, 1, 22, 17], 5
    doc["profiler"] = {"output": prof, "area": area, "props": {"depth": depth}}
    json.dump(doc, open(path, "w"))
    w, h, spp = 24, 18, 2
    out_raw = str(tmp_path / "room.krraw")
    res = subprocess.run([host.CR_PATH, path, "-o", str(tmp_path / "room.png"), "-s", str(spp),
                          f"--resolution={w}x{h}", "--seed", "5", "--dump-raw", out_raw],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "Enabled in [ 2, 1, 22, 17 ]" in res.stdout and "Written to '" + prof + "'" in res.stdout
    _, _, mean, _ = _read_raw(out_raw)
    inside = np.zeros((h, w), bool)
    inside[1:17, 2:22] = True
    assert not mean[~inside.ravel()].any()
    # the same slices through the C ABI
    ls = host.LoadedScene(path)
    gs = kr.GpuScene(device, ls.to_desc())
    r = kr.Render(gs, ls.camera(0), kr.Params(w, h, spp, ls.max_bounce, 5, 0, 0))
    r.set_profiler(area, depth)
    r.run(0, spp)
    L, bounce, recs = r.read_slices()
    r.free()
    gs.free()
    ls.free()
    lines = open(prof).read().split("\n")
    assert lines[-1] == ""
    i, s_idx, fireflies = 0, 0, 0
    aw = area[2] - area[0]
    while i < len(lines) - 1:
        px, k = s_idx // spp, s_idx % spp
        assert lines[i] == f"at [ {area[0] + px % aw}, {area[1] + px // aw}, {k} ]: ", lines[i]
        i += 1
        if np.all(L[s_idx] >= 100.0):
            fireflies += 1
            assert lines[i] == f"bounce: {bounce[s_idx]} / {depth}"
            i += 1
            for j in range(min(int(bounce[s_idx]), depth)):
                rec = recs[s_idx][j]
                assert lines[i].startswith("ray: [ ") and lines[i + 1].startswith("f: [ ")
                assert lines[i + 2].startswith("theta: ") and lines[i + 3].startswith("L: [ ")
                assert lines[i + 4] == f"shader: {rec['shader']}"
                got = [float(x) for x in lines[i + 3][len("L: [ "):-2].split(", ")]
                assert np.allclose(got, rec["L"], rtol=1e-5)
                i += 5
        s_idx += 1
    assert s_idx == (area[2] - area[0]) * (area[3] - area[1]) * spp
    assert fireflies > 0
