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
