"""Host side (libkoishi.so, `cr`): the reference's scene JSON + OBJ import,
the 8-bit PNG output and the command line -- on the CPU, no device calls."""
import json
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

import oracle_lib as ol
from koishi_amd import host, kr, scenegen as sg


def _tri_geometry(pos, nrm, idx):
    return pos[idx.astype(np.int64)], nrm[idx.astype(np.int64)]


@pytest.mark.parametrize("maker", [lambda: sg.sphere_scene(), lambda: sg.mesh_scene(freq=12)])
def test_scene_file_round_trip(tmp_path, maker):
    s = maker()
    path = s.write(str(tmp_path))
    ls = host.LoadedScene(path)
    pos, nrm, idx, prims, shaders = ls.arrays()
    # triangle geometry identical bit for bit (the OBJ reader re-indexes)
    gp, gn = _tri_geometry(s.positions, s.normals, s.indices)
    lp, ln = _tri_geometry(pos, nrm, idx)
    assert np.array_equal(gp, lp) and np.array_equal(gn, ln)
    # primitives in the reference order: imported meshes, then spheres
    assert len(prims) == len(s.meshes) + len(s.spheres)
    names = ls.shader_names()
    for p, (_, shader, b, c) in zip(prims, s.meshes):
        assert p.type == kr.PRIM_MESH and p.tri_begin == b and p.tri_count == c and names[p.shader] == shader
    for p, (cen, r, shader) in zip(prims[len(s.meshes):], s.spheres):
        assert p.type == kr.PRIM_SPHERE and list(p.center) == cen and p.radius == r and names[p.shader] == shader
    for i, n in enumerate(names):
        want = sg.shader_desc_from_config(s.shaders[n])
        got = shaders[i]
        assert (got.type, got.distribution, list(got.color), list(got.R), got.roughness, got.strength) == \
            (want.type, want.distribution, list(want.color), list(want.R), want.roughness, want.strength)
    cam, want = ls.camera(0), s.camera_struct()
    assert list(cam.position) == list(want.position) and list(cam.target) == list(want.target)
    assert cam.fovx == want.fovx and list(cam.upaxis) == list(want.upaxis)
    assert ls.max_bounce == s.max_bounce and ls.tracer == "HIPWavefront" and ls.kernel == "Radiance"
    assert ls.desc.flags == 0
    # the oracle renders both descriptions identically
    p = s.params(seed=3, width=24, height=16, spp=4)
    a = ol.OracleScene(s, "f64", traversal=1).render(p)
    b = ol.OracleScene(ls, "f64", traversal=1).render(p)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def _write_scene(tmp_path, doc, obj=None):
    if obj is not None:
        (tmp_path / "m.obj").write_text(obj)
    p = tmp_path / "s.json"
    p.write_text(json.dumps(doc) if not isinstance(doc, str) else doc)
    return str(p)


QUAD_OBJ = """o floor
v -1 -1 0
v 1 -1 0
v 1 1 0
v -1 1 0
usemtl floor
f 1 2 3 4
o lamp
v -0.5 -0.5 2
v 0.5 -0.5 2
v 0.5 0.5 2
usemtl lamp
f -3 -1 -2
"""


def _base_doc(tmp_path, shaders=None, assets=None):
    return {"scene": {"camera": [{"position": [0, -3, 1], "target": [0, 1, -0.3]}],
                      "assets": assets if assets is not None else [["import", {"path": str(tmp_path / "m.obj")}]],
                      "shaders": shaders if shaders is not None else
                      {"floor": ["Lambert", {}], "lamp": ["Emission", {"strength": 7}]}}}


def test_obj_polygons_negative_indices_and_generated_normals(tmp_path):
    ls = host.LoadedScene(_write_scene(tmp_path, _base_doc(tmp_path), QUAD_OBJ))
    pos, nrm, idx, prims, shaders = ls.arrays()
    assert [p.tri_count for p in prims] == [2, 1]  # quad fanned into 2 triangles
    assert np.allclose(nrm[idx[:2].ravel()], [0, 0, 1])  # smooth normals of a flat quad
    assert np.allclose(nrm[idx[2]], [0, 0, -1])  # the lamp face is wound clockwise
    names = ls.shader_names()
    assert names == ["floor", "lamp"]
    assert shaders[1].type == kr.SHADER_EMISSION and shaders[1].strength == 7
    assert shaders[0].type == kr.SHADER_LAMBERT and list(shaders[0].color) == [1, 1, 1]
    cam = ls.camera(0)
    assert cam.fovx == 90 and list(cam.upaxis) == [0, 0, 1] and cam.lens == kr.LENS_PINHOLE
    assert ls.tracer == "CPUMulticore" and ls.max_bounce == 16  # reference defaults


@pytest.mark.parametrize("doc_fn,msg", [
    (lambda t: {"scene": {"assets": [["teapot", {}]]}}, "unknown asset type: teapot"),
    (lambda t: {"scene": {"assets": [["areaLight", {}]]}}, "unknown light type: areaLight"),
    (lambda t: {"scene": {"assets": [["import", {"path": str(t / "m.obj")}], ["import", {"path": str(t / "m.obj")}]]}},
     "importing multiple scenes is not allowed"),
    (lambda t: _base_doc(t, shaders={"floor": ["Velvet", {}], "lamp": ["Emission", {}]}), "no such type: Velvet"),
    (lambda t: _base_doc(t, shaders={"floor": ["GlossyBSDF", {}], "lamp": ["Emission", {}]}),
     'No such key named "distribution"'),
    (lambda t: _base_doc(t, shaders={"floor": ["GlassBSDF", {}], "lamp": ["Emission", {}]}), 'No such key named "ior"'),
    (lambda t: _base_doc(t, shaders={"floor": ["Mixed", {"fac": ["Constant", {"value": 0.5}],
                                                          "shaders": [["Lambert", {}]]}],
                                     "lamp": ["Emission", {}]}), "Mixed needs"),
    (lambda t: _base_doc(t, shaders={"floor": ["Mixed", {"fac": ["Noise", {}], "shaders": [["Lambert", {}]] * 2}],
                                     "lamp": ["Emission", {}]}), "no such type: Noise"),
    (lambda t: {"scene": {"camera": [{"lens": "fisheye"}], "assets": []}}, "unknown lens type: fisheye"),
    (lambda t: {"scene": {"assets": [["sphere", {"o": [0, 0, 0], "r": 1}]]}}, 'No such key named "shader"'),
    (lambda t: "{\"scene\": [1, 2,,]}", "json parse error at line 1"),
])
def test_loader_errors(tmp_path, doc_fn, msg):
    path = _write_scene(tmp_path, doc_fn(tmp_path), QUAD_OBJ)
    with pytest.raises(host.HostError) as e:
        host.LoadedScene(path)
    assert msg in str(e.value)


def test_missing_shader_config_disables_light_promotion(tmp_path):
    """A shader name with no configuration makes the reference skip light
    promotion (src/core/meta/scene.cu:193-216); a primitive that uses such a
    shader is an error here (a null shader in the reference)."""
    doc = _base_doc(tmp_path)
    doc["scene"]["assets"].append(["sphere", {"o": [0, 0, 5], "r": 0.5, "shader": "ghost"}])
    with pytest.raises(host.HostError, match="ghost"):
        host.LoadedScene(_write_scene(tmp_path, doc, QUAD_OBJ))
    # faces before any usemtl belong to DefaultMaterial
    no_mtl = QUAD_OBJ.replace("usemtl floor\n", "")
    doc = _base_doc(tmp_path)
    with pytest.raises(host.HostError, match="DefaultMaterial"):
        host.LoadedScene(_write_scene(tmp_path, doc, no_mtl))
    doc["scene"]["shaders"]["DefaultMaterial"] = ["Lambert", {}]
    ls = host.LoadedScene(_write_scene(tmp_path, doc, no_mtl))
    assert ls.shader_names() == ["DefaultMaterial", "lamp"] and ls.desc.flags == 0
    # a material mentioned without faces and without configuration: no lights
    ls = host.LoadedScene(_write_scene(tmp_path, _base_doc(tmp_path), QUAD_OBJ + "usemtl unused_name\n"))
    assert "unused_name" in ls.shader_names()
    assert ls.desc.flags == kr.KR_SCENE_NO_AREA_LIGHTS
    assert ol.OracleScene(ls, "f64").counts()["lights"] == 0


def _read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w, h = 8, b"", 0, 0
    while pos < len(data):
        n, typ = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        crc = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
        assert crc == zlib.crc32(typ + body) & 0xFFFFFFFF
        if typ == b"IHDR":
            w, h = struct.unpack(">II", body[:8])
            assert body[8:10] == b"\x08\x02"
        elif typ == b"IDAT":
            idat += body
        pos += 12 + n
    raw = zlib.decompress(idat)
    rows = np.frombuffer(raw, np.uint8).reshape(h, 3 * w + 1)
    assert (rows[:, 0] == 0).all()
    return rows[:, 1:].reshape(h, w, 3)


def test_png_matches_reference_byte_conversion(tmp_path):
    w, h = 5, 3
    rng = np.random.default_rng(0)
    valid = rng.integers(0, 4, w * h).astype(np.uint32)
    mean = rng.uniform(-0.5, 1.5, (w * h, 3))
    mean[0] = [np.nan, 0.25, 2.0]
    rgb = mean * valid[:, None]
    host.write_png(str(tmp_path / "a.png"), w, h, rgb, valid)
    img = _read_png(str(tmp_path / "a.png")).reshape(-1, 3)
    with np.errstate(invalid="ignore", divide="ignore"):
        m = (rgb / valid[:, None]).astype(np.float32)
    m = np.where(m > 0, m, 0).astype(np.float32)
    m = np.where(m < 1, m, 1).astype(np.float32)
    want = (m * np.float32(255)).astype(np.uint8)
    assert np.array_equal(img, want)


def test_cr_cli_list_help_and_errors(tmp_path):
    cr = host.CR_PATH
    out = subprocess.run([cr, "-l"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "HIPWavefront" in out.stdout and "Radiance" in out.stdout
    out = subprocess.run([cr, "-h"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "--resolution" in out.stdout
    out = subprocess.run([cr, str(tmp_path / "missing.json"), "-o", str(tmp_path / "x.png")], capture_output=True,
                         text=True, timeout=60)
    assert out.returncode == 1 and "[ fatal ]" in out.stdout and "cannot open" in out.stdout
    path = _write_scene(tmp_path, _base_doc(tmp_path), QUAD_OBJ)
    out = subprocess.run([cr, path, "-k", "Normal"], capture_output=True, text=True, timeout=60,
                         env=dict(os.environ, HIP_VISIBLE_DEVICES="-1"))
    assert out.returncode == 1 and "[ fatal ]" in out.stdout
    out = subprocess.run([cr, path, "-t", "OptiX"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 1 and "no such type: OptiX" in out.stdout
    out = subprocess.run([cr, path, "--bogus"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 1 and "does not exist" in out.stdout
