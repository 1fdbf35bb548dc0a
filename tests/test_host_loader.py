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


# ---------------------------------------------------------------- PLY import --
def _ply_bytes(fmt, pos, faces, nrm=None, extra=False):
    """A PLY file of the given format with float vertices (x y z [nx ny nz]
    [uchar red green blue]), uchar/int face lists and, with `extra`, an
    unused edge element after the faces."""
    head = ["ply", f"format {fmt} 1.0", "comment test", f"element vertex {len(pos)}",
            "property float x", "property float y", "property float z"]
    if nrm is not None:
        head += ["property float nx", "property float ny", "property float nz"]
    if extra:
        head += ["property uchar red", "property uchar green", "property uchar blue"]
    head += [f"element face {len(faces)}", "property list uchar int vertex_indices"]
    if extra:
        head += ["element edge 1", "property int vertex1", "property int vertex2"]
    head += ["end_header"]
    out = ("\n".join(head) + "\n").encode()
    if fmt == "ascii":
        rows = []
        for i, p in enumerate(pos):
            r = [repr(float(np.float32(x))) for x in p]
            if nrm is not None:
                r += [repr(float(np.float32(x))) for x in nrm[i]]
            if extra:
                r += ["10", "20", "30"]
            rows.append(" ".join(r))
        for f in faces:
            rows.append(" ".join([str(len(f))] + [str(i) for i in f]))
        if extra:
            rows.append("0 1")
        return out + ("\n".join(rows) + "\n").encode()
    e = "<" if fmt == "binary_little_endian" else ">"
    for i, p in enumerate(pos):
        out += struct.pack(e + "3f", *p)
        if nrm is not None:
            out += struct.pack(e + "3f", *nrm[i])
        if extra:
            out += bytes([10, 20, 30])
    for f in faces:
        out += struct.pack(e + "B" + "i" * len(f), len(f), *f)
    if extra:
        out += struct.pack(e + "2i", 0, 1)
    return out


def _load_mesh(tmp_path, name, payload):
    (tmp_path / name).write_bytes(payload) if isinstance(payload, bytes) else (tmp_path / name).write_text(payload)
    doc = {"scene": {"camera": [{"lens": "pinhole", "fovx": 60, "position": [0, -5, 1], "target": [0, 1, 0]}],
                     "assets": [["import", {"path": name}]],
                     "shaders": {"DefaultMaterial": ["Lambert", {"color": [0.5, 0.5, 0.5]}]}},
           "tracer": ["CPUMulticore", {"kernel": ["Radiance", {"maxBounce": 2}]}]}
    p = tmp_path / "s.json"
    p.write_text(json.dumps(doc))
    ls = host.LoadedScene(str(p))
    return ls


def _corner_geometry(ls):
    """Triangle corners and normals, and the primitives copied out of the
    loaded scene (whose arrays die with it)."""
    pos, nrm, idx, prims, _ = ls.arrays()
    pr = [(int(p.type), int(p.shader), int(p.tri_begin), int(p.tri_count)) for p in prims]
    return pos[idx.astype(np.int64)], nrm[idx.astype(np.int64)], pr, ls.shader_names()


def test_ply_formats_and_obj_agree(tmp_path):
    """The same quad + pentagon with normals as ascii, little- and big-endian
    PLY (with extra color properties and an extra element to skip) and as OBJ:
    identical triangle corners and normals, one mesh primitive on the shader
    "DefaultMaterial", polygons split into fans."""
    rng = np.random.default_rng(3)
    pos = rng.uniform(-1, 1, (9, 3)).astype(np.float32)
    nrm = rng.normal(size=(9, 3))
    nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    faces = [[0, 1, 2, 3], [4, 5, 6, 7, 8]]
    got = []
    for fmt in ("ascii", "binary_little_endian", "binary_big_endian"):
        ls = _load_mesh(tmp_path, f"m_{fmt}.ply", _ply_bytes(fmt, pos, faces, nrm, extra=True))
        got.append(_corner_geometry(ls))
    obj = "".join("v " + " ".join(repr(float(x)) for x in p) + "\n" for p in pos)
    obj += "".join("vn " + " ".join(repr(float(x)) for x in n) + "\n" for n in nrm)
    obj += "".join("f " + " ".join(f"{i + 1}//{i + 1}" for i in f) + "\n" for f in faces)
    got.append(_corner_geometry(_load_mesh(tmp_path, "m.obj", obj)))
    cp, cn, prims, names = got[0]
    assert cp.shape == (5, 3, 3)  # 2 + 3 fan triangles
    fan = [[0, 1, 2], [0, 2, 3], [4, 5, 6], [4, 6, 7], [4, 7, 8]]
    assert np.array_equal(cp, pos[np.array(fan)]) and np.array_equal(cn, nrm[np.array(fan)])
    assert len(prims) == 1 and prims[0][3] == 5 and names[prims[0][1]] == "DefaultMaterial"
    for p2, n2, pr2, nm2 in got[1:]:
        assert np.array_equal(p2, cp) and np.array_equal(n2, cn)
        assert len(pr2) == 1 and nm2[pr2[0][1]] == "DefaultMaterial"


@pytest.mark.parametrize("fmt", ["ply", "obj"])
def test_generated_normals_are_unit_face_normal_sums(tmp_path, fmt):
    """Without normals in the file, a vertex gets the normalised sum of the
    UNIT normals of the faces touching its position (assimp's
    GenSmoothNormals), not an area-weighted sum: a tent of one large and one
    small triangle sharing an edge has normalize(n1 + n2) on the shared
    corners, where area weighting would lean towards the large face's normal;
    the other corners keep their own face's normal."""
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 4, 2], [0.3, -0.5, 0.4]], np.float32)
    faces = [[0, 1, 2], [1, 0, 3]]
    if fmt == "ply":
        ls = _load_mesh(tmp_path, "t.ply", _ply_bytes("ascii", pos, faces))
    else:
        obj = "".join("v " + " ".join(repr(float(x)) for x in p) + "\n" for p in pos)
        obj += "".join("f " + " ".join(str(i + 1) for i in f) + "\n" for f in faces)
        ls = _load_mesh(tmp_path, "t.obj", obj)
    cp, cn, _, _ = _corner_geometry(ls)
    p64 = pos.astype(np.float64)

    def unit(f):
        n = np.cross(p64[f[1]] - p64[f[0]], p64[f[2]] - p64[f[0]])
        return n / np.linalg.norm(n)
    n1, n2 = unit(faces[0]), unit(faces[1])
    shared = (n1 + n2) / np.linalg.norm(n1 + n2)
    want = {0: shared, 1: shared, 2: n1, 3: n2}
    for t, f in enumerate(faces):
        for k, v in enumerate(f):
            assert np.allclose(cp[t, k], pos[v])
            assert np.allclose(cn[t, k], want[v], atol=1e-6), (v, cn[t, k], want[v])
    a1 = np.linalg.norm(np.cross(p64[1] - p64[0], p64[2] - p64[0]))
    a2 = np.linalg.norm(np.cross(p64[0] - p64[1], p64[3] - p64[1]))
    area = (a1 * n1 + a2 * n2) / np.linalg.norm(a1 * n1 + a2 * n2)
    assert np.abs(area - shared).max() > 1e-2  # the two rules differ here


def test_bad_ply_is_rejected(tmp_path):
    pos = np.zeros((3, 3), np.float32)
    for payload, word in ((b"plx\n", "not a PLY"), (_ply_bytes("ascii", pos, [[0, 1, 5]]), "out of range"),
                          (_ply_bytes("binary_little_endian", pos, [[0, 1, 2]])[:-3], "end of data")):
        with pytest.raises(RuntimeError, match=word):
            _load_mesh(tmp_path, "b.ply", payload)
