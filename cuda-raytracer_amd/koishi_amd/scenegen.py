"""Seeded synthetic scenes for the benchmark workloads, in the reference's
scene format.

Every scene is built Y-up as the workload describes it and then rotated by
R_x(+90 deg): (x, y, z) -> (x, -z, y), into the reference's Z-up convention
(camera upaxis (0, 0, 1), src/util/config.hpp:32). Without the rotation every
normal along +-Y would give a NaN tangent frame
(u = normalize(cross(n, (0, 1, 0))), src/core/meta/scene.hpp:49).

Randomness is SplitMix64 (Steele, Lea, Flood, OOPSLA 2014) so that the same
seed gives the same scene on any machine.

A SceneData converts to the flat kr_scene_desc of include/kr_api.h
(`to_desc`) and writes the reference's on-disk form (`write`): a scene JSON
with an `import` of an OBJ file (one object per mesh primitive, `usemtl`
naming its shader) followed by `sphere` assets, a `shaders` table and the
`tracer` entry (src/core/renderer/renderer.hpp:17-23, src/core/meta/scene.cu:25-172).
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import kr

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def coin(self) -> bool:
        return (self.next_u64() >> 63) == 1

    def choice(self, seq):
        return seq[int(self.uniform() * len(seq)) % len(seq)]


def splitmix64_np(x: np.ndarray) -> np.ndarray:
    """Vectorised SplitMix64 finaliser over uint64 (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        z = x.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def rot_yup_to_zup(p):
    p = np.asarray(p, np.float64)
    return np.stack([p[..., 0], -p[..., 2], p[..., 1]], axis=-1)


@dataclass
class SceneData:
    positions: np.ndarray  # (V, 3) float32
    normals: np.ndarray  # (V, 3) float32
    indices: np.ndarray  # (T, 3) uint32
    meshes: list  # [(object name, shader name, tri_begin, tri_count)]
    spheres: list  # [(center xyz, radius, shader name)]
    shaders: dict  # name -> reference shader config [type, props]
    camera: dict  # CameraConfig fields
    width: int
    height: int
    spp: int
    max_bounce: int
    name: str = "scene"
    light_seed: int = 0x4B4F49534849  # fixed key of the light sample tables
    point_lights: list = field(default_factory=list)  # [(position xyz, color rgb, strength)]
    _keep: list = field(default_factory=list, repr=False)

    # ---- flat descriptor -------------------------------------------------
    def shader_names(self):
        # the reference's order: imported materials first, then sphere shaders
        names = []
        for _, s, _, _ in self.meshes:
            if s not in names:
                names.append(s)
        for _, _, s in self.spheres:
            if s not in names:
                names.append(s)
        return names

    def shader_descs(self):
        """Named shaders first (prims index them), then the inline children
        of Mixed shaders, as the C++ loader lays them out."""
        names = self.shader_names()
        table = [None] * len(names)
        for i, n in enumerate(names):
            fill_shader(self.shaders[n], table, i)
        arr = (kr.ShaderDesc * max(1, len(table)))()
        for i, d in enumerate(table):
            arr[i] = d
        return names, arr, len(table)

    def to_desc(self, bvh_width: int = 0, leaf_size: int = 0, flags: int = 0):
        names, sh, nshaders = self.shader_descs()
        index = {n: i for i, n in enumerate(names)}
        prims = (kr.PrimDesc * (len(self.meshes) + len(self.spheres)))()
        k = 0
        for _, s, b, c in self.meshes:
            prims[k] = kr.PrimDesc(kr.PRIM_MESH, index[s], b, c, (kr.f32 * 3)(0, 0, 0), 0.0)
            k += 1
        for cen, r, s in self.spheres:
            prims[k] = kr.PrimDesc(kr.PRIM_SPHERE, index[s], 0, 0, (kr.f32 * 3)(*cen), r)
            k += 1
        pos = np.ascontiguousarray(self.positions, np.float32)
        nrm = np.ascontiguousarray(self.normals, np.float32)
        idx = np.ascontiguousarray(self.indices, np.uint32)
        desc = kr.SceneDesc()
        desc.positions = pos.ctypes.data_as(C.POINTER(kr.f32))
        desc.normals = nrm.ctypes.data_as(C.POINTER(kr.f32))
        desc.num_vertices = pos.shape[0]
        desc.indices = idx.ctypes.data_as(C.POINTER(kr.u32))
        desc.num_triangles = idx.shape[0]
        desc.prims = prims
        desc.num_prims = k
        desc.shaders = sh
        desc.num_shaders = nshaders
        desc.light_seed = self.light_seed
        desc.bvh_width = bvh_width
        desc.leaf_size = leaf_size
        desc.flags = flags
        pls = (kr.PointLight * max(1, len(self.point_lights)))()
        for i, (pos_, col, strength) in enumerate(self.point_lights):
            pls[i] = kr.PointLight((kr.f32 * 3)(*pos_), (kr.f32 * 3)(*[c * strength for c in col]))
        desc.point_lights = C.cast(pls, C.c_void_p) if self.point_lights else None
        desc.num_point_lights = len(self.point_lights)
        holder = DescHolder(desc, [pos, nrm, idx, prims, sh, pls])
        return holder

    def camera_struct(self):
        c = self.camera
        lens = kr.LENS_PINHOLE if c.get("lens", "pinhole") == "pinhole" else kr.LENS_ORTHOGRAPHIC
        return kr.Camera(lens, float(c.get("fovx", 90.0)), (kr.f32 * 3)(*c["position"]),
                         (kr.f32 * 3)(*c["target"]), (kr.f32 * 3)(*c.get("upaxis", [0, 0, 1])))

    def params(self, seed: int = 1, spp: int | None = None, flags: int = 0, chunk_paths: int = 0,
               width: int | None = None, height: int | None = None):
        return kr.Params(width or self.width, height or self.height, spp or self.spp, self.max_bounce, seed, flags,
                         chunk_paths)

    @property
    def num_triangles(self):
        return int(self.indices.shape[0])

    # ---- reference on-disk form --------------------------------------------
    def write(self, directory: str, stem: str | None = None) -> str:
        """Writes <stem>.json (+ <stem>.obj when there are meshes); returns the JSON path."""
        stem = stem or self.name
        os.makedirs(directory, exist_ok=True)
        assets = []
        if self.meshes:
            obj_path = os.path.join(directory, stem + ".obj")
            write_obj(obj_path, self)
            assets.append(["import", {"path": obj_path}])
        for cen, r, s in self.spheres:
            assets.append(["sphere", {"o": [float(x) for x in cen], "r": float(r), "shader": s}])
        for pos_, col, strength in self.point_lights:
            assets.append(["pointLight", {"position": [float(x) for x in pos_], "color": [float(x) for x in col],
                                          "strength": float(strength)}])
        doc = {
            "scene": {
                "camera": [self.camera],
                "assets": assets,
                "shaders": self.shaders,
            },
            "tracer": ["HIPWavefront", {"kernel": ["Radiance", {"maxBounce": self.max_bounce}]}],
        }
        path = os.path.join(directory, stem + ".json")
        with open(path, "w") as f:
            json.dump(doc, f, indent=1)
        return path


class DescHolder:
    """Keeps the numpy arrays behind a kr_scene_desc alive."""

    def __init__(self, desc, keep):
        self.desc = desc
        self.keep = keep


def _distribution(props, d):
    dist = props["distribution"]
    dname = dist if isinstance(dist, str) else dist[0]
    dprops = {} if isinstance(dist, str) or len(dist) < 2 else dist[1]
    if dname == "Beckmann":
        d.distribution = kr.DIST_BECKMANN
        d.roughness = float(dprops.get("roughness", 0.5))
    elif dname == "Cosine":
        d.distribution = kr.DIST_COSINE
    else:
        raise ValueError("no such type: " + dname)


def fill_shader(cfg, table: list, slot: int, depth: int = 0):
    """Reference shader config [type, props] -> kr_shader_desc at table[slot]
    with the reference's defaults (lambert.hpp:12-19, glossy.hpp:11-16,
    emission.hpp:12-19, beckmann.hpp:11-17, glass.hpp:12-19,
    refraction.hpp:12-19, mixed.hpp:12-18); a Mixed's two inline shaders are
    appended to the table."""
    if isinstance(cfg, str):
        cfg = [cfg]
    typ, props = cfg[0], (cfg[1] if len(cfg) > 1 else {})
    d = kr.ShaderDesc()
    d.color = (kr.f32 * 3)(*props.get("color", [1, 1, 1]))
    d.R = (kr.f32 * 3)(*props.get("R", [1, 1, 1]))
    d.roughness = 0.5
    d.strength = float(props.get("strength", 2.0))
    if typ == "Lambert":
        d.type = kr.SHADER_LAMBERT
    elif typ == "GlossyBSDF":
        d.type = kr.SHADER_GLOSSY
        _distribution(props, d)
    elif typ == "Emission":
        d.type = kr.SHADER_EMISSION
    elif typ in ("GlassBSDF", "RefractionBSDF"):
        d.type = kr.SHADER_GLASS if typ == "GlassBSDF" else kr.SHADER_REFRACTION
        d.ior = float(props["ior"])
        _distribution(props, d)
    elif typ == "Transparent":
        d.type = kr.SHADER_TRANSPARENT
    elif typ == "Blackhole":
        d.type = kr.SHADER_BLACKHOLE
    elif typ == "Mixed":
        if depth >= 3:
            raise ValueError("Mixed shaders nest deeper than 3 levels")
        d.type = kr.SHADER_MIXED
        fname, fprops = props["fac"][0], (props["fac"][1] if len(props["fac"]) > 1 else {})
        if fname == "Constant":
            d.fac_type, d.fac_value = kr.SCALA_CONSTANT, float(fprops["value"])
        elif fname == "Fresnel":
            d.fac_type, d.fac_value = kr.SCALA_FRESNEL, float(fprops["ior"])
        else:
            raise ValueError("no such type: " + fname)
        table[slot] = d
        for k in range(2):
            c = len(table)
            table.append(None)
            table[slot].child[k] = c
            fill_shader(props["shaders"][k], table, c, depth + 1)
        return
    else:
        raise ValueError("no such type: " + typ)
    table[slot] = d


def shader_desc_from_config(cfg) -> kr.ShaderDesc:
    """A leaf shader config -> kr_shader_desc (Mixed needs a table: fill_shader)."""
    table = [None]
    fill_shader(cfg, table, 0)
    if len(table) > 1:
        raise ValueError("Mixed shaders need a shader table")
    return table[0]


def fmt(x) -> str:
    return repr(float(np.float32(x)))


def write_obj(path: str, s: SceneData):
    """OBJ with explicit per-vertex normals; one `o` + `usemtl` per mesh primitive."""
    pos = np.asarray(s.positions, np.float32)
    nrm = np.asarray(s.normals, np.float32)
    with open(path, "w") as f:
        f.write("# koishi-mi355x synthetic scene\n")
        f.write("".join("v %r %r %r\n" % (float(a), float(b), float(c)) for a, b, c in pos))
        f.write("".join("vn %r %r %r\n" % (float(a), float(b), float(c)) for a, b, c in nrm))
        idx = np.asarray(s.indices, np.int64) + 1
        for name, shader, b, c in s.meshes:
            f.write(f"o {name}\nusemtl {shader}\n")
            f.write("".join(f"f {i}//{i} {j}//{j} {k}//{k}\n" for i, j, k in idx[b:b + c]))


# ----------------------------------------------------------------- geometry --
ICO_V = None


def icosahedron():
    t = (1 + math.sqrt(5)) / 2
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    return v, f


def geodesic_sphere(freq: int, rotation=None):
    """Unit geodesic sphere: every icosahedron face split into freq^2 triangles.
    Shared edge/corner points are merged through canonical (corner, weight) keys.
    `rotation` (3x3) turns the base icosahedron first, so that no edges lie in
    a coordinate plane (the icosahedron's own symmetry planes are x, y, z = 0,
    where a camera's centre row or column of rays would graze them exactly)."""
    V, F = icosahedron()
    if rotation is not None:
        V = V @ np.asarray(rotation).T
    n = freq
    ij = [(i, j) for i in range(n + 1) for j in range(n + 1 - i)]
    ij = np.array(ij, np.int64)  # weights of corner B (i) and C (j)
    nf = F.shape[0]
    npt = ij.shape[0]
    wa = n - ij[:, 0] - ij[:, 1]
    ids = np.stack([np.broadcast_to(F[:, 0:1], (nf, npt)), np.broadcast_to(F[:, 1:2], (nf, npt)),
                    np.broadcast_to(F[:, 2:3], (nf, npt))], axis=-1)  # (nf, npt, 3)
    ws = np.stack([np.broadcast_to(wa, (nf, npt)), np.broadcast_to(ij[:, 0], (nf, npt)),
                   np.broadcast_to(ij[:, 1], (nf, npt))], axis=-1)
    ids = np.where(ws == 0, 99, ids)  # zero weight: corner irrelevant
    order = np.argsort(ids, axis=-1, kind="stable")
    ids_s = np.take_along_axis(ids, order, -1)
    ws_s = np.take_along_axis(ws, order, -1)
    key = np.concatenate([ids_s, ws_s], axis=-1).reshape(-1, 6)
    uniq, inv = np.unique(key, axis=0, return_inverse=True)
    inv = inv.reshape(nf, npt)
    # positions from the canonical key (identical floats for shared points)
    Vp = np.concatenate([V, np.zeros((1, 3))], axis=0)  # index 99 -> clamp below
    cid = np.where(uniq[:, :3] == 99, 12, uniq[:, :3])
    pts = (Vp[cid] * uniq[:, 3:, None].astype(np.float64)).sum(axis=1) / n
    pts /= np.linalg.norm(pts, axis=1, keepdims=True)
    # triangles per face
    lut = -np.ones((n + 1, n + 1), np.int64)
    for k, (i, j) in enumerate(ij):
        lut[i, j] = k
    tri_local = []
    for i in range(n):
        for j in range(n - i):
            tri_local.append((lut[i, j], lut[i + 1, j], lut[i, j + 1]))
            if i + j < n - 1:
                tri_local.append((lut[i + 1, j], lut[i + 1, j + 1], lut[i, j + 1]))
    tri_local = np.array(tri_local, np.int64)
    tris = inv[:, tri_local].reshape(-1, 3)
    return pts, tris


def value_noise(p: np.ndarray, seed: int, octaves: int = 4, base_freq: float = 2.0, gain: float = 0.5,
                lacunarity: float = 2.0) -> np.ndarray:
    """4-octave value noise in [-1, 1]: lattice values hashed with SplitMix64,
    smoothstep-weighted trilinear interpolation."""
    total = np.zeros(p.shape[0])
    amp, freq, norm = 1.0, base_freq, 0.0
    for o in range(octaves):
        q = p * freq
        b = np.floor(q)
        f = q - b
        w = f * f * (3 - 2 * f)
        bi = b.astype(np.int64)
        acc = np.zeros(p.shape[0])
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    x = (bi[:, 0] + dx).astype(np.uint64)
                    y = (bi[:, 1] + dy).astype(np.uint64)
                    z = (bi[:, 2] + dz).astype(np.uint64)
                    with np.errstate(over="ignore"):
                        h = splitmix64_np(np.uint64(seed & MASK64) ^ (x * np.uint64(0x8DA6B343))
                                          ^ (y * np.uint64(0xD8163841)) ^ (z * np.uint64(0xCB1AB31F))
                                          ^ np.uint64(o * 0x9E3779B1))
                    val = (h >> np.uint64(11)).astype(np.float64) * (2.0 / (1 << 53)) - 1.0
                    wx = w[:, 0] if dx else 1 - w[:, 0]
                    wy = w[:, 1] if dy else 1 - w[:, 1]
                    wz = w[:, 2] if dz else 1 - w[:, 2]
                    acc += val * wx * wy * wz
        total += amp * acc
        norm += amp
        amp *= gain
        freq *= lacunarity
    return total / norm


def vertex_normals(pos: np.ndarray, tris: np.ndarray) -> np.ndarray:
    a, b, c = pos[tris[:, 0]], pos[tris[:, 1]], pos[tris[:, 2]]
    fn = np.cross(b - a, c - a)  # area weighted
    n = np.zeros_like(pos)
    for k in range(3):
        np.add.at(n, tris[:, k], fn)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    return n


def displaced_icosphere(freq: int, seed: int, amplitude: float = 0.15):
    pts, tris = geodesic_sphere(freq, quat_matrix(SplitMix64(seed ^ 0x1C05)))
    h = value_noise(pts, seed)
    pos = pts * (1.0 + amplitude * h)[:, None]
    nrm = vertex_normals(pos, tris)
    # orient faces outward (normal of the first vertex)
    a, b, c = pos[tris[:, 0]], pos[tris[:, 1]], pos[tris[:, 2]]
    flip = (np.cross(b - a, c - a) * a).sum(axis=1) < 0
    tris[flip] = tris[flip][:, [0, 2, 1]]
    return pos, nrm, tris


def quad(center, sx, sz, normal_sign=-1.0):
    """Axis-aligned quad in the y = center.y plane (Y-up), 2 triangles, normal +-y."""
    cx, cy, cz = center
    v = np.array([[cx - sx / 2, cy, cz - sz / 2], [cx + sx / 2, cy, cz - sz / 2], [cx + sx / 2, cy, cz + sz / 2],
                  [cx - sx / 2, cy, cz + sz / 2]], np.float64)
    n = np.tile([0.0, normal_sign, 0.0], (4, 1))
    t = np.array([[0, 1, 2], [0, 2, 3]], np.int64)
    a, b, c = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    if (np.cross(b - a, c - a)[:, 1] * normal_sign < 0).any():
        t = t[:, [0, 2, 1]]
    return v, n, t


def rect(p0, e1, e2, normal):
    """Parallelogram p0, p0+e1, p0+e1+e2, p0+e2 (Y-up), 2 triangles wound so
    that their geometric normal points along `normal`."""
    p0, e1, e2, normal = (np.asarray(x, np.float64) for x in (p0, e1, e2, normal))
    v = np.array([p0, p0 + e1, p0 + e1 + e2, p0 + e2])
    n = np.tile(normal / np.linalg.norm(normal), (4, 1))
    t = np.array([[0, 1, 2], [0, 2, 3]], np.int64)
    if np.dot(np.cross(e1, e2), normal) < 0:
        t = t[:, [0, 2, 1]]
    return v, n, t


class Builder:
    """Accumulates Y-up meshes, then rotates everything into Z-up."""

    def __init__(self):
        self.pos, self.nrm, self.idx, self.meshes = [], [], [], []
        self.nv = 0
        self.nt = 0
        self.spheres = []
        self.shaders = {}
        self.point_lights = []

    def add_mesh(self, name, shader, pos, nrm, tris):
        self.pos.append(np.asarray(pos, np.float64))
        self.nrm.append(np.asarray(nrm, np.float64))
        self.idx.append(np.asarray(tris, np.int64) + self.nv)
        self.meshes.append((name, shader, self.nt, int(tris.shape[0])))
        self.nv += pos.shape[0]
        self.nt += tris.shape[0]

    def add_sphere(self, center, r, shader):
        self.spheres.append((center, r, shader))

    def add_point_light(self, position, color, strength):
        self.point_lights.append((position, color, strength))

    def finish(self, camera_eye, camera_look, fovx, width, height, spp, max_bounce, name) -> SceneData:
        if self.pos:
            pos = rot_yup_to_zup(np.concatenate(self.pos)).astype(np.float32)
            nrm = rot_yup_to_zup(np.concatenate(self.nrm))
            nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
            idx = np.concatenate(self.idx).astype(np.uint32)
        else:
            pos = np.zeros((0, 3), np.float32)
            nrm = np.zeros((0, 3), np.float32)
            idx = np.zeros((0, 3), np.uint32)
        spheres = [([float(x) for x in rot_yup_to_zup(np.array(c)).astype(np.float32)], float(np.float32(r)), s)
                   for c, r, s in self.spheres]
        eye = rot_yup_to_zup(np.array(camera_eye, np.float64))
        look = rot_yup_to_zup(np.array(camera_look, np.float64))
        tgt = look - eye
        tgt /= np.linalg.norm(tgt)
        cam = {"lens": "pinhole", "fovx": float(fovx), "position": [float(x) for x in np.float32(eye)],
               "target": [float(x) for x in np.float32(tgt)], "upaxis": [0.0, 0.0, 1.0]}
        pls = [([float(x) for x in rot_yup_to_zup(np.array(p_)).astype(np.float32)], list(c), float(k))
               for p_, c, k in self.point_lights]
        return SceneData(pos, nrm, idx, self.meshes, spheres, self.shaders, cam, width, height, spp, max_bounce,
                         name=name, point_lights=pls)


def _rand_color(rng):
    return [round(rng.uniform(0.2, 0.7), 6) for _ in range(3)]


def _glossy(color, roughness):
    return ["GlossyBSDF", {"distribution": ["Beckmann", {"roughness": float(roughness)}], "color": color}]


# ------------------------------------------------------------------ configs --
def sphere_scene(seed: int = 1, width: int = 256, height: int = 256, spp: int = 16, max_bounce: int = 5,
                 n_spheres: int = 8) -> SceneData:
    """Plain sphere scene: 8 seeded spheres (centres U[-5,5]^3, r U[0.2,1]),
    ground sphere r = 1000 at y = -1000, 2x2 quad light at y = 6 facing -y with
    emission (50,50,50); materials Lambert or Beckmann (alpha in {.1,.3,.5,.6})
    with albedo U[0.2,0.7]^3. 256x256, 16 spp, max depth 5."""
    rng = SplitMix64(seed)
    b = Builder()
    b.shaders["light"] = ["Emission", {"color": [1.0, 1.0, 1.0], "strength": 50.0}]
    v, n, t = quad((0.0, 6.0, 0.0), 2.0, 2.0, -1.0)
    b.add_mesh("light", "light", v, n, t)
    for i in range(n_spheres):
        c = [rng.uniform(-5, 5) for _ in range(3)]
        r = rng.uniform(0.2, 1.0)
        name = f"s{i}"
        if rng.coin():
            b.shaders[name] = ["Lambert", {"color": _rand_color(rng)}]
        else:
            alpha = rng.choice([0.1, 0.3, 0.5, 0.6])
            b.shaders[name] = _glossy(_rand_color(rng), alpha)
        b.add_sphere(c, r, name)
    b.shaders["ground"] = ["Lambert", {"color": [0.5, 0.5, 0.5]}]
    b.add_sphere([0.0, -1000.0, 0.0], 1000.0, "ground")
    return b.finish((0.0, 3.0, -14.0), (0.0, 0.0, 0.0), 60.0, width, height, spp, max_bounce, "spheres")


def mesh_scene(seed: int = 2, freq: int = 71, width: int = 1024, height: int = 512, spp: int = 64,
               max_bounce: int = 8) -> SceneData:
    """Triangle-mesh BVH scene: displaced geodesic icosphere (20 freq^2
    triangles; freq 71 -> 100,820), 0.15 x 4-octave value noise along the
    normal, on a 2-triangle ground at y = -1.2, one 1x1 quad light at y = 3
    with emission 30. Beckmann alpha 0.3 on the mesh, Lambert ground.
    1024x512, 64 spp, max depth 8."""
    b = Builder()
    pos, nrm, tris = displaced_icosphere(freq, seed)
    b.shaders["mesh"] = _glossy([0.8, 0.8, 0.8], 0.3)
    b.shaders["ground"] = ["Lambert", {"color": [0.5, 0.5, 0.5]}]
    b.shaders["light"] = ["Emission", {"color": [1.0, 1.0, 1.0], "strength": 30.0}]
    b.add_mesh("mesh", "mesh", pos, nrm, tris)
    v, n, t = quad((0.0, -1.2, 0.0), 20.0, 20.0, 1.0)
    b.add_mesh("ground", "ground", v, n, t)
    v, n, t = quad((0.0, 3.0, 0.0), 1.0, 1.0, -1.0)
    b.add_mesh("light", "light", v, n, t)
    return b.finish((0.0, 1.0, -4.0), (0.0, 0.0, 0.0), 50.0, width, height, spp, max_bounce, "mesh")


def room_scene(seed: int = 6, freq: int = 20, width: int = 1024, height: int = 768, spp: int = 64,
               max_bounce: int = 8, turn_deg: float = 25.0) -> SceneData:
    """Closed room (4 x 3 x 4, Y-up): white floor, ceiling and back wall, red
    and green side walls, a white wall behind the camera, a 1 x 1 ceiling
    light of emission 15 and a displaced icosphere (Beckmann 0.4). Indirect
    light carries a large share of the image, unlike the open workloads.
    Everything is turned by turn_deg about the vertical, so that no wall
    normal ends up along +-Y after the Z-up rotation (the reference's shading
    frame is NaN there)."""
    c, s_ = math.cos(math.radians(turn_deg)), math.sin(math.radians(turn_deg))
    turn = np.array([[c, 0.0, s_], [0.0, 1.0, 0.0], [-s_, 0.0, c]])
    b = Builder()
    b.shaders["white"] = ["Lambert", {"color": [0.75, 0.75, 0.75]}]
    b.shaders["red"] = ["Lambert", {"color": [0.7, 0.15, 0.12]}]
    b.shaders["green"] = ["Lambert", {"color": [0.15, 0.6, 0.15]}]
    b.shaders["mesh"] = _glossy([0.8, 0.7, 0.5], 0.4)
    b.shaders["light"] = ["Emission", {"color": [1.0, 1.0, 1.0], "strength": 15.0}]

    def add(name, shader, v, n, t):
        b.add_mesh(name, shader, np.asarray(v) @ turn.T, np.asarray(n) @ turn.T, t)

    X, Y, Z = 2.0, 3.0, 2.0
    walls = [("floor", "white", (-X, 0, -Z), (2 * X, 0, 0), (0, 0, 2 * Z), (0, 1, 0)),
             ("ceiling", "white", (-X, Y, -Z), (2 * X, 0, 0), (0, 0, 2 * Z), (0, -1, 0)),
             ("back", "white", (-X, 0, Z), (2 * X, 0, 0), (0, Y, 0), (0, 0, -1)),
             ("front", "white", (-X, 0, -Z), (2 * X, 0, 0), (0, Y, 0), (0, 0, 1)),
             ("left", "red", (-X, 0, -Z), (0, 0, 2 * Z), (0, Y, 0), (1, 0, 0)),
             ("right", "green", (X, 0, -Z), (0, 0, 2 * Z), (0, Y, 0), (-1, 0, 0))]
    for name, shader, p0, e1, e2, nn in walls:
        add(name, shader, *rect(p0, e1, e2, nn))
    pos, nrm, tris = displaced_icosphere(freq, seed)
    add("mesh", "mesh", pos * 0.8 + np.array([0.3, 0.95, 0.4]), nrm, tris)
    add("light", "light", *quad((0.0, Y - 0.01, 0.0), 1.0, 1.0, -1.0))
    eye = turn @ np.array([0.0, 1.5, -1.9])
    look = turn @ np.array([0.0, 1.2, 0.0])
    return b.finish(tuple(eye), tuple(look), 70.0, width, height, spp, max_bounce, "room")


def materials_scene(width: int = 160, height: int = 96, spp: int = 16, max_bounce: int = 8) -> SceneData:
    """Every shader type, light type and Mixed fac of the Radiance path on
    one ground: GlassBSDF, RefractionBSDF (Beckmann and Cosine), Transparent,
    Blackhole, Mixed with a Constant and a Fresnel fac (nested two levels,
    with an Emission leaf), a Mixed displaced icosphere, one quad area light
    and two point lights."""
    b = Builder()
    b.shaders["ground"] = ["Lambert", {"color": [0.6, 0.6, 0.55]}]
    b.shaders["light"] = ["Emission", {"color": [1.0, 0.9, 0.8], "strength": 12.0}]
    b.shaders["glass"] = ["GlassBSDF", {"ior": 1.5, "color": [0.95, 0.95, 1.0],
                                        "distribution": ["Beckmann", {"roughness": 0.15}]}]
    b.shaders["water"] = ["RefractionBSDF", {"ior": 1.33, "color": [0.8, 0.9, 1.0],
                                             "distribution": ["Beckmann", {"roughness": 0.3}]}]
    b.shaders["frosted"] = ["RefractionBSDF", {"ior": 1.2, "distribution": ["Cosine", {}]}]
    b.shaders["clear"] = ["Transparent", {"color": [0.7, 0.9, 0.7]}]
    b.shaders["void"] = ["Blackhole", {}]
    b.shaders["mix"] = ["Mixed", {"fac": ["Constant", {"value": 0.3}],
                                  "shaders": [["Lambert", {"color": [0.7, 0.3, 0.2]}],
                                              _glossy([0.9, 0.9, 0.9], 0.2)]}]
    b.shaders["coated"] = ["Mixed", {"fac": ["Fresnel", {"ior": 1.6}],
                                     "shaders": [["Lambert", {"color": [0.2, 0.4, 0.7]}],
                                                 ["Mixed", {"fac": ["Constant", {"value": 0.5}],
                                                            "shaders": [["GlassBSDF", {"ior": 1.4, "distribution":
                                                                                       ["Beckmann", {"roughness": 0.05}]}],
                                                                        ["Emission", {"strength": 3.0}]]}]]}]
    v, n, t = quad((0.0, 0.0, 0.0), 14.0, 14.0, 1.0)
    b.add_mesh("ground", "ground", v, n, t)
    v, n, t = quad((0.0, 4.0, 0.0), 1.5, 1.5, -1.0)
    b.add_mesh("light", "light", v, n, t)
    pos, nrm, tris = displaced_icosphere(12, 9)
    b.add_mesh("blob", "coated", pos * 0.7 + np.array([0.0, 0.8, 1.6]), nrm, tris)
    for c, r, sh in (((-2.4, 0.6, 0.0), 0.6, "glass"), ((-1.0, 0.5, -0.6), 0.5, "water"),
                     ((0.3, 0.45, -1.0), 0.45, "frosted"), ((1.5, 0.5, -0.4), 0.5, "clear"),
                     ((2.6, 0.4, 0.6), 0.4, "void"), ((1.2, 0.55, 1.2), 0.55, "mix")):
        b.add_sphere(c, r, sh)
    b.add_point_light((2.0, 3.0, -2.0), [1.0, 1.0, 1.0], 1.5)
    b.add_point_light((-3.0, 2.0, 1.0), [0.4, 0.6, 1.0], 0.8)
    return b.finish((0.0, 2.2, -6.5), (0.0, 0.6, 0.0), 55.0, width, height, spp, max_bounce, "materials")


def quat_matrix(rng: SplitMix64):
    """Uniform random rotation (Shoemake, Graphics Gems III, 1992)."""
    u1, u2, u3 = rng.uniform(), rng.uniform(), rng.uniform()
    a, b = math.sqrt(1 - u1), math.sqrt(u1)
    x, y = a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2)
    z, w = b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def large_scene(seed: int = 3, instances: int = 40, freq: int = 36, width: int = 1920, height: int = 1080,
                spp: int = 64, max_bounce: int = 8, n_lights: int = 4) -> SceneData:
    """Large scene: `instances` seeded copies of the displaced icosphere
    (freq 36 -> 25,920 triangles each; 40 -> 1,036,800), random rotation and
    scale U[0.5, 2] over a 20x20 ground, 4 quad lights with emission U[10, 50];
    50% Lambert (albedo U[0.2,0.7]^3) / 50% Beckmann (alpha U[0.05, 0.8]).
    1920x1080, 64 spp, max depth 8."""
    rng = SplitMix64(seed)
    b = Builder()
    base_pos, base_nrm, base_tris = displaced_icosphere(freq, seed)
    for k in range(instances):
        R = quat_matrix(rng)
        s = rng.uniform(0.5, 2.0)
        x, z = rng.uniform(-9, 9), rng.uniform(-9, 9)
        name = f"inst{k}"
        if rng.coin():
            b.shaders[name] = ["Lambert", {"color": _rand_color(rng)}]
        else:
            b.shaders[name] = _glossy(_rand_color(rng), rng.uniform(0.05, 0.8))
        pos = (base_pos @ R.T) * s + np.array([x, s, z])
        nrm = base_nrm @ R.T
        b.add_mesh(name, name, pos, nrm, base_tris.copy())
    b.shaders["ground"] = ["Lambert", {"color": [0.5, 0.5, 0.5]}]
    v, n, t = quad((0.0, 0.0, 0.0), 20.0, 20.0, 1.0)
    b.add_mesh("ground", "ground", v, n, t)
    for k in range(n_lights):
        e = rng.uniform(10, 50)
        name = f"light{k}"
        b.shaders[name] = ["Emission", {"color": [1.0, 1.0, 1.0], "strength": float(e)}]
        v, n, t = quad((rng.uniform(-6, 6), 8.0, rng.uniform(-6, 6)), 2.0, 2.0, -1.0)
        b.add_mesh(name, name, v, n, t)
    return b.finish((0.0, 9.0, -20.0), (0.0, 0.0, 0.0), 60.0, width, height, spp, max_bounce, "large")


BATCH_RESOLUTIONS = ((256, 256), (512, 512), (1024, 512), (1920, 1080))


def batch_workload(seed: int = 4, jobs: int = 16, max_bounce: int = 8):
    """Batched uneven workload: `jobs` independent scenes from the large-scene
    generator (scene k seeded 4000 + k), each with a triangle count T
    log-uniform in [10k, 2M] (instances = clamp(round(T / 25920), 1, 40),
    icosphere frequency round(sqrt(T / (20 instances)))), a resolution drawn
    from 256^2, 512^2, 1024x512, 1920x1080 and spp uniform in {16..256};
    depth 8. Returns (scenes, specs) with specs[k] = dict(scene=k, width,
    height, spp, max_bounce, seed)."""
    rng = SplitMix64(seed)
    scenes, specs = [], []
    for k in range(jobs):
        t = math.exp(rng.uniform(math.log(1e4), math.log(2e6)))
        inst = int(min(40, max(1, round(t / 25920))))
        freq = int(max(2, round(math.sqrt(t / (20 * inst)))))
        w, h = BATCH_RESOLUTIONS[int(rng.uniform() * len(BATCH_RESOLUTIONS)) % len(BATCH_RESOLUTIONS)]
        spp = 16 + int(rng.uniform() * 241) % 241
        scenes.append(large_scene(seed=4000 + k, instances=inst, freq=freq, width=w, height=h, spp=spp,
                                  max_bounce=max_bounce))
        specs.append(dict(scene=k, width=w, height=h, spp=spp, max_bounce=max_bounce, seed=4000 + k))
    return scenes, specs


WORKLOADS = {
    "spheres": sphere_scene,
    "mesh": mesh_scene,
    "large": large_scene,
    "room": room_scene,
    "materials": materials_scene,
}
