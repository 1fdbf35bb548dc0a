"""Pins of the CPU oracle (oracle/koishi_oracle.hpp).

Each oracle function is checked against something it does not restate:
  * the reference's own known-answer tests (test/intersect.cu, test/sphere.cu);
  * the reference's own headers, compiled from /root/reference into
    oracle/_ref/libkoishi_ref.so (skipped where that library is absent);
  * the published Philox4x32-10 known-answer vectors (Random123 kat_vectors);
  * closed forms and invariants (cosine-weighted moments, Beckmann
    normalisation, tan^2 distribution, reflection identities).
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

import oracle_lib as ol

D = C.c_double
REALS = ["f64", "f32"]


def d3(v):
    return ol.d3(v)


# ------------------------------------------------------------------ Philox --
def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32_10
    assert ol.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert ol.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert ol.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_u01_is_24_bit_and_half_open():
    f = ol.oracle().oracle_u01
    assert f(0) == 0.0
    assert f(0xFFFFFFFF) == pytest.approx(1 - 2 ** -24, abs=0)
    assert f(0x80000000) == 0.5


# --------------------------------------------- reference known-answer tests --
# test/intersect.cu:7-71 -- segment vs unit box (vmin 0, vmax 1)
SEG_CASES = [
    ((2, 0, 0), (-1, 0, 1), True), ((2, 0, 0), (-1, 0, 1.01), False), ((2, 0, 0), (-1.01, 0, 1), True),
    ((2, 0, 0), (-0.99, 0, 1), False), ((2, 2, 0), (-1, -1, 1), True), ((2, 2, 0), (-1, -1, 1.01), False),
    ((2, 2, 0), (-1, -1, 0), True), ((2, 2, 0), (-0.99, -0.99, 0), False), ((0, 0, 2), (0, 0, -2), True),
    ((0, 0, 2), (0, 0, -1), True), ((0, 0, 2), (0, 0, -0.99), False), ((0, 0, 1e5), (0, 0, -2e5), True),
    ((0.5, 0.5, 1e5), (0, 0, -2e5), True), ((0.5, 0.5, 0.5), (0, 0, 0.1), True),
    ((1e-5, 1e-5, 1e-5), (1 - 2e-5, 1 - 2e-5, 1 - 2e-5), True),
]


def _seg_case(o, v):
    v = np.array(v, np.float64)
    # the test normalises in float and takes length(v) in float
    vf = v.astype(np.float32)
    n = np.float32(1.0) / np.sqrt(np.float32((vf * vf).sum()))
    return o, (vf * n).astype(np.float64), float(np.sqrt(np.float32((vf * vf).sum())))


# Cases whose segment ends exactly on a box edge: true in the reference's
# float arithmetic, decided by rounding only; the FP64 oracle is not held to them.
TOUCHING = {((2, 0, 0), (-1, 0, 1)), ((2, 2, 0), (-1, -1, 1)), ((2, 2, 0), (-1, -1, 0))}


@pytest.mark.parametrize("real", REALS)
def test_reference_segment_box_cases(real):
    f = ol.fn("oracle_seg_bbox", real)
    for o, v, expect in SEG_CASES:
        if real == "f64" and (o, v) in TOUCHING:
            continue
        o_, d_, t_ = _seg_case(o, v)
        got = f(d3(o_), d3(d_), t_, d3((0, 0, 0)), d3((1, 1, 1)))
        assert bool(got) == expect, (real, o, v)


@pytest.mark.parametrize("real", REALS)
def test_reference_sphere_cases(real):
    # test/sphere.cu:7-18: unit sphere at the origin, ray from (2,0,0)
    f = ol.fn("oracle_sphere_intersect", real)
    t, n = D(), d3((0, 0, 0))
    assert f(d3((0, 0, 0)), 1.0, d3((2, 0, 0)), d3((-1, 0, 0)), C.byref(t), n) == 1
    assert t.value == pytest.approx(1.0)
    assert list(n) == pytest.approx([1, 0, 0])
    dd = np.array([-1, 0, 1.0]) / math.sqrt(2)
    assert f(d3((0, 0, 0)), 1.0, d3((2, 0, 0)), d3(dd), C.byref(t), n) == 0


# -------------------------------------------- against the reference headers --
class _LazyRef:
    """oracle/_ref/libkoishi_ref.so, loaded by the first test that calls into
    it (so that collecting this module, e.g. for a GPU-only run, maps
    nothing)."""

    def __getattr__(self, name):
        return getattr(ol.ref(), name)


ref = _LazyRef()
needs_ref = pytest.mark.skipif(not os.path.exists(ol.REF_SO), reason="oracle/_ref not built (no reference sources)")


def _unit(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


@needs_ref
@pytest.mark.parametrize("real", REALS)
def test_ray_box_and_triangle_match_reference(real):
    rng = np.random.default_rng(1)
    fb, ft = ol.fn("oracle_ray_bbox", real), ol.fn("oracle_ray_triangle", real)
    agree_b = agree_t = 0
    N = 3000
    for i in range(N):
        o = rng.uniform(-3, 3, 3)
        d = _unit(rng, 1)[0].astype(np.float32).astype(np.float64)
        lo = rng.uniform(-1, 0, 3)
        hi = lo + rng.uniform(0.1, 2, 3)
        agree_b += fb(d3(o), d3(d), d3(lo), d3(hi)) == ref.ref_ray_bbox(d3(o), d3(d), d3(lo), d3(hi))
        v0 = rng.uniform(-1, 1, 3)
        e1, e2 = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
        a, b = d3((0, 0, 0)), d3((0, 0, 0))
        ha = ft(d3(o), d3(d), d3(v0), d3(e1), d3(e2), a)
        hb = ref.ref_ray_triangle(d3(o), d3(d), d3(v0), d3(e1), d3(e2), b)
        agree_t += ha == hb
        if ha and hb:
            assert np.allclose(list(a), list(b), rtol=2e-4, atol=2e-5)
    assert agree_b >= N - 3 and agree_t >= N - 3


@needs_ref
@pytest.mark.parametrize("real", REALS)
def test_segment_tests_match_reference(real):
    rng = np.random.default_rng(2)
    fb, ft = ol.fn("oracle_seg_bbox", real), ol.fn("oracle_seg_triangle", real)
    agree = 0
    N = 3000
    for i in range(N):
        o = rng.uniform(-2, 2, 3)
        d = _unit(rng, 1)[0].astype(np.float32).astype(np.float64)
        t = rng.uniform(0.1, 4)
        lo = rng.uniform(-1, 0, 3)
        hi = lo + rng.uniform(0.1, 2, 3)
        agree += fb(d3(o), d3(d), t, d3(lo), d3(hi)) == ref.ref_seg_bbox(d3(o), d3(d), t, d3(lo), d3(hi))
        v0, e1, e2 = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
        agree += ft(d3(o), d3(d), t, d3(v0), d3(e1), d3(e2)) == \
            ref.ref_seg_triangle(d3(o), d3(d), t, d3(v0), d3(e1), d3(e2))
    assert agree >= 2 * N - 4


@needs_ref
@pytest.mark.parametrize("real", REALS)
def test_sphere_intersect_and_sample_match_reference(real):
    rng = np.random.default_rng(3)
    fi, fs = ol.fn("oracle_sphere_intersect", real), ol.fn("oracle_sphere_sample", real)
    for i in range(500):
        c = rng.uniform(-3, 3, 3)
        r = rng.uniform(0.2, 2)
        o = rng.uniform(-5, 5, 3)
        d = _unit(rng, 1)[0].astype(np.float32).astype(np.float64)
        ta, tb, na, nb = D(), D(), d3((0, 0, 0)), d3((0, 0, 0))
        ha = fi(d3(c), r, d3(o), d3(d), C.byref(ta), na)
        hb = ref.ref_sphere_intersect(d3(c), r, d3(o), d3(d), C.byref(tb), nb)
        assert ha == hb
        if ha:
            assert ta.value == pytest.approx(tb.value, rel=1e-4, abs=1e-4)
            assert np.allclose(list(na), list(nb), atol=1e-4)
        u0, u1 = rng.uniform(), rng.uniform()
        pa, pb, qa, qb, pda, pdb = d3((0,) * 3), d3((0,) * 3), d3((0,) * 3), d3((0,) * 3), D(), D()
        fs(d3(c), r, u0, u1, pa, qa, C.byref(pda))
        ref.ref_sphere_sample(d3(c), r, u0, u1, pb, qb, C.byref(pdb))
        assert np.allclose(list(pa), list(pb), rtol=1e-5, atol=1e-5)
        assert np.allclose(list(qa), list(qb), atol=1e-5)
        assert pda.value == pytest.approx(pdb.value, rel=1e-5)


@needs_ref
@pytest.mark.parametrize("real", REALS)
def test_bsdf_sampling_matches_reference(real):
    rng = np.random.default_rng(4)
    fc, fbs, fbf, fr = (ol.fn(n, real) for n in
                        ("oracle_sample_cos", "oracle_beckmann_sample", "oracle_beckmann_f", "oracle_reflect"))
    for i in range(1000):
        u0, u1 = float(np.float32(rng.uniform())), float(np.float32(rng.uniform()))
        a, b = d3((0,) * 3), d3((0,) * 3)
        fc(u0, u1, a)
        ref.ref_sample_cos(u0, u1, b)
        assert np.allclose(list(a), list(b), atol=2e-6)
        rough = rng.uniform(0.05, 0.9)
        pdf = D()
        fbs(u0 if u0 > 0 else 0.5, u1, float(np.float32(rough)) ** 2, a)
        ref.ref_beckmann_sample(u0 if u0 > 0 else 0.5, u1, rough, b, C.byref(pdf))
        assert np.allclose(list(a), list(b), atol=2e-5)
        w = _unit(rng, 1)[0]
        w[2] = abs(w[2]) + 0.05
        w /= np.linalg.norm(w)
        w = w.astype(np.float32).astype(np.float64)
        va = fbf(d3(w), float(np.float32(rough)) ** 2)
        vb = ref.ref_beckmann_f(d3(w), rough)
        assert va == pytest.approx(vb, rel=2e-4, abs=1e-6)
        n = _unit(rng, 1)[0].astype(np.float32).astype(np.float64)
        fr(d3(w), d3(n), a)
        ref.ref_reflect(d3(w), d3(n), b)
        assert np.allclose(list(a), list(b), atol=1e-6)


@needs_ref
@pytest.mark.parametrize("real", REALS)
def test_lens_matches_reference(real):
    from koishi_amd import scenegen as sg
    s = sg.sphere_scene()
    cam = s.camera_struct()
    f = ol.fn("oracle_lens_ray", real)
    rng = np.random.default_rng(5)
    for lens in (0, 1):
        cam.lens = lens
        for i in range(300):
            w, h, spp = 256, 128, int(rng.integers(1, 70))
            x, y, k = int(rng.integers(0, w)), int(rng.integers(0, h)), int(rng.integers(0, spp))
            oa, da, ob, db = d3((0,) * 3), d3((0,) * 3), d3((0,) * 3), d3((0,) * 3)
            f(C.byref(cam), w, h, spp, x, y, k, oa, da)
            ref.ref_lens_ray(lens, cam.fovx, d3(cam.position), d3(cam.target), d3(cam.upaxis), w, h, spp, x, y, k,
                             ob, db)
            assert np.allclose(list(oa), list(ob), rtol=1e-5, atol=1e-5)
            assert np.allclose(list(da), list(db), atol=2e-6)


@needs_ref
@pytest.mark.parametrize("real", REALS)
def test_frame_and_spawn_match_reference(real):
    rng = np.random.default_rng(6)
    ff, fs = ol.fn("oracle_frame", real), ol.fn("oracle_emit_seg", real)
    for i in range(500):
        n = _unit(rng, 1)[0].astype(np.float32).astype(np.float64)
        w = _unit(rng, 1)[0].astype(np.float32).astype(np.float64)
        outs_a = [d3((0,) * 3) for _ in range(4)]
        outs_b = [d3((0,) * 3) for _ in range(4)]
        ff(d3(n), d3(w), *outs_a)
        ref.ref_frame(d3(n), d3(w), *outs_b)
        for a, b in zip(outs_a, outs_b):
            assert np.allclose(list(a), list(b), atol=1e-5)
        p, q = rng.uniform(-3, 3, 3), rng.uniform(-3, 3, 3)
        oa, da, ob, db, ta, tb = d3((0,) * 3), d3((0,) * 3), d3((0,) * 3), d3((0,) * 3), D(), D()
        fs(d3(p), d3(q), oa, da, C.byref(ta))
        ref.ref_emit_seg(d3(p), d3(q), ob, db, C.byref(tb))
        assert np.allclose(list(oa), list(ob), atol=1e-5) and np.allclose(list(da), list(db), atol=1e-6)
        assert ta.value == pytest.approx(tb.value, rel=1e-5)


def test_frame_is_nan_along_y():
    """K6: the reference frame u = normalize(n x (0,1,0)) is NaN for n = +-Y."""
    ff = ol.fn("oracle_frame", "f64")
    outs = [d3((0,) * 3) for _ in range(4)]
    ff(d3((0, 1, 0)), d3((0, 0, 1)), *outs)
    assert all(math.isnan(x) for x in outs[0])
    if ref is not None:
        ref.ref_frame(d3((0, 1, 0)), d3((0, 0, 1)), *outs)
        assert all(math.isnan(x) for x in outs[0])


# ----------------------------------------------------------- closed forms --
def _grid(n):
    g = (np.arange(n) + 0.5) / n
    return np.stack(np.meshgrid(g, g, indexing="ij"), -1).reshape(-1, 2)


@pytest.mark.parametrize("real", REALS)
def test_cosine_sampling_moments(real):
    """Concentric-disk cosine sampling: E[cos] = 2/3, E[cos^2] = 1/2, and
    the disk radius^2 is uniform (the map preserves area)."""
    f = ol.fn("oracle_sample_cos", real)
    w = d3((0,) * 3)
    zs, r2 = [], []
    for u0, u1 in _grid(200):
        f(u0, u1, w)
        x, y, z = list(w)
        zs.append(z)
        r2.append(x * x + y * y)
        assert abs(x * x + y * y + z * z - 1) < 1e-5
    zs = np.array(zs)
    assert zs.mean() == pytest.approx(2 / 3, abs=2e-4)
    assert (zs ** 2).mean() == pytest.approx(0.5, abs=2e-4)
    # area preservation: r^2 of random samples is uniform (chi-square, 20 bins)
    from scipy.stats import chi2
    rng = np.random.default_rng(8)
    r2 = []
    for u0, u1 in rng.uniform(size=(60000, 2)):
        f(u0, u1, w)
        r2.append(w[0] ** 2 + w[1] ** 2)
    hist, _ = np.histogram(r2, bins=20, range=(0, 1))
    stat = ((hist - hist.mean()) ** 2 / hist.mean()).sum()
    assert chi2.sf(stat, 19) > 1e-4


@pytest.mark.parametrize("real", REALS)
@pytest.mark.parametrize("rough", [0.1, 0.3, 0.6])
def test_beckmann_normalisation_and_sampling(real, rough):
    """int D'(w) dw over the hemisphere = 1 (D' = D cos), and the sampled
    half vectors have P(tan^2 <= x) = 1 - exp(-x / alpha^2)."""
    a2 = rough * rough
    ff, fs = ol.fn("oracle_beckmann_f", real), ol.fn("oracle_beckmann_sample", real)
    th = (np.arange(20000) + 0.5) / 20000 * (math.pi / 2)
    vals = np.array([ff(d3((math.sin(t), 0.0, math.cos(t))), a2) for t in th])
    integral = (vals * np.sin(th)).sum() * (math.pi / 2 / 20000) * 2 * math.pi
    assert integral == pytest.approx(1.0, abs=2e-3 if real == "f32" else 1e-4)
    w = d3((0,) * 3)
    tan2 = []
    for u0, u1 in _grid(60):
        fs(u0, u1, a2, w)
        x, y, z = list(w)
        tan2.append((x * x + y * y) / (z * z))
    tan2 = np.sort(tan2)
    emp = (np.arange(tan2.size) + 0.5) / tan2.size
    assert np.abs(emp - (1 - np.exp(-tan2 / a2))).max() < 0.01


def test_reflect_identities():
    fr = ol.fn("oracle_reflect", "f64")
    rng = np.random.default_rng(7)
    out = d3((0,) * 3)
    for i in range(200):
        wo, h = _unit(rng, 2)
        fr(d3(wo), d3(h), out)
        r = np.array(list(out))
        assert np.linalg.norm(r) == pytest.approx(1.0, abs=1e-12)
        assert r @ h == pytest.approx(wo @ h, abs=1e-12)
        assert np.allclose(r + wo, 2 * (wo @ h) * h, atol=1e-12)


# ------------------------------------------ refraction, Fresnel, new BSDFs --
def _pd(a):
    return a.ctypes.data_as(ol.PD)


def _refract(real, i, n, eta):
    out = np.zeros(3)
    ok = ol.fn("oracle_refract", real)(d3(i), d3(n), eta, _pd(out))
    return bool(ok), out


@pytest.mark.parametrize("real", REALS)
def test_refract_and_fresnel_match_reference_known_answers(real):
    # test/math.cu:6-29: refract through n = z at 1/1.33 and back, Fresnel(1.4)
    wo = np.array([1.0, 0.0, 1.0]) / math.sqrt(2.0)
    ok, wi = _refract(real, wo, [0, 0, 1], 1 / 1.33)
    assert ok and np.allclose(wi, [-0.531659, 0.0, -0.846958], atol=1e-4)
    ok, wi2 = _refract(real, wi, [0, 0, -1], 1.33)
    assert ok and np.allclose(wi2, [0.707107, 0.0, 0.707107], atol=1e-4)
    assert ol.fn("oracle_fresnel", real)(1.4, d3(wo)) == pytest.approx(0.0365785, abs=1e-4)


def test_refract_fresnel_total_internal_reflection_and_limits():
    # beyond the critical angle refract fails and Fresnel reflects everything
    ok, _ = _refract("f64", [math.sin(1.2), 0, -math.cos(1.2)], [0, 0, -1], 1.5)
    assert not ok
    assert ol.fn("oracle_fresnel", "f64")(1.5, d3([math.sin(1.2), 0, -math.cos(1.2)])) == 1.0
    # normal incidence: ((n - 1) / (n + 1))^2
    assert ol.fn("oracle_fresnel", "f64")(1.5, d3([0, 0, 1])) == pytest.approx((0.5 / 2.5) ** 2, rel=1e-9)
    # Snell: n1 sin(i) = n2 sin(t) for a refracted direction
    for th in (0.1, 0.5, 0.9):
        ok, t = _refract("f64", [math.sin(th), 0, math.cos(th)], [0, 0, 1], 1 / 1.5)
        assert ok and math.sin(th) == pytest.approx(1.5 * abs(t[0]), rel=1e-9) and t[2] < 0


def test_refract_fresnel_vs_reference_headers():
    ref = ol.ref()
    if ref is None:
        pytest.skip("reference shim not built")
    ref.ref_refract.argtypes = [ol.PD, ol.PD, D, ol.PD]
    ref.ref_refract.restype = C.c_int
    ref.ref_fresnel.argtypes = [D, ol.PD]
    ref.ref_fresnel.restype = D
    rng = np.random.default_rng(7)
    for _ in range(400):
        i = rng.normal(size=3)
        i /= np.linalg.norm(i)
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        eta = float(rng.choice([1 / 1.5, 1.5, 1 / 1.33, 1.33]))
        a = np.zeros(3)
        oka = ref.ref_refract(d3(i), d3(n), eta, _pd(a))
        okb, b = _refract("f32", i, n, eta)
        assert bool(oka) == okb
        if okb:
            assert np.allclose(a, b, atol=2e-6)
        ior = float(rng.uniform(1.1, 2.4))
        assert ref.ref_fresnel(ior, d3(i)) == pytest.approx(ol.fn("oracle_fresnel", "f32")(ior, d3(i)), abs=2e-6)


def _shader(typ, **kw):
    from koishi_amd import kr
    d = kr.ShaderDesc()
    d.type = typ
    d.color = (kr.f32 * 3)(*kw.get("color", (0.8, 0.6, 0.4)))
    d.R = (kr.f32 * 3)(1, 1, 1)
    d.roughness = kw.get("roughness", 0.5)
    d.distribution = kw.get("dist", 0)
    d.ior = kw.get("ior", 1.5)
    return d


def _sample(real, sh, wo, u0, u1, uc, wi_in=(9.0, 9.0, 9.0)):
    wi, f = np.zeros(3), np.zeros(3)
    ol.fn("oracle_shader_sample", real)(C.byref(sh), d3(wo), u0, u1, 0.0, uc, d3(wi_in), _pd(wi), _pd(f))
    return wi, f


def test_transparent_blackhole_refraction_glass_closed_forms():
    from koishi_amd import kr
    wo = np.array([0.3, -0.2, 0.9])
    wo /= np.linalg.norm(wo)
    # Transparent (transparent.hpp:20-23): straight through with f = color
    wi, f = _sample("f64", _shader(kr.SHADER_TRANSPARENT), wo, 0.3, 0.7, 0.0)
    assert np.allclose(wi, -wo) and np.allclose(f, [0.8, 0.6, 0.4], atol=1e-7)
    # Blackhole (blackhole.hpp:14-21): wi = z, f = 0
    wi, f = _sample("f64", _shader(kr.SHADER_BLACKHOLE), wo, 0.3, 0.7, 0.0)
    assert np.allclose(wi, [0, 0, 1]) and not f.any()
    # Beckmann with u0 = 1 samples the macro normal exactly (tan^2 = -a^2 ln 1 = 0):
    # RefractionBSDF then refracts through z (Snell) with f = color / e^2 (e < 1)
    ok, t = _refract("f64", wo, [0, 0, 1], 1 / 1.5)
    wi, f = _sample("f64", _shader(kr.SHADER_REFRACTION), wo, 1.0, 0.25, 0.0)
    assert ok and np.allclose(wi, t, atol=1e-12)
    assert np.allclose(f, np.array([0.8, 0.6, 0.4]) / (1 / 1.5) ** 2, rtol=1e-6)
    # GlassBSDF: reflection when uc < Fresnel(ior) (glass.hpp:22-46), refraction otherwise
    fr = ol.fn("oracle_fresnel", "f64")(1.5, d3(wo))
    wi, f = _sample("f64", _shader(kr.SHADER_GLASS), wo, 1.0, 0.25, fr * 0.5)
    assert np.allclose(wi, [-wo[0], -wo[1], wo[2]], atol=1e-12) and np.allclose(f, [0.8, 0.6, 0.4], atol=1e-7)
    wi, f = _sample("f64", _shader(kr.SHADER_GLASS), wo, 1.0, 0.25, min(1.0, fr * 1.5 + 0.01))
    assert np.allclose(wi, t, atol=1e-12)
    # total internal reflection leaves wi as it was, with f = 0
    wo_in = np.array([math.sin(1.2), 0, -math.cos(1.2)])
    wi, f = _sample("f64", _shader(kr.SHADER_REFRACTION), wo_in, 1.0, 0.25, 0.0, wi_in=(0.1, 0.2, 0.3))
    assert np.allclose(wi, [0.1, 0.2, 0.3]) and not f.any()


def test_glass_eval_and_zero_eval_shaders():
    from koishi_amd import kr
    wo = np.array([0.2, 0.1, 0.95])
    wo /= np.linalg.norm(wo)
    wi = np.array([-0.4, 0.3, 0.8])
    wi /= np.linalg.norm(wi)
    f, e = np.zeros(3), np.zeros(3)
    # Cosine distribution: D(h) = 1/pi for h in the upper hemisphere
    sh = _shader(kr.SHADER_GLASS, dist=kr.DIST_COSINE)
    ol.fn("oracle_shader_eval", "f64")(C.byref(sh), d3(wo), d3(wi), 0.0, _pd(f), _pd(e))
    assert np.allclose(f, np.array([0.8, 0.6, 0.4]) / math.pi, rtol=1e-6)
    for typ in (kr.SHADER_REFRACTION, kr.SHADER_TRANSPARENT, kr.SHADER_BLACKHOLE):
        ol.fn("oracle_shader_eval", "f64")(C.byref(_shader(typ)), d3(wo), d3(wi), 0.5, _pd(f), _pd(e))
        assert not f.any() and not e.any()
