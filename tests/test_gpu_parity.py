"""Parity of the gfx950 path (through the C ABI in libkr.so) with the FP64
oracle (oracle/liboracle.so) on identical inputs.

Bars (the hot path's contract):
  * closest hit: primitive (and triangle) agreement >= 99.99% on identical
    rays; where both agree, hit t within 1e-5 relative to t on camera rays
    and to max(t, 1) on rays from surfaces or anywhere in the scene, on every
    hit;
  * radiance: the GPU draws the oracle's Philox numbers, so a path's radiance
    is compared path by path; FP32 shading vs FP64 makes a small fraction of
    paths branch differently (Russian roulette threshold, grazing hits), so
    >= 99.9% of paths must agree to 1e-3 relative and the images' means to
    0.5% / within Monte Carlo error.
"""
import numpy as np
import pytest

import oracle_lib as ol
from koishi_amd import kr, scenegen as sg

pytestmark = pytest.mark.gpu

REL_T = 1e-5


def _camera_rays(scene, params, k=0, pixels=None):
    o, d = ol.lens_rays(scene, params, "f32", k, pixels)
    return o.astype(np.float32), d.astype(np.float32)


def _surface_rays(scene, orc, n, seed):
    """Rays leaving surface points in random directions, spawned like
    LocalVaryings::emitRay (origin offset 4e-3 along the new direction)."""
    rng = np.random.default_rng(seed)
    p = scene.params(width=64, height=32)
    o, d = _camera_rays(scene, p)
    t, prim, _, _, _ = orc.trace_closest(o, d)
    hit = prim != kr.KR_MISS
    pts = (o[hit] + d[hit] * t[hit, None]).astype(np.float64)
    idx = rng.integers(0, pts.shape[0], n)
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    org = pts[idx] + dirs * 4e-3
    return org.astype(np.float32), dirs.astype(np.float32)


def _random_rays(scene, n, seed, spread=6.0):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-spread, spread, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o.astype(np.float32), d.astype(np.float32)


def _compare_closest(g, c, min_agree=0.9999, t_floor=1.0):
    gt, gp, gtri, gu, gv = g
    ct, cp, ctri, cu, cv = c
    n = gt.shape[0]
    agree = (gp == cp) & ((gp == kr.KR_MISS) | (gtri == ctri))
    frac = agree.mean()
    assert frac >= min_agree, f"closest-hit agreement {frac:.6f} over {n} rays"
    both = agree & (gp != kr.KR_MISS)
    # relative to max(t, t_floor): camera rays (t_floor 0) are held to t
    # itself; rays that start on a surface or anywhere in the scene often hit
    # at t << 1, where the FP32 resolution of O(1-10) world coordinates
    # (~1e-6 absolute) bounds any FP32 intersection, so they are held to
    # max(t, 1)
    rel = np.abs(gt[both].astype(np.float64) - ct[both]) / np.maximum(np.abs(ct[both]), t_floor)
    bad = (rel > REL_T).mean() if both.any() else 0.0
    return frac, rel, bad


@pytest.fixture(scope="module")
def spheres():
    return sg.sphere_scene()


@pytest.fixture(scope="module")
def mesh():
    return sg.mesh_scene()


def test_closest_hit_sphere_scene(device, spheres):
    gs = kr.GpuScene(device, spheres.to_desc())
    orc = ol.OracleScene(spheres, "f64")
    p = spheres.params()
    rays = [_camera_rays(spheres, p, 0), _camera_rays(spheres, p, 15), _random_rays(spheres, 100000, 1)]
    for k, (o, d) in enumerate(rays):
        frac, rel, bad = _compare_closest(gs.trace_closest(o, d), orc.trace_closest(o, d), t_floor=0.0 if k < 2 else 1.0)
        assert bad == 0, f"t beyond {REL_T} relative on {bad:.2e} of hits (max {rel.max():.2e})"
    gs.free()


def test_closest_hit_mesh_scene(device, mesh):
    gs = kr.GpuScene(device, mesh.to_desc())
    orc = ol.OracleScene(mesh, "f64", traversal=1)
    p = mesh.params(width=256, height=128)
    rays = [_camera_rays(mesh, p, 0), _surface_rays(mesh, orc, 50000, 2), _random_rays(mesh, 50000, 3, 2.0)]
    for k, (o, d) in enumerate(rays):
        g = gs.trace_closest(o, d)
        c = orc.trace_closest(o, d)
        frac, rel, bad = _compare_closest(g, c, t_floor=0.0 if k == 0 else 1.0)
        assert bad == 0, f"t beyond {REL_T} relative on {bad:.2e} of hits (max {rel.max():.2e})"
        both = (g[1] == c[1]) & (g[1] != kr.KR_MISS)
        assert np.abs(g[3][both] - c[3][both]).max() < 1e-3
        assert np.abs(g[4][both] - c[4][both]).max() < 1e-3
    gs.free()


@pytest.mark.parametrize("width", [2, 4, 8])
@pytest.mark.parametrize("nodes", ["compressed", "full"])
def test_closest_hit_bvh_widths(device, mesh, width, nodes):
    flags = kr.KR_SCENE_FULL_NODES if nodes == "full" else 0
    gs = kr.GpuScene(device, mesh.to_desc(bvh_width=width, flags=flags))
    info = gs.info()
    assert info.bvh_width == width
    assert info.node_bytes == ((7 * width + 3) // 4 * 16 if nodes == "full" else (16 + 10 * width + 15) // 16 * 16)
    orc = ol.OracleScene(mesh, "f64", traversal=1)
    o, d = _surface_rays(mesh, orc, 30000, 4)
    _compare_closest(gs.trace_closest(o, d), orc.trace_closest(o, d))
    gs.free()


@pytest.mark.parametrize("leaf", [2, 4, 16])
def test_leaf_sizes_give_identical_hits(device, mesh, leaf):
    """Leaves of several triangles (kr_scene_desc.leaf_size) make another
    tree, but a closest hit is the nearest triangle with exact ties to the
    lowest index and the boxes are conservative, so the hits are the same
    bit for bit as with the default one triangle per leaf."""
    orc = ol.OracleScene(mesh, "f64", traversal=1)
    o, d = _surface_rays(mesh, orc, 30000, 7)
    tmax = np.random.default_rng(7).uniform(0.01, 8.0, o.shape[0]).astype(np.float32)
    one = kr.GpuScene(device, mesh.to_desc())
    many = kr.GpuScene(device, mesh.to_desc(leaf_size=leaf))
    assert many.info().bvh_leaves < one.info().bvh_leaves
    g1, gm = one.trace_closest(o, d), many.trace_closest(o, d)
    for a, b in zip(g1, gm):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(one.trace_shadow(o, d, tmax), many.trace_shadow(o, d, tmax))
    _compare_closest(gm, orc.trace_closest(o, d))
    one.free()
    many.free()


@pytest.mark.parametrize("width,nodes", [(4, "full"), (4, "compressed"), (8, "compressed"), (2, "compressed")])
def test_shadow_segments(device, mesh, spheres, width, nodes):
    flags = kr.KR_SCENE_FULL_NODES if nodes == "full" else 0
    for scene, seed in ((mesh, 5), (spheres, 6)):
        gs = kr.GpuScene(device, scene.to_desc(bvh_width=width, flags=flags))
        orc = ol.OracleScene(scene, "f64", traversal=1)
        o, d = _surface_rays(scene, orc, 40000, seed)
        tmax = np.random.default_rng(seed).uniform(0.01, 8.0, o.shape[0]).astype(np.float32)
        g = gs.trace_shadow(o, d, tmax)
        c = orc.trace_shadow(o, d, tmax)
        assert (g == c).all(), f"shadow agreement {(g == c).mean():.6f}"
        gs.free()


def test_light_tables_match_oracle(device, mesh):
    gs = kr.GpuScene(device, mesh.to_desc())
    orc = ol.OracleScene(mesh, "f64")
    assert gs.info().num_lights == orc.counts()["lights"] == 1
    g, gpdf = gs.light_samples(0)
    c, cpdf = orc.light_samples(0)
    assert np.abs(g - c).max() < 1e-5
    assert abs(gpdf - cpdf) <= 1e-6 * cpdf
    gs.free()


def _per_path(device, scene, w, h, samples, seed, flags=0):
    gs = kr.GpuScene(device, scene.to_desc())
    orc = ol.OracleScene(scene, "f64", traversal=1)
    p = scene.params(seed=seed, width=w, height=h, flags=flags)
    rend = kr.Render(gs, scene.camera_struct(), p)
    out = []
    for k in samples:
        rend.reset()
        rend.run(k, 1)
        rend.sync()
        grgb, gval = rend.read()
        crgb, cval, _ = orc.render(p, k, 1)
        out.append((grgb, gval, crgb, cval))
    rend.free()
    gs.free()
    return out


def _check_paths(out, min_frac=0.999):
    g = np.concatenate([o[0] for o in out])
    gv = np.concatenate([o[1] for o in out])
    c = np.concatenate([o[2] for o in out])
    cv = np.concatenate([o[3] for o in out])
    assert (gv == cv).mean() >= 0.999, "valid-sample flags disagree"
    ok = (gv == 1) & (cv == 1)
    diff = np.abs(g[ok] - c[ok]).max(axis=1)
    tol = 1e-3 * np.maximum(1.0, np.abs(c[ok]).max(axis=1))
    frac = (diff <= tol).mean()
    assert frac >= min_frac, f"only {frac:.4f} of paths agree with the oracle"
    gm, cm = g[ok].mean(axis=0), c[ok].mean(axis=0)
    assert np.all(np.abs(gm - cm) <= 0.005 * np.maximum(cm, 1e-3)), (gm, cm)
    return frac


def test_radiance_per_path_sphere_scene(device, spheres):
    out = _per_path(device, spheres, 128, 128, [0, 5, 11], seed=11)
    _check_paths(out)


def test_radiance_per_path_mesh_scene(device, mesh):
    out = _per_path(device, mesh, 128, 64, [0, 33], seed=12)
    _check_paths(out)


def test_material_sort_does_not_change_results(device, spheres):
    gs = kr.GpuScene(device, spheres.to_desc())
    imgs = []
    for flags in (0, kr.KR_RENDER_NO_MATERIAL_SORT):
        p = spheres.params(seed=3, width=96, height=96, flags=flags)
        r = kr.Render(gs, spheres.camera_struct(), p)
        r.run(0, p.spp)
        r.sync()
        imgs.append(r.read())
        r.free()
    assert np.array_equal(imgs[0][0], imgs[1][0]) and np.array_equal(imgs[0][1], imgs[1][1])
    gs.free()


def test_chunking_does_not_change_results(device, spheres):
    gs = kr.GpuScene(device, spheres.to_desc())
    imgs = []
    for chunk in (0, 5000, 96 * 96 + 7):
        p = spheres.params(seed=4, width=96, height=96, chunk_paths=chunk)
        r = kr.Render(gs, spheres.camera_struct(), p)
        r.run(0, p.spp)
        r.sync()
        imgs.append(r.read())
        r.free()
    for rgb, val in imgs[1:]:
        assert np.array_equal(val, imgs[0][1])
        assert np.allclose(rgb, imgs[0][0], rtol=1e-12, atol=1e-12)
    gs.free()


def test_radiance_converged_vs_oracle(device, spheres):
    """Independent seeds: GPU at 512 spp vs the FP64 oracle at 128 spp on a
    48x48 frame; the mean radiance must agree within Monte Carlo error
    (variance estimated from 4 independent GPU batches)."""
    w = h = 48
    gs = kr.GpuScene(device, spheres.to_desc())
    means = []
    for seed in (21, 22, 23, 24):
        p = spheres.params(seed=seed, width=w, height=h, spp=128)
        r = kr.Render(gs, spheres.camera_struct(), p)
        r.run(0, 128)
        r.sync()
        rgb, val = r.read()
        means.append((rgb / np.maximum(val, 1)[:, None]).mean(axis=0))
        r.free()
    means = np.array(means)
    orc = ol.OracleScene(spheres, "f64")
    p = spheres.params(seed=99, width=w, height=h, spp=128)
    crgb, cval, _ = orc.render(p)
    cm = (crgb / np.maximum(cval, 1)[:, None]).mean(axis=0)
    gm = means.mean(axis=0)
    se = means.std(axis=0, ddof=1) / np.sqrt(4) * np.sqrt(5)  # oracle batch ~ one GPU batch
    assert np.all(np.abs(gm - cm) <= 4 * se + 1e-3 * cm), (gm, cm, se)
    gs.free()


def test_render_stats_and_profiling(device, mesh):
    gs = kr.GpuScene(device, mesh.to_desc())
    p = mesh.params(seed=5, width=64, height=32, spp=4)
    r = kr.Render(gs, mesh.camera_struct(), p)
    r.set_profiling(kr.KR_PROFILE_EVENTS | kr.KR_PROFILE_COUNTERS)
    r.run(0, 4)
    r.sync()
    st = r.stats()
    assert st.paths == 64 * 32 * 4
    assert st.closest_rays >= st.paths
    assert st.node_visits > st.closest_rays and st.tri_tests > 0
    assert st.trace_launches == (p.max_bounce + 1) and st.trace_ms > 0
    with pytest.raises(kr.KrError):
        r.run(0, 5)  # beyond the render's spp
    r.free()
    gs.free()


@pytest.mark.parametrize("workload,part", [("spheres", ""), ("mesh", ""), ("large", ""), ("mesh", "_crop"),
                                           ("large", "_crop"), ("spheres", "_full")])
def test_unbiased_against_converged_oracle(device, workload, part):
    """Full-size frame vs the committed FP64 oracle image (same lens strata,
    65536 samples per pixel on a fixed 4096-pixel subset, on the 64 x 64
    block centred in the frame, or on every pixel of the sphere frame): the
    squared error against the oracle must equal the render's own Monte Carlo
    variance, estimated from independent seeds (bias^2 ~ 0)."""
    import os
    path = os.path.join(ol.ROOT, "tests", "golden", f"oracle_{workload}{part}.npz")
    if not os.path.exists(path):
        pytest.skip(f"oracle image {os.path.basename(path)} not generated (tests/golden/make_oracle_images.py)")
    f = np.load(path)
    import sys
    sys.path.insert(0, os.path.join(ol.ROOT, "tests", "golden"))
    from make_oracle_images import scene_hash
    scene = sg.WORKLOADS[workload]()
    assert scene_hash(scene) == str(f["scene_hash"]), "oracle fixture is stale: regenerate it"
    gs = kr.GpuScene(device, scene.to_desc())
    pix = f["pixels"].astype(np.int64)
    means = []
    for seed in (101, 202, 303, 404):
        p = scene.params(seed=seed)
        r = kr.Render(gs, scene.camera_struct(), p)
        r.run(0, p.spp)
        r.sync()
        rgb, val = r.read()
        r.free()
        assert (val[pix] > 0).all()
        means.append(rgb[pix] / val[pix][:, None])
    gs.free()
    ref = f["mean"]
    err = np.mean([((m - ref) ** 2).mean() for m in means])
    var = np.mean([((means[i] - means[j]) ** 2).mean() / 2 for i, j in ((0, 1), (2, 3), (0, 2), (1, 3))])
    ratio = err / var
    assert 0.8 < ratio < 1.25, f"error^2 / variance = {ratio:.3f} (err {err:.3e}, var {var:.3e})"
