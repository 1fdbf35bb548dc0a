"""The radiance cache's gfx950 kernels (MFMA BF16 forward / backward, Adam)
against the FP64 oracle (oracle/nrc_oracle.py), through the C ABI."""
import numpy as np
import pytest

from koishi_amd import kr
from nrc_util import BOX_MAX, BOX_MIN, nrc_oracle as O, random_records, smooth_target

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nrc(device):
    n = kr.Nrc(device, BOX_MIN, BOX_MAX, seed=7)
    yield n
    n.free()


def _oracle_params(seed=11, bias_scale=0.05):
    ws, bs = O.init_params(seed)
    rng = np.random.default_rng(seed)
    bs = [rng.normal(0, bias_scale, b.shape) for b in bs]
    flat = O.flatten(ws, bs).astype(np.float32)
    return flat


def test_encoding_matches_oracle(nrc):
    raw = random_records(1000, 1)
    g = nrc.encode(raw)
    c = O.encode(raw[:, 0:3], raw[:, 3:6], raw[:, 6:9], raw[:, 9], raw[:, 10:13], BOX_MIN, BOX_MAX)
    # FP32 sin/cos of arguments up to 2^11 pi: |error| ~ 1e-3 at the top frequency
    assert np.abs(g[:, :72] - c[:, :72]).max() < 2e-3
    assert np.abs(g[:, 72:] - c[:, 72:]).max() < 1e-5


def test_forward_matches_oracle(nrc):
    flat = _oracle_params()
    nrc.set_params(flat)
    assert np.array_equal(nrc.get_params(), flat)
    raw = random_records(777, 2)  # not a multiple of 64: exercises padding
    y = nrc.predict(raw)
    x = nrc.encode(raw).astype(np.float64)  # the GPU's own features isolate the MLP
    ws, bs = O.unflatten(flat.astype(np.float64))
    ref, _ = O.forward(ws, bs, x, quantize=True)
    # BF16 activations: an FP32 accumulation that lands on the other side of
    # a BF16 rounding boundary than the FP64 oracle's moves one activation by
    # one BF16 ulp (2^-8 relative) and that propagates. Such flips are rare,
    # so nearly all outputs agree to FP32 accumulation error and the few
    # others stay within a few BF16 ulps of the output scale.
    scale = max(1.0, np.abs(ref[:, :3]).max())
    err = np.abs(y - ref[:, :3]) / scale
    assert np.quantile(err, 0.95) <= 1e-5, np.quantile(err, 0.95)
    assert err.max() <= 2e-2, err.max()


def test_gradient_matches_oracle(nrc):
    flat = _oracle_params(12)
    nrc.set_params(flat)
    raw = random_records(640, 3)
    t = smooth_target(raw)
    g, loss = nrc.gradient(raw, t)
    x = nrc.encode(raw).astype(np.float64)
    ws, bs = O.unflatten(flat.astype(np.float64))
    y, acts = O.forward(ws, bs, x, quantize=True)
    lref, gout = O.loss_and_grad_out(y, t.astype(np.float64))
    gws, gbs = O.backward(ws, bs, acts, gout, quantize=True)
    gref = O.flatten(gws, gbs)
    assert loss == pytest.approx(lref, rel=1e-4)
    rel = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    assert rel < 1e-3, rel
    # per layer, including the first (96 -> 64) and the output layer
    o = 0
    for fi, fo in O.LAYERS:
        n = fi * fo + fo
        a, r = g[o:o + n], gref[o:o + n]
        assert np.linalg.norm(a - r) <= 2e-3 * np.linalg.norm(r) + 1e-9, (fi, fo)
        o += n


def test_adam_step_matches_oracle(nrc):
    flat = _oracle_params(13)
    nrc.set_params(flat)
    raw = random_records(256, 4)
    t = smooth_target(raw)
    g, _ = nrc.gradient(raw, t)
    nrc.train(raw, t, steps=1)
    after = nrc.get_params()
    ws, bs = O.unflatten(flat.astype(np.float64))
    params = [flat.astype(np.float64)]
    O.adam_step(params, [g.astype(np.float64)], [np.zeros_like(params[0])], [np.zeros_like(params[0])], 1)
    assert np.abs(after - params[0]).max() < 1e-5


def test_training_reduces_loss(nrc):
    nrc.set_params(_oracle_params(14, 0.0))
    raw = random_records(16384, 5)
    t = smooth_target(raw)
    first = nrc.train(raw, t, steps=1)
    for _ in range(19):
        last = nrc.train(raw, t, steps=10)
    held = random_records(4096, 6)
    tgt = smooth_target(held)
    rel = np.abs(nrc.predict(held) - tgt).mean() / np.abs(tgt).mean()
    assert last < 0.1 * first, (first, last)
    # held-out records: the fitted network generalises (an untrained one is
    # off by ~100 %, predicting the per-channel mean is off by 54 %)
    assert rel < 0.35, rel


# ----------------------------------------------------------- in the render --
def _box(scene):
    p = np.asarray(scene.positions, np.float64)
    lo, hi = p.min(axis=0) - 0.01, p.max(axis=0) + 0.01
    for c, r, _ in scene.spheres:
        if r < 100:
            lo = np.minimum(lo, np.array(c) - r)
            hi = np.maximum(hi, np.array(c) + r)
    return lo, hi


def test_all_training_paths_leave_the_image_unchanged(device):
    """With every path a training path nothing is cut short: the image must
    be bit-identical to the plain render (the record keeping only reads)."""
    from koishi_amd import scenegen as sg
    s = sg.mesh_scene(freq=20)
    gs = kr.GpuScene(device, s.to_desc())
    p = s.params(seed=9, width=96, height=48, spp=4)
    imgs = []
    for with_cache in (False, True):
        r = kr.Render(gs, s.camera_struct(), p)
        n = None
        if with_cache:
            n = kr.Nrc(device, *_box(s), train_log2=kr.NRC_TRAIN_ALL)
            r.attach_nrc(n)
        r.run(0, p.spp)
        r.sync()
        imgs.append(r.read())
        if n is not None:
            raw, tgt = n.read_pool(1 << 18)
            assert raw.shape[0] > 0 and np.isfinite(tgt).all()
            n.free()
        r.free()
    assert np.array_equal(imgs[0][0], imgs[1][0]) and np.array_equal(imgs[0][1], imgs[1][1])
    gs.free()


def test_training_targets_are_path_radiance(device):
    """maxBounce 0 and 1 spp: a training path's only vertex carries exactly the
    path's radiance (NEE, no emission on the floor), so the pool's targets and
    the image's pixels are the same multiset."""
    from koishi_amd import scenegen as sg
    b = sg.Builder()
    b.shaders["floor"] = ["Lambert", {"color": [0.6, 0.5, 0.4]}]
    b.shaders["light"] = ["Emission", {"color": [1.0, 1.0, 1.0], "strength": 20.0}]
    v, nn, t = sg.quad((0.0, 0.0, 0.0), 8.0, 8.0, 1.0)
    b.add_mesh("floor", "floor", v, nn, t)
    v, nn, t = sg.quad((0.7, 2.0, -0.3), 1.5, 1.0, -1.0)
    b.add_mesh("light", "light", v, nn, t)
    s = b.finish((0.3, 1.0, 0.2), (0.35, 0.0, 0.25), 60.0, 48, 32, 1, 0, "nee")
    gs = kr.GpuScene(device, s.to_desc())
    p = s.params(seed=3)
    r = kr.Render(gs, s.camera_struct(), p)
    n = kr.Nrc(device, *_box(s), train_log2=kr.NRC_TRAIN_ALL, steps_per_frame=1)
    r.attach_nrc(n)
    r.run(0, 1)
    r.sync()
    rgb, val = r.read()
    raw, tgt = n.read_pool(1 << 16)
    lit = (val == 1) & (rgb.max(axis=1) > 0)
    a = np.sort(rgb[lit].astype(np.float32), axis=0)
    bb = np.sort(tgt[tgt.max(axis=1) > 0], axis=0)
    assert a.shape == bb.shape
    assert np.allclose(a, bb, rtol=1e-6, atol=1e-7)
    n.free()
    r.free()
    gs.free()


def _floor_light_scene(width=48, height=32, max_bounce=0):
    """A Lambert floor lit by an Emission quad: every vertex is diffuse."""
    from koishi_amd import scenegen as sg
    b = sg.Builder()
    b.shaders["floor"] = ["Lambert", {"color": [0.6, 0.5, 0.4]}]
    b.shaders["light"] = ["Emission", {"color": [1.0, 1.0, 1.0], "strength": 20.0}]
    v, nn, t = sg.quad((0.0, 0.0, 0.0), 8.0, 8.0, 1.0)
    b.add_mesh("floor", "floor", v, nn, t)
    v, nn, t = sg.quad((0.7, 2.0, -0.3), 1.5, 1.0, -1.0)
    b.add_mesh("light", "light", v, nn, t)
    return b.finish((0.3, 1.0, 0.2), (0.35, 0.0, 0.25), 60.0, width, height, 1, max_bounce, "nee")


def _one_spp(device, gs, s, p, nrc=None, stats=None):
    r = kr.Render(gs, s.camera_struct(), p)
    if nrc is not None:
        r.attach_nrc(nrc)
    r.run(0, 1)
    r.sync()
    out = r.read()
    if stats is not None:
        stats.append(r.stats())
    r.free()
    return out


def test_indirect_targets_are_the_continuation(device):
    """KR_NRC_INDIRECT targets leave out the vertex's own light sample and
    emission. maxBounce 1, 1 spp, camera throughput 1: the first vertex's
    target is what the second vertex added, i.e. the pixel of the 2-vertex
    render minus the pixel of the 1-vertex render (same RNG streams); the
    second vertex's target is 0."""
    s = _floor_light_scene(max_bounce=1)
    gs = kr.GpuScene(device, s.to_desc())
    p1 = s.params(seed=3)
    p0 = s.params(seed=3)
    p0.max_bounce = 0
    rgb0, _ = _one_spp(device, gs, s, p0)
    n = kr.Nrc(device, *_box(s), train_log2=kr.NRC_TRAIN_ALL, steps_per_frame=1, flags=kr.NRC_INDIRECT)
    rgb1, val1 = _one_spp(device, gs, s, p1, n)
    raw, tgt = n.read_pool(1 << 16)
    cont = (rgb1 - rgb0)[val1 == 1]
    a = np.sort(cont[cont.max(axis=1) > 0].astype(np.float32), axis=0)
    bb = np.sort(tgt[tgt.max(axis=1) > 0], axis=0)
    assert a.shape == bb.shape and a.shape[0] > 50
    assert np.allclose(a, bb, rtol=1e-5, atol=1e-5)
    # as many records as vertices: first vertices and the second ones
    assert raw.shape[0] >= 2 * a.shape[0] - 1
    n.free()
    gs.free()


def test_indirect_paths_end_at_the_first_diffuse_vertex(device):
    """With KR_NRC_INDIRECT and an all-zero frozen cache a path ends after its
    first diffuse vertex's light sample and emission: on a scene whose
    vertices are all diffuse the image is the 1-vertex render's, except for
    the training paths (one per group of 2^30 paths at most: 2 here)."""
    s = _floor_light_scene(max_bounce=8)
    gs = kr.GpuScene(device, s.to_desc())
    p0 = s.params(seed=5)
    p0.max_bounce = 0
    rgb0, val0 = _one_spp(device, gs, s, p0)
    n = kr.Nrc(device, *_box(s), train_log2=30, flags=kr.NRC_FROZEN | kr.NRC_INDIRECT)
    n.set_params(np.zeros(kr.NRC_PARAMS, np.float32))
    st = []
    rgb, val = _one_spp(device, gs, s, s.params(seed=5), n, st)
    assert np.array_equal(val, val0)
    differ = np.any(rgb != rgb0, axis=1)
    assert differ.sum() <= 2, differ.sum()
    # one query per camera ray that hits the scene, training paths aside
    # (every hit is lit: the light sees the whole floor and itself emits)
    st0 = []
    _one_spp(device, gs, s, p0, None, st0)
    hits = int((rgb0.max(axis=1) > 0).sum())
    assert st0[0].cache_queries == 0 and hits > 100
    assert hits - 2 <= st[0].cache_queries <= min(hits, st0[0].closest_rays)
    # without the flag the same cache ends paths at their second vertex
    m = kr.Nrc(device, *_box(s), train_log2=30, flags=kr.NRC_FROZEN)
    m.set_params(np.zeros(kr.NRC_PARAMS, np.float32))
    rgb2, _ = _one_spp(device, gs, s, s.params(seed=5), m)
    assert np.any(rgb2 != rgb0, axis=1).sum() > 10
    n.free()
    m.free()
    gs.free()


def _mean_image(rgb, val):
    return rgb / np.maximum(val, 1)[:, None]


@pytest.mark.parametrize("flags", [0, kr.NRC_INDIRECT])
def test_cached_render_learns_and_gains(device, flags):
    """What the trained cache gains over the method without it. Paths end at
    their second diffuse or rough vertex in both cached renders; a frozen
    all-zero network stands for "no cache" (the cut-off radiance is dropped),
    the trained network supplies it. Against a converged plain render the
    trained cache must remove most of the truncation's bias and lower the
    per-pixel error, with every pixel valid and finite. (The remaining bias
    is the cache's error after 16 short frames of online training.)"""
    from koishi_amd import scenegen as sg
    s = sg.room_scene(width=256, height=192, spp=16)  # closed room: most radiance arrives after the second vertex
    gs = kr.GpuScene(device, s.to_desc())
    p = s.params(seed=2)
    ref = 0
    for k in range(8):
        plain = kr.Render(gs, s.camera_struct(), s.params(seed=100 + k))
        plain.run(0, p.spp)
        ref = ref + _mean_image(*plain.read())
        plain.free()
    ref /= 8

    zero = kr.Nrc(device, *_box(s), train_log2=5, flags=kr.NRC_FROZEN | flags)
    zero.set_params(np.zeros(kr.NRC_PARAMS, np.float32))
    trained = kr.Nrc(device, *_box(s), train_log2=5, steps_per_frame=16, flags=flags)
    out = {}
    for name, n in (("zero", zero), ("trained", trained)):
        r = kr.Render(gs, s.camera_struct(), p)
        r.attach_nrc(n)
        for frame in range(16 if name == "trained" else 1):
            r.reset()
            r.run(0, p.spp)
            r.sync()
        rgb, val = r.read()
        assert (val > 0).all() and np.isfinite(rgb).all()
        out[name] = _mean_image(rgb, val)
        r.free()
    m = ref.mean()
    bias_zero = abs(out["zero"].mean() - m)
    bias_trained = abs(out["trained"].mean() - m)
    err = {k: np.sqrt(((v - ref) ** 2).mean()) for k, v in out.items()}
    assert bias_zero > 0.05 * m, (bias_zero, m)  # the cut-off radiance matters here
    assert bias_trained < 0.5 * bias_zero, (bias_trained, bias_zero, m)
    assert err["trained"] < err["zero"], err
    zero.free()
    trained.free()
    gs.free()


@pytest.mark.parametrize("decay", [0.9, -1.0])
def test_renders_query_the_moving_average(device, decay):
    """Renders query the bias-corrected moving average of the trained
    parameters, m_t = m_(t-1) + (p_t - m_(t-1)) (1 - d) / (1 - d^t) over the
    Adam steps t (m_1 = p_1), recomputed here on the host from the trained
    parameters after every step; with a negative decay they query the trained
    parameters. The proof that a render queries exactly these: its image
    equals the image of a frozen cache set to them."""
    from koishi_amd import scenegen as sg
    s = sg.room_scene(width=96, height=64, spp=4, freq=6)
    gs = kr.GpuScene(device, s.to_desc())
    a = kr.Nrc(device, *_box(s), train_log2=3, steps_per_frame=1, ema_decay=decay)
    p0 = a.get_params()
    assert np.array_equal(a.get_render_params(), p0)
    r = kr.Render(gs, s.camera_struct(), s.params(seed=4))
    r.attach_nrc(a)
    m = p0.copy()
    for t in range(1, 7):
        r.reset()
        r.run(0, s.spp)
        r.sync()
        p = a.get_params()
        assert not np.array_equal(p, p0)  # one Adam step per run
        if decay > 0:
            k = np.float32((1.0 - decay) / (1.0 - decay ** t))
            m = (m + (p - m) * k).astype(np.float32)
        else:
            m = p
        np.testing.assert_allclose(a.get_render_params(), m, rtol=1e-5, atol=1e-7)
        p0 = p
    if decay > 0:
        assert np.abs(m - p).max() > 1e-4  # the average lags the trained parameters
    # the next run queries the current average (training follows the run)
    b = kr.Nrc(device, *_box(s), train_log2=3, flags=kr.NRC_FROZEN)
    b.set_params(a.get_render_params())
    out = []
    for n in (a, b):
        rn = kr.Render(gs, s.camera_struct(), s.params(seed=9))
        rn.attach_nrc(n)
        rn.run(0, s.spp)
        out.append(rn.read())
        rn.free()
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][0], out[1][0])
    r.free()
    a.free()
    b.free()
    gs.free()
