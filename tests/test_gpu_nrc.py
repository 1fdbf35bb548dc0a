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
    err = np.abs(y - ref[:, :3])
    assert err.max() <= 1e-3 * max(1.0, np.abs(ref[:, :3]).max()), err.max()


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
    pred = nrc.predict(held)
    rel = np.abs(pred - smooth_target(held)).mean() / np.abs(smooth_target(held)).mean()
    assert last < 0.1 * first, (first, last)
    assert rel < 0.1, rel
