"""Pins of the FP64 radiance-cache oracle (oracle/nrc_oracle.py): its
backward pass against central finite differences of its forward pass, the
encoding against closed forms, and one Adam step against the published
update rule written out by hand."""
import numpy as np

from nrc_util import BOX_MAX, BOX_MIN, nrc_oracle as O, random_records, smooth_target


def _loss(ws, bs, x, t):
    y, _ = O.forward(ws, bs, x)
    return O.loss_and_grad_out(y, t)[0]


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    raw = random_records(24, 2)
    x = O.encode(raw[:, 0:3], raw[:, 3:6], raw[:, 6:9], raw[:, 9], raw[:, 10:13], BOX_MIN, BOX_MAX)
    t = smooth_target(raw).astype(np.float64)
    ws, bs = O.init_params(3)
    bs = [b + rng.normal(0, 0.1, b.shape) for b in bs]  # nonzero biases exercise the bias gradient
    y, acts = O.forward(ws, bs, x)
    loss, g = O.loss_and_grad_out(y, t)
    # the loss treats the denominator as a constant (stop-gradient): check against that
    den = y[:, :3] ** 2 + 0.01

    def loss_sg(ws_, bs_):
        yy, _ = O.forward(ws_, bs_, x)
        return float(((yy[:, :3] - t) ** 2 / den).sum() / (3 * x.shape[0]))

    gws, gbs = O.backward(ws, bs, acts, g)
    h = 1e-6
    for li in range(len(ws)):
        for _ in range(6):
            i, o = rng.integers(0, ws[li].shape[0]), rng.integers(0, ws[li].shape[1])
            wp = [w.copy() for w in ws]
            wm = [w.copy() for w in ws]
            wp[li][i, o] += h
            wm[li][i, o] -= h
            fd = (loss_sg(wp, bs) - loss_sg(wm, bs)) / (2 * h)
            assert abs(fd - gws[li][i, o]) <= 1e-6 + 1e-4 * abs(fd), (li, i, o, fd, gws[li][i, o])
        o = rng.integers(0, bs[li].shape[0])
        bp = [b.copy() for b in bs]
        bm = [b.copy() for b in bs]
        bp[li][o] += h
        bm[li][o] -= h
        fd = (loss_sg(ws, bp) - loss_sg(ws, bm)) / (2 * h)
        assert abs(fd - gbs[li][o]) <= 1e-6 + 1e-4 * abs(fd)


def test_encoding_closed_forms():
    raw = random_records(50, 3)
    x = O.encode(raw[:, 0:3], raw[:, 3:6], raw[:, 6:9], raw[:, 9], raw[:, 10:13], BOX_MIN, BOX_MAX)
    assert x.shape == (50, 96) and np.all(x[:, 80:] == 0)
    p = (raw[:, 0:3] - BOX_MIN) / (BOX_MAX - BOX_MIN)
    # sin^2 + cos^2 = 1 per frequency and axis; k = 3 is sin(8 pi p)
    for k in range(12):
        assert np.allclose(x[:, 6 * k:6 * k + 3] ** 2 + x[:, 6 * k + 3:6 * k + 6] ** 2, 1.0)
    assert np.allclose(x[:, 18:21], np.sin(8 * np.pi * p))
    # direction (0, 0, 1) -> theta 0; (1, 0, 0) -> theta 1/2, phi 1/2
    a = O.sphere_angles(np.array([[0, 0, 1.0], [1, 0, 0], [0, 1, 0], [0, 0, -1]]))
    assert np.allclose(a, [[0, 0.5], [0.5, 0.5], [0.5, 0.75], [1, 0.5]])
    assert np.allclose(x[:, 76], raw[:, 9]) and np.allclose(x[:, 77:80], raw[:, 10:13])


def test_adam_step_rule():
    p = [np.array([1.0, -2.0])]
    g = [np.array([0.5, 0.25])]
    m, v = [np.zeros(2)], [np.zeros(2)]
    O.adam_step(p, g, m, v, 1)
    # first step: m_hat = g, v_hat = g^2 -> p -= lr * g / (|g| + eps)
    assert np.allclose(p[0], [1.0 - 1e-2, -2.0 - 1e-2], atol=1e-7)


def test_bf16_rounding():
    assert O.bf16(np.array([1.0]))[0] == 1.0
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: round to even -> 1
    assert O.bf16(np.array([1.0 + 2 ** -8]))[0] == 1.0
    assert O.bf16(np.array([1.0 + 3 * 2 ** -8]))[0] == 1.0 + 2 ** -6
    assert np.isnan(O.bf16(np.array([np.nan]))[0])
