"""FP64 restatement of the neural radiance cache (TEST INFRASTRUCTURE).

The reference has no learned component; the cache follows the published
method (Mueller, Rousselle, Novak, Keller, "Real-time Neural Radiance
Caching for Path Tracing", SIGGRAPH 2021) at the sizes this build uses:

  input   80 features padded to 96:
            position p in [0,1]^3 (scene box): sin / cos(2^k pi p), k < 12  (72)
            outgoing direction (theta/pi, phi/2pi + 1/2)                     (2)
            shading normal     (theta/pi, phi/2pi + 1/2)                     (2)
            roughness (0 for Lambert / Emission)                             (1)
            albedo (shader color)                                            (3)
  layers  96->64, 4 x (64->64), 64->16 (outputs 0..2 used), ReLU between,
          linear output; BF16 weights and activations, FP32 accumulation
  loss    relative L2: mean over samples and RGB of (y - t)^2 / (sg(y)^2 + 0.01)
  update  Adam (lr 1e-2, beta1 0.9, beta2 0.99, eps 1e-8) on FP32 masters

Used only by tests/ to check the GPU kernels (forward values, gradients,
one Adam step) and pinned itself by central finite differences.
"""
from __future__ import annotations

import numpy as np

IN_FEATURES = 80
IN = 96
WIDTH = 64
OUT = 16
LAYERS = [(IN, WIDTH), (WIDTH, WIDTH), (WIDTH, WIDTH), (WIDTH, WIDTH), (WIDTH, WIDTH), (WIDTH, OUT)]
FREQS = 12
LR, BETA1, BETA2, EPS = 1e-2, 0.9, 0.99, 1e-8


def bf16(x):
    """Round to bfloat16 (round to nearest even), returned as float64."""
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    nan = np.isnan(a)
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    out[nan] = np.nan
    return out


def sphere_angles(w):
    w = np.asarray(w, np.float64)
    th = np.arccos(np.clip(w[:, 2], -1.0, 1.0)) / np.pi
    ph = np.arctan2(w[:, 1], w[:, 0]) / (2 * np.pi) + 0.5
    return np.stack([th, ph], axis=1)


def encode(pos, wo, normal, roughness, albedo, box_min, box_max):
    """Feature vector of a path vertex, (n, 96), float64."""
    pos = np.asarray(pos, np.float64)
    p = (pos - np.asarray(box_min)) / (np.asarray(box_max) - np.asarray(box_min))
    feats = []
    for k in range(FREQS):
        a = (2.0 ** k) * np.pi * p
        feats.append(np.sin(a))
        feats.append(np.cos(a))
    enc = np.concatenate(feats, axis=1)  # (n, 72) ordered [sin k0 xyz, cos k0 xyz, sin k1 ...]
    x = np.concatenate([enc, sphere_angles(wo), sphere_angles(normal), np.asarray(roughness)[:, None],
                        np.asarray(albedo, np.float64)], axis=1)
    return np.concatenate([x, np.zeros((x.shape[0], IN - IN_FEATURES))], axis=1)


def init_params(seed=5):
    """Uniform(-sqrt(6/(fan_in+fan_out)), +...) weights (Glorot), zero biases."""
    rng = np.random.default_rng(seed)
    ws, bs = [], []
    for fi, fo in LAYERS:
        lim = np.sqrt(6.0 / (fi + fo))
        ws.append(rng.uniform(-lim, lim, (fi, fo)))
        bs.append(np.zeros(fo))
    return ws, bs


def forward(ws, bs, x, quantize=False):
    """Returns (outputs (n, 16), list of layer inputs). With quantize=True the
    weights, the input and every hidden activation are rounded to BF16 where
    the GPU kernels round them."""
    q = bf16 if quantize else (lambda a: a)
    h = q(x)
    acts = [h]
    for li, (w, b) in enumerate(zip(ws, bs)):
        z = h @ q(w) + b
        if li < len(ws) - 1:
            h = q(np.maximum(z, 0.0))
            acts.append(h)
        else:
            h = z
    return h, acts


def loss_and_grad_out(y, target):
    """Relative L2 on the 3 used outputs; returns (loss, dL/dy (n, 16))."""
    n = y.shape[0]
    yy = y[:, :3]
    den = yy * yy + 0.01
    diff = yy - target
    loss = float((diff * diff / den).sum() / (3 * n))
    g = np.zeros_like(y)
    g[:, :3] = 2.0 * diff / den / (3 * n)
    return loss, g


def backward(ws, bs, acts, gout, quantize=False):
    """Gradients of the loss for each weight and bias, given dL/dy."""
    q = bf16 if quantize else (lambda a: a)
    gws, gbs = [None] * len(ws), [None] * len(ws)
    d = gout
    for li in range(len(ws) - 1, -1, -1):
        h = acts[li]
        gws[li] = h.T @ d
        gbs[li] = d.sum(axis=0)
        if li > 0:
            dh = q(d) @ q(ws[li]).T
            d = dh * (acts[li] > 0)
    return gws, gbs


def adam_step(params, grads, m, v, t):
    """One Adam update in place; t counts from 1."""
    for k in range(len(params)):
        g = grads[k]
        m[k] = BETA1 * m[k] + (1 - BETA1) * g
        v[k] = BETA2 * v[k] + (1 - BETA2) * g * g
        mh = m[k] / (1 - BETA1 ** t)
        vh = v[k] / (1 - BETA2 ** t)
        params[k] -= LR * mh / (np.sqrt(vh) + EPS)


def flatten(ws, bs):
    """Parameter order of the GPU buffers: W0 (row-major in x out), b0, W1, b1, ..."""
    return np.concatenate([np.concatenate([w.ravel(), b]) for w, b in zip(ws, bs)])


def unflatten(flat):
    ws, bs, o = [], [], 0
    for fi, fo in LAYERS:
        ws.append(flat[o:o + fi * fo].reshape(fi, fo))
        o += fi * fo
        bs.append(flat[o:o + fo])
        o += fo
    return ws, bs


NUM_PARAMS = sum(fi * fo + fo for fi, fo in LAYERS)
