"""Shared helpers of the radiance-cache tests: loads oracle/nrc_oracle.py
(the FP64 restatement of the cache, test infrastructure) and makes records."""
import importlib.util
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_spec = importlib.util.spec_from_file_location("nrc_oracle", os.path.join(ROOT, "oracle", "nrc_oracle.py"))
nrc_oracle = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(nrc_oracle)

BOX_MIN = np.array([-2.0, -2.0, -1.5])
BOX_MAX = np.array([2.0, 2.0, 1.5])


def random_records(n, seed=0):
    """Vertex records: position in the box, unit outgoing direction and
    normal, roughness in {0} u [0.05, 0.8], albedo in [0.1, 0.9]."""
    rng = np.random.default_rng(seed)
    raw = np.zeros((n, 16), np.float32)
    raw[:, 0:3] = rng.uniform(BOX_MIN, BOX_MAX, (n, 3))
    for a in (3, 6):
        v = rng.normal(size=(n, 3))
        raw[:, a:a + 3] = v / np.linalg.norm(v, axis=1, keepdims=True)
    raw[:, 9] = np.where(rng.uniform(size=n) < 0.5, 0.0, rng.uniform(0.05, 0.8, n))
    raw[:, 10:13] = rng.uniform(0.1, 0.9, (n, 3))
    return raw


def smooth_target(raw):
    """A smooth radiance field of the record (for learning tests)."""
    p = raw[:, 0:3].astype(np.float64)
    base = 0.5 + 0.3 * np.sin(1.3 * p[:, 0]) * np.cos(0.7 * p[:, 1]) + 0.2 * p[:, 2] / 1.5
    return (raw[:, 10:13] * base[:, None] * (1.0 + 0.5 * raw[:, 5:6] ** 2)).astype(np.float32)
