"""bench.py's command line on a machine without a GPU: more than one GPU is
refused with a message and a nonzero exit before any device is touched, and
the help lists the workloads and the cache options."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2")
    assert r.returncode != 0
    assert "one MI355X" in (r.stdout + r.stderr)


def test_help_lists_workloads_and_cache_modes():
    r = _bench("--help")
    assert r.returncode == 0
    for word in ("mesh", "spheres", "large", "room", "batch", "--cache-mode", "--cache-ladder"):
        assert word in r.stdout, word


def test_noise_bias_fit_recovers_known_terms():
    """The fit of mse(m) = b^2 + sigma^2 / m behind config.rmse_fit: on exact
    data it returns b and sigma, with the oracle's own noise sigma^2 / 1024
    taken off b^2; on data with b = 0 it reports no bias."""
    sys.path.insert(0, ROOT)
    import bench
    ms = [1, 2, 4, 8, 16]
    b, sigma = 0.03, 0.09
    mse = [b * b + sigma * sigma / m + sigma * sigma / 1024 for m in ms]
    fit = bench.fit_noise_bias(ms, mse, 1024)
    assert abs(fit["bias"] - b) < 1e-9 and abs(fit["sigma_per_frame"] - sigma) < 1e-9
    fit0 = bench.fit_noise_bias(ms, [sigma * sigma / m for m in ms], 1024)
    assert abs(fit0["sigma_per_frame"] - sigma) < 1e-9 and abs(fit0["bias"]) < 3e-3
