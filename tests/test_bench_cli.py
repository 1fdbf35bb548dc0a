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
