import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "cuda-raytracer_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X / gfx950); calls through the C ABI")


@pytest.fixture(scope="session")
def device():
    """The first HIP device, opened through kr_device_open. Fails (does not
    skip) when libkr.so or the device is missing: GPU tests only run where a
    GPU is expected."""
    from koishi_amd import kr
    dev = kr.Device(0)
    yield dev
    dev.close()
