import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def golden_dir():
    return os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="session")
def gpu_api():
    """The HIP library, loaded and probed; GPU tests fail (not skip) when the
    device or the library is missing."""
    from sovereign_amd import api

    info = api.device_info()
    print(f"[gpu] {info['name']} {info['arch']} available={info['available']} {info['reason']}", flush=True)
    assert info["available"], info["reason"]
    return api
