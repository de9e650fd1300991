import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (gfx950) and libimpec_hip.so")
    config.addinivalue_line("markers", "slow: longer CPU-side test")


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library with at least one visible device; fails loudly otherwise."""
    from impec2d_amd import capi
    lib = capi.load()
    if lib.impec_device_count() < 1:
        pytest.fail("no HIP device visible to libimpec_hip.so")
    return lib
