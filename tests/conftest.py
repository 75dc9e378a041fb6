import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); calls librz through its C ABI")


@pytest.fixture(scope="session")
def ctx():
    """One librz context on device 0 for the GPU tests."""
    import rasteriser_amd as rz
    c = rz.Context(0)
    yield c
    c.close()
