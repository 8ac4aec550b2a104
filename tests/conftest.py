import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library with a visible GPU; fails (never skips) when missing."""
    from astroflow_amd import capi
    n = capi.device_count()
    assert n >= 1, "no GPU visible to libastroflow_amd.so"
    return capi
