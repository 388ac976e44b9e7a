import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def have_gpu():
    try:
        import ctypes
        from fd_adi_amd import _lib
        n = ctypes.c_int()
        return _lib.lib.fdadi_device_count(ctypes.byref(n)) == 0 and n.value > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not have_gpu():
        pytest.fail("GPU test run without a HIP device (libfdadi.so reports none)")
    return True
