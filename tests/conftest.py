import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def _has_gpu():
    try:
        import dualres_amd._lib as L
        import ctypes
        lib = L.load()
        n = ctypes.c_int(0)
        return lib.dr_device_count(ctypes.byref(n)) == 0 and n.value > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu_available():
    if not _has_gpu():
        pytest.fail("GPU test selected but no HIP device / engine library is available")
    return True
