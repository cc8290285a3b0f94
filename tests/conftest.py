import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "pdf-raster_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long CPU test")


def _have_gpu():
    try:
        import rr_hip
        import ctypes
        h = ctypes.c_void_p()
        rc = rr_hip.lib().rr_hip_create(0, ctypes.byref(h))
        if rc == 0:
            rr_hip.lib().rr_hip_destroy(h)
            return True
    except Exception:
        pass
    return False


@pytest.fixture(scope="session")
def backend():
    import rr_hip
    be = rr_hip.Backend(0)  # raises loudly when no device / no library
    yield be
    be.close()
