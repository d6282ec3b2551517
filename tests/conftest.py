import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "dispersion-shear_d3q19-opencl_amd")
for p in (PKG, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    config.addinivalue_line("markers", "slow: CPU test that takes more than ~10 s")


@pytest.fixture(scope="session")
def gpu_available():
    import d3q19
    import ctypes
    n = ctypes.c_int(0)
    try:
        rc = d3q19.lib().d3q19_device_count(ctypes.byref(n))
    except Exception:
        return False
    return rc == 0 and n.value > 0
