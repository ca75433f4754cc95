import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libtoucan_rt.so on the device)")


@pytest.fixture(scope="session")
def gpu_available():
    from toucangpurt_amd import _native
    import ctypes
    lib = _native.rt_lib()
    h = ctypes.c_void_p()
    rc = lib.tgs_create(0, ctypes.byref(h))
    if rc != 0:
        pytest.fail(f"GPU test on a machine without a usable GPU (tgs_create -> {rc})")
    lib.tgs_destroy(h)
    return True
