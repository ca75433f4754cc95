import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libtoucan_rt.so on the device)")
    config.addinivalue_line("markers", "subprocess_gpu: GPU work runs in a child process (collected first)")


def pytest_collection_modifyitems(session, config, items):
    """Run the tests that drive the GPU from a child process first.

    The persistent kernels are cooperative launches; once this process has
    made one it holds the GPU's cooperative queue for its lifetime, and a
    cooperative launch from another process on the same GPU is then refused
    (hipErrorIllegalState). Tests whose GPU work runs in a subprocess
    (the C++ simulator driver) therefore go before any in-process GPU test.
    """
    first = [it for it in items if it.get_closest_marker("subprocess_gpu")]
    if first:
        rest = [it for it in items if not it.get_closest_marker("subprocess_gpu")]
        items[:] = first + rest


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
