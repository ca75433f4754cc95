"""A context on the caller's HIP stream (tgs_create_on_stream): results equal
the oracle's, and the stream stays usable after the context is destroyed."""
import ctypes

import numpy as np
import pytest

from simharness import OracleRun, assert_pools_equal, gpu_pools
from toucangpurt_amd import KERNEL_BATCH, KERNEL_STREAM, gen, rt

pytestmark = pytest.mark.gpu


def _hip():
    return ctypes.CDLL("libamdhip64.so")


@pytest.mark.parametrize("kernel", [KERNEL_STREAM, KERNEL_BATCH])
def test_context_on_caller_stream(gpu_available, kernel):
    d = gen.generate(None, n_luts=3000, depth=10, n_regs=200, n_inputs=40, part_luts=500, seed=2)
    hip = _hip()
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    print(f"case caller stream kernel={kernel}", flush=True)
    try:
        sim = rt.Simulator(0, stream=stream)
        assert sim.init(0, d.design, d.symbols, True, False) == 0
        sim.set_kernel(kernel)
        if kernel == KERNEL_BATCH:
            sim.set_instances(3)
        sim.set_stimulus(d.inputs, 21)
        sim.eval_free_running(17)
        for k in ([0, 2] if kernel == KERNEL_BATCH else [0]):
            orc = OracleRun(d.design, d.symbols, True, d.inputs, seed=21, instance=k)
            orc.run(17)
            assert_pools_equal(gpu_pools(sim, k), orc.pools(), f"instance {k}")
        sim.cleanup()
        # the caller's stream outlives the context
        assert hip.hipStreamSynchronize(stream) == 0
    finally:
        assert hip.hipStreamDestroy(stream) == 0


def test_create_on_stream_rejects_null(gpu_available):
    lib = rt._native.rt_lib()
    h = ctypes.c_void_p()
    assert lib.tgs_create_on_stream(0, None, ctypes.byref(h)) != 0
