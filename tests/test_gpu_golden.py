"""Every kernel family against the golden fixtures (tests/golden/*.npz).

The fixture design images were written by the reference's serializer and the
randomized initial pools by its Init + Randomize(0); the expected pools, stop
flag and print log are the pinned oracle's. Randomized fixtures go through
init(shouldRandomize=True), so the device-side initial state is checked
against the reference's too.
"""
import glob
import os

import numpy as np
import pytest

from simharness import gpu_sim
from toucangpurt_amd import KERNEL_BATCH, KERNEL_PLAIN, KERNEL_STREAM, KERNEL_WAVE64, POOL_MEM, POOL_REG

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = sorted(os.path.basename(f)[:-4] for f in glob.glob(os.path.join(GOLDEN, "*.npz")))
KERNELS = {"plain": KERNEL_PLAIN, "wave64": KERNEL_WAVE64, "stream": KERNEL_STREAM, "batch": KERNEL_BATCH}


@pytest.mark.parametrize("kname", list(KERNELS))
@pytest.mark.parametrize("name", NAMES)
def test_kernel_reproduces_fixture(gpu_available, name, kname):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    print(f"case golden {name} kernel={kname}", flush=True)
    randomized = bool(z["randomized"])
    sim = gpu_sim(z["design"].tobytes(), z["symbols"].tobytes() if z["symbols"].size else None, randomized,
                  kernel=KERNELS[kname], enable_print=bool(z["enable_print"]))
    try:
        instances = 70 if kname == "batch" else 1
        if kname == "batch":
            sim.set_instances(instances)
        for i in range(instances):
            if randomized:
                np.testing.assert_array_equal(sim.read_pool(POOL_REG, i), z["init_reg"])
                np.testing.assert_array_equal(sim.read_pool(POOL_MEM, i), z["init_mem"])
            else:
                sim.write_pool(POOL_REG, z["init_reg"], i)
                sim.write_pool(POOL_MEM, z["init_mem"], i)
        if z["inputs"].size:
            sim.set_stimulus(z["inputs"], int(z["seed"]))
        cycles = int(z["cycles"])
        if kname == "batch":
            stop = sim.run_batch(np.full(instances, cycles, np.uint32))
        else:
            stop = sim.eval_free_running(cycles)
        np.testing.assert_array_equal(sim.read_pool(POOL_REG, 0), z["want_reg"])
        np.testing.assert_array_equal(sim.read_pool(POOL_MEM, 0), z["want_mem"])
        assert sim.stopped(0) == bool(z["want_stop"])
        if kname != "batch":
            assert stop == bool(z["want_stop"])
        if z["want_stop"]:
            assert sim.stop_cycle(0) == int(z["want_stop_cycle"])
        want_log = sorted(zip(z["want_log_cycles"].tolist(), z["want_log_msgs"].tolist()))
        assert sim.print_log(0) == want_log
    finally:
        sim.cleanup()
