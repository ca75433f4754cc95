"""The streamed kernel with several partitions per workgroup and region.

Few workgroups force rounds: partitions of a region run one after another
in a workgroup, or side by side in its slots (split steps, each partition
on a share of the compute waves, a round with one partition on all of
them). Every shape must end where the oracle ends. The same image layouts
are interpreted on the host by tests/test_stream_image.py.
"""
import os

import numpy as np
import pytest

from simharness import OracleRun
from toucangpurt_amd import KERNEL_STREAM, POOL_MEM, POOL_REG, gen, rt

pytestmark = pytest.mark.gpu

KW = dict(n_luts=12000, depth=10, n_regs=400, n_inputs=90, n_vecops=500, n_srams=3, sram_depth=1024, sram_width=32,
          n_prints=3, n_stops=1, part_luts=600, word_part_ops=40, seed=61)
_d = {}


def design():
    if "d" not in _d:
        _d["d"] = gen.generate(None, **KW)
    return _d["d"]


@pytest.mark.parametrize("block,groups,slots", [(704, 3, 1), (704, 3, 2), (704, 7, 2), (704, 5, 4), (448, 4, 2),
                                                (448, 2, 4)])
def test_rounds_match_oracle(gpu_available, block, groups, slots):
    d = design()
    print(f"case stream rounds block={block} groups={groups} slots={slots} parts={d.stats['n_parts']}", flush=True)
    os.environ["TGS_STREAM_SLOTS"] = str(slots)
    try:
        sim = rt.Simulator(0)
        assert sim.init(0, d.design, d.symbols, False, True) == 0
        sim.set_kernel(KERNEL_STREAM)
        sim.set_launch(KERNEL_STREAM, block, groups)
        sim.set_stimulus(d.inputs, 13)
        sim.eval_free_running(7)
        prof = sim.profile()
        assert prof["grid"] == groups and prof["block"] == block
        orc = OracleRun(d.design, inputs=d.inputs, seed=13)
        orc.run(7, enable_print=True)
        reg, mem, _ = orc.pools()
        np.testing.assert_array_equal(sim.read_pool(POOL_REG), reg)
        np.testing.assert_array_equal(sim.read_pool(POOL_MEM), mem)
        assert sim.stopped() == orc.st.stopped
        assert sim.print_log() == sorted(orc.st.print_log())
        sim.cleanup()
    finally:
        del os.environ["TGS_STREAM_SLOTS"]


@pytest.mark.parametrize("block,groups", [(704, 3), (704, 7), (448, 4)])
def test_rounds_auto_slots_match_oracle(gpu_available, block, groups):
    """Slots chosen by the runtime (per region, by the estimated critical path)."""
    d = design()
    print(f"case stream rounds auto slots block={block} groups={groups}", flush=True)
    assert "TGS_STREAM_SLOTS" not in os.environ
    sim = rt.Simulator(0)
    assert sim.init(0, d.design, d.symbols, False, True) == 0
    sim.set_kernel(KERNEL_STREAM)
    sim.set_launch(KERNEL_STREAM, block, groups)
    sim.set_stimulus(d.inputs, 29)
    sim.eval_free_running(9)
    prof = sim.profile()
    assert prof["grid"] == groups and prof["block"] == block
    orc = OracleRun(d.design, inputs=d.inputs, seed=29)
    orc.run(9, enable_print=True)
    reg, mem, _ = orc.pools()
    np.testing.assert_array_equal(sim.read_pool(POOL_REG), reg)
    np.testing.assert_array_equal(sim.read_pool(POOL_MEM), mem)
    assert sim.print_log() == sorted(orc.st.print_log())
    sim.cleanup()
