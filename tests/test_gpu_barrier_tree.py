"""The persistent kernels' two-level grid barrier on grids around its group count.

Workgroup g arrives at the counter of group g % 16 (kBarGroups): grids of 16,
17, 31 and 33 workgroups give groups of one and of two members, and grids
below 16 fewer groups than 16. Two launches in a row on one grid carry the
barrier index on (the tree is cleared only when the grid size changes), a
third on another grid clears it. Every run must end where the oracle ends.
"""
import numpy as np
import pytest

from simharness import OracleRun
from toucangpurt_amd import KERNEL_STREAM, KERNEL_WAVE64, POOL_MEM, POOL_REG, gen, rt

pytestmark = pytest.mark.gpu

KW = dict(n_luts=24000, depth=8, n_regs=600, n_inputs=64, n_vecops=400, n_srams=2, sram_depth=512, sram_width=32,
          n_prints=0, n_stops=0, part_luts=300, word_part_ops=40, seed=73)
_d = {}


def design():
    if "d" not in _d:
        _d["d"] = gen.generate(None, **KW)
    return _d["d"]


@pytest.mark.parametrize("kernel,block", [(KERNEL_STREAM, 704), (KERNEL_WAVE64, 512)])
@pytest.mark.parametrize("groups", [16, 17, 31, 33])
def test_barrier_tree_grids_match_oracle(gpu_available, kernel, block, groups):
    d = design()
    print(f"case barrier tree kernel={kernel} block={block} groups={groups} parts={d.stats['n_parts']}", flush=True)
    sim = rt.Simulator(0)
    assert sim.init(0, d.design, d.symbols, False, False) == 0
    sim.set_kernel(kernel)
    sim.set_launch(kernel, block, groups)
    sim.set_stimulus(d.inputs, 5)
    sim.eval_free_running(6)
    sim.eval_free_running(5)
    assert sim.profile()["grid"] == groups
    sim.set_launch(kernel, block, groups + 1)  # another grid: the tree is cleared first
    sim.eval_free_running(4)
    assert sim.profile()["grid"] == groups + 1
    orc = OracleRun(d.design, inputs=d.inputs, seed=5)
    orc.run(15)
    reg, mem, _ = orc.pools()
    np.testing.assert_array_equal(sim.read_pool(POOL_REG), reg)
    np.testing.assert_array_equal(sim.read_pool(POOL_MEM), mem)
    sim.cleanup()


def test_wave64_several_resident_programs_per_workgroup_repeated(gpu_available):
    """Regression: with several partitions per workgroup the wave64 kernel keeps
    their programs side by side in LDS, each copied in 16-byte pieces; programs
    must start at whole pieces, or a copy's padded last piece overwrites the
    next program's first words in a race (intermittently wrong registers: about
    one run in ten on 16 workgroups before the fix)."""
    d = design()
    orc = OracleRun(d.design, inputs=d.inputs, seed=5)
    orc.run(15)
    reg, mem, _ = orc.pools()
    for rep in range(12):
        print(f"case wave64 16 workgroups repetition {rep}", flush=True)
        sim = rt.Simulator(0)
        assert sim.init(0, d.design, d.symbols, False, False) == 0
        sim.set_kernel(KERNEL_WAVE64)
        sim.set_launch(KERNEL_WAVE64, 512, 16)
        sim.set_stimulus(d.inputs, 5)
        sim.eval_free_running(15)
        np.testing.assert_array_equal(sim.read_pool(POOL_REG), reg)
        np.testing.assert_array_equal(sim.read_pool(POOL_MEM), mem)
        sim.cleanup()
