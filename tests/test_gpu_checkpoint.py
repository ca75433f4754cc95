"""Checkpoint and restore of one instance's state through the C-ABI.

A run split by save/restore into a fresh context (of any kernel layout) must
end exactly where the uninterrupted run ends, and both where the oracle ends
after the same number of cycles. Images of another design and damaged images
are refused.
"""
import numpy as np
import pytest

from simharness import OracleRun, gpu_sim
from toucangpurt_amd import KERNEL_BATCH, KERNEL_PLAIN, KERNEL_STREAM, KERNEL_WAVE64, POOL_MEM, POOL_REG, gen
from toucangpurt_amd._native import TGS_E_ARG, TGS_E_FORMAT
from toucangpurt_amd.rt import SimError

pytestmark = pytest.mark.gpu

KW = dict(n_luts=4000, depth=8, n_regs=200, n_inputs=50, n_vecops=200, n_srams=2, sram_depth=512, sram_width=32,
          n_prints=3, n_stops=1, part_luts=700, seed=41)
_d = {}


def design():
    if "d" not in _d:
        _d["d"] = gen.generate(None, **KW)
    return _d["d"]


def oracle_pools(d, cycles, seed, instance=0):
    o = OracleRun(d.design, d.symbols, True, d.inputs, seed, instance)
    o.run(cycles)
    return o.st.pool(0), o.st.pool(1), o.st


@pytest.mark.parametrize("src,dst", [(KERNEL_STREAM, KERNEL_STREAM), (KERNEL_WAVE64, KERNEL_PLAIN),
                                     (KERNEL_BATCH, KERNEL_STREAM), (KERNEL_WAVE64, KERNEL_BATCH)])
def test_split_run_equals_uninterrupted(gpu_available, src, dst):
    d = design()
    seed, first, second = 17, 7, 9
    print(f"case checkpoint src={src} dst={dst}", flush=True)
    inst = 3 if src == KERNEL_BATCH else 0
    a = gpu_sim(d.design, d.symbols, True, kernel=src)
    if src == KERNEL_BATCH:
        a.set_instances(6)
    a.set_stimulus(d.inputs, seed)
    if src == KERNEL_BATCH:
        a.run_batch(np.full(6, first, np.uint32))
    else:
        a.eval_free_running(first)
    image = a.checkpoint(inst)
    a.cleanup()

    b = gpu_sim(d.design, d.symbols, False, kernel=dst)
    slot = 0
    if dst == KERNEL_BATCH:
        b.set_instances(4)
        slot = 2
    # the stimulus input list is configuration, not state: set it (any seed),
    # then the restore brings back the generator state
    b.set_stimulus(d.inputs, 99)
    b.restore(image, slot)
    assert b.cycles(slot) == first
    if dst == KERNEL_BATCH:
        cyc = np.zeros(4, np.uint32)
        cyc[slot] = second
        b.run_batch(cyc)
    else:
        b.eval_free_running(second)
    reg, mem, st = oracle_pools(d, first + second, seed, inst)
    np.testing.assert_array_equal(b.read_pool(POOL_REG, slot), reg)
    np.testing.assert_array_equal(b.read_pool(POOL_MEM, slot), mem)
    assert b.cycles(slot) == first + second
    assert b.stopped(slot) == st.stopped
    if st.stopped:
        assert b.stop_cycle(slot) == st.stop_cycle
    b.cleanup()


def test_restore_refuses_other_design_and_damage(gpu_available):
    d = design()
    other = gen.generate(None, **dict(KW, seed=42))
    a = gpu_sim(d.design, d.symbols, True, kernel=KERNEL_WAVE64)
    image = a.checkpoint(0)
    b = gpu_sim(other.design, other.symbols, True, kernel=KERNEL_WAVE64)
    with pytest.raises(SimError) as e:
        b.restore(image)
    assert e.value.code == TGS_E_ARG
    with pytest.raises(SimError) as e:
        a.restore(image[:len(image) - 1])
    assert e.value.code == TGS_E_FORMAT
    with pytest.raises(SimError) as e:
        a.restore(b"XXXXXXXX" + image[8:])
    assert e.value.code == TGS_E_FORMAT
    a.restore(image)  # intact image still restores
    a.cleanup()
    b.cleanup()
