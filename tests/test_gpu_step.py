"""Testbench stepping through tgs_step (write register nibbles -> run cycles ->
read register nibbles, every instance in one call) against the oracle run
step by step with the same writes, on every kernel family (one or several
single-instance contexts, a batched context); plus its argument checks."""
import numpy as np
import pytest

from simharness import OracleRun, assert_pools_equal, gpu_pools, gpu_sim
from toucangpurt_amd import KERNEL_BATCH, KERNEL_PLAIN, KERNEL_STREAM, KERNEL_WAVE64, gen
from toucangpurt_amd.rt import SimError

pytestmark = pytest.mark.gpu

_cache = {}


def design():
    if "d" not in _cache:
        _cache["d"] = gen.generate(None, n_luts=6000, depth=8, n_regs=300, n_inputs=70, n_vecops=600, n_srams=3,
                                   sram_depth=1024, sram_width=32, part_luts=600, word_part_ops=48, seed=5)
    return _cache["d"]


def _ids(d, rng):
    regs = OracleRun(d.design).st.pool(0).size
    free = np.setdiff1d(np.arange(regs, dtype=np.uint32), np.asarray(d.inputs, np.uint32))
    w = rng.choice(free, size=9, replace=False).astype(np.uint32)
    r = np.concatenate([w[:4], rng.choice(free, size=12, replace=False)]).astype(np.uint32)
    return w, r


@pytest.mark.parametrize("kernel,n", [(KERNEL_STREAM, 1), (KERNEL_STREAM, 3), (KERNEL_WAVE64, 1), (KERNEL_PLAIN, 2),
                                      (KERNEL_BATCH, 70)])
def test_step_matches_oracle(gpu_available, kernel, n):
    d = design()
    rng = np.random.default_rng(11 + n)
    wids, rids = _ids(d, rng)
    print(f"case step kernel={kernel} n={n}", flush=True)
    sim = gpu_sim(d.design, d.symbols, True, None, kernel=kernel)
    sim.set_instances(n)
    sim.set_stimulus(d.inputs, 99)
    check = [0, n - 1] if n > 2 else list(range(n))
    orcs = {k: OracleRun(d.design, d.symbols, True, d.inputs, seed=99, instance=k) for k in check}
    for step, cycles in enumerate((0, 1, 3, 2, 5)):
        vals = rng.integers(0, 16, size=(n, wids.size)).astype(np.uint8)
        got, _ = sim.step(wids, vals, cycles, rids)
        assert got.shape == (n, rids.size)
        for k, orc in orcs.items():
            reg = orc.st.pool(0)
            reg[wids] = vals[k]
            orc.st.set_pool(0, reg)
            orc.run(cycles)
            want = orc.st.pool(0)[rids]
            assert np.array_equal(got[k], want), f"step {step} instance {k}: {got[k].tolist()} != {want.tolist()}"
    for k, orc in orcs.items():
        assert_pools_equal(gpu_pools(sim, k), orc.pools(), f"instance {k} after the steps")
    sim.cleanup()


def test_one_launch_steps_across_kernels_and_free_runs(gpu_available):
    """A one-instance step on the streamed or wave64 kernel is one launch that
    applies the writes before its first cycle and takes the reads after its
    last; the grid-barrier counter goes on from launch to launch and is
    cleared when the grid changes (a kernel switch). Steps, free runs and
    switches interleaved stay equal to the oracle, and the profile's barrier
    and loop tick totals grow with every launch (barrier <= total)."""
    d = design()
    rng = np.random.default_rng(23)
    wids, rids = _ids(d, rng)
    print("case one-launch steps across kernels", flush=True)
    sim = gpu_sim(d.design, d.symbols, True, None, kernel=KERNEL_STREAM)
    sim.set_stimulus(d.inputs, 7)
    orc = OracleRun(d.design, d.symbols, True, d.inputs, seed=7)
    last = (0, 0)
    plan = [(KERNEL_STREAM, "step", 1), (KERNEL_STREAM, "run", 5), (KERNEL_STREAM, "step", 2),
            (KERNEL_WAVE64, "step", 1), (KERNEL_WAVE64, "run", 3), (KERNEL_WAVE64, "step", 2),
            (KERNEL_STREAM, "step", 4), (KERNEL_STREAM, "run", 1), (KERNEL_STREAM, "step", 1)]
    for i, (kernel, what, cycles) in enumerate(plan):
        sim.set_kernel(kernel)
        if what == "step":
            vals = rng.integers(0, 16, size=(1, wids.size)).astype(np.uint8)
            got, _ = sim.step(wids, vals, cycles, rids)
            reg = orc.st.pool(0)
            reg[wids] = vals[0]
            orc.st.set_pool(0, reg)
            orc.run(cycles)
            want = orc.st.pool(0)[rids]
            assert np.array_equal(got[0], want), f"{i}: {got[0].tolist()} != {want.tolist()}"
        else:
            sim.eval_free_running(cycles)
            orc.run(cycles)
        assert_pools_equal(gpu_pools(sim), orc.pools(), f"after plan entry {i}")
        prof = sim.profile()
        assert prof["total_ticks"] > last[1] and prof["barrier_ticks"] >= last[0], f"{i}: {prof}"
        assert prof["barrier_ticks"] <= prof["total_ticks"]
        last = (prof["barrier_ticks"], prof["total_ticks"])
    sim.cleanup()


def test_step_equals_register_calls(gpu_available):
    """Same results as tgs_write_regs / tgs_run_batch / tgs_read_regs per instance."""
    d = design()
    rng = np.random.default_rng(5)
    wids, rids = _ids(d, rng)
    n = 66
    print("case step vs register calls, batched n=66", flush=True)
    a = gpu_sim(d.design, d.symbols, False, None, kernel=KERNEL_BATCH)
    b = gpu_sim(d.design, d.symbols, False, None, kernel=KERNEL_BATCH)
    for s in (a, b):
        s.set_instances(n)
        s.set_stimulus(d.inputs, 5)
    for cycles in (2, 0, 4):
        vals = rng.integers(0, 16, size=(n, wids.size)).astype(np.uint8)
        got, _ = a.step(wids, vals, cycles, rids)
        for k in range(n):
            b._check(b._lib.tgs_write_regs(b._h, k, wids.ctypes.data_as(_u32p()), wids.size,
                                           vals[k].ctypes.data_as(_u8p())))
        b.run_batch(np.full(n, cycles, np.uint32))
        for k in (0, 1, 40, n - 1):
            out = np.zeros(rids.size, np.uint8)
            b._check(b._lib.tgs_read_regs(b._h, k, rids.ctypes.data_as(_u32p()), rids.size,
                                          out.ctypes.data_as(_u8p())))
            assert np.array_equal(got[k], out), f"instance {k}"
    a.cleanup()
    b.cleanup()


def test_step_rejects_bad_ids(gpu_available):
    d = design()
    sim = gpu_sim(d.design, d.symbols, False, None, kernel=KERNEL_STREAM)
    regs = sim.pool_size(0)
    with pytest.raises(SimError):
        sim.step([regs], [[1]], 1, [0])
    with pytest.raises(SimError):
        sim.step([0], [[1]], 1, [regs + 5])
    with pytest.raises(SimError):
        sim.step([3, 3], [[1, 2]], 1, [3])
    # no writes and no reads is a plain run
    got, _ = sim.step([], np.zeros((1, 0), np.uint8), 2, [])
    assert got.shape == (1, 0) and sim.cycles(0) == 2
    sim.cleanup()


def _u32p():
    import ctypes
    return ctypes.POINTER(ctypes.c_uint32)


def _u8p():
    import ctypes
    return ctypes.POINTER(ctypes.c_uint8)


@pytest.mark.parametrize("stops", [2, 0])
@pytest.mark.parametrize("kernel,n", [(KERNEL_STREAM, 1), (KERNEL_BATCH, 5)])
def test_step_reports_stop(gpu_available, kernel, n, stops):
    """The stop flag a step returns is the oracle's: set by a design whose stop
    op fires, never set by one without stop ops."""
    d = gen.generate(None, n_luts=2000, depth=6, n_regs=100, n_inputs=20, n_vecops=50, n_prints=2, n_stops=stops,
                     stop_rarity=3, part_luts=400, seed=13)
    print(f"case step stop kernel={kernel} n={n} stops={stops}", flush=True)
    sim = gpu_sim(d.design, d.symbols, False, None, kernel=kernel)
    sim.set_instances(n)
    sim.set_stimulus(d.inputs, 3)
    orcs = [OracleRun(d.design, d.symbols, False, d.inputs, seed=3, instance=k) for k in range(n)]
    seen = False
    for cycles in (0, 1, 2, 4, 8):
        _, stopped = sim.step([], np.zeros((n, 0), np.uint8), cycles, [])
        for orc in orcs:
            orc.run(cycles)
        want = any(bool(orc.st.stopped) for orc in orcs)
        assert stopped == want, f"after {cycles} more cycles: step says {stopped}, oracle {want}"
        seen |= want
    assert seen == (stops > 0)
    sim.cleanup()


@pytest.mark.parametrize("kernel", [KERNEL_STREAM, KERNEL_WAVE64])
def test_one_launch_steps_keep_the_print_log(gpu_available, kernel):
    """One-instance steps with printing on (each step one launch that also
    writes and reads registers): the print log after steps of 1..7 cycles
    with register writes in between, and its cycle numbers, equal the
    oracle's run the same way."""
    d = gen.generate(None, n_luts=2000, depth=6, n_regs=100, n_inputs=20, n_vecops=50, n_prints=3, n_stops=0,
                     part_luts=400, seed=29)
    rng = np.random.default_rng(4)
    wids, rids = _ids(d, rng)
    print(f"case one-launch steps with prints kernel={kernel}", flush=True)
    sim = gpu_sim(d.design, d.symbols, False, d.inputs, seed=8, kernel=kernel, enable_print=True)
    orc = OracleRun(d.design, d.symbols, False, d.inputs, seed=8)
    for cycles in (1, 3, 2, 7, 1, 5):
        vals = rng.integers(0, 16, size=(1, wids.size)).astype(np.uint8)
        got, _ = sim.step(wids, vals, cycles, rids)
        reg = orc.st.pool(0)
        reg[wids] = vals[0]
        orc.st.set_pool(0, reg)
        orc.run(cycles, enable_print=True)
        assert np.array_equal(got[0], orc.st.pool(0)[rids])
    want = sorted(orc.st.print_log())
    assert len(want) > 0
    assert sim.print_log() == want
    assert_pools_equal(gpu_pools(sim), orc.pools(), "after the steps")
    sim.cleanup()
