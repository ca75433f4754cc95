"""Batched kernel (instance per lane) on the GPU: every instance equals the
oracle run alone with its own seed and cycle count, results do not depend on
which other instances share the launch, and state survives switching
between the per-instance and instance-minor layouts."""
import numpy as np
import pytest

from simharness import OracleRun, assert_pools_equal, gpu_pools, gpu_sim
from toucangpurt_amd import KERNEL_BATCH, KERNEL_PLAIN, KERNEL_WAVE64, gen

pytestmark = pytest.mark.gpu

_cache = {}


def design(name):
    if name not in _cache:
        if name == "mixed":
            _cache[name] = gen.generate(None, n_luts=6000, depth=8, n_regs=300, n_inputs=70, n_vecops=600, n_srams=3,
                                        sram_depth=1024, sram_width=32, part_luts=600, word_part_ops=48, seed=5)
        else:
            _cache[name] = gen.generate("lut20k", seed=3, part_luts=512)
    return _cache[name]


@pytest.mark.parametrize("dname", ["mixed", "lut20k"])
def test_batched_instances_match_oracle(gpu_available, dname):
    d = design(dname)
    rng = np.random.default_rng(4)
    n = 70  # two 64-lane groups, the second partly padding
    cycles = rng.integers(0, 40, size=n).astype(np.uint32)
    print(f"case batched {dname} n={n}", flush=True)
    sim = gpu_sim(d.design, d.symbols, True, None, kernel=KERNEL_BATCH)
    sim.set_instances(n)
    sim.set_stimulus(d.inputs, 1234)
    sim.run_batch(cycles)
    for k in list(range(0, n, 7)) + [n - 1]:
        orc = OracleRun(d.design, d.symbols, True, d.inputs, seed=1234, instance=k)
        orc.run(int(cycles[k]))
        assert_pools_equal(gpu_pools(sim, k), orc.pools(), f"instance {k} after {cycles[k]} cycles")
        assert sim.cycles(k) == cycles[k]
    # a second uneven run continues every instance from where it stopped
    more = rng.integers(0, 20, size=n).astype(np.uint32)
    sim.run_batch(more)
    for k in (0, 33, 69):
        orc = OracleRun(d.design, d.symbols, True, d.inputs, seed=1234, instance=k)
        orc.run(int(cycles[k]) + int(more[k]))
        assert_pools_equal(gpu_pools(sim, k), orc.pools(), f"instance {k} after two runs")
    sim.cleanup()


def test_alone_equals_in_mix(gpu_available):
    d = design("mixed")
    print("case batched alone vs mix", flush=True)
    alone = gpu_sim(d.design, d.symbols, False, None, kernel=KERNEL_BATCH)
    alone.set_instances(1)
    alone.set_stimulus(d.inputs, 77)
    alone.run_batch([25])
    mix = gpu_sim(d.design, d.symbols, False, None, kernel=KERNEL_BATCH)
    mix.set_instances(130)
    mix.set_stimulus(d.inputs, 77)
    mix.run_batch([25] + [int(x) for x in np.random.default_rng(1).integers(1, 60, size=129)])
    assert_pools_equal(gpu_pools(mix, 0), gpu_pools(alone, 0), "instance 0 alone vs in a mix of 130")
    alone.cleanup()
    mix.cleanup()


def test_layout_switch_keeps_state(gpu_available):
    d = design("mixed")
    print("case layout switch", flush=True)
    sim = gpu_sim(d.design, d.symbols, True, None, kernel=KERNEL_WAVE64)
    sim.set_instances(3)
    sim.set_stimulus(d.inputs, 5)
    sim.run_batch([10, 4, 7])
    sim.set_kernel(KERNEL_BATCH)
    sim.run_batch([3, 9, 1])
    sim.set_kernel(KERNEL_PLAIN)
    sim.run_batch([2, 2, 2])
    for k, total in enumerate([15, 15, 10]):
        orc = OracleRun(d.design, d.symbols, True, d.inputs, seed=5, instance=k)
        orc.run(total)
        assert_pools_equal(gpu_pools(sim, k), orc.pools(), f"instance {k}")
    sim.cleanup()


def test_batched_trace_and_prints(gpu_available):
    d = gen.generate(None, n_luts=3000, depth=6, n_regs=100, n_inputs=20, n_prints=4, n_stops=2, part_luts=600,
                     cluster_luts=100, seed=21)
    print("case batched trace/prints", flush=True)
    sim = gpu_sim(d.design, d.symbols, False, None, kernel=KERNEL_BATCH, enable_print=True)
    sim.set_instances(5)
    sim.set_stimulus(d.inputs, 3)
    orcs = [OracleRun(d.design, d.symbols, False, d.inputs, seed=3, instance=k) for k in range(5)]
    for c in range(6):
        sim.eval()
        for k, o in enumerate(orcs):
            o.run(1, enable_print=True)
            assert_pools_equal(gpu_pools(sim, k), o.pools(), f"instance {k} cycle {c}")
    for k, o in enumerate(orcs):
        assert sim.print_log(k) == sorted(o.st.print_log())
        assert sim.stopped(k) == o.st.stopped
        assert sim.stop_cycle(k) == o.st.stop_cycle
    sim.cleanup()


def test_fast_lut_tasks_equal_bounds_checked_ones(gpu_available, monkeypatch):
    """On a nibble program the batched image evaluates LUT pairs without index
    bounds checks (T_LUTFAST, batch_fast_tasks > 0); TGS_BATCH_FAST=0 builds
    the bounds-checked image. Both agree on every instance and with the oracle."""
    d = design("mixed")
    n = 40
    cycles = np.random.default_rng(8).integers(1, 30, size=n).astype(np.uint32)
    pools = {}
    for fast in (True, False):
        if fast:
            monkeypatch.delenv("TGS_BATCH_FAST", raising=False)
        else:
            monkeypatch.setenv("TGS_BATCH_FAST", "0")
        print(f"case batched fast={fast}", flush=True)
        sim = gpu_sim(d.design, d.symbols, True, None, kernel=KERNEL_BATCH)
        sim.set_instances(n)
        sim.set_stimulus(d.inputs, 61)
        sim.run_batch(cycles)
        assert (sim.profile()["batch_fast_tasks"] > 0) == fast
        pools[fast] = [gpu_pools(sim, k) for k in range(n)]
        sim.cleanup()
    for k in range(n):
        assert_pools_equal(pools[True][k], pools[False][k], f"instance {k}: fast vs bounds-checked")
    for k in (0, 17, n - 1):
        orc = OracleRun(d.design, d.symbols, True, d.inputs, seed=61, instance=k)
        orc.run(int(cycles[k]))
        assert_pools_equal(pools[True][k], orc.pools(), f"instance {k}")


def test_nibble_rows_equal_byte_rows(gpu_available, monkeypatch):
    """On a nibble program whose LUT tasks are all T_LUTFAST the batched rows
    hold a nibble per instance (batch_nibble_rows); TGS_BATCH_NIB=0 keeps a
    byte per instance and TGS_BATCH_FAST=0 the bounds-checked tasks. All three
    agree on every instance, and with the oracle, for instance counts that
    give different shapes (nibble rows of 16 and 32 instances per wave)."""
    d = design("mixed")
    for n in (70, 1024):
        cycles = np.random.default_rng(n).integers(1, 24, size=n).astype(np.uint32)
        pools = {}
        for mode in ("nib", "byte", "rows"):
            monkeypatch.delenv("TGS_BATCH_NIB", raising=False)
            monkeypatch.delenv("TGS_BATCH_FAST", raising=False)
            if mode == "byte":
                monkeypatch.setenv("TGS_BATCH_NIB", "0")
            if mode == "rows":
                monkeypatch.setenv("TGS_BATCH_FAST", "0")
            print(f"case batched rows={mode} n={n}", flush=True)
            sim = gpu_sim(d.design, d.symbols, True, None, kernel=KERNEL_BATCH)
            sim.set_instances(n)
            sim.set_stimulus(d.inputs, 23)
            sim.run_batch(cycles)
            prof = sim.profile()
            assert prof["batch_nibble_rows"] == (1 if mode == "nib" else 0), prof
            assert (prof["batch_fast_tasks"] > 0) == (mode != "rows"), prof
            pools[mode] = [gpu_pools(sim, k) for k in range(n)]
            sim.cleanup()
        for k in range(n):
            assert_pools_equal(pools["nib"][k], pools["byte"][k], f"n={n} instance {k}: nibble vs byte rows")
            assert_pools_equal(pools["nib"][k], pools["rows"][k], f"n={n} instance {k}: nibble vs bounds-checked")
        for k in (0, n // 2, n - 1):
            orc = OracleRun(d.design, d.symbols, True, d.inputs, seed=23, instance=k)
            orc.run(int(cycles[k]))
            assert_pools_equal(pools["nib"][k], orc.pools(), f"n={n} instance {k}")


def test_wide_register_bytes_switch_off_fast_tasks(gpu_available):
    """Register bytes above 15 written through the API end the nibble state:
    the next run rebuilds the image with bounds-checked LUT tasks, and the
    instance still follows the oracle given the same bytes."""
    from toucangpurt_amd._native import POOL_REG
    d = design("mixed")
    print("case batched wide register bytes", flush=True)
    sim = gpu_sim(d.design, d.symbols, True, None, kernel=KERNEL_BATCH)
    sim.set_instances(3)
    sim.set_stimulus(d.inputs, 19)
    sim.run_batch([5, 5, 5])
    assert sim.profile()["batch_fast_tasks"] > 0
    orc = OracleRun(d.design, d.symbols, True, d.inputs, seed=19, instance=1)
    orc.run(5)
    reg = sim.read_pool(POOL_REG, 1).copy()
    wide = np.random.default_rng(2).integers(0, 256, size=reg.size).astype(np.uint8)
    inputs = np.asarray(d.inputs, np.int64)
    wide[inputs] = reg[inputs]
    sim.write_pool(POOL_REG, wide, instance=1)
    orc.st.set_pool(0, wide)
    sim.run_batch([6, 6, 6])
    assert sim.profile()["batch_fast_tasks"] == 0
    orc.run(6)
    assert_pools_equal(gpu_pools(sim, 1), orc.pools(), "instance 1 after wide register bytes")
    sim.cleanup()
