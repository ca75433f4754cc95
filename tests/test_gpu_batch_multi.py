"""Several batched contexts (different designs) in shared launches
(tgs_run_batch_multi): every instance ends where tgs_run_batch on its own
context would leave it, and where the oracle run alone does."""
import numpy as np
import pytest

from simharness import OracleRun, assert_pools_equal, gpu_pools, gpu_sim
from toucangpurt_amd import KERNEL_BATCH, KERNEL_STREAM, gen, rt

pytestmark = pytest.mark.gpu

_cache = {}


def designs():
    if not _cache:
        _cache["a"] = gen.generate(None, n_luts=6000, depth=8, n_regs=300, n_inputs=70, n_vecops=600, n_srams=3,
                                   sram_depth=1024, sram_width=32, part_luts=600, word_part_ops=48, seed=5)
        _cache["b"] = gen.generate("lut20k", seed=3, part_luts=256)
        _cache["c"] = gen.generate(None, n_luts=3000, depth=6, n_regs=100, n_inputs=20, n_prints=4, n_stops=2,
                                   part_luts=300, cluster_luts=100, seed=21)
    return _cache


def batched(d, n, seed):
    sim = gpu_sim(d.design, d.symbols, False, None, kernel=KERNEL_BATCH)
    sim.set_instances(n)
    sim.set_stimulus(d.inputs, seed)
    return sim


def test_multi_equals_alone_and_oracle(gpu_available):
    ds = designs()
    rng = np.random.default_rng(12)
    plan = [("a", 70, 31), ("b", 9, 32), ("c", 130, 33)]
    cycles = {k: rng.integers(0, 40, size=n).astype(np.uint32) for k, n, _ in plan}
    more = {k: rng.integers(0, 15, size=n).astype(np.uint32) for k, n, _ in plan}
    print("case multi a/b/c", flush=True)
    sims = [batched(ds[k], n, seed) for k, n, seed in plan]
    rt.run_batch_multi(sims, [cycles[k] for k, _, _ in plan])
    stopped = rt.run_batch_multi(sims, [more[k] for k, _, _ in plan])
    for (k, n, seed), sim in zip(plan, sims):
        alone = batched(ds[k], n, seed)
        alone.run_batch(cycles[k])
        st = alone.run_batch(more[k])
        assert st == stopped[plan.index((k, n, seed))]
        for i in range(n):
            assert_pools_equal(gpu_pools(sim, i), gpu_pools(alone, i), f"design {k} instance {i}: shared vs alone")
            assert sim.cycles(i) == cycles[k][i] + more[k][i]
        for i in (0, n - 1):
            orc = OracleRun(ds[k].design, ds[k].symbols, False, ds[k].inputs, seed=seed, instance=i)
            orc.run(int(cycles[k][i]) + int(more[k][i]))
            assert_pools_equal(gpu_pools(sim, i), orc.pools(), f"design {k} instance {i}")
        alone.cleanup()
    for sim in sims:
        sim.cleanup()


def test_multi_with_different_instance_groups(gpu_available, monkeypatch):
    """Contexts built for different instances per wave run in separate
    launches of one call; results are unchanged."""
    ds = designs()
    print("case multi mixed groups", flush=True)
    a = batched(ds["a"], 20, 41)
    a.run_batch(np.ones(20, np.uint32))
    monkeypatch.setenv("TGS_BATCH_INST", "16")
    b = batched(ds["b"], 20, 42)
    b.run_batch(np.ones(20, np.uint32))
    monkeypatch.delenv("TGS_BATCH_INST")
    assert a.profile()["batch_group"] != b.profile()["batch_group"]
    cyc = [np.full(20, 12, np.uint32), np.arange(20, dtype=np.uint32)]
    rt.run_batch_multi([a, b], cyc)
    for sim, key, seed, c in ((a, "a", 41, cyc[0]), (b, "b", 42, cyc[1])):
        for i in (0, 7, 19):
            orc = OracleRun(ds[key].design, ds[key].symbols, False, ds[key].inputs, seed=seed, instance=i)
            orc.run(1 + int(c[i]))
            assert_pools_equal(gpu_pools(sim, i), orc.pools(), f"design {key} instance {i}")
    a.cleanup()
    b.cleanup()


def test_multi_rejects_bad_calls(gpu_available):
    ds = designs()
    print("case multi errors", flush=True)
    a = batched(ds["a"], 4, 1)
    s = gpu_sim(ds["b"].design, ds["b"].symbols, False, None, kernel=KERNEL_STREAM)
    with pytest.raises(rt.SimError):
        rt.run_batch_multi([a, a], [np.ones(4, np.uint32)] * 2)
    with pytest.raises(rt.SimError):
        rt.run_batch_multi([a, s], [np.ones(4, np.uint32), np.ones(1, np.uint32)])
    assert rt.run_batch_multi([a], [np.ones(4, np.uint32)]) == [False]
    a.cleanup()
    s.cleanup()
