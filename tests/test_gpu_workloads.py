"""Bit-exact parity of the GPU kernels with the oracle on the benchmark
workloads themselves (the sizes bench.py measures), not only on small
designs:

* lut1m (1,000,000 LUT4 + 20,000 word ops, depth 64, 16 SRAMs 4096 x 64 bit,
  128 inputs): 600 cycles, all three pools compared every 100 cycles, on the
  streamed, wave64 and plain kernels;
* lut10m (10,000,000 LUT4, depth 128, 64 SRAMs, 256 inputs): 60 cycles on the
  streamed and plain kernels, pools compared at 20 and 60 cycles;
* lut200k x 1,024 batched instances (partitions of 256 LUT4, the shape the
  autotuner picks) with uneven cycle counts: every 16th instance and the
  first and last of every 64-instance group against single-instance oracle
  runs, all three pools, stop flag and cycle count.

The oracle runs the partitions of each region on the host's CPU share
(orc_run_threads, equal to the single-threaded run on every golden fixture).
Cycle counts are scaled down from the benchmark's so the file runs in about a
minute; the per-cycle semantics do not depend on the count.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle as O
from simharness import OracleRun, assert_pools_equal, gpu_pools, gpu_sim
from toucangpurt_amd import KERNEL_BATCH, KERNEL_PLAIN, KERNEL_STREAM, KERNEL_WAVE64, gen, rt
from toucangpurt_amd._native import POOL_EXCHANGE, POOL_MEM, POOL_REG

pytestmark = pytest.mark.gpu

SEED = 11
_designs = {}
_oracle_pools = {}


def threads():
    env = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, int(env)) if env.isdigit() else max(1, min(16, os.cpu_count() or 1))


def design(name, **kw):
    key = (name, tuple(sorted(kw.items())))
    if key not in _designs:
        _designs[key] = gen.generate(name, seed=1, **kw)
    return _designs[key]


def oracle_trace(name, points):
    """Oracle pools of `name` after each cycle count in `points` (computed once per session)."""
    if name not in _oracle_pools:
        d = design(name)
        orc = OracleRun(d.design, inputs=d.inputs, seed=SEED)
        out, done = {}, 0
        for p in points:
            orc.st.run(p - done, threads=threads())
            done = p
            out[p] = orc.pools()
        _oracle_pools[name] = out
    return _oracle_pools[name]


def run_trace(name, kernel, points):
    d = design(name)
    want = oracle_trace(name, points)
    sim = gpu_sim(d.design, inputs=d.inputs, seed=SEED, kernel=kernel)
    done = 0
    for p in points:
        sim.eval_free_running(p - done)
        done = p
        assert_pools_equal(gpu_pools(sim), want[p], f"{name} after {p} cycles")
    assert sim.cycles() == points[-1]
    prof = sim.profile()
    sim.cleanup()
    return prof


LUT1M_POINTS = [100, 200, 300, 400, 500, 600]


@pytest.mark.parametrize("kname,kernel", [("stream", KERNEL_STREAM), ("wave64", KERNEL_WAVE64),
                                          ("plain", KERNEL_PLAIN)])
def test_lut1m_600_cycles(gpu_available, kname, kernel):
    print(f"case lut1m kernel={kname}", flush=True)
    prof = run_trace("lut1m", kernel, LUT1M_POINTS)
    assert prof["kernel"] == kernel


@pytest.mark.parametrize("kname,kernel", [("stream", KERNEL_STREAM), ("plain", KERNEL_PLAIN)])
def test_lut10m_60_cycles(gpu_available, kname, kernel):
    print(f"case lut10m kernel={kname}", flush=True)
    prof = run_trace("lut10m", kernel, [20, 60])
    assert prof["kernel"] == kernel


def checked_instances(n):
    ids = set(range(0, n, 16))
    for g in range(0, n, 64):
        ids.add(g)
        ids.add(min(g + 63, n - 1))
    return sorted(ids)


def test_lut200k_1024_batched_instances(gpu_available):
    d = design("lut200k", part_luts=256)
    n = 1024
    cycles = np.random.default_rng(4).integers(40, 241, size=n).astype(np.uint32)
    print(f"case lut200k batched {n} instances, cycles {cycles.min()}-{cycles.max()}", flush=True)
    sim = gpu_sim(d.design, inputs=None, kernel=KERNEL_BATCH)
    sim.set_instances(n)
    sim.set_stimulus(d.inputs, SEED)
    sim.run_batch(cycles)
    prof = sim.profile()
    assert prof["kernel"] == KERNEL_BATCH and prof["cycles"] == int(cycles.sum())
    ids = checked_instances(n)
    got = {i: gpu_pools(sim, i) for i in ids}
    state = {i: (sim.stopped(i), sim.cycles(i)) for i in ids}
    sim.cleanup()

    def oracle(i):
        orc = OracleRun(d.design, inputs=d.inputs, seed=SEED, instance=i)
        orc.st.run(int(cycles[i]))
        return orc.pools(), orc.st.stopped

    with ThreadPoolExecutor(threads()) as ex:
        want = dict(zip(ids, ex.map(oracle, ids)))
    for i in ids:
        assert_pools_equal(got[i], want[i][0], f"instance {i} after {cycles[i]} cycles")
        assert state[i] == (want[i][1], int(cycles[i])), i


def test_serving_mix_with_a_1m_lut4_design(gpu_available):
    """Serving over designs of 50K, 400K and 1.2M LUT4 (partitions of 256):
    staggered jobs of 1-64 instances share batched launches; every job equals
    the same job run alone, and one job per design equals the oracle."""
    from toucangpurt_amd.serve import Engine, Job, run_alone
    sizes = {"d50k": 50_000, "d400k": 400_000, "d1200k": 1_200_000}
    designs = {k: gen.generate(None, n_luts=v, depth=32, n_regs=max(64, v // 20), n_inputs=64, part_luts=256,
                               seed=40 + i) for i, (k, v) in enumerate(sizes.items())}
    rng = np.random.default_rng(8)
    jobs, t = [], 0.0
    for k in range(9):
        name = list(sizes)[k % 3]
        jobs.append(Job(k, name, int(rng.integers(1, 65)), int(rng.integers(100, 401)), int(rng.integers(1 << 30)),
                        arrival=t))
        t += 0.003
    print(f"case serving {len(jobs)} jobs over 50K/400K/1.2M LUT4", flush=True)
    eng = Engine(designs, capacity=128, slice_cycles=64)
    rep = eng.run(jobs)
    eng.close()
    assert rep["jobs"] == len(jobs)
    checked = set()
    for j in jobs:
        reg, mem = run_alone(designs[j.design], j)
        for i in range(j.instances):
            assert np.array_equal(j.reg[i], reg[i]), (j.job_id, i)
            assert np.array_equal(j.mem[i], mem[i]), (j.job_id, i)
        if j.design not in checked:
            checked.add(j.design)
            d = designs[j.design]
            orc = OracleRun(d.design, inputs=d.inputs, seed=j.seed, instance=j.instances - 1)
            orc.st.run(j.cycles, threads=threads())
            assert np.array_equal(j.reg[-1], orc.pools()[0]), j.job_id
            assert np.array_equal(j.mem[-1], orc.pools()[1]), j.job_id
    assert checked == set(sizes)
