"""Serving on the GPU: jobs of several designs with uneven sizes, staggered
arrivals and a slice much shorter than the jobs, so jobs join running
batches and reuse freed slots. Every job's result must equal the same job
run alone in its own context, and (for one job per design) the oracle."""
import numpy as np
import pytest

from simharness import OracleRun
from toucangpurt_amd import gen
from toucangpurt_amd.serve import Engine, Job, gpu_backend, run_alone

pytestmark = pytest.mark.gpu

DESIGNS = {
    "small": dict(n_luts=1500, depth=6, n_regs=90, n_inputs=20, part_luts=300, seed=21),
    "mixed": dict(n_luts=5000, depth=8, n_regs=250, n_inputs=70, n_vecops=300, n_srams=2, sram_depth=512,
                  sram_width=32, part_luts=600, word_part_ops=32, seed=22),
    "wide": dict(n_luts=12000, depth=10, n_regs=400, n_inputs=130, n_vecops=200, word_bits_mix=100,
                 part_luts=512, word_part_ops=24, seed=23),
}


def test_serving_results_alone_equal_mix(gpu_available):
    designs = {k: gen.generate(None, **v) for k, v in DESIGNS.items()}
    rng = np.random.default_rng(5)
    jobs = []
    t = 0.0
    for k in range(12):
        name = list(DESIGNS)[k % 3]
        jobs.append(Job(k, name, int(rng.integers(1, 40)), int(rng.integers(60, 400)), int(rng.integers(1 << 30)),
                        arrival=t))
        t += 0.002
    print(f"case serving {len(jobs)} jobs", flush=True)
    eng = Engine(designs, capacity=64, slice_cycles=48, small_jobs=8, small_capacity=32)
    rep = eng.run(jobs)
    eng.close()
    assert rep["jobs"] == len(jobs)
    assert rep["launches"] > len(jobs)
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
            orc.run(j.cycles)
            assert np.array_equal(j.reg[-1], orc.pools()[0]), j.job_id
    assert rep["p99_latency_s"] >= rep["p50_latency_s"] > 0


def context_device_bytes(d, capacity):
    """Device bytes of a batched context of `capacity` slots after one run."""
    sim = gpu_backend(d)
    sim.set_instances(capacity)
    sim.set_stimulus(d.inputs, 0)
    sim.run_batch(np.ones(capacity, np.uint32))
    n = sim.device_bytes()
    sim.cleanup()
    return n


def test_parked_jobs_resume_exactly(gpu_available):
    """HBM budget for one design at a time: the second design's job parks the
    first design's running jobs (device checkpoints to host memory), and they
    resume later in fresh slots of a new context; every result equals the job
    run alone."""
    designs = {k: gen.generate(None, **DESIGNS[k]) for k in ("small", "mixed")}
    t = [0.0]
    eng = Engine(designs, capacity=64, slice_cycles=40, clock=lambda: t[0], park_after=0.5)
    # room for the larger context as the runtime reports it, not for both,
    # and estimates calibrated on both
    measured = {k: context_device_bytes(d, 64) for k, d in designs.items()}
    for k, d in designs.items():
        eng.calibrate(d, measured[k], 64)
    eng.budget = max(measured.values()) + 1
    jobs = [Job(0, "small", 5, 300, 31, arrival=0.0), Job(1, "small", 3, 220, 32, arrival=0.0),
            Job(2, "mixed", 4, 150, 33, arrival=0.1)]
    print("case serving parking", flush=True)
    for j in jobs:
        eng.submit(j)
    eng._t0 = 0.0
    for _ in range(400):
        t[0] += 0.1
        if eng.step() == 0:
            break
    eng.close()
    assert len(eng.done) == 3 and eng.parked >= 1
    for j in jobs:
        reg, mem = run_alone(designs[j.design], j)
        for i in range(j.instances):
            assert np.array_equal(j.reg[i], reg[i]), (j.job_id, i)
            assert np.array_equal(j.mem[i], mem[i]), (j.job_id, i)
