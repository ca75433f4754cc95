"""Serving engine logic on the host: admission by slots and HBM budget,
slicing, slot reuse, eviction, and the alone-vs-mix property, with the
oracle standing in for the device (each instance slot an oracle state)."""
import numpy as np
import pytest

import oracle as O
from simharness import initial_pools
from toucangpurt_amd import gen
from toucangpurt_amd.serve import POOL_MEM, POOL_REG, Engine, Job, poisson_jobs, run_alone


class OracleBackend:
    """Batched-context stand-in: one oracle state per instance slot."""

    def __init__(self, design, kernel=None):
        self.prog, self.reg0, self.mem0 = initial_pools(design.design, None, False)
        self.inputs = None
        self.st = []

    def set_instances(self, n):
        self.st = [O.State(self.prog, self.reg0, self.mem0) for _ in range(n)]

    def set_stimulus(self, inputs, seed):
        self.inputs = np.asarray(inputs, np.uint32)
        for k, s in enumerate(self.st):
            s.set_stimulus(self.inputs, seed, k)

    def reset_instance(self, slot, seed, index):
        s = O.State(self.prog, self.reg0, self.mem0)
        s.set_stimulus(self.inputs, seed, index)
        self.st[slot] = s

    def run_batch(self, cycles):
        assert len(cycles) == len(self.st)
        for s, c in zip(self.st, cycles):
            if c:
                s.run(int(c))

    def read_pool(self, pool, slot):
        return self.st[slot].pool(pool).copy()

    def cleanup(self):
        self.st = []


DESIGNS = {
    "a": dict(n_luts=800, depth=5, n_regs=60, n_inputs=12, part_luts=200, seed=3),
    "b": dict(n_luts=1500, depth=6, n_regs=90, n_inputs=20, n_vecops=60, n_srams=1, sram_depth=128,
              sram_width=16, part_luts=300, seed=4),
}


@pytest.fixture(scope="module")
def designs():
    return {k: gen.generate(None, **v) for k, v in DESIGNS.items()}


def test_jobs_alone_equal_mixed(designs):
    jobs = [Job(0, "a", 3, 40, 11, arrival=0.0), Job(1, "b", 2, 25, 12, arrival=0.0),
            Job(2, "a", 4, 17, 13, arrival=0.0), Job(3, "a", 2, 33, 14, arrival=0.0),
            Job(4, "b", 5, 9, 15, arrival=0.0)]
    eng = Engine(designs, capacity=8, slice_cycles=7, backend=OracleBackend, small_jobs=0)
    rep = eng.run(jobs)
    assert rep["jobs"] == len(jobs)
    # slot reuse: 11 instances of design a through 8 slots
    assert eng.launches > 3
    for j in jobs:
        reg, mem = run_alone(designs[j.design], j, backend=OracleBackend)
        for i in range(j.instances):
            np.testing.assert_array_equal(j.reg[i], reg[i], err_msg=f"job {j.job_id} instance {i}")
            np.testing.assert_array_equal(j.mem[i], mem[i])
    assert rep["instance_cycles"] == sum(j.instances * j.cycles for j in jobs)


def test_budget_evicts_idle_context(designs):
    eng = Engine(designs, capacity=4, slice_cycles=50, backend=OracleBackend, small_jobs=0)
    one = max(eng.context_bytes(d) for d in designs.values())
    eng.budget = one + 1  # room for one resident design at a time
    jobs = [Job(0, "a", 2, 10, 1), Job(1, "b", 2, 10, 2), Job(2, "a", 1, 10, 3)]
    rep = eng.run(jobs)
    assert rep["jobs"] == 3
    assert len(eng.ctx) == 1 and eng.used <= eng.budget


def test_rejects_oversized_jobs(designs):
    eng = Engine(designs, capacity=4, backend=OracleBackend, small_jobs=0)
    with pytest.raises(ValueError):
        eng.submit(Job(0, "a", 5, 10, 1))
    with pytest.raises(ValueError):
        eng.submit(Job(1, "a", 1, 0, 1))


def test_poisson_jobs_shape():
    rng = np.random.default_rng(0)
    jobs = poisson_jobs(["x", "y"], 200, 200.0, rng)
    t = np.array([j.arrival for j in jobs])
    assert np.all(np.diff(t) >= 0)
    # mean inter-arrival 5 ms within sampling error
    assert 0.0035 < np.diff(t).mean() < 0.0065
    assert all(1 <= j.instances <= 64 and 1000 <= j.cycles <= 100_000 for j in jobs)


class CheckpointingOracleBackend(OracleBackend):
    """Adds checkpoint/restore: an instance's state object moves out and back."""

    def checkpoint(self, slot):
        st = self.st[slot]
        self.st[slot] = O.State(self.prog, self.reg0, self.mem0)
        return st

    def restore(self, image, slot):
        self.st[slot] = image


def test_parking_under_budget_pressure(designs):
    """Room for one design: a job of the other design parks the running jobs
    (checkpoints), runs, and the parked jobs resume; results equal alone runs."""
    t = [0.0]
    eng = Engine(designs, capacity=4, slice_cycles=5, backend=CheckpointingOracleBackend, clock=lambda: t[0],
                 park_after=0.5, small_jobs=0)
    eng.budget = max(eng.context_bytes(d) for d in designs.values()) + 1
    jobs = [Job(0, "a", 2, 60, 21, arrival=0.0), Job(1, "a", 1, 45, 22, arrival=0.0),
            Job(2, "b", 3, 20, 23, arrival=0.1)]
    for j in jobs:
        eng.submit(j)
    eng._t0 = 0.0
    for _ in range(200):
        t[0] += 0.1
        if eng.step() == 0:
            break
    assert len(eng.done) == 3 and eng.parked >= 1
    assert jobs[0].parks >= 1
    assert max(j.parks for j in jobs) <= 3  # time slices of park_after, no ping-pong every step
    for j in jobs:
        reg, mem = run_alone(designs[j.design], j, backend=OracleBackend)
        for i in range(j.instances):
            np.testing.assert_array_equal(j.reg[i], reg[i], err_msg=f"job {j.job_id} instance {i}")
            np.testing.assert_array_equal(j.mem[i], mem[i])


def test_small_jobs_use_their_own_contexts(designs):
    """Jobs of at most small_jobs instances run in a design's second context
    (the streamed kernel on the device); results are unchanged."""
    jobs = [Job(0, "a", 1, 30, 41), Job(1, "a", 6, 22, 42), Job(2, "b", 2, 17, 43), Job(3, "a", 2, 9, 44)]
    eng = Engine(designs, capacity=8, slice_cycles=6, backend=OracleBackend, small_jobs=2, small_capacity=4)
    rep = eng.run(jobs)
    assert rep["jobs"] == 4
    assert {k for k in eng.ctx} == {("a", "small"), ("a", "batch"), ("b", "small")}
    for j in jobs:
        reg, mem = run_alone(designs[j.design], j, backend=OracleBackend)
        for i in range(j.instances):
            np.testing.assert_array_equal(j.reg[i], reg[i], err_msg=f"job {j.job_id} instance {i}")
            np.testing.assert_array_equal(j.mem[i], mem[i])


def test_small_jobs_beyond_small_capacity_rejected(designs):
    """A small-job bound above the small contexts' slot count would route jobs
    to a context that can never hold them; the engine refuses it up front."""
    with pytest.raises(ValueError, match="small_capacity"):
        Engine(designs, capacity=8, backend=OracleBackend, small_jobs=5, small_capacity=4)


class SizedOracleBackend(OracleBackend):
    """Reports device bytes (tgs_device_bytes) of three times the estimate."""
    estimate = {}

    def __init__(self, design, kernel=None):
        super().__init__(design, kernel)
        self.design = design

    def device_bytes(self):
        return 3 * self.estimate[id(self.design)]


def test_admission_charges_reported_device_bytes(designs):
    """After a context's first run the engine holds what the runtime reports,
    and estimates for later contexts scale by the ratio seen."""
    eng = Engine(designs, capacity=4, slice_cycles=10, backend=SizedOracleBackend, small_jobs=0)
    for d in designs.values():
        SizedOracleBackend.estimate[id(d)] = eng.context_bytes(d)
    eng.run([Job(0, "a", 2, 10, 1)])
    assert eng.used == 3 * SizedOracleBackend.estimate[id(designs["a"])]
    assert eng.context_bytes(designs["b"]) == 3 * SizedOracleBackend.estimate[id(designs["b"])]
