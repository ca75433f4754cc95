"""Serving: continuous batching of simulation jobs on one GPU.

A job is (design, instances, cycles, seed): run `instances` stimulus instances
of one design for `cycles` cycles each and return their final register and
memory pools. Instance i of a job is driven by the xorshift64 generator
seeded with splitmix64(seed + i), so a job's result is exactly the result of
running it alone in its own context (tgs_set_instances(n) +
tgs_set_stimulus(inputs, seed)).

The engine keeps one batched context per design resident in HBM (program
image plus instance-minor state for a fixed number of instance slots) and,
optionally, a second context per design on the streamed single-instance
kernel for jobs of at most `small_jobs` instances. A batched launch-cycle
of a few instances costs a good part of one of many (each wave walks a
partition's tasks in series for its group of instances: about 0.17 ms for
64 instances and 1.0 ms for 1,024 on 200K LUT4 partitioned by 256), while
the streamed kernel spreads one instance's partition over a workgroup
(about 52 us per cycle on the same design): on a lightly loaded GPU a small
job finishes sooner on it. Under load the batched context, which
shares every program read among all its instances, gives the higher
throughput, so `small_jobs` is 0 (off) by default. The engine runs in
slices: every step makes one run per context with active slots, each active
slot advancing min(remaining, slice) cycles; the slices of all batched
contexts run in shared launches (tgs_run_batch_multi: every launch cycle
runs region 0 of every design, then region 1), so the small batches of
different designs occupy the GPU together.
Requests that arrive while others run are admitted at the next step into
slots that finished jobs freed, so jobs of one design overlap in the same
launch, and so do jobs of different designs. Admission
is by slots and by an HBM budget: a design context is charged an estimate
of its program image and slot state when it is created and, from its
first run on, the device memory the runtime reports it holds
(tgs_device_bytes; the estimate for later contexts is scaled by the
largest ratio seen); when the budget is short, idle design
contexts are evicted (least recently used) before a job waits. A job that
has waited longer than `park_after` seconds for a context that does not fit
parks the jobs of the least recently used busy context: each of their
instances is checkpointed to host memory (pools, cycle count, stop flag,
stimulus generator), the context is evicted, and the parked jobs resume
later from their checkpoints in whatever slots are free, so parking never
changes a result.

The reference has no serving path (one simulator per process, global device
state, src/ToucanSim.cu:102-204); this module is the batching/serving layer
the C-ABI contexts make possible.
"""
import time
from dataclasses import dataclass, field

import numpy as np

from . import _native

KERNEL_BATCH = _native.KERNEL_BATCH
KERNEL_STREAM = _native.KERNEL_STREAM
POOL_REG, POOL_MEM, POOL_EXCHANGE = _native.POOL_REG, _native.POOL_MEM, _native.POOL_EXCHANGE


@dataclass
class Job:
    job_id: int
    design: str          # key into the engine's design table
    instances: int
    cycles: int
    seed: int
    arrival: float = 0.0  # seconds after the engine's start
    # filled by the engine
    admitted: float = None
    finished: float = None
    slots: list = field(default_factory=list)
    remaining: np.ndarray = None
    images: list = None   # per instance checkpoint while parked
    parks: int = 0
    waiting_since: float = None  # arrival, or the time it was last parked
    running_since: float = None  # admission, or the time it last resumed
    reg: list = None      # per instance: final register pool
    mem: list = None      # per instance: final memory pool

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


def gpu_backend(design, kernel=KERNEL_BATCH):
    """A simulator context for a generated design (gen.Design) on `kernel`."""
    from . import rt
    sim = rt.Simulator(0)
    if sim.init(0, design.design, design.symbols, False, False) != 0:
        raise RuntimeError(f"cannot load design: {sim._err()}")
    sim.set_kernel(kernel)
    return sim


def run_shared(sims, cycles):
    """One slice of several batched contexts: in shared launches
    (tgs_run_batch_multi) when they are device contexts, else one by one."""
    from . import rt
    if len(sims) > 1 and all(isinstance(s, rt.Simulator) for s in sims):
        rt.run_batch_multi(sims, cycles)
        return 1
    for s, c in zip(sims, cycles):
        s.run_batch(c)
    return len(sims)


class _Ctx:
    """One resident context of a design: `capacity` instance slots of one kind
    ("batch": the batched kernel; "small": the streamed kernel)."""

    def __init__(self, key, design, sim, capacity, nbytes):
        self.key, self.design, self.sim, self.capacity, self.nbytes = key, design, sim, capacity, nbytes
        self.design_key, self.kind = key
        self.free = list(range(capacity))
        self.jobs = []
        self.last_used = 0.0


class Engine:
    """Continuous-batching job engine.

    designs:   {name: gen.Design} (program, symbols, stimulus input ids)
    budget:    HBM bytes the engine may hold resident (programs + state)
    capacity:  instance slots per design context (a multiple of 64 is best)
    slice:     cycles per step for each active slot
    backend:   factory design -> simulator (gpu_backend by default; tests
               pass a host stand-in with the same methods)
    """

    def __init__(self, designs, budget=250 << 30, capacity=128, slice_cycles=2000, backend=gpu_backend,
                 clock=time.perf_counter, park_after=1.0, small_jobs=0, small_capacity=64, shared=True):
        if small_jobs > small_capacity:
            # a job routed to the small context could never find enough free slots
            raise ValueError(f"small_jobs ({small_jobs}) exceeds small_capacity ({small_capacity})")
        self.designs = designs
        self.shared = shared  # batched contexts' slices in shared launches
        self.lut_cycles = 0   # instance-cycles weighted by the design's LUT4 count
        self.budget = budget
        self.park_after = park_after
        self.small_jobs = small_jobs
        self.small_capacity = small_capacity
        self.capacity = capacity
        self.slice = slice_cycles
        self.backend = backend
        self.clock = clock
        self.ctx = {}
        self.used = 0
        self.queue = []          # admitted-order queue of waiting jobs
        self.done = []
        self.instance_cycles = 0
        self.steps = 0
        self.launches = 0
        self.parked = 0
        self._ratio = 1.0  # device bytes / estimate, largest seen
        self._t0 = None

    # ---- HBM accounting -------------------------------------------------
    @staticmethod
    def state_bytes(design):
        """Per-instance resident state: register, memory and exchange pools."""
        st = design.stats
        return int(st["reg_pool"]) + int(st["mem_pool"]) + int(st["exchange_pool"]) + 64

    def context_bytes(self, design, capacity=None):
        """Estimate before a context exists: program image (about the
        serialized size) plus all slots' state, scaled by the largest ratio
        of device bytes to this estimate seen on the contexts so far."""
        return int(self._estimate(design, capacity) * self._ratio)

    def _estimate(self, design, capacity=None):
        cap = self.capacity if capacity is None else capacity
        return int(design.stats["design_bytes"]) + cap * self.state_bytes(design)

    def calibrate(self, design, device_bytes, capacity=None):
        """Scale later estimates by a context of `design` measured at
        `device_bytes` (tgs_device_bytes), as after a context's first run."""
        self._ratio = max(self._ratio, device_bytes / max(self._estimate(design, capacity), 1))

    def _charge(self, c):
        """Charge a context the device memory it holds (tgs_device_bytes),
        once the runtime reports it; the estimate stands until then."""
        q = getattr(c.sim, "device_bytes", None)
        if q is None:
            return
        actual = q()
        self.calibrate(c.design, actual, c.capacity)
        self.used += actual - c.nbytes
        c.nbytes = actual

    def kind_of(self, job):
        return "small" if self.small_jobs and job.instances <= self.small_jobs else "batch"

    def _make_room(self, need, now):
        while self.used + need > self.budget:
            idle = [c for c in self.ctx.values() if not c.jobs]
            if not idle:
                return False
            victim = min(idle, key=lambda c: c.last_used)
            victim.sim.cleanup()
            self.used -= victim.nbytes
            del self.ctx[victim.key]
        return True

    def _context(self, name, kind, now):
        key = (name, kind)
        c = self.ctx.get(key)
        if c is not None:
            return c
        d = self.designs[name]
        cap = self.small_capacity if kind == "small" else self.capacity
        if self._estimate(d, cap) > self.budget:
            raise ValueError(f"design {name} alone needs {self._estimate(d, cap)} B, more than the budget "
                             f"{self.budget} B")
        # the scaled estimate, at most the whole budget (a context that
        # alone fills it runs when everything else is evicted or parked)
        nbytes = min(self.context_bytes(d, cap), self.budget)
        if not self._make_room(nbytes, now):
            return None
        sim = self.backend(d, KERNEL_STREAM if kind == "small" else KERNEL_BATCH)
        sim.set_instances(cap)
        sim.set_stimulus(d.inputs, 0)
        c = _Ctx(key, d, sim, cap, nbytes)
        self.ctx[key] = c
        self.used += nbytes
        return c

    # ---- scheduling ---------------------------------------------------------
    def submit(self, job):
        if job.instances < 1 or job.instances > self.capacity:
            raise ValueError(f"job {job.job_id}: 1..{self.capacity} instances per job")
        if job.cycles < 1:
            raise ValueError(f"job {job.job_id}: at least one cycle")
        self.queue.append(job)

    def _park(self, c, now):
        """Checkpoint every running job of context `c`, evict it; returns the jobs."""
        jobs = c.jobs
        for job in jobs:
            job.images = [c.sim.checkpoint(s) for s in job.slots]
            job.slots = []
            job.parks += 1
            job.waiting_since = now
            self.parked += 1
        c.jobs = []
        c.sim.cleanup()
        self.used -= c.nbytes
        del self.ctx[c.key]
        return jobs

    def _admit(self, now):
        waiting, parked = [], []
        for job in self.queue:
            if job.arrival > now:
                waiting.append(job)
                continue
            kind = self.kind_of(job)
            c = self._context(job.design, kind, now)
            since = job.arrival if job.waiting_since is None else job.waiting_since
            if c is None and now - since > self.park_after:
                # only contexts whose jobs have all run at least park_after since
                # they started or resumed (no ping-pong between two designs)
                busy = [x for x in self.ctx.values() if x.jobs and x.design_key != job.design and
                        all(now - j.running_since >= self.park_after for j in x.jobs)]
                if busy:
                    parked += self._park(min(busy, key=lambda x: x.last_used), now)
                    c = self._context(job.design, kind, now)
            if c is None or len(c.free) < job.instances:
                waiting.append(job)
                continue
            job.slots = [c.free.pop(0) for _ in range(job.instances)]
            job.running_since = now
            if job.images is not None:  # a parked job resumes where it stopped
                for img, s in zip(job.images, job.slots):
                    c.sim.restore(img, s)
                job.images = None
            else:
                for i, s in enumerate(job.slots):
                    c.sim.reset_instance(s, job.seed, i)
                job.remaining = np.full(job.instances, job.cycles, np.int64)
                job.admitted = now
            c.jobs.append(job)
        self.queue = parked + waiting

    def step(self):
        """Admit arrived jobs, run one slice of every design with active slots,
        retire finished jobs. Returns the number of jobs still pending."""
        if self._t0 is None:
            self._t0 = self.clock()
        now = self.clock() - self._t0
        self._admit(now)
        live = [c for c in self.ctx.values() if c.jobs]
        cycs = []
        for c in live:
            cyc = np.zeros(c.capacity, np.uint32)
            for job in c.jobs:
                n = np.minimum(job.remaining, self.slice)
                cyc[job.slots] = n
                job.remaining -= n
            cycs.append(cyc)
            self.instance_cycles += int(cyc.sum())
            self.lut_cycles += int(cyc.sum()) * int(c.design.stats.get("n_luts", 0))
        # every batched context's slice in shared launches, small-job
        # (streamed-kernel) contexts one by one
        batch = [k for k, c in enumerate(live) if c.key[1] == "batch"] if self.shared else []
        if batch:
            self.launches += run_shared([live[k].sim for k in batch], [cycs[k] for k in batch])
        for k, c in enumerate(live):
            if k not in batch:
                c.sim.run_batch(cycs[k])
                self.launches += 1
        for c in live:
            t = self.clock() - self._t0
            c.last_used = t
            self._charge(c)  # images are built on a context's first run
            keep = []
            for job in c.jobs:
                if job.remaining.max() > 0:
                    keep.append(job)
                    continue
                job.reg = [c.sim.read_pool(POOL_REG, s) for s in job.slots]
                job.mem = [c.sim.read_pool(POOL_MEM, s) for s in job.slots]
                job.finished = self.clock() - self._t0
                c.free.extend(job.slots)
                c.free.sort()
                self.done.append(job)
            c.jobs = keep
        self.steps += 1
        return len(self.queue) + sum(len(c.jobs) for c in self.ctx.values())

    def run(self, jobs, idle_sleep=0.0005):
        """Serve `jobs` (arrival times relative to the start) to completion."""
        for j in sorted(jobs, key=lambda j: j.arrival):
            self.submit(j)
        self._t0 = self.clock()
        while True:
            pending = self.step()
            if pending == 0:
                break
            if not any(c.jobs for c in self.ctx.values()):
                # nothing on the device: wait for the next arrival
                nxt = min((j.arrival for j in self.queue), default=None)
                now = self.clock() - self._t0
                if nxt is not None and nxt > now:
                    time.sleep(min(nxt - now, 0.05))
                elif nxt is not None:
                    # arrived but not admissible (budget held by running jobs is impossible here)
                    time.sleep(idle_sleep)
        return self.report()

    def report(self):
        lat = np.array([j.latency for j in self.done]) if self.done else np.zeros(0)
        wall = (self.clock() - self._t0) if self._t0 is not None else 0.0
        return {
            "jobs": len(self.done),
            "p50_latency_s": float(np.percentile(lat, 50)) if lat.size else None,
            "p99_latency_s": float(np.percentile(lat, 99)) if lat.size else None,
            "mean_latency_s": float(lat.mean()) if lat.size else None,
            "instance_cycles": self.instance_cycles,
            "seconds": wall,
            "cycles_per_s": self.instance_cycles / wall if wall > 0 else None,
            "lut4_cycles_per_s": self.lut_cycles / wall if wall > 0 else None,
            "shared_launches": self.shared,
            "launches": self.launches,
            "parked": self.parked,
            "steps": self.steps,
            "resident_bytes": self.used,
        }

    def close(self):
        for c in self.ctx.values():
            c.sim.cleanup()
        self.ctx.clear()
        self.used = 0


def run_alone(design, job, backend=gpu_backend):
    """The reference result of a job: its instances alone in a fresh context."""
    sim = backend(design)
    sim.set_instances(job.instances)
    sim.set_stimulus(design.inputs, job.seed)
    sim.run_batch(np.full(job.instances, job.cycles, np.uint32))
    reg = [sim.read_pool(POOL_REG, i) for i in range(job.instances)]
    mem = [sim.read_pool(POOL_MEM, i) for i in range(job.instances)]
    sim.cleanup()
    return reg, mem


def poisson_jobs(design_names, n_jobs, rate, rng, min_inst=1, max_inst=64, min_cycles=1000, max_cycles=100_000):
    """Poisson arrivals at `rate` jobs/s over the designs, instances and cycle
    counts uniform in the given ranges."""
    t = 0.0
    jobs = []
    for k in range(n_jobs):
        t += rng.exponential(1.0 / rate)
        jobs.append(Job(job_id=k, design=design_names[int(rng.integers(len(design_names)))],
                        instances=int(rng.integers(min_inst, max_inst + 1)),
                        cycles=int(rng.integers(min_cycles, max_cycles + 1)), seed=int(rng.integers(1 << 31)),
                        arrival=t))
    return jobs


__all__ = ["Job", "Engine", "gpu_backend", "run_alone", "poisson_jobs"]
