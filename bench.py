#!/usr/bin/env python3
"""Benchmark: simulated cycles/s of the 1M-LUT4 design on one MI355X, and the
batched stimulus-cycles/s of the 200K-LUT4 design, each beside the plain
one-instance-per-launch kernel.

Single leg (the headline `value`): seeded synthetic design of 1,000,000 LUT4s
plus 20,000 word ops at depth 64, 50,000 registers, 16 SRAMs of 4096 x 64
bit, 128 xorshift64-driven inputs, 100,000 cycles. A step is one
eval_free_running call of --chunk cycles (default 10,000); the default 10
steps cover the 100,000 cycles. Inputs are resident in HBM before timing.

Batched leg: the 200,000-LUT4 design (depth 48, 8 SRAMs of 1024 x 32 bit)
run as --instances instances (default 1,024) with distinct xorshift64 seeds
and cycle counts drawn uniformly from [1,000, 50,000], in one persistent
launch (instances end at their own counts inside it; --slice N cuts the run
into launches of at most N cycles per instance instead). An autotuner first sweeps the partition
size (256..4,096 LUT4s) and the workgroup size on a short run; the logical design is identical for every
partition size, so results do not depend on its choice.

Large leg: the 10,000,000-LUT4 design (depth 128, 500,000 registers, 64
SRAMs) on the streamed kernel for --large-cycles cycles (default 2,000, a
reduced count of its 10,000-cycle workload: the rate is per cycle), with its
own roofline, the plain kernel beside it and parity of all three pools
against the oracle after --large-parity cycles.

Serving leg: --serve-designs seeded designs of log-uniform size, jobs of 1-64
instances with Poisson arrivals, served by the continuous-batching engine
(toucangpurt_amd.serve) under an HBM budget; reports p50/p99 job latency
(arrival to result) and aggregate instance-cycles/s. Designs span the
serving workload's 50K-5M LUT4; job count (48) and cycles per job
(100-2,000) are reduced from its 1K-100K so the leg runs in about a minute,
and the flags restore the full range. The batched contexts' slices run in
shared launches (tgs_run_batch_multi).

The single leg picks the faster of the streamed and the wave64 persistent
kernels on a short run (--kernel names one).

Baselines: the plain kernel (reference execution model) on the same design,
timed on a bounded cycle count (its rate is per cycle), and the C oracle on
the host (bounded samples: all of the process's CPU share as cpu_baseline,
1 thread beside it). The oracle is the
checker; it is timed, never used to produce a GPU result.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", choices=["single", "batched", "large", "serve", "testbench", "both", "all"], default="all",
                    help="all = single + batched + large + serving + testbench legs")
    ap.add_argument("--tb-steps", type=int, default=300, help="testbench steps timed per path")
    ap.add_argument("--serve-jobs", type=int, default=48)
    ap.add_argument("--serve-rate", type=float, default=200.0, help="Poisson arrivals, jobs/s")
    ap.add_argument("--serve-designs", type=int, default=8)
    ap.add_argument("--serve-min-luts", type=int, default=50_000)
    ap.add_argument("--serve-max-luts", type=int, default=5_000_000)
    ap.add_argument("--serve-min-cycles", type=int, default=100)
    ap.add_argument("--serve-max-cycles", type=int, default=2_000)
    ap.add_argument("--serve-slice", type=int, default=500)
    ap.add_argument("--serve-capacity", type=int, default=128)
    ap.add_argument("--serve-budget-gb", type=float, default=250.0)
    ap.add_argument("--serve-separate", action="store_true",
                    help="serving: one launch per design context per step instead of shared launches")
    ap.add_argument("--serve-small-jobs", type=int, default=0,
                    help="jobs of at most this many instances run on the streamed kernel (0: all batched)")
    ap.add_argument("--chunk", type=int, default=10_000, help="cycles per step (single leg)")
    ap.add_argument("--kernel", choices=["auto", "stream", "wave64"], default="auto", help="single-leg kernel")
    ap.add_argument("--tune-single", type=int, default=2_000, help="cycles per kernel in the single-leg choice")
    ap.add_argument("--plain-cycles", type=int, default=1_000)
    ap.add_argument("--large-cycles", type=int, default=2_000, help="timed cycles of the 10M-LUT4 leg")
    ap.add_argument("--large-parity", type=int, default=40, help="cycles of the 10M-LUT4 leg checked against the oracle")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--instances", type=int, default=1024)
    ap.add_argument("--min-cycles", type=int, default=1_000)
    ap.add_argument("--max-cycles", type=int, default=50_000)
    ap.add_argument("--tune-cycles", type=int, default=100)
    ap.add_argument("--part-luts", type=int, default=0,
                    help="batched leg: partition size in LUT4s (0: the autotuner picks)")
    ap.add_argument("--slice", type=int, default=0,
                    help="cycles per instance per batched launch (0: the whole run in one launch)")
    ap.add_argument("--seed", type=int, default=1)
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_threads():
    """Host threads for the all-cores CPU baseline: the process's CPU share
    (OMP_NUM_THREADS where the host sets it), else the CPU count."""
    env = os.environ.get("OMP_NUM_THREADS")
    return max(1, int(env)) if env and env.isdigit() else max(1, os.cpu_count() or 1)


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), for the CPU baseline."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def oracle_rate(st, seconds, threads=1):
    """Oracle cycles/s on the host with `threads` threads, continuing state `st`,
    bounded to about `seconds`. Returns (cycles timed, seconds)."""
    t0 = time.perf_counter()
    st.run(2, threads=threads)
    per = max((time.perf_counter() - t0) / 2, 1e-6)
    n = max(2, int(seconds / per))
    t0 = time.perf_counter()
    st.run(n, threads=threads)
    return n, time.perf_counter() - t0


def plain_rate(d, cycles, seed):
    from toucangpurt_amd import KERNEL_PLAIN, rt
    ps = rt.Simulator(0)
    assert ps.init(0, d.design, d.symbols, False, False) == 0
    ps.set_kernel(KERNEL_PLAIN)
    ps.set_stimulus(d.inputs, seed)
    ps.eval_free_running(10)
    ps.reset_profile()
    t1 = time.perf_counter()
    ps.eval_free_running(cycles)
    dt = time.perf_counter() - t1
    prof = ps.profile()
    ps.cleanup()
    return {"value": cycles / dt, "unit": "simulated cycles/s", "cycles": cycles, "seconds": dt,
            "kernel_ms": prof["kernel_ms"], "grid": prof["grid"], "block": prof["block"]}


# How the generator draws LUT fan-in (stated with every workload).
FANIN = ("each LUT4 input drawn uniformly from the earlier levels of the LUT's own 256-LUT4 cluster, "
         "level 0 being the cluster's 32 register/input (region 0) or exchange (region 1) sources")

KERNEL_NAMES = {0: "plain", 1: "wave64", 3: "stream"}
# device function names of the kernels, as rocprofv3 reports them (profile file names)
KERNEL_SYMBOLS = {0: "plainCycles", 1: "w64Cycles", 3: "streamCycles"}
SOURCES = ["global", "lds-staged", "lds-resident", "streamed"]


def measured_traffic(kernel_name):
    """HBM bytes per simulated cycle measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, calibrated on a copy kernel) for this kernel on this workload, from
    the committed profile summary; None when that profile is absent."""
    path = os.path.join(ROOT, "profiles", f"{kernel_name}_lut1m_hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_cycle:"):
                return float(line.split(":", 1)[1])
    return None


def run_single(args):
    import numpy as np
    from toucangpurt_amd import KERNEL_STREAM, KERNEL_WAVE64, POOL_EXCHANGE, POOL_MEM, POOL_REG, gen, rt
    d = gen.generate("lut1m", seed=args.seed)
    log(f"single: design generated {d.stats['n_parts']} partitions")
    # pick the faster of the persistent kernels on a short run
    kernels = [KERNEL_STREAM, KERNEL_WAVE64] if args.kernel == "auto" else \
        [{"stream": KERNEL_STREAM, "wave64": KERNEL_WAVE64}[args.kernel]]
    best = (0.0, kernels[0]) if len(kernels) == 1 else None
    for k in (kernels if len(kernels) > 1 else []):
        sim = rt.Simulator(0)
        assert sim.init(0, d.design, d.symbols, False, False) == 0
        sim.set_kernel(k)
        sim.set_stimulus(d.inputs, args.seed)
        sim.eval_free_running(100)
        t0 = time.perf_counter()
        sim.eval_free_running(args.tune_single)
        rate = args.tune_single / (time.perf_counter() - t0)
        sim.cleanup()
        log(f"single: {KERNEL_NAMES[k]} {rate:.0f} cycles/s on {args.tune_single} cycles")
        if best is None or rate > best[0]:
            best = (rate, k)
    kernel = best[1]
    sim = rt.Simulator(0)
    assert sim.init(0, d.design, d.symbols, False, False) == 0
    sim.set_kernel(kernel)
    sim.set_stimulus(d.inputs, args.seed)
    stats = sim.program_stats()
    for _ in range(args.warmup):
        sim.eval_free_running(args.chunk)
    sim.reset_profile()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sim.eval_free_running(args.chunk)
    dt = time.perf_counter() - t0
    prof = sim.profile()
    sim.cleanup()
    cycles = args.steps * args.chunk
    launch_ms = prof["kernel_ms"] / max(prof["launches"], 1)
    cyc_per_launch = prof["cycles"] / max(prof["launches"], 1)
    bytes_per_cycle = stats["b_prog"] + stats["b_state"]
    achieved = bytes_per_cycle * cyc_per_launch / (launch_ms / 1e3) / 1e9
    tpc = measured_traffic(KERNEL_SYMBOLS[kernel])
    log(f"single: {KERNEL_NAMES[kernel]} {cycles / dt:.0f} cycles/s, launch {launch_ms:.1f} ms")
    out = {
        "value": cycles / dt, "dt": dt, "cycles": cycles,
        "config": {"workload": "lut1m: 1,000,000 LUT4 + 20,000 word ops, depth 64, 50,000 regs, "
                               "16 SRAM 4096x64b, 128 inputs, 100,000 cycles; " + FANIN,
                   "cycles": cycles, "cycles_per_step": args.chunk, "kernel": KERNEL_NAMES[kernel],
                   "grid": prof["grid"], "block": prof["block"], "lds_bytes": prof["lds_bytes"],
                   "program_source": SOURCES[prof["resident"]]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": tpc * cyc_per_launch if tpc is not None else None,
                     "bytes_per_cycle": bytes_per_cycle, "traffic_bytes_per_cycle": tpc,
                     "b_prog": stats["b_prog"], "b_state": stats["b_state"], "launch_ms": launch_ms,
                     "cycles_per_launch": cyc_per_launch},
        "barrier_stall_frac": (prof["barrier_ticks"] / prof["total_ticks"]) if prof["total_ticks"] else None,
    }
    if not args.no_plain:
        out["plain"] = plain_rate(d, args.plain_cycles, args.seed)
        log(f"single: plain {out['plain']['value']:.0f} cycles/s")
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        st = O.State(O.Program(d.design))
        st.set_stimulus(d.inputs, args.seed)
        n1, dt1 = oracle_rate(st, args.cpu_seconds / 2, 1)
        nt = cpu_threads()
        n2, dt2 = oracle_rate(st, args.cpu_seconds, nt)
        out["cpu"] = {"value": n2 / dt2, "unit": "simulated cycles/s", "cores": nt, "kind": "port", "cpu": cpu_model(),
                      "sample": f"{n2} cycles of the 1M-LUT4 design with the same stimulus (after {n1 + 4} "
                                f"others), C oracle, partitions of each region over {nt} threads"}
        out["cpu_1thread"] = {"value": n1 / dt1, "unit": "simulated cycles/s", "cores": 1, "kind": "port",
                              "cpu": cpu_model(),
                              "sample": f"{n1} cycles of the same run, C oracle, 1 thread"}
        total = n1 + n2 + 4
        chk = rt.Simulator(0)
        assert chk.init(0, d.design, d.symbols, False, False) == 0
        chk.set_kernel(kernel)
        chk.set_stimulus(d.inputs, args.seed)
        chk.eval_free_running(total)
        out["parity"] = all(bool(np.array_equal(chk.read_pool(pool), st.pool(k)))
                            for k, pool in enumerate((POOL_REG, POOL_MEM, POOL_EXCHANGE)))
        out["parity_cycles"] = total
        chk.cleanup()
        log(f"single: cpu {out['cpu']['value']:.1f} cycles/s, parity {out['parity']}")
    return out


def run_large(args):
    """The 10M-LUT4 design on the streamed kernel: --large-cycles timed cycles
    (a reduced count of the 10,000-cycle workload; the rate is per cycle),
    roofline from the program's byte model, parity of all three pools against
    the multi-threaded oracle after --large-parity cycles."""
    import numpy as np
    from toucangpurt_amd import KERNEL_STREAM, POOL_EXCHANGE, POOL_MEM, POOL_REG, gen, rt
    d = gen.generate("lut10m", seed=args.seed)
    log(f"large: design generated {d.stats['n_parts']} partitions")
    sim = rt.Simulator(0)
    assert sim.init(0, d.design, d.symbols, False, False) == 0
    sim.set_kernel(KERNEL_STREAM)
    sim.set_stimulus(d.inputs, args.seed)
    stats = sim.program_stats()
    sim.eval_free_running(20)
    sim.reset_profile()
    t0 = time.perf_counter()
    sim.eval_free_running(args.large_cycles)
    dt = time.perf_counter() - t0
    prof = sim.profile()
    sim.cleanup()
    launch_ms = prof["kernel_ms"] / max(prof["launches"], 1)
    cyc_per_launch = prof["cycles"] / max(prof["launches"], 1)
    bytes_per_cycle = stats["b_prog"] + stats["b_state"]
    achieved = bytes_per_cycle * cyc_per_launch / (launch_ms / 1e3) / 1e9
    out = {"value": args.large_cycles / dt, "unit": "simulated cycles/s", "cycles": args.large_cycles,
           "seconds": dt, "kernel": "stream", "grid": prof["grid"], "block": prof["block"],
           "lds_bytes": prof["lds_bytes"],
           "workload": f"lut10m: 10,000,000 LUT4, depth 128, 500,000 regs, 64 SRAM 8192x64b, 256 inputs; "
                       f"{args.large_cycles} timed cycles (reduced from 10,000; the rate is per cycle); " + FANIN,
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": None, "bytes_per_cycle": bytes_per_cycle,
                        "launch_ms": launch_ms, "cycles_per_launch": cyc_per_launch},
           "barrier_stall_frac": (prof["barrier_ticks"] / prof["total_ticks"]) if prof["total_ticks"] else None}
    log(f"large: stream {out['value']:.0f} cycles/s, launch {launch_ms:.1f} ms")
    if not args.no_plain:
        out["plain"] = plain_rate(d, max(20, args.plain_cycles // 10), args.seed)
        out["speedup_vs_plain"] = out["value"] / out["plain"]["value"]
        log(f"large: plain {out['plain']['value']:.0f} cycles/s")
    if not args.no_cpu and args.large_parity > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        st = O.State(O.Program(d.design))
        st.set_stimulus(d.inputs, args.seed)
        nt = cpu_threads()
        t0 = time.perf_counter()
        st.run(args.large_parity, threads=nt)
        cdt = time.perf_counter() - t0
        out["cpu"] = {"value": args.large_parity / cdt, "unit": "simulated cycles/s", "cores": nt, "kind": "port",
                      "cpu": cpu_model(),
                      "sample": f"the first {args.large_parity} cycles of the 10M-LUT4 design, C oracle, "
                                f"partitions of each region over {nt} threads"}
        chk = rt.Simulator(0)
        assert chk.init(0, d.design, d.symbols, False, False) == 0
        chk.set_kernel(KERNEL_STREAM)
        chk.set_stimulus(d.inputs, args.seed)
        chk.eval_free_running(args.large_parity)
        out["parity"] = all(bool(np.array_equal(chk.read_pool(pool), st.pool(k)))
                            for k, pool in enumerate((POOL_REG, POOL_MEM, POOL_EXCHANGE)))
        out["parity_cycles"] = args.large_parity
        chk.cleanup()
        log(f"large: cpu {out['cpu']['value']:.1f} cycles/s, parity {out['parity']}")
    return out


def run_batched(args):
    import numpy as np
    from toucangpurt_amd import KERNEL_BATCH, POOL_EXCHANGE, POOL_MEM, POOL_REG, gen, rt
    from toucangpurt_amd.rt import SimError
    n = args.instances
    rng = np.random.default_rng(args.seed)
    cycles = rng.integers(args.min_cycles, args.max_cycles + 1, size=n).astype(np.uint32)
    # autotune the partition size on a short uniform run (the runtime's
    # default workgroup shape: two 64-instance groups per workgroup)
    best = None
    tuning = []
    designs = {}
    for part in ((args.part_luts,) if args.part_luts else (256, 512, 1024, 2048, 4096)):
        d = gen.generate("lut200k", seed=args.seed, part_luts=part)
        designs[part] = d
        for block in (0,):
            sim = rt.Simulator(0)
            assert sim.init(0, d.design, d.symbols, False, False) == 0
            try:
                sim.set_launch(KERNEL_BATCH, block, 0)
                sim.set_kernel(KERNEL_BATCH)
                sim.set_instances(n)
                sim.set_stimulus(d.inputs, args.seed)
                sim.run_batch(np.full(n, 10, np.uint32))
                t0 = time.perf_counter()
                sim.run_batch(np.full(n, args.tune_cycles, np.uint32))
                rate = n * args.tune_cycles / (time.perf_counter() - t0)
            except SimError as e:
                tuning.append({"part_luts": part, "error": str(e)[:80]})
                sim.cleanup()
                continue
            tp = sim.profile()
            tuning.append({"part_luts": part, "block": tp["block"], "grid": tp["grid"],
                           "instances_per_wave": tp["batch_group"], "rate": rate})
            log(f"batched tune part_luts={part} block={block}: {rate:.3e} stimulus-cycles/s")
            if best is None or rate > best[0]:
                best = (rate, part, block)
            sim.cleanup()
    _, part, block = best
    d = designs[part]
    sim = rt.Simulator(0)
    assert sim.init(0, d.design, d.symbols, False, False) == 0
    sim.set_launch(KERNEL_BATCH, block, 0)
    sim.set_kernel(KERNEL_BATCH)
    sim.set_instances(n)
    sim.set_stimulus(d.inputs, args.seed)
    sim.reset_profile()
    # one persistent launch for the whole run (or slices of --slice cycles per
    # instance, each one launch)
    if args.slice <= 0:
        log(f"batched: one launch over {int(cycles.sum())} instance-cycles, longest {int(cycles.max())} cycles")
    left = cycles.astype(np.int64).copy()
    launch_cycles = 0
    t0 = time.perf_counter()
    while left.max() > 0:
        step = (left if args.slice <= 0 else np.minimum(left, args.slice)).astype(np.uint32)
        sim.run_batch(step)
        launch_cycles += int(step.max())
        left -= step
        log(f"batched: {int((cycles - left).sum())} of {int(cycles.sum())} instance-cycles done, "
            f"{int((left > 0).sum())} instances live, {time.perf_counter() - t0:.1f} s")
    dt = time.perf_counter() - t0
    prof = sim.profile()
    total = int(cycles.sum())
    stats = sim.program_stats()
    value = total / dt
    log(f"batched: {value:.3e} stimulus-cycles/s over {total} instance-cycles in {dt:.2f} s")
    out = {"value": value, "unit": "stimulus-cycles/s", "instances": n, "total_cycles": total, "seconds": dt,
           "kernel_ms": prof["kernel_ms"], "launches": prof["launches"], "part_luts": part, "block": prof["block"], "grid": prof["grid"],
           "instances_per_wave": prof["batch_group"], "lds_bytes": prof["lds_bytes"], "tuning": tuning,
           "workload": "lut200k: 200,000 LUT4, depth 48, 8 SRAM 1024x32b, 1,024 instances, "
                       f"cycles uniform in [{args.min_cycles}, {args.max_cycles}]; " + FANIN,
           "bytes_per_cycle_shared": stats["b_prog"], "bytes_per_cycle_per_instance": stats["b_state"],
           "lut4_cycles_per_s": value * 200_000}
    # roofline: program bytes once per launch-cycle, state bytes per live instance-cycle
    algo = stats["b_prog"] * launch_cycles + stats["b_state"] * total
    achieved = algo / (prof["kernel_ms"] / 1e3) / 1e9
    out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS}
    # spot parity: the instance with the fewest cycles against a fresh single-instance run
    k = int(np.argmin(cycles))
    got = [sim.read_pool(pool, k) for pool in (POOL_REG, POOL_MEM, POOL_EXCHANGE)]
    sim.cleanup()
    if not args.no_plain:
        out["plain"] = plain_rate(d, args.plain_cycles, args.seed)
        out["speedup_vs_plain"] = value / out["plain"]["value"]
        log(f"batched: plain {out['plain']['value']:.0f} cycles/s")
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        st = O.State(O.Program(d.design))
        st.set_stimulus(d.inputs, args.seed, k)
        nt = cpu_threads()
        t0 = time.perf_counter()
        st.run(int(cycles[k]), threads=nt)
        cdt = time.perf_counter() - t0
        out["cpu"] = {"value": int(cycles[k]) / cdt, "unit": "stimulus-cycles/s", "cores": nt, "kind": "port",
                      "cpu": cpu_model(),
                      "sample": f"instance {k}: {int(cycles[k])} cycles of the 200K-LUT4 design, C oracle, "
                                f"partitions of each region over {nt} threads"}
        out["parity_instance"] = k
        out["parity"] = all(bool(np.array_equal(g, st.pool(i))) for i, g in enumerate(got))
        log(f"batched: cpu {out['cpu']['value']:.1f} cycles/s, parity {out['parity']}")
    return out


def run_serve(args):
    """Serving leg: Poisson arrivals over seeded designs of log-uniform size,
    continuous batching over resident batched contexts, HBM-budget admission."""
    import numpy as np
    from toucangpurt_amd import gen
    from toucangpurt_amd.serve import Engine, poisson_jobs
    rng = np.random.default_rng(args.seed + 7)
    sizes = np.exp(rng.uniform(np.log(args.serve_min_luts), np.log(args.serve_max_luts), args.serve_designs))
    designs = {}
    t0 = time.perf_counter()
    for k, n in enumerate(sizes):
        n = int(n)
        designs[f"d{k}"] = gen.generate(None, n_luts=n, depth=32, n_regs=max(64, n // 20), n_inputs=64,
                                        part_luts=256, seed=args.seed + 100 + k)
    log(f"serve: {len(designs)} designs generated in {time.perf_counter() - t0:.1f} s "
        f"({', '.join(str(int(s)) for s in sorted(sizes))} LUT4)")
    jobs = poisson_jobs(list(designs), args.serve_jobs, args.serve_rate, rng,
                        min_cycles=args.serve_min_cycles, max_cycles=args.serve_max_cycles)
    eng = Engine(designs, budget=int(args.serve_budget_gb * (1 << 30)), capacity=args.serve_capacity,
                 slice_cycles=args.serve_slice, small_jobs=args.serve_small_jobs, shared=not args.serve_separate)
    done = [0]

    def progress():
        n = len(eng.done)
        if n >= done[0] + 8:
            done[0] = n
            log(f"serve: {n} of {len(jobs)} jobs done, {eng.instance_cycles} instance-cycles")

    orig = eng.step

    def step():
        r = orig()
        progress()
        return r
    eng.step = step
    rep = eng.run(jobs)
    eng.close()
    rep.update({"workload": f"{args.serve_designs} designs log-uniform {args.serve_min_luts}-{args.serve_max_luts} "
                            f"LUT4 (depth 32, partitions of 256), {args.serve_jobs} jobs, Poisson "
                            f"{args.serve_rate}/s, 1-64 instances x {args.serve_min_cycles}-{args.serve_max_cycles} "
                            f"cycles, {args.serve_budget_gb} GB budget, {args.serve_capacity} slots per design, "
                            f"slice {args.serve_slice} cycles, jobs of <= {args.serve_small_jobs} instances on the streamed kernel, "
                            f"{'one launch per design' if args.serve_separate else 'shared launches'}",
                "unit": "instance-cycles/s"})
    log(f"serve: p50 {rep['p50_latency_s']:.3f} s p99 {rep['p99_latency_s']:.3f} s, "
        f"{rep['cycles_per_s']:.3e} instance-cycles/s, {rep['lut4_cycles_per_s']:.3e} LUT4-instance-cycles/s")
    return rep


def run_testbench(args):
    """Testbench leg: setRegister -> eval -> readRegister per step (4 registers
    written, 4 read, one cycle) through the reference-style per-register calls
    against tgs_step, on the 20K-LUT4 design (one instance on the streamed
    kernel, 1,024 instances on the batched kernel)."""
    import numpy as np
    from toucangpurt_amd import KERNEL_BATCH, KERNEL_STREAM, gen, rt

    def regs_of(sim):
        names = [n for n in sorted(sim.symbols["reg"]) if len(sim.symbols["reg"][n]) <= 16][:8]
        ids = [np.array([l[0] for n in ns for l in sim.symbols["reg"][n]], np.uint32) for ns in (names[:4], names[4:])]
        return names[:4], names[4:], ids[0], ids[1]

    out = {"design": "lut20k: 20,000 LUT4, depth 32", "registers_written": 4, "registers_read": 4, "cycles_per_step": 1}
    d = gen.generate("lut20k", seed=args.seed)
    sim = rt.Simulator(0)
    assert sim.init(0, d.design, d.symbols, False, False) == 0
    sim.set_kernel(KERNEL_STREAM)
    win, rout, wids, rids = regs_of(sim)
    vals = np.zeros((1, wids.size), np.uint8)

    def slow_step(i):
        for n in win:
            sim.setRegister(n, i & 1)
        sim.eval()
        return [sim.readRegister(n) for n in rout]

    for i in range(20):
        slow_step(i)
        sim.step(wids, vals, 1, rids)
    n = args.tb_steps
    t0 = time.perf_counter()
    for i in range(n):
        slow_step(i)
    slow = (time.perf_counter() - t0) / n
    t0 = time.perf_counter()
    for i in range(n):
        vals[:] = i & 1
        sim.step(wids, vals, 1, rids)
    fast = (time.perf_counter() - t0) / n
    sim.cleanup()
    out["single"] = {"instances": 1, "kernel": "stream", "per_register_calls_us": slow * 1e6, "tgs_step_us": fast * 1e6}
    log(f"testbench: 1 instance {slow * 1e6:.0f} us per step with per-register calls, {fast * 1e6:.0f} us with tgs_step")
    db = gen.generate("lut20k", seed=args.seed, part_luts=512)
    sim = rt.Simulator(0)
    assert sim.init(0, db.design, db.symbols, False, False) == 0
    sim.set_kernel(KERNEL_BATCH)
    ni = 1024
    sim.set_instances(ni)
    sim.set_stimulus(db.inputs, args.seed)
    win, rout, wids, rids = regs_of(sim)
    vals = np.zeros((ni, wids.size), np.uint8)
    for _ in range(5):
        sim.step(wids, vals, 1, rids)
    m = max(args.tb_steps // 10, 5)
    t0 = time.perf_counter()
    for i in range(m):
        vals[:] = i & 1
        sim.step(wids, vals, 1, rids)
    fast = (time.perf_counter() - t0) / m
    t0 = time.perf_counter()
    for k in range(ni):
        for n in win:
            sim.setRegister(n, 1, k)
    sim.run_batch(np.ones(ni, np.uint32))
    for k in range(ni):
        for n in rout:
            sim.readRegister(n, k)
    slow = time.perf_counter() - t0
    sim.cleanup()
    out["batched"] = {"instances": ni, "kernel": "batched", "design": "lut20k with partitions of 512 LUT4",
                      "per_register_calls_us": slow * 1e6, "tgs_step_us": fast * 1e6}
    log(f"testbench: {ni} instances {slow * 1e3:.1f} ms per step with per-register calls, "
        f"{fast * 1e6:.0f} us with tgs_step")
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): sharding by instance or by "
              "partition across GPUs is not built", file=sys.stderr)
        sys.exit(2)
    single = run_single(args) if args.mode in ("single", "both", "all") else None
    batched = run_batched(args) if args.mode in ("batched", "both", "all") else None
    large = run_large(args) if args.mode in ("large", "all") else None
    serving = run_serve(args) if args.mode in ("serve", "all") else None
    testbench = run_testbench(args) if args.mode in ("testbench", "all") else None
    if single is None and batched is None and large is not None:
        out = {"metric": "simulated cycles/sec @10M-LUT design", "value": large["value"],
               "unit": "simulated cycles/s", "n_gpus": 1, "steps": 1, "warmup": 1,
               "ms_per_step": large["seconds"] * 1e3, "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "u8", "data": "synthetic", "config": {"workload": large["workload"]},
               "roofline": large["roofline"], "cpu_baseline": large.get("cpu"), "large": large,
               "serving": serving, "testbench_step": testbench}
        print(json.dumps(out))
        return
    if single is None and batched is None and serving is None:
        out = {"metric": "testbench step latency (write 4 registers, 1 cycle, read 4 registers)",
               "value": testbench["single"]["tgs_step_us"], "unit": "us per step", "n_gpus": 1,
               "steps": args.tb_steps, "warmup": 20, "ms_per_step": testbench["single"]["tgs_step_us"] / 1e3,
               "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
               "data": "synthetic", "config": {"workload": testbench["design"]}, "testbench_step": testbench}
        print(json.dumps(out))
        return
    if single is None and batched is None:
        out = {"metric": "serving p50/p99 job latency and instance-cycles/s", "value": serving["cycles_per_s"],
               "unit": "instance-cycles/s", "n_gpus": 1, "steps": serving["steps"], "warmup": 0,
               "ms_per_step": serving["seconds"] / max(serving["steps"], 1) * 1e3, "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
               "config": {"workload": serving["workload"]}, "serving": serving}
    elif single is not None:
        out = {"metric": "simulated cycles/sec @1M-LUT design; batched stimulus-cycles/sec vs plain",
               "value": single["value"], "unit": "simulated cycles/s", "n_gpus": 1, "steps": args.steps,
               "warmup": args.warmup, "ms_per_step": single["dt"] / args.steps * 1e3, "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
               "config": single["config"], "roofline": single["roofline"],
               "barrier_stall_frac": single["barrier_stall_frac"],
               "plain_kernel": single.get("plain"),
               "speedup_vs_plain": (single["value"] / single["plain"]["value"]) if single.get("plain") else None,
               "cpu_baseline": single.get("cpu"), "cpu_baseline_1thread": single.get("cpu_1thread"),
               "parity_vs_oracle": single.get("parity"), "parity_cycles": single.get("parity_cycles"),
               "batched": batched, "large": large, "serving": serving, "testbench_step": testbench}
    else:
        out = {"metric": "batched stimulus-cycles/sec @200K-LUT design x instances", "value": batched["value"],
               "unit": "stimulus-cycles/s", "n_gpus": 1, "steps": 1, "warmup": 1,
               "ms_per_step": batched["seconds"] * 1e3, "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "u8", "data": "synthetic",
               "config": {"workload": batched["workload"], "part_luts": batched["part_luts"]},
               "roofline": batched["roofline"], "cpu_baseline": batched.get("cpu"),
               "speedup_vs_plain": batched.get("speedup_vs_plain"), "batched": batched, "large": large,
               "serving": serving,
               "testbench_step": testbench}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
