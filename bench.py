#!/usr/bin/env python3
"""Benchmark: simulated cycles/s of the 1M-LUT4 design on one MI355X.

Workload (single mode, default): seeded synthetic design of 1,000,000 LUT4s
plus 20,000 word ops at depth 64, 50,000 registers, 16 SRAMs of 4096 x 64
bit, 128 xorshift64-driven inputs, 100,000 cycles. A step is one
eval_free_running call of --chunk cycles (default 10,000); --steps 10 covers
the 100,000 cycles. Inputs are resident in HBM before timing starts.

Batched mode (--mode batched): the 200,000-LUT4 design run as --instances
stimulus instances with uneven cycle counts; reports aggregate
stimulus-cycles/s.

Baselines reported beside the value:
  * plain: the reference-execution-model kernel on the same design (a
    bounded cycle count, rate is per cycle);
  * cpu_baseline: the C oracle on the GPU box's host, 1 thread, on a bounded
    number of cycles of the same workload (checker code, measured only).
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
    ap.add_argument("--mode", choices=["single", "batched"], default="single")
    ap.add_argument("--chunk", type=int, default=10_000, help="cycles per step (single mode)")
    ap.add_argument("--plain-cycles", type=int, default=1_000)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--instances", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=1)
    return ap.parse_args()


def cpu_baseline(design, inputs, seed, seconds):
    """Oracle cycles/s on the host (1 thread), bounded to about `seconds`."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    prog = O.Program(design)
    st = O.State(prog)
    st.set_stimulus(inputs, seed)
    t0 = time.perf_counter()
    st.run(2)
    per = max((time.perf_counter() - t0) / 2, 1e-6)
    n = max(2, int(seconds / per))
    t0 = time.perf_counter()
    st.run(n)
    dt = time.perf_counter() - t0
    return n, dt, st


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU sharding is described "
              "in DESIGN.md and not built", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    from toucangpurt_amd import KERNEL_PLAIN, KERNEL_WAVE64, POOL_REG, gen, rt

    if args.mode == "single":
        d = gen.generate("lut1m", seed=args.seed)
        workload = "lut1m: 1,000,000 LUT4 + 20,000 word ops, depth 64, 50,000 regs, 16 SRAM 4096x64b, 128 inputs"
        sim = rt.Simulator(0)
        assert sim.init(0, d.design, d.symbols, False, False) == 0
        sim.set_kernel(KERNEL_WAVE64)
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
        cycles = args.steps * args.chunk
        value = cycles / dt
        launch_ms = prof["kernel_ms"] / max(prof["launches"], 1)
        cyc_per_launch = prof["cycles"] / max(prof["launches"], 1)
        bytes_per_cycle = stats["b_prog"] + stats["b_state"]
        achieved = bytes_per_cycle * cyc_per_launch / (launch_ms / 1e3) / 1e9
        stall = prof["barrier_ticks"] / prof["total_ticks"] if prof["total_ticks"] else None
        sim.cleanup()

        plain = None
        if not args.no_plain:
            ps = rt.Simulator(0)
            assert ps.init(0, d.design, d.symbols, False, False) == 0
            ps.set_kernel(KERNEL_PLAIN)
            ps.set_stimulus(d.inputs, args.seed)
            ps.eval_free_running(10)
            t1 = time.perf_counter()
            ps.eval_free_running(args.plain_cycles)
            pdt = time.perf_counter() - t1
            plain = {"value": args.plain_cycles / pdt, "cycles": args.plain_cycles, "seconds": pdt,
                     "kernel_ms": ps.profile()["kernel_ms"]}
            ps.cleanup()

        cpu = None
        parity = None
        if not args.no_cpu:
            n, cdt, st = cpu_baseline(d.design, d.inputs, args.seed, args.cpu_seconds)
            cpu = {"value": n / cdt, "unit": "simulated cycles/s", "cores": 1, "kind": "port",
                   "sample": f"{n + 2} cycles of the same design and stimulus on the C oracle, 1 thread"}
            chk = rt.Simulator(0)
            assert chk.init(0, d.design, d.symbols, False, False) == 0
            chk.set_kernel(KERNEL_WAVE64)
            chk.set_stimulus(d.inputs, args.seed)
            chk.eval_free_running(n + 2)
            parity = bool(np.array_equal(chk.read_pool(POOL_REG), st.pool(0)))
            chk.cleanup()

        out = {
            "metric": "simulated cycles/sec @1M-LUT design",
            "value": value,
            "unit": "simulated cycles/s",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": dt / args.steps * 1e3,
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "u8",
            "data": "synthetic",
            "config": {"workload": workload, "cycles": cycles, "cycles_per_step": args.chunk,
                       "kernel": "wave64", "grid": prof["grid"], "block": prof["block"],
                       "lds_bytes": prof["lds_bytes"], "program_resident_in_lds": bool(prof["resident"])},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                         "bytes_per_cycle": bytes_per_cycle, "b_prog": stats["b_prog"],
                         "b_state": stats["b_state"], "launch_ms": launch_ms,
                         "cycles_per_launch": cyc_per_launch},
            "barrier_stall_frac": stall,
            "plain_kernel": plain,
            "speedup_vs_plain": (value / plain["value"]) if plain else None,
            "cpu_baseline": cpu,
            "parity_vs_oracle": parity,
        }
        print(json.dumps(out))
        return

    # ---------------- batched ----------------
    raise SystemExit("batched mode is not built yet")


if __name__ == "__main__":
    main()
