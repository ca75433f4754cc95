"""Per-partition profile (the reference's ENABLE_SIM_PROFILE instrument,
src/SimEval.cu:894-983) and the print log across long runs, probe changes
and slot reuse.

* Profile: every partition of the design is reported by every
  single-instance kernel with 0 < useful <= total ticks; the barrier-stall
  share covers every workgroup; turning the profile off leaves nothing
  collected.
* Print log: more print events than one device log holds (65,536), fired
  over several eval_free_running calls, all reach the host in the oracle's
  order; print events survive a probe-list change mid-run (readSignal /
  enableVCDDump add probes); a batched slot reset for a new job starts with
  an empty log.
"""
import numpy as np
import pytest

from simharness import OracleRun, assert_pools_equal, gpu_pools, gpu_sim
from toucangpurt_amd import KERNEL_BATCH, KERNEL_PLAIN, KERNEL_STREAM, KERNEL_WAVE64, gen

pytestmark = pytest.mark.gpu

KERNELS = [("plain", KERNEL_PLAIN), ("wave64", KERNEL_WAVE64), ("stream", KERNEL_STREAM)]


@pytest.mark.parametrize("kname,kernel", KERNELS)
def test_every_partition_profiled(gpu_available, kname, kernel):
    d = gen.generate("lut200k", seed=1)
    n_parts = sum(d.stats["n_parts"])
    print(f"case profile lut200k kernel={kname}", flush=True)
    sim = gpu_sim(d.design, inputs=d.inputs, seed=3, kernel=kernel)
    u, t, k = sim.part_profile()
    assert k == 0 and not u.any()  # off by default
    sim.set_part_profile(True, 10, 50)
    sim.eval_free_running(100)
    u, t, k = sim.part_profile()
    assert k == 50 and u.size == n_parts
    assert (u > 0).all() and (u <= t).all(), (u[u > t][:5], t[u > t][:5])
    # partitions of one workgroup in one region share the region's total
    prof = sim.profile()
    if kernel != KERNEL_PLAIN:
        assert 0 < prof["barrier_ticks"] < prof["total_ticks"]
    # a launch that does not reach the window keeps the last averages
    sim.eval_free_running(5)
    u2, t2, k2 = sim.part_profile()
    assert k2 == 50 and np.array_equal(u, u2) and np.array_equal(t, t2)
    sim.set_part_profile(False)
    assert sim.part_profile()[2] == 0
    sim.cleanup()


def print_design():
    return gen.generate(None, n_luts=3000, depth=6, n_regs=100, n_inputs=20, n_prints=6, n_stops=0,
                        part_luts=600, cluster_luts=100, seed=21)


@pytest.mark.parametrize("kname,kernel", KERNELS)
def test_print_log_beyond_one_device_log(gpu_available, kname, kernel):
    d = print_design()
    print(f"case print log 40000 cycles kernel={kname}", flush=True)
    sim = gpu_sim(d.design, d.symbols, False, d.inputs, seed=2, kernel=kernel, enable_print=True)
    orc = OracleRun(d.design, d.symbols, False, d.inputs, seed=2)
    for _ in range(2):
        sim.eval_free_running(20000)
    orc.run(40000, enable_print=True)
    want = sorted(orc.st.print_log())
    assert len(want) > 65536, len(want)
    assert sim.print_log() == want
    assert sim.clear_print_log() == 0  # nothing was dropped
    assert sim.print_log() == []
    sim.eval_free_running(10)
    orc.run(10, enable_print=True)
    assert sim.print_log() == sorted(orc.st.print_log())[len(want):]
    assert_pools_equal(gpu_pools(sim), orc.pools(), "after 40010 cycles")
    sim.cleanup()


@pytest.mark.parametrize("kname,kernel", KERNELS + [("batched", KERNEL_BATCH)])
def test_prints_survive_probe_changes(gpu_available, kname, kernel):
    d = print_design()
    print(f"case prints + probes kernel={kname}", flush=True)
    sim = gpu_sim(d.design, d.symbols, False, d.inputs, seed=4, kernel=kernel, enable_print=True)
    orc = OracleRun(d.design, d.symbols, False, d.inputs, seed=4)
    names = sorted(sim.symbols["signal"])
    sim.eval_free_running(60)
    sim.probe_signals(names[:3])  # new probe list mid-run
    sim.eval_free_running(60)
    sim.probe_signals(names[3:6])
    sim.eval_free_running(30)
    orc.run(150, enable_print=True)
    assert sim.print_log() == sorted(orc.st.print_log())
    assert_pools_equal(gpu_pools(sim), orc.pools(), "after 150 cycles")
    sim.cleanup()


def test_reset_slot_starts_with_an_empty_log(gpu_available):
    d = print_design()
    print("case batched slot reset clears its print log", flush=True)
    sim = gpu_sim(d.design, d.symbols, False, None, kernel=KERNEL_BATCH, enable_print=True)
    sim.set_instances(3)
    sim.set_stimulus(d.inputs, 6)
    sim.run_batch([40, 40, 40])
    assert sim.print_log(1)
    sim.reset_instance(1, 9, 0)  # slot 1 takes a new job: generator splitmix64(9 + 0)
    assert sim.print_log(1) == []
    sim.run_batch([0, 25, 0])
    orc = OracleRun(d.design, d.symbols, False, d.inputs, seed=9, instance=0)
    orc.run(25, enable_print=True)
    assert sim.print_log(1) == sorted(orc.st.print_log())
    other = OracleRun(d.design, d.symbols, False, d.inputs, seed=6, instance=2)
    other.run(40, enable_print=True)
    assert sim.print_log(2) == sorted(other.st.print_log())
    sim.cleanup()
