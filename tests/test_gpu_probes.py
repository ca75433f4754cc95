"""Signal probes (tgs_set_probes / tgs_read_probes) on every kernel family.

Probes name (partition, value slot) pairs; after a run each holds the value
its partition left in that slot in the last cycle. Probed slots are ones every
visit writes (oracle written-slot listing), so the value is defined; the
oracle's own probes give the expected values.
"""
import numpy as np
import pytest

import oracle as O
from simharness import OracleRun, gpu_sim
from toucangpurt_amd import KERNEL_BATCH, KERNEL_PLAIN, KERNEL_STREAM, KERNEL_WAVE64, gen
from toucangpurt_amd._native import TGS_E_ARG
from toucangpurt_amd.rt import SimError

pytestmark = pytest.mark.gpu

KW = dict(n_luts=6000, depth=8, n_regs=300, n_inputs=70, n_vecops=400, n_srams=2, sram_depth=1024, sram_width=32,
          word_bits_mix=50, part_luts=800, seed=51)
_d = {}


def design():
    if "d" not in _d:
        _d["d"] = gen.generate(None, **KW)
    return _d["d"]


def probe_set(d, n, seed):
    prog = O.Program(d.design)
    rng = np.random.default_rng(seed)
    parts, slots = [], []
    for _ in range(n):
        p = int(rng.integers(prog.num_parts()))
        w = prog.written_slots(p)
        parts.append(p)
        slots.append(int(w[rng.integers(w.size)]))
    return np.array(parts, np.uint32), np.array(slots, np.uint32)


def oracle_probes(d, parts, slots, cycles, seed, instance=0):
    o = OracleRun(d.design, d.symbols, True, d.inputs, seed, instance)
    o.st.set_probes(parts, slots)
    o.run(cycles)
    return o.st.probes()


@pytest.mark.parametrize("kernel", [KERNEL_PLAIN, KERNEL_WAVE64, KERNEL_STREAM],
                         ids=["plain", "wave64", "stream"])
def test_probes_match_oracle(gpu_available, kernel):
    d = design()
    parts, slots = probe_set(d, 400, 3)
    print(f"case probes kernel={kernel}", flush=True)
    sim = gpu_sim(d.design, d.symbols, True, d.inputs, 11, kernel=kernel)
    sim.set_probes(parts, slots)
    assert not sim.read_probes().any()  # zeroed until a cycle runs
    sim.eval_free_running(5)
    np.testing.assert_array_equal(sim.read_probes(), oracle_probes(d, parts, slots, 5, 11))
    sim.eval()
    np.testing.assert_array_equal(sim.read_probes(), oracle_probes(d, parts, slots, 6, 11))
    sim.cleanup()


def test_batched_probes_per_instance(gpu_available):
    d = design()
    parts, slots = probe_set(d, 300, 4)
    print("case probes batch", flush=True)
    n = 70
    cycles = (np.arange(n) % 5 + 2).astype(np.uint32)  # lanes stop at different cycles
    sim = gpu_sim(d.design, d.symbols, True, kernel=KERNEL_BATCH)
    sim.set_instances(n)
    sim.set_stimulus(d.inputs, 21)
    sim.set_probes(parts, slots)
    sim.run_batch(cycles)
    for i in (0, 1, 4, 33, 69):
        np.testing.assert_array_equal(sim.read_probes(i), oracle_probes(d, parts, slots, int(cycles[i]), 21, i),
                                      err_msg=f"instance {i}")
    sim.cleanup()


def test_probes_survive_layout_switch_and_reject_bad_slots(gpu_available):
    d = design()
    parts, slots = probe_set(d, 50, 5)
    print("case probes switch", flush=True)
    sim = gpu_sim(d.design, d.symbols, True, d.inputs, 8, kernel=KERNEL_WAVE64)
    sim.set_probes(parts, slots)
    sim.eval_free_running(3)
    want = oracle_probes(d, parts, slots, 3, 8)
    sim.set_kernel(KERNEL_BATCH)  # instance-minor layout: the values move with the state
    np.testing.assert_array_equal(sim.read_probes(), want)
    prog = O.Program(d.design)
    with pytest.raises(SimError) as e:
        sim.set_probes([prog.num_parts()], [20])
    assert e.value.code == TGS_E_ARG
    with pytest.raises(SimError) as e:
        sim.set_probes([0], [65535])
    assert e.value.code == TGS_E_ARG
    sim.cleanup()


def test_read_signal_python(gpu_available):
    d = design()
    print("case readSignal", flush=True)
    sim = gpu_sim(d.design, d.symbols, True, d.inputs, 2, kernel=KERNEL_STREAM)
    sigs = sim.symbols["signal"]
    names = sorted(sigs)
    assert names
    assert sim.readSignal(names[0]) == (1 << 64) - 1  # first named: observable from the next cycle
    sim.probe_signals(names)
    sim.eval_free_running(4)
    vals = oracle_probes(d, [p for n in names for p, _, _ in sigs[n]], [v for n in names for _, v, _ in sigs[n]], 4, 2)
    k = 0
    for n in names:
        want = 0
        for _, _, w in sigs[n]:
            want = (want << w) | (int(vals[k]) & ((1 << w) - 1))
            k += 1
        assert sim.readSignal(n) == want, n
    sim.cleanup()
