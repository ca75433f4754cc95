"""Asynchronous snapshots (register pool + signal probes) and the VCD writer on the GPU."""
import numpy as np
import pytest

from simharness import OracleRun
from toucangpurt_amd import KERNEL_BATCH, KERNEL_STREAM, POOL_REG, gen, rt
from toucangpurt_amd.rt import SimError

pytestmark = pytest.mark.gpu

KW = dict(n_luts=3000, depth=6, n_regs=120, n_inputs=16, n_vecops=120, n_srams=2, sram_depth=256, sram_width=32,
          part_luts=600, seed=31)


@pytest.mark.parametrize("kernel", [KERNEL_STREAM, KERNEL_BATCH], ids=["stream", "batch"])
def test_snapshots_match_pool_reads(gpu_available, kernel):
    d = gen.generate(None, **KW)
    print(f"case snapshots kernel={kernel}", flush=True)
    sim = rt.Simulator(0)
    assert sim.init(0, d.design, d.symbols, False, False) == 0
    sim.set_kernel(kernel)
    if kernel == KERNEL_BATCH:
        sim.set_instances(3)
    sim.set_stimulus(d.inputs, 7)
    tickets, want = [], []
    for k in range(6):
        sim.eval_free_running(k + 1)
        tickets.append((sim.snapshot_begin(1 if kernel == KERNEL_BATCH else 0), sim.cycles(0)))
        want.append(sim.read_pool(POOL_REG, 1 if kernel == KERNEL_BATCH else 0))
    # only the last four are still held
    with pytest.raises(SimError):
        sim.snapshot_wait(tickets[0][0])
    for (t, cyc), w in list(zip(tickets, want))[2:]:
        pool, c = sim.snapshot_wait(t)
        assert c == cyc
        assert np.array_equal(pool, w)
    sim.cleanup()


def parse_vcd(text):
    ids, stamps, cur, vals = {}, [], None, {}
    for line in text.splitlines():
        if line.startswith("$var"):
            _, _, width, ident, name, _ = line.split()
            ids[ident] = (name, int(width))
        elif line.startswith("#"):
            cur = int(line[1:])
            stamps.append(cur)
            vals[cur] = {}
        elif cur is not None and line:
            if line.startswith("b"):
                bits, ident = line[1:].split()
                vals[cur][ids[ident][0]] = int(bits, 2)
            else:
                vals[cur][ids[line[1:]][0]] = int(line[0])
    return stamps, vals


def test_python_vcd_trace_matches_oracle(gpu_available, tmp_path):
    d = gen.generate(None, **KW)
    print("case python vcd", flush=True)
    sim = rt.Simulator(0)
    assert sim.init(0, d.design, d.symbols, False, False) == 0
    sim.set_kernel(KERNEL_STREAM)
    sim.set_stimulus(d.inputs, 3)
    path = tmp_path / "w.vcd"
    sim.enableVCDDump(str(path))
    for _ in range(5):
        sim.eval()
    sim.eval_free_running(5)
    sim.cleanup()
    stamps, vals = parse_vcd(path.read_text())
    assert stamps == [0, 1, 2, 3, 4, 10]
    syms = sim.symbols["reg"]
    sigs = sim.symbols["signal"]
    assert sigs and not set(sigs) & set(syms)
    orc = OracleRun(d.design, inputs=d.inputs, seed=3)
    names = sorted(sigs)
    orc.st.set_probes([p for n in names for p, _, _ in sigs[n]], [v for n in names for _, v, _ in sigs[n]])
    state, seen = {}, {}
    for stamp in stamps:
        orc.run((stamp + 1 if stamp < 5 else 10) - orc.st.cycles)
        reg = orc.pools()[0]
        for name, frags in syms.items():
            v = 0
            for rid, w in frags:
                v = (v << w) | (int(reg[rid]) & ((1 << w) - 1))
            state[name] = v
        probes, k = orc.st.probes(), 0
        for name in names:
            v = 0
            for _, _, w in sigs[name]:
                v = (v << w) | (int(probes[k]) & ((1 << w) - 1))
                k += 1
            state[name] = v
        # change-only dumps: the accumulated waveform equals the oracle's registers
        # and signals (of the last cycle run) at every stamp
        seen.update(vals[stamp])
        assert set(seen) == set(syms) | set(sigs)
        for name in state:
            assert seen[name] == state[name], (stamp, name)
