"""The drop-in C++ ToucanSimulator (libtoucan_sim.so) on the GPU: init error
code, setRegister/readRegister round trip, eval + eval_free_running against
the oracle by register name, readSignal against the oracle's signal probes,
VCD output structure."""
import os
import subprocess

import numpy as np
import pytest

from simharness import OracleRun
from toucangpurt_amd import gen
from toucangpurt_amd.rt import parse_symbols

pytestmark = [pytest.mark.gpu, pytest.mark.subprocess_gpu]
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("kernel", [0, 1, 2, 3], ids=["plain", "wave64", "batch", "stream"])
def test_cpp_simulator_class(gpu_available, tmp_path, kernel):
    d = gen.generate(None, n_luts=3000, depth=6, n_regs=120, n_inputs=16, n_vecops=120, n_srams=2, sram_depth=256,
                     sram_width=32, part_luts=600, seed=31)
    (tmp_path / "d.bin").write_bytes(d.design)
    (tmp_path / "s.bin").write_bytes(d.symbols)
    vcd = tmp_path / "w.vcd"
    print(f"case cpp api kernel={kernel}", flush=True)
    out = subprocess.run([os.path.join(ROOT, "toucangpurt_amd", "sim_api_check"), str(tmp_path / "d.bin"),
                          str(tmp_path / "s.bin"), str(vcd), "5", str(kernel)], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr[-2000:]
    got, sig = {}, {}
    for line in out.stdout.splitlines():
        if line.startswith("reg "):
            _, name, val = line.split()
            got[name] = int(val)
        if line.startswith("sig "):
            _, name, val = line.split()
            sig[name] = int(val)
    assert "cycles 10" in out.stdout
    # ToucanSimulator::step equals setRegister / eval / readRegister (0+1+2+3 cycles)
    assert "step ok 6" in out.stdout
    syms = parse_symbols(d.symbols)["reg"]
    orc = OracleRun(d.design, d.symbols, True)
    reg = orc.st.pool(0)
    for name, frags in syms.items():
        if len(frags) < 2:
            continue
        v = 0x0123456789ABCDEF & ((1 << (4 * len(frags))) - 1)
        for k in range(len(frags) - 1, -1, -1):
            reg[frags[k][0]] = v & 15
            v >>= 4
    orc.st.set_pool(0, reg)
    signals = parse_symbols(d.symbols)["signal"]
    names = sorted(signals)
    parts = [p for n in names for p, _, _ in signals[n]]
    slots = [v for n in names for _, v, _ in signals[n]]
    orc.st.set_probes(parts, slots)
    orc.run(10)
    reg = orc.st.pool(0)
    vals = orc.st.probes()
    k = 0
    assert names and sorted(sig) == names
    for n in names:
        want = 0
        for _, _, w in signals[n]:
            want = (want << w) | (int(vals[k]) & ((1 << w) - 1))
            k += 1
        assert sig[n] == want, n
    for name, frags in syms.items():
        want = 0
        for rid, _ in frags:
            want = (want << 4) | int(reg[rid])
        assert got[name] == want, name
    text = vcd.read_text()
    assert "$enddefinitions $end" in text and "$dumpvars" in text
    # signals first, in (partition, value slot) order of their first fragment
    order = sorted(names, key=lambda n: (signals[n][0][0], signals[n][0][1]))
    decl = [l.split()[4] for l in text.splitlines() if l.startswith("$var")]
    assert decl[:len(order)] == order
    stamps = [l for l in text.splitlines() if l.startswith("#")]
    assert stamps == ["#0", "#1", "#2", "#3", "#4", "#10"]
