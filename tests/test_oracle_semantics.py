"""Pins the oracle: hand-built programs with closed-form answers, published
known-answer tests (xorshift64, SplitMix64, LFSR, counter), and agreement
with the netlist-level simulator on generated designs."""
import numpy as np
import pytest

import edge_programs as E
import oracle as O
from netlist_sim import Netlist, NetlistSim, splitmix64, xorshift64
from toucangpurt_amd import gen


def run_case(case):
    prog = O.Program(case.prog)
    st = O.State(prog, case.reg, case.mem)
    stop = st.run(case.cycles, enable_print=True)
    return st, stop


@pytest.mark.parametrize("case", E.all_cases(), ids=lambda c: c.name)
def test_edge_case_closed_form(case):
    st, stop = run_case(case)
    reg, mem = st.pool(0), st.pool(1)
    for k, v in case.want_reg.items():
        assert reg[k] == v, f"{case.name}: reg[{k}] = {reg[k]}, want {v}"
    for k, v in case.want_mem.items():
        assert mem[k] == v, f"{case.name}: mem[{k}] = {mem[k]}, want {v}"
    if case.want_log is not None:
        assert sorted(st.print_log()) == case.want_log
    if case.want_stop is not None:
        assert stop == case.want_stop
        assert st.stop_cycle == (0 if case.want_stop else 2 ** 64 - 1)
    assert st.violations()[0] == 0, "program read a slot it did not write"


def test_oob_memory_is_counted():
    st, _ = run_case(E.memory())
    assert st.violations()[2] == 1  # base 50 + 19 is past the pool


def test_unwritten_slot_is_detected():
    """A planted mistake: the register copy covers a slot nothing writes."""
    from progbuild import Part, Prog, lut_mp
    p = Prog(16, lut=E.TABLE)
    P = p.add(0, Part(64))
    P.regread = [(0, 20)]
    P.level(lut_mp(top=[], last=[(0, 24)]))
    P.regwrite = (4, 20, 8)  # slots 20..28: 21..23 and 25..27 never written
    st = O.State(O.Program(p.serialize()))
    st.run(1)
    assert st.violations()[0] == 6


def test_published_generators():
    # SplitMix64 first output for state 0 (Vigna's reference implementation)
    assert O.splitmix64(0) == 0xE220A8397B1DCDAF
    assert splitmix64(0) == 0xE220A8397B1DCDAF
    # xorshift64 (13, 7, 17) from state 1, worked by hand: 0x2001 -> 0x2041 -> 0x40822041
    assert O.xorshift64(1) == 0x40822041
    assert xorshift64(1) == 0x40822041
    x = 88172645463325252  # Marsaglia's xorshift example seed
    for _ in range(5):
        assert O.xorshift64(x) == xorshift64(x)
        x = xorshift64(x)


def test_stimulus_bits_follow_draw_order():
    """Input i takes bit i%64 of draw i//64 of the cycle's xorshift64 sequence."""
    from progbuild import Part, Prog, lut_mp
    p = Prog(200, lut=E.TABLE)
    P = p.add(0, Part(64))
    P.regread = [(199, 20)]
    P.level(lut_mp(top=[], last=[(0, 24)]))
    st = O.State(O.Program(p.serialize()))
    ids = np.arange(130, dtype=np.uint32)
    st.set_stimulus(ids, 7, 3)
    s = splitmix64(7 + 3)
    for _ in range(3):
        st.run(1)
        draws = []
        for _ in range(3):
            s = xorshift64(s)
            draws.append(s)
        want = [(draws[i // 64] >> (i % 64)) & 1 for i in range(130)]
        assert st.pool(0)[:130].tolist() == want


DESIGNS = [
    dict(preset="lut20k", seed=3),
    dict(n_luts=4000, depth=8, n_regs=100, n_inputs=70, n_vecops=400, n_srams=3, sram_depth=1024, sram_width=32,
         part_luts=1000, word_part_ops=40, seed=11),
    dict(n_luts=4000, depth=6, n_regs=200, n_inputs=10, n_vecops=300, n_srams=2, sram_depth=16, sram_width=64,
         word_bits_mix=100, part_luts=300, cluster_luts=100, seed=12),
]


@pytest.mark.parametrize("kw", DESIGNS, ids=["lut20k", "words_srams", "wide_small_parts"])
@pytest.mark.parametrize("randomize", [False, True])
def test_oracle_matches_netlist_simulator(kw, randomize):
    """The oracle evaluates only the serialized program; the netlist simulator only the
    logical design. Per-cycle register and memory equality pins the program semantics."""
    from simharness import initial_pools
    kw = dict(kw)
    d = gen.generate(kw.pop("preset", None), **kw)
    prog, reg, mem = initial_pools(d.design, d.symbols, randomize)
    st = O.State(prog, reg, mem)
    st.set_stimulus(d.inputs, 5)
    ns = NetlistSim(Netlist(d.netlist), reg, mem)
    ns.set_stimulus(5)
    for c in range(25):
        st.run(1)
        ns.step()
        assert np.array_equal(st.pool(0), ns.reg), f"registers differ after cycle {c}"
        assert np.array_equal(st.pool(1), ns.mem), f"memory differs after cycle {c}"
    assert st.violations() == [0, 0, 0, 0]


def test_partitioning_does_not_change_results():
    """The same logical design packed into different partition sizes gives identical
    register values by name on the oracle."""
    from toucangpurt_amd.rt import parse_symbols
    out = []
    for part_luts in (300, 1200, 5000):
        d = gen.generate(None, n_luts=6000, depth=10, n_regs=300, n_inputs=40, part_luts=part_luts,
                         cluster_luts=150, seed=4)
        st = O.State(O.Program(d.design))
        st.set_stimulus(d.inputs, 9)
        st.run(30)
        reg = st.pool(0)
        syms = parse_symbols(d.symbols)["reg"]
        out.append({name: [int(reg[i]) for i, _ in frags] for name, frags in syms.items()})
    assert out[0] == out[1] == out[2]


def test_probes_closed_form():
    """Probes hold a partition's slot values after its exec levels: the L0 reads
    of registers 0-7 into slots 20-27 and the LUT results in slots 64-71 of the
    top_operands program, whose values are known in closed form."""
    c = E.top_operands()
    st = O.State(O.Program(c.prog), c.reg, c.mem)
    st.set_probes([0] * 16, list(range(20, 28)) + list(range(64, 72)))
    assert st.probes().tolist() == [0] * 16
    st.run(1)
    v = st.probes().tolist()
    assert v[:8] == [int(c.reg[k]) for k in range(8)]
    assert v[8:] == [c.want_reg[16 + k] for k in range(8)]
    # written-slot listing: L0 destinations and the last-level results
    assert set(O.Program(c.prog).written_slots(0).tolist()) == set(range(20, 28)) | set(range(64, 72))


def test_probes_threads_agree():
    d = gen.generate(None, n_luts=4000, depth=8, n_regs=200, n_inputs=50, n_vecops=200, n_srams=2, sram_depth=512,
                     sram_width=32, part_luts=500, seed=12)
    prog = O.Program(d.design)
    rng = np.random.default_rng(1)
    parts = rng.integers(prog.num_parts(), size=200)
    slots = [int(prog.written_slots(int(p))[rng.integers(prog.written_slots(int(p)).size)]) for p in parts]
    got = []
    for threads in (1, 4):
        st = O.State(prog)
        st.set_stimulus(d.inputs, 3)
        st.set_probes(parts, slots)
        st.run(7, threads=threads)
        got.append(st.probes())
    np.testing.assert_array_equal(got[0], got[1])
    with pytest.raises(ValueError):
        O.State(prog).set_probes([prog.num_parts()], [0])
