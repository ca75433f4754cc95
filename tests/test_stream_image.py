"""The stream image (streamed kernel) interpreted on the host against the oracle.

stream_check walks every workgroup's step stream the way the streamed kernel
does (lanes, clamped row addressing, L0 rows, last-level lists) with every
index bounds-checked, so a layout mistake in the image builder fails here on
the CPU before any kernel runs. The interpreter starts from Init() state
(no Randomize) and drives the same xorshift64 stimulus as the oracle.
"""
import os
import subprocess

import numpy as np
import pytest

from simharness import OracleRun
from toucangpurt_amd import gen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "toucangpurt_amd", "stream_check")

DESIGNS = {
    "lut": dict(n_luts=3000, depth=10, n_regs=200, n_inputs=40, part_luts=500, seed=2),
    "mixed": dict(n_luts=6000, depth=8, n_regs=300, n_inputs=70, n_vecops=600, n_srams=3, sram_depth=1024,
                  sram_width=32, part_luts=1200, word_part_ops=48, seed=5),
    "wide": dict(n_luts=8000, depth=12, n_regs=500, n_inputs=130, n_vecops=400, word_bits_mix=100, n_srams=2,
                 sram_depth=4096, sram_width=64, part_luts=700, cluster_luts=100, seed=9),
    "prints": dict(n_luts=2000, depth=6, n_regs=100, n_inputs=20, n_vecops=50, n_prints=6, n_stops=2,
                   stop_rarity=3, part_luts=400, seed=13),
}
_cache = {}


def design(name):
    if name not in _cache:
        _cache[name] = gen.generate(None, **DESIGNS[name])
    return _cache[name]


@pytest.mark.skipif(not os.path.exists(EXE), reason="stream_check not built")
@pytest.mark.parametrize("dname", list(DESIGNS))
@pytest.mark.parametrize("nc,groups,slots", [(2, 1, 1), (4, 3, 1), (4, 1000, 1), (2, 7, 1), (8, 3, 2), (8, 1000, 2),
                                             (4, 2, 2), (8, 5, 4), (8, 3, -2), (8, 2, -4)])
def test_stream_image_matches_oracle(tmp_path, dname, nc, groups, slots):
    """slots < 0: the builder picks one or -slots per region, as the runtime does."""
    d = design(dname)
    (tmp_path / "d.bin").write_bytes(d.design)
    (tmp_path / "in.u32").write_bytes(np.asarray(d.inputs, np.uint32).tobytes())
    cycles, seed = 6, 29
    out = subprocess.run([EXE, str(tmp_path / "d.bin"), str(tmp_path / "in.u32"), str(nc), str(groups), str(cycles),
                          str(seed), str(tmp_path / "out.bin"), "-", str(slots)], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr
    got = np.frombuffer((tmp_path / "out.bin").read_bytes(), np.uint8)
    orc = OracleRun(d.design, inputs=d.inputs, seed=seed)
    orc.run(cycles)
    reg, mem, _ = orc.pools()
    assert got.size == reg.size + mem.size
    np.testing.assert_array_equal(got[:reg.size], reg)
    np.testing.assert_array_equal(got[reg.size:], mem)


@pytest.mark.skipif(not os.path.exists(EXE), reason="stream_check not built")
@pytest.mark.parametrize("nc,slots", [(2, 1), (4, 1), (8, 2)])
def test_stream_image_edge_programs(tmp_path, nc, slots):
    import edge_programs as E
    import oracle as O
    for case in E.all_cases():
        (tmp_path / "d.bin").write_bytes(case.prog)
        (tmp_path / "in.u32").write_bytes(b"")
        (tmp_path / "init.bin").write_bytes(case.reg.tobytes() + case.mem.tobytes())
        out = subprocess.run([EXE, str(tmp_path / "d.bin"), str(tmp_path / "in.u32"), str(nc), "2",
                              str(case.cycles), "0", str(tmp_path / "out.bin"), str(tmp_path / "init.bin"), str(slots)],
                             capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, (case.name, out.stderr)
        got = np.frombuffer((tmp_path / "out.bin").read_bytes(), np.uint8)
        st = O.State(O.Program(case.prog), case.reg, case.mem)
        st.run(case.cycles)
        np.testing.assert_array_equal(got[:case.reg.size], st.pool(0), err_msg=case.name)
        np.testing.assert_array_equal(got[case.reg.size:], st.pool(1), err_msg=case.name)
