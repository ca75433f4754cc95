"""The batched image's LUT tasks checked on the host.

batch_check builds the batched image of a design twice for one instance
count per wave, with every LUT pair as a bounds-checked T_LUTROWS task and
with T_LUTFAST tasks where the builder allows them, runs both on a model of
one wave's LDS rows the way the kernel's lane rounds do, from the same random
nibble state (the design's LUT table, then random nibble tables), and requires
equal value pools after every task; every T_LUTFAST offset is bounds-checked
and every operand it reads must be a nibble. T_LUTROWS itself is checked
against the oracle by the GPU tests, so a layout mistake in T_LUTFAST fails
here on the CPU first.
"""
import os
import subprocess

import pytest

from toucangpurt_amd import gen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "toucangpurt_amd", "batch_check")

DESIGNS = {
    "lut": dict(n_luts=3000, depth=10, n_regs=200, n_inputs=40, part_luts=500, seed=2),
    "mixed": dict(n_luts=6000, depth=8, n_regs=300, n_inputs=70, n_vecops=600, n_srams=3, sram_depth=1024,
                  sram_width=32, part_luts=600, word_part_ops=48, seed=5),
    "wide": dict(n_luts=8000, depth=12, n_regs=500, n_inputs=130, n_vecops=400, word_bits_mix=100, n_srams=2,
                 sram_depth=4096, sram_width=64, part_luts=700, cluster_luts=100, seed=9),
    "deep": dict(n_luts=4000, depth=40, n_regs=100, n_inputs=20, part_luts=256, seed=17),
}
_cache = {}


def design_file(tmp_path, name):
    if name not in _cache:
        _cache[name] = gen.generate(None, **DESIGNS[name]).design
    p = tmp_path / f"{name}.bin"
    p.write_bytes(_cache[name])
    return p


@pytest.mark.skipif(not os.path.exists(EXE), reason="batch_check not built")
@pytest.mark.parametrize("dname", list(DESIGNS))
@pytest.mark.parametrize("inst", [8, 16, 32, 64])
def test_fast_lut_tasks_equal_bounds_checked(tmp_path, dname, inst):
    p = design_file(tmp_path, dname)
    out = subprocess.run([EXE, str(p), str(inst), "7", "2"], capture_output=True, text=True, timeout=300)
    if "do not fit 16-bit offsets" in out.stderr:
        pytest.skip(f"value pool x {inst} instances exceeds 16-bit offsets (the runtime keeps T_LUTROWS)")
    assert out.returncode == 0, out.stderr
    fast, rows = map(int, out.stdout.split()[1:3])
    assert fast > 0


@pytest.mark.skipif(not os.path.exists(EXE), reason="batch_check not built")
@pytest.mark.parametrize("dname", list(DESIGNS))
@pytest.mark.parametrize("inst", [8, 16, 32, 64])
def test_fast_lut_tasks_on_nibble_rows(tmp_path, dname, inst):
    """The same check with the fast image laid out for nibble rows (two
    instances per byte, offsets row * I / 2), compared per instance with the
    byte rows of the bounds-checked tasks; constant rows 16..31 are not
    representable in a nibble and hold random nibbles, so a task that read
    them would differ."""
    p = design_file(tmp_path, dname)
    out = subprocess.run([EXE, str(p), str(inst), "11", "2", "nib"], capture_output=True, text=True, timeout=300)
    if "do not fit 16-bit offsets" in out.stderr:
        pytest.skip(f"value pool x {inst // 2} bytes per row exceeds 16-bit offsets")
    if "nibble rows need every LUT task fast" in out.stderr:
        pytest.skip("some LUT pair needs the bounds-checked task (the runtime keeps byte rows)")
    assert out.returncode == 0, out.stderr
    fast, rows = map(int, out.stdout.split()[1:3])
    assert fast > 0 and rows == 0
