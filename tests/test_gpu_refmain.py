"""The reference's own caller on the GPU: main.cpp (main.cpp:33-68), unchanged,
compiled against include/ and linked to libtoucan_sim.so (oracle/_ref/ref_main,
built in the build container where the reference sources lie), loads
GPUSimDesign.bin / GPUSimIOSymbols.bin from its working directory and free
runs. The ENABLE_SIM_PROFILE build (oracle/_ref/ref_main_profile) prints the
reference's per-partition profile lines (src/ToucanSim.cu:306-325). Also the
class's error paths after a real load (tests/cpp/api_errors.cpp).

GPU work runs in child processes, so these tests are collected first."""
import os
import re
import subprocess

import pytest

from toucangpurt_amd import gen

pytestmark = [pytest.mark.gpu, pytest.mark.subprocess_gpu]

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")


def _design_dir(tmp_path):
    d = gen.generate("lut20k", seed=3)
    (tmp_path / "GPUSimDesign.bin").write_bytes(d.design)
    (tmp_path / "GPUSimIOSymbols.bin").write_bytes(d.symbols)
    return d


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "ref_main")), reason="oracle/_ref/ref_main not built")
def test_reference_main_runs_on_the_gpu(tmp_path):
    _design_dir(tmp_path)
    print("case reference main.cpp, lut20k, 2000 cycles", flush=True)
    r = subprocess.run([os.path.join(REF, "ref_main"), "2000"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    assert lines[:2] == ["Loading design", "Loading done"], r.stdout
    assert re.fullmatch(r"Free running for 2000 cycles in \d+ ms", lines[2]), r.stdout
    assert "failed" not in r.stderr and "Error" not in r.stderr, r.stderr


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "ref_main_profile")),
                    reason="oracle/_ref/ref_main_profile not built")
def test_reference_main_profile_lines(tmp_path):
    d = _design_dir(tmp_path)
    n_parts = sum(d.stats["n_parts"])
    print("case reference main.cpp with ENABLE_SIM_PROFILE, lut20k, 1000 cycles", flush=True)
    r = subprocess.run([os.path.join(REF, "ref_main_profile")], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    parts = {}
    for line in r.stdout.splitlines():
        m = re.fullmatch(r"Part (\d+), useful ([0-9.e+-]+)%, ticks (\d+), total ticks (\d+)", line)
        if m:
            parts[int(m.group(1))] = (float(m.group(2)), int(m.group(3)), int(m.group(4)))
    assert sorted(parts) == list(range(n_parts)), r.stdout
    for p, (ratio, useful, total) in parts.items():
        assert 0 < useful <= total, (p, useful, total)
        assert abs(ratio - 100.0 * useful / total) < 0.01, (p, ratio)


def test_class_error_paths_after_a_real_load(tmp_path):
    exe = str(tmp_path / "api_errors")
    cmd = ["g++", "-std=c++20", "-O1", "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "api_errors.cpp"), "-o", exe,
           "-L" + os.path.join(ROOT, "toucangpurt_amd"), "-ltoucan_sim", "-ltoucan_rt",
           "-Wl,-rpath," + os.path.join(ROOT, "toucangpurt_amd")]
    c = subprocess.run(cmd, capture_output=True, text=True)
    assert c.returncode == 0, c.stderr[-2000:]
    d = _design_dir(tmp_path)
    print("case api_errors on the GPU", flush=True)
    r = subprocess.run([exe, str(tmp_path / "GPUSimDesign.bin"), str(tmp_path / "GPUSimIOSymbols.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.splitlines()[-2:] == ["loaded", "ok"], r.stdout
