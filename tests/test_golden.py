"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

The design images in the fixtures were written by the reference's own
serializer and the randomized initial pools by its Init + Randomize(0), so
these checks pin the writer and Randomize without the reference present:
the generator and the hand-built programs must reproduce the fixture bytes,
and tgs_initial_state must reproduce the reference's pools. The oracle (one
thread and several) and the host interpreter of the streamed kernel's image
must reproduce the fixtures' expected pools, stop flag and print log.
"""
import glob
import os
import subprocess
import sys

import numpy as np
import pytest

import edge_programs as E
import oracle as O
from toucangpurt_amd import _native, gen

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
sys.path.insert(0, GOLDEN)
from make_golden import GENERATED  # noqa: E402

FIXTURES = sorted(glob.glob(os.path.join(GOLDEN, "*.npz")))
NAMES = [os.path.basename(f)[:-4] for f in FIXTURES]
STREAM_CHECK = os.path.join(os.path.dirname(HERE), "toucangpurt_amd", "stream_check")


def load(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def test_fixtures_present():
    assert len(FIXTURES) == len(E.all_cases()) + len(GENERATED)


@pytest.mark.parametrize("name", NAMES)
def test_writer_reproduces_reference_serializer_bytes(name):
    z = load(name)
    if name.startswith("edge_"):
        case = {f"edge_{c.name}": c for c in E.all_cases()}[name]
        mine = case.prog
    else:
        mine = gen.generate(None, **GENERATED[name][0]).design
    assert mine == z["design"].tobytes()


@pytest.mark.parametrize("name", [n for n in NAMES if not n.startswith("edge_")])
def test_randomize_reproduces_reference(name):
    z = load(name)
    design, symbols = z["design"].tobytes(), z["symbols"].tobytes()
    reg = np.zeros(z["init_reg"].size, np.uint8)
    mem = np.zeros(z["init_mem"].size, np.uint8)
    u8 = _native._u8p
    rc = _native.rt_lib().tgs_initial_state(_native.as_u8p(design), len(design), _native.as_u8p(symbols),
                                            len(symbols), 1, 0, reg.ctypes.data_as(u8), reg.size,
                                            mem.ctypes.data_as(u8), mem.size)
    assert rc == 0
    np.testing.assert_array_equal(reg, z["init_reg"])
    np.testing.assert_array_equal(mem, z["init_mem"])


@pytest.mark.parametrize("threads", [1, 3])
@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name, threads):
    z = load(name)
    st = O.State(O.Program(z["design"].tobytes()), z["init_reg"], z["init_mem"])
    if z["inputs"].size:
        st.set_stimulus(z["inputs"], int(z["seed"]))
    st.run(int(z["cycles"]), bool(z["enable_print"]), threads=threads)
    np.testing.assert_array_equal(st.pool(0), z["want_reg"])
    np.testing.assert_array_equal(st.pool(1), z["want_mem"])
    assert st.stopped == bool(z["want_stop"]) and st.stop_cycle == int(z["want_stop_cycle"])
    log = st.print_log()
    if threads > 1:
        log = sorted(log)
        want = sorted(zip(z["want_log_cycles"].tolist(), z["want_log_msgs"].tolist()))
    else:
        want = list(zip(z["want_log_cycles"].tolist(), z["want_log_msgs"].tolist()))
    assert log == want


@pytest.mark.parametrize("name", [n for n in NAMES if n.startswith("edge_")])
def test_edge_fixture_closed_forms(name):
    """The frozen expectations of the hand-built programs still meet their closed forms."""
    case = {f"edge_{c.name}": c for c in E.all_cases()}[name]
    z = load(name)
    for k, v in case.want_reg.items():
        assert z["want_reg"][k] == v, (name, k)
    for k, v in case.want_mem.items():
        assert z["want_mem"][k] == v, (name, k)


@pytest.mark.skipif(not os.path.exists(STREAM_CHECK), reason="stream_check not built")
@pytest.mark.parametrize("nc,slots", [(4, 1), (8, 2)])
@pytest.mark.parametrize("name", NAMES)
def test_stream_image_reproduces_fixture(tmp_path, name, nc, slots):
    z = load(name)
    (tmp_path / "d.bin").write_bytes(z["design"].tobytes())
    (tmp_path / "in.u32").write_bytes(z["inputs"].astype(np.uint32).tobytes())
    (tmp_path / "init.bin").write_bytes(z["init_reg"].tobytes() + z["init_mem"].tobytes())
    out = subprocess.run([STREAM_CHECK, str(tmp_path / "d.bin"), str(tmp_path / "in.u32"), str(nc), "3",
                          str(int(z["cycles"])), str(int(z["seed"])), str(tmp_path / "out.bin"),
                          str(tmp_path / "init.bin"), str(slots)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    got = np.frombuffer((tmp_path / "out.bin").read_bytes(), np.uint8)
    n = z["want_reg"].size
    np.testing.assert_array_equal(got[:n], z["want_reg"])
    np.testing.assert_array_equal(got[n:], z["want_mem"])
