"""The C-ABI library: loads without a GPU, exports every entry point that
include/toucan_rt.h declares, and turns format / invariant violations into
error codes (the reference asserts or exits, src/SimEval.cu:92-97,1183,
1292-1311; src/ToucanSim.cu:153)."""
import ctypes
import os
import re

import pytest

import edge_programs as E
from progbuild import Part, Prog, lut_mp, mid, vecop_mp
from toucangpurt_amd import _native
from toucangpurt_amd.rt import SimError, program_stats

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"\b(tgs_\w+)\s*\(", text)))


def test_runtime_exports_every_declared_symbol():
    lib = _native.rt_lib()
    names = declared("toucan_rt.h")
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), f"libtoucan_rt.so does not export {n}"
    assert sorted(_native.RT_SYMBOLS) == names


def test_generator_exports_every_declared_symbol():
    lib = _native.gen_lib()
    for n in declared("toucan_gen.h"):
        assert hasattr(lib, n), f"libtoucan_gen.so does not export {n}"


def test_create_without_gpu_fails_cleanly():
    lib = _native.rt_lib()
    h = ctypes.c_void_p()
    rc = lib.tgs_create(0, ctypes.byref(h))
    if rc == 0:
        lib.tgs_destroy(h)
        pytest.skip("a GPU is present")
    assert rc == _native.TGS_E_NODEV and not h.value


def _base():
    p = Prog(32, lut=E.TABLE)
    P = p.add(0, Part(64))
    P.regread = [(0, 20)]
    P.level(lut_mp(top=[(0, 0, 0, 20)], last=[(32, 24)]))
    return p, P


def _code(blob):
    try:
        program_stats(blob)
    except SimError as e:
        return e.code
    return 0


def test_valid_base_program():
    p, _ = _base()
    assert _code(p.serialize()) == 0


def _mutate(fn):
    p, P = _base()
    fn(p, P)
    return _code(p.serialize())


INVALID = {
    "both_l0_kinds": lambda p, P: setattr(P, "exgread", [(0, 21)]),
    "no_l0": lambda p, P: setattr(P, "regread", []),
    "slot_out_of_range": lambda p, P: setattr(P, "regread", [(0, 64)]),
    "reg_out_of_range": lambda p, P: setattr(P, "regread", [(32, 20)]),
    "top_33_lanes": lambda p, P: P.level(lut_mp(top=[(0, 0, 0, 1)] * 33, last=[(32, 30)])),
    "mid_33_lanes": lambda p, P: P.level(lut_mp(top=[(0, 0, 0, 1)], mids=[[mid(0, 0, 0, 32)] * 33], last=[(32, 30)])),
    "159_mid_levels": lambda p, P: P.level(lut_mp(top=[(0, 0, 0, 1)], mids=[[mid(0, 0, 0, 32)]] * 159,
                                                  last=[(32, 30)])),
    "vecop_len_33": lambda p, P: P.level(vecop_mp([(20, 20, 33, 0, 0, 40)])),
    "regwrite_past_pool": lambda p, P: setattr(P, "regwrite", (28, 20, 8)),
    "same_slot_written_twice_in_level": lambda p, P: P.level(lut_mp(top=[], last=[(0, 30)]),
                                                             lut_mp(top=[], last=[(1, 30)])),
    "read_write_same_level": lambda p, P: P.level(lut_mp(top=[], last=[(0, 30)]),
                                                  lut_mp(top=[(0, 0, 0, 30)], last=[(32, 31)])),
    "vecop_lanes_overlap": lambda p, P: P.level(vecop_mp([(20, 20, 1, 0, 0, 40), (40, 40, 1, 0, 0, 41)])),
    "print_msg_missing": lambda p, P: setattr(P, "prints", [(1, 0)]),
}


@pytest.mark.parametrize("name", list(INVALID))
def test_invalid_programs_are_rejected(name):
    assert _code_mut(name) == _native.TGS_E_INVALID


def _code_mut(name):
    return _mutate(INVALID[name])


def test_value_pool_over_limit():
    p = Prog(32, lut=E.TABLE)
    P = p.add(0, Part(70000))
    P.regread = [(0, 20)]
    assert _code(p.serialize()) == _native.TGS_E_LIMIT


def test_three_regions_rejected():
    p, _ = _base()
    blob = bytearray(p.serialize())
    # region count sits after the partitions: rewrite "2 regions" as 3 by appending an empty one
    import struct
    tail = struct.pack("<Q", 2) + struct.pack("<Q", 1) + struct.pack("<I", 0) + struct.pack("<Q", 0) + struct.pack("<Q", 0)
    assert blob.endswith(tail)
    blob = blob[:-len(tail)] + struct.pack("<Q", 3) + struct.pack("<Q", 1) + struct.pack("<I", 0) + \
        struct.pack("<Q", 0) + struct.pack("<Q", 0) + struct.pack("<Q", 0)
    assert _code(bytes(blob)) == _native.TGS_E_INVALID


def test_wrong_lut_size_rejected():
    p, _ = _base()
    p.lut = bytes(100)
    assert _code(p.serialize()) == _native.TGS_E_INVALID
