"""Write the golden fixtures under tests/golden/ (run in a checkout where the
reference sources are present, after `make -C oracle ref`).

The reference ships no tests, golden vectors or known-answer files, so the
fixtures are made here from two sources:

* bytes and state produced by the reference's own host code, compiled from
  its sources into oracle/_ref/libref.so: every design image is the output of
  the reference's serializeSimDesignInfo after its deserializeSimDesignInfo
  (src/Serializer.cpp:199-259), and the randomized initial pools are its
  SimDesignInfo::Init + Randomize(0) (src/SimInit.cpp:16-64), as
  ToucanSimulator::init does with shouldRandomize (src/ToucanSim.cu:128-134);
* expected pools after a number of cycles from the C oracle, which is pinned
  separately (closed forms, known-answer designs, an independent netlist
  simulator); the fixtures freeze those outputs so the GPU box, where neither
  the reference nor a compiler for it exists, checks against fixed data.

Each fixture is one .npz of plain integer arrays (loadable with
allow_pickle=False). Usage: python tests/golden/make_golden.py
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    sys.path.insert(0, p)

import edge_programs as E  # noqa: E402
import oracle as O  # noqa: E402
from toucangpurt_amd import _native, gen  # noqa: E402
from toucangpurt_amd.rt import program_stats  # noqa: E402

_u8p = ctypes.POINTER(ctypes.c_uint8)

# generated designs: generator parameters, cycles, stimulus seed, print enable
GENERATED = {
    "gen_mixed": (dict(n_luts=2500, depth=8, n_regs=120, n_inputs=40, n_vecops=150, n_srams=2, sram_depth=256,
                       sram_width=32, n_prints=4, n_stops=1, part_luts=500, seed=28), 12, 5),
    "gen_wide": (dict(n_luts=3000, depth=10, n_regs=200, n_inputs=70, n_vecops=200, word_bits_mix=100, n_srams=1,
                      sram_depth=1024, sram_width=64, part_luts=600, cluster_luts=80, seed=22), 10, 9),
}


def ref_lib():
    path = os.path.join(ROOT, "oracle", "_ref", "libref.so")
    if not os.path.exists(path):
        sys.exit("oracle/_ref/libref.so is not built: make -C oracle ref (needs the reference sources)")
    L = ctypes.CDLL(path)
    L.ref_roundtrip_design.argtypes = [_u8p, ctypes.c_size_t, ctypes.POINTER(_u8p), ctypes.POINTER(ctypes.c_size_t)]
    L.ref_randomize.argtypes = [_u8p, ctypes.c_size_t, _u8p, ctypes.c_size_t, ctypes.c_uint32, _u8p, ctypes.c_size_t,
                                _u8p, ctypes.c_size_t]
    L.ref_free.argtypes = [_u8p]
    return L


def ref_serialized(L, blob):
    out = _u8p()
    n = ctypes.c_size_t()
    if L.ref_roundtrip_design(_native.as_u8p(blob), len(blob), ctypes.byref(out), ctypes.byref(n)) != 0:
        raise RuntimeError("reference deserializer rejected a design")
    data = ctypes.string_at(out, n.value)
    L.ref_free(out)
    return data


def ref_randomized(L, design, symbols):
    st = program_stats(design)
    reg = np.zeros(st["reg_pool"], np.uint8)
    mem = np.zeros(st["mem_pool"], np.uint8)
    rc = L.ref_randomize(_native.as_u8p(design), len(design), _native.as_u8p(symbols), len(symbols), 0,
                         reg.ctypes.data_as(_u8p), reg.size, mem.ctypes.data_as(_u8p), mem.size)
    if rc != 0:
        raise RuntimeError(f"reference Randomize failed ({rc})")
    return reg, mem


def expected(design, reg, mem, cycles, inputs, seed, enable_print):
    st = O.State(O.Program(design), reg, mem)
    if len(inputs):
        st.set_stimulus(inputs, seed)
    st.run(cycles, enable_print)
    log = st.print_log()
    return dict(want_reg=st.pool(0), want_mem=st.pool(1), want_stop=np.uint8(st.stopped),
                want_stop_cycle=np.uint64(st.stop_cycle),
                want_log_cycles=np.array([c for c, _ in log], np.uint64),
                want_log_msgs=np.array([m for _, m in log], np.uint32))


def main():
    L = ref_lib()
    made = []
    for case in E.all_cases():
        prog = ref_serialized(L, case.prog)
        out = dict(design=np.frombuffer(prog, np.uint8), symbols=np.zeros(0, np.uint8), randomized=np.uint8(0),
                   init_reg=case.reg, init_mem=case.mem, cycles=np.uint32(case.cycles), inputs=np.zeros(0, np.uint32),
                   seed=np.uint64(0), enable_print=np.uint8(1))
        out.update(expected(prog, case.reg, case.mem, case.cycles, [], 0, True))
        np.savez_compressed(os.path.join(HERE, f"edge_{case.name}.npz"), **out)
        made.append(f"edge_{case.name}")
    for name, (kw, cycles, seed) in GENERATED.items():
        d = gen.generate(None, **kw)
        prog = ref_serialized(L, d.design)
        reg, mem = ref_randomized(L, prog, d.symbols)
        out = dict(design=np.frombuffer(prog, np.uint8), symbols=np.frombuffer(d.symbols, np.uint8),
                   randomized=np.uint8(1), init_reg=reg, init_mem=mem, cycles=np.uint32(cycles),
                   inputs=np.asarray(d.inputs, np.uint32), seed=np.uint64(seed), enable_print=np.uint8(1),
                   gen_seed=np.uint64(kw["seed"]))
        out.update(expected(prog, reg, mem, cycles, d.inputs, seed, True))
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        made.append(name)
    print("wrote", ", ".join(made))


if __name__ == "__main__":
    main()
