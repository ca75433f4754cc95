"""Shared helpers for parity tests: run a design on the GPU and on the oracle."""
import numpy as np

import oracle as O
from toucangpurt_amd import rt
from toucangpurt_amd._native import POOL_EXCHANGE, POOL_MEM, POOL_REG


def initial_pools(design, symbols, randomize):
    """Initial register / memory pools as the runtime computes them (Init + Randomize(0))."""
    import ctypes
    from toucangpurt_amd import _native
    lib = _native.rt_lib()
    prog = O.Program(design)
    reg = np.zeros(prog.pool_size(0), np.uint8)
    mem = np.zeros(prog.pool_size(1), np.uint8)
    rc = lib.tgs_initial_state(_native.as_u8p(design), len(design), _native.as_u8p(symbols) if symbols else None,
                               len(symbols) if symbols else 0, 1 if randomize else 0, 0,
                               reg.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), reg.size,
                               mem.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), mem.size)
    assert rc == 0, rc
    return prog, reg, mem


class OracleRun:
    def __init__(self, design, symbols=None, randomize=False, inputs=None, seed=0, instance=0):
        self.prog, reg, mem = initial_pools(design, symbols, randomize)
        self.st = O.State(self.prog, reg, mem)
        if inputs is not None and len(inputs):
            self.st.set_stimulus(inputs, seed, instance)

    def run(self, cycles, enable_print=False):
        return self.st.run(cycles, enable_print)

    def pools(self):
        return self.st.pool(0), self.st.pool(1), self.st.pool(2)


def gpu_sim(design, symbols=None, randomize=False, inputs=None, seed=0, kernel=None, enable_print=False):
    sim = rt.Simulator(0)
    assert sim.init(0, design, symbols, randomize, enable_print) == 0
    if kernel is not None:
        sim.set_kernel(kernel)
    if inputs is not None and len(inputs):
        sim.set_stimulus(inputs, seed)
    return sim


def gpu_pools(sim, instance=0):
    return sim.read_pool(POOL_REG, instance), sim.read_pool(POOL_MEM, instance), sim.read_pool(POOL_EXCHANGE, instance)


def assert_pools_equal(got, want, what=""):
    names = ("regPool", "memPool", "exchangePool")
    for g, w, n in zip(got, want, names):
        if not np.array_equal(g, w):
            bad = np.nonzero(g != w)[0]
            raise AssertionError(f"{what}: {n} differs at {bad.size} nibbles, first {bad[:8].tolist()}: "
                                 f"gpu {g[bad[:8]].tolist()} oracle {w[bad[:8]].tolist()}")
