"""ctypes binding of the C oracle (liboracle.so). TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg use
this module, always as the checker / CPU baseline, never as the product.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        L.orc_prog_load.argtypes = [_u8p, ctypes.c_size_t, ctypes.POINTER(vp), ctypes.c_char_p, ctypes.c_size_t]
        L.orc_prog_free.argtypes = [vp]
        L.orc_prog_free.restype = None
        L.orc_pool_size.argtypes = [vp, ctypes.c_int]
        L.orc_pool_size.restype = ctypes.c_uint32
        L.orc_num_parts.argtypes = [vp, ctypes.c_int]
        L.orc_num_parts.restype = ctypes.c_uint32
        L.orc_state_new.argtypes = [vp, ctypes.POINTER(vp)]
        L.orc_state_free.argtypes = [vp]
        L.orc_state_free.restype = None
        L.orc_state_copy_pool.argtypes = [vp, ctypes.c_int, ctypes.c_int, _u8p, ctypes.c_size_t]
        L.orc_set_stimulus.argtypes = [vp, _u32p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32]
        L.orc_run.argtypes = [vp, ctypes.c_uint32, ctypes.c_int]
        L.orc_run_threads.argtypes = [vp, ctypes.c_uint32, ctypes.c_int, ctypes.c_int]
        L.orc_cycles.argtypes = [vp]
        L.orc_cycles.restype = ctypes.c_uint64
        L.orc_stopped.argtypes = [vp]
        L.orc_stop_cycle.argtypes = [vp]
        L.orc_stop_cycle.restype = ctypes.c_uint64
        L.orc_violations.argtypes = [vp, _u64p]
        L.orc_violations.restype = None
        L.orc_print_log.argtypes = [vp, _u64p, _u32p, ctypes.c_uint64]
        L.orc_print_log.restype = ctypes.c_uint64
        L.orc_set_probes.argtypes = [vp, _u32p, _u32p, ctypes.c_uint32]
        L.orc_read_probes.argtypes = [vp, _u8p]
        L.orc_read_probes.restype = None
        L.orc_written_slots.argtypes = [vp, ctypes.c_uint32, _u32p, ctypes.c_uint32]
        L.orc_written_slots.restype = ctypes.c_uint32
        L.orc_xorshift64.argtypes = [ctypes.c_uint64]
        L.orc_xorshift64.restype = ctypes.c_uint64
        L.orc_splitmix64.argtypes = [ctypes.c_uint64]
        L.orc_splitmix64.restype = ctypes.c_uint64
        _lib = L
    return _lib


class Program:
    def __init__(self, design: bytes):
        L = lib()
        self._buf = design
        self.h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(256)
        rc = L.orc_prog_load(ctypes.cast(ctypes.c_char_p(design), _u8p), len(design), ctypes.byref(self.h),
                             err, len(err))
        if rc != 0:
            raise ValueError("oracle rejected program: " + err.value.decode())

    def pool_size(self, which):
        return lib().orc_pool_size(self.h, which)

    def num_parts(self):
        return lib().orc_num_parts(self.h, 0) + lib().orc_num_parts(self.h, 1)

    def written_slots(self, part):
        n = lib().orc_written_slots(self.h, part, None, 0)
        out = np.zeros(max(n, 1), np.uint32)
        lib().orc_written_slots(self.h, part, out.ctypes.data_as(_u32p), n)
        return out[:n]

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_prog_free(self.h)
            self.h = None


class State:
    """One simulated instance: register, memory and exchange pools."""

    def __init__(self, prog: Program, reg=None, mem=None):
        L = lib()
        self.prog = prog
        self.h = ctypes.c_void_p()
        L.orc_state_new(prog.h, ctypes.byref(self.h))
        if reg is not None:
            self.set_pool(0, reg)
        if mem is not None:
            self.set_pool(1, mem)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_state_free(self.h)
            self.h = None

    def set_pool(self, which, data):
        a = np.ascontiguousarray(np.frombuffer(bytes(data), dtype=np.uint8) if isinstance(data, (bytes, bytearray))
                                 else data, dtype=np.uint8)
        rc = lib().orc_state_copy_pool(self.h, which, 1, a.ctypes.data_as(_u8p), a.size)
        if rc != 0:
            raise ValueError("pool size mismatch")

    def pool(self, which):
        n = self.prog.pool_size(which)
        a = np.zeros(n, dtype=np.uint8)
        lib().orc_state_copy_pool(self.h, which, 0, a.ctypes.data_as(_u8p), n)
        return a

    def set_stimulus(self, ids, seed, instance=0):
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        rc = lib().orc_set_stimulus(self.h, ids.ctypes.data_as(_u32p), ids.size, seed, instance)
        if rc != 0:
            raise ValueError("stimulus input id out of range")

    def run(self, cycles, enable_print=False, threads=1):
        if threads > 1:
            return bool(lib().orc_run_threads(self.h, cycles, 1 if enable_print else 0, threads))
        return bool(lib().orc_run(self.h, cycles, 1 if enable_print else 0))

    def set_probes(self, parts, slots):
        self._probe_parts = np.ascontiguousarray(parts, dtype=np.uint32)
        self._probe_slots = np.ascontiguousarray(slots, dtype=np.uint32)
        rc = lib().orc_set_probes(self.h, self._probe_parts.ctypes.data_as(_u32p),
                                  self._probe_slots.ctypes.data_as(_u32p), self._probe_parts.size)
        if rc != 0:
            raise ValueError("probe partition or slot out of range")

    def probes(self):
        n = getattr(self, "_probe_parts", np.zeros(0)).size
        out = np.zeros(max(n, 1), np.uint8)
        lib().orc_read_probes(self.h, out.ctypes.data_as(_u8p))
        return out[:n]

    @property
    def cycles(self):
        return lib().orc_cycles(self.h)

    @property
    def stopped(self):
        return bool(lib().orc_stopped(self.h))

    @property
    def stop_cycle(self):
        return lib().orc_stop_cycle(self.h)

    def violations(self):
        v = (ctypes.c_uint64 * 4)()
        lib().orc_violations(self.h, v)
        return list(v)

    def print_log(self):
        n = lib().orc_print_log(self.h, None, None, 0)
        c = np.zeros(max(n, 1), dtype=np.uint64)
        m = np.zeros(max(n, 1), dtype=np.uint32)
        lib().orc_print_log(self.h, c.ctypes.data_as(_u64p), m.ctypes.data_as(_u32p), n)
        return list(zip(c[:n].tolist(), m[:n].tolist()))


def xorshift64(x):
    return lib().orc_xorshift64(x)


def splitmix64(x):
    return lib().orc_splitmix64(x)
