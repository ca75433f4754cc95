"""Python mirror of toucanGPUSim::ToucanSimulator over the C-ABI (libtoucan_rt.so).

Method names and meanings follow include/ToucanGPUSim.h:16-34 of the
reference (init / eval / eval_free_running / setRegister / readRegister /
readSignal / cleanup), so parity tests read like the reference's own usage
in main.cpp:43-60. Errors the reference reports by return code keep that
form (``init`` returns -1 on a missing file or an LDS limit); HIP and format
errors raise ``SimError`` instead of exiting the process.
"""
import ctypes
import os

import numpy as np

from . import _native
from ._native import (KERNEL_BATCH, KERNEL_PLAIN, KERNEL_WAVE64, POOL_EXCHANGE, POOL_MEM, POOL_REG,
                      TGS_E_IO, TGS_E_LIMIT, TGS_OK)

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)


class SimError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def program_stats(design: bytes) -> dict:
    """Record counts and the per-cycle byte model of a serialized design (no GPU needed)."""
    lib = _native.rt_lib()
    st = _native.ProgramStats()
    err = ctypes.create_string_buffer(512)
    rc = lib.tgs_program_stats_from_bytes(_native.as_u8p(design), len(design), ctypes.byref(st), err, len(err))
    if rc != TGS_OK:
        raise SimError(rc, err.value.decode())
    return {n: (list(getattr(st, n)) if n == "n_parts" else getattr(st, n)) for n, _ in _native.ProgramStats._fields_}


class Simulator:
    """One design resident on one GPU, with any number of stimulus instances."""

    def __init__(self, gpu_id=0, stream=None):
        """stream: a hipStream_t of gpu_id (as an int or c_void_p) to run on;
        the caller keeps it alive and destroys it (tgs_create_on_stream)."""
        self._lib = _native.rt_lib()
        self._h = ctypes.c_void_p()
        if stream is not None:
            rc = self._lib.tgs_create_on_stream(gpu_id, ctypes.c_void_p(int(getattr(stream, "value", stream) or 0)),
                                                ctypes.byref(self._h))
        else:
            rc = self._lib.tgs_create(gpu_id, ctypes.byref(self._h))
        if rc != TGS_OK:
            raise SimError(rc, "no usable GPU (cooperative launch required)")
        self.symbols = None
        self.cycle_cnt = 0
        self._n_inst = 1
        self._reset_probes()
        self._design = None
        self._symbols_blob = None
        self._vcd = None

    # -- helpers -------------------------------------------------------------
    def _reset_probes(self):
        self._probe_first = {}
        self._probe_parts, self._probe_slots = [], []
        self._probe_epoch = None
        self._n_probes = 0

    def _err(self):
        buf = ctypes.create_string_buffer(1024)
        self._lib.tgs_last_error(self._h, buf, len(buf))
        return buf.value.decode(errors="replace")

    def _check(self, rc):
        if rc != TGS_OK:
            raise SimError(rc, self._err())

    # -- reference entry points ---------------------------------------------
    def init(self, gpu_id, designBin, symbolBin, shouldRandomize=False, enablePrint=False):
        """ToucanSimulator::init (src/ToucanSim.cu:102-204). Paths or bytes; returns 0 / -1."""
        design = designBin
        symbols = symbolBin
        try:
            if isinstance(design, (str, os.PathLike)):
                with open(design, "rb") as f:
                    design = f.read()
            if isinstance(symbols, (str, os.PathLike)):
                with open(symbols, "rb") as f:
                    symbols = f.read()
        except OSError:
            return -1
        rc = self._lib.tgs_load(self._h, _native.as_u8p(design), len(design),
                                _native.as_u8p(symbols) if symbols else None, len(symbols) if symbols else 0,
                                1 if shouldRandomize else 0, 1 if enablePrint else 0)
        if rc in (TGS_E_IO, TGS_E_LIMIT):
            return -1
        self._check(rc)
        self._design = design
        self._symbols_blob = symbols
        self.symbols = parse_symbols(symbols) if symbols else None
        self.cycle_cnt = 0
        self._n_inst = 1
        self._reset_probes()
        return 0

    def eval(self):
        """One cycle (src/ToucanSim.cu:206-230); returns the sticky stop flag."""
        st = ctypes.c_int()
        self._check(self._lib.tgs_eval(self._h, ctypes.byref(st)))
        if self._vcd is not None:
            self._vcd.dump(self.cycle_cnt)
        self.cycle_cnt += 1
        return bool(st.value)

    def eval_free_running(self, max_cycles):
        """N cycles in one launch (src/ToucanSim.cu:232-267); returns the stop flag."""
        st = ctypes.c_int()
        self._check(self._lib.tgs_eval_free_running(self._h, int(max_cycles), ctypes.byref(st)))
        self.cycle_cnt += int(max_cycles)
        if self._vcd is not None:
            self._vcd.dump(self.cycle_cnt)
        return bool(st.value)

    def setRegister(self, name, value, instance=0):
        """write_reg_to_gpu (src/SimEval.cu:1072-1081): 4 bits per fragment, last fragment first."""
        locs = self.symbols["reg"][name]
        if len(locs) > 16:
            raise ValueError("register wider than 64 bits")
        ids = np.array([l[0] for l in locs], dtype=np.uint32)
        vals = np.zeros(len(locs), dtype=np.uint8)
        v = int(value)
        for k in range(len(locs) - 1, -1, -1):
            vals[k] = v & 0xF
            v >>= 4
        if v != 0:
            raise ValueError("Given value is wider than register")
        self._check(self._lib.tgs_write_regs(self._h, instance, ids.ctypes.data_as(_u32p), len(ids),
                                             vals.ctypes.data_as(_u8p)))

    def readRegister(self, name, instance=0):
        """read_reg_from_gpu (src/SimEval.cu:1058-1069): r = r << 4 | nibble, in list order."""
        locs = self.symbols["reg"][name]
        if len(locs) > 16:
            raise ValueError("register wider than 64 bits")
        ids = np.array([l[0] for l in locs], dtype=np.uint32)
        out = np.zeros(len(locs), dtype=np.uint8)
        self._check(self._lib.tgs_read_regs(self._h, instance, ids.ctypes.data_as(_u32p), len(ids),
                                            out.ctypes.data_as(_u8p)))
        r = 0
        for x in out:
            r = (r << 4) | int(x)
        return r

    def readSignal(self, name, instance=0):
        """Value a signal (SimDebugInfo signal locations) had in the last cycle run.

        The reference's readSignal is a stub returning UINT64_MAX
        (src/ToucanSim.cu:291-300). Here a signal is observed through signal
        probes from the first call that names it (or enableVCDDump) on;
        until a cycle has run after that, UINT64_MAX as in the reference.
        Fragments are concatenated most significant first, `width` bits each
        (the VCD writer's order, unsupport/SimVCD.cpp:222-233)."""
        locs = self.symbols["signal"][name]
        if sum(w for _, _, w in locs) > 64:
            raise ValueError("signal wider than 64 bits")
        if name not in self._probe_first:
            self.probe_signals([name])
        if self._probe_epoch == self.cycle_cnt:
            return (1 << 64) - 1
        vals = self.read_probes(instance)
        r, k = 0, self._probe_first[name]
        for _, _, w in locs:
            r = (r << w) | (int(vals[k]) & ((1 << w) - 1))
            k += 1
        return r

    def probe_signals(self, names):
        """Add signals to the probe list (one probe per fragment)."""
        added = False
        for n in names:
            if n in self._probe_first:
                continue
            self._probe_first[n] = len(self._probe_parts)
            for part, slot, _ in self.symbols["signal"][n]:
                self._probe_parts.append(part)
                self._probe_slots.append(slot)
            added = True
        if added:
            self.set_probes(self._probe_parts, self._probe_slots)
            self._probe_epoch = self.cycle_cnt
        return added

    def set_probes(self, parts, slots):
        """Probe value-pool slots: (partition, slot) pairs, values read back in this order."""
        p = np.ascontiguousarray(parts, dtype=np.uint32)
        s = np.ascontiguousarray(slots, dtype=np.uint32)
        if p.size != s.size:
            raise ValueError("one slot per partition")
        self._check(self._lib.tgs_set_probes(self._h, p.ctypes.data_as(_u32p), s.ctypes.data_as(_u32p), p.size))
        self._n_probes = int(p.size)

    def read_probes(self, instance=0):
        out = np.zeros(max(self._n_probes, 1), np.uint8)
        self._check(self._lib.tgs_read_probes(self._h, int(instance), out.ctypes.data_as(_u8p), self._n_probes))
        return out[:self._n_probes]

    def enableVCDDump(self, path):
        """VCD writer (unsupport/SimVCD.cpp:99-241) of the symbol table's signals
        (through signal probes) followed by its registers, over asynchronous
        snapshots: one per eval() (stamped with the cycle before it) and one
        after each eval_free_running() (stamped with the cycle count after it),
        as the C++ class."""
        from .vcd import VcdWriter
        self._vcd = VcdWriter(self, path)

    def cleanup(self):
        if self._vcd is not None:
            self._vcd.close()
            self._vcd = None
        if getattr(self, "_h", None):
            self._lib.tgs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.cleanup()
        except Exception:
            pass

    # -- extensions over the reference -------------------------------------
    def set_kernel(self, kernel):
        self._check(self._lib.tgs_set_kernel(self._h, kernel))

    def set_launch(self, kernel, block=0, groups=0):
        """Threads per workgroup and workgroup count of a kernel family (0 = automatic)."""
        self._check(self._lib.tgs_set_launch(self._h, kernel, int(block), int(groups)))

    def set_print(self, enable):
        self._check(self._lib.tgs_set_print(self._h, 1 if enable else 0))

    def set_instances(self, n):
        self._check(self._lib.tgs_set_instances(self._h, int(n)))
        self._n_inst = int(n)

    def instances(self):
        return self._n_inst

    def set_stimulus(self, input_ids, seed):
        ids = np.ascontiguousarray(input_ids, dtype=np.uint32)
        self._check(self._lib.tgs_set_stimulus(self._h, ids.ctypes.data_as(_u32p), ids.size, int(seed)))

    def device_bytes(self):
        """Device memory the context holds now (tgs_device_bytes)."""
        v = ctypes.c_uint64()
        self._check(self._lib.tgs_device_bytes(self._h, ctypes.byref(v)))
        return int(v.value)

    def run_batch(self, cycles):
        c = np.ascontiguousarray(cycles, dtype=np.uint32)
        st = ctypes.c_int()
        self._check(self._lib.tgs_run_batch(self._h, c.ctypes.data_as(_u32p), c.size, ctypes.byref(st)))
        return bool(st.value)

    def step(self, write_ids, write_vals, cycles, read_ids):
        """One testbench step of every instance (tgs_step): write register
        nibbles write_ids (write_vals shaped [instances, len(write_ids)]),
        run `cycles` cycles, read register nibbles read_ids. Returns
        (values [instances, len(read_ids)], any instance stopped)."""
        wi = np.ascontiguousarray(write_ids, dtype=np.uint32).reshape(-1)
        ri = np.ascontiguousarray(read_ids, dtype=np.uint32).reshape(-1)
        n = self.instances()
        wv = np.ascontiguousarray(write_vals, dtype=np.uint8).reshape(n, wi.size)
        out = np.zeros((n, ri.size), dtype=np.uint8)
        st = ctypes.c_int()
        self._check(self._lib.tgs_step(self._h, wi.ctypes.data_as(_u32p), wi.size, wv.ctypes.data_as(_u8p),
                                       int(cycles), ri.ctypes.data_as(_u32p), ri.size, out.ctypes.data_as(_u8p),
                                       ctypes.byref(st)))
        self.cycle_cnt += int(cycles)
        return out, bool(st.value)

    def pool_size(self, pool):
        n = ctypes.c_uint64()
        self._check(self._lib.tgs_pool_size(self._h, pool, ctypes.byref(n)))
        return n.value

    def read_pool(self, pool, instance=0):
        n = self.pool_size(pool)
        out = np.zeros(n, dtype=np.uint8)
        self._check(self._lib.tgs_read_pool(self._h, instance, pool, 0, n, out.ctypes.data_as(_u8p)))
        return out

    def write_pool(self, pool, data, instance=0, offset=0):
        a = np.ascontiguousarray(data, dtype=np.uint8)
        self._check(self._lib.tgs_write_pool(self._h, instance, pool, offset, a.size, a.ctypes.data_as(_u8p)))

    def stopped(self, instance=0):
        st = ctypes.c_int()
        self._check(self._lib.tgs_get_stop(self._h, instance, ctypes.byref(st)))
        return bool(st.value)

    def stop_cycle(self, instance=0):
        c = ctypes.c_uint64()
        self._check(self._lib.tgs_get_stop_cycle(self._h, instance, ctypes.byref(c)))
        return c.value

    def reset_instance(self, instance, seed, index):
        """Restart one instance from the initial state with stimulus generator
        splitmix64(seed + index) (a slot reused for a new job)."""
        self._check(self._lib.tgs_reset_instance(self._h, int(instance), int(seed), int(index)))

    def snapshot_begin(self, instance=0):
        """Start an asynchronous register-pool snapshot; returns its ticket."""
        t = ctypes.c_uint64()
        self._check(self._lib.tgs_snapshot_begin(self._h, int(instance), ctypes.byref(t)))
        return t.value

    def snapshot_wait(self, ticket):
        """(register pool copy, cycle) of a started snapshot; the signal-probe
        values, when probes are set, follow the register pool in list order."""
        ptr = _u8p()
        cyc = ctypes.c_uint64()
        self._check(self._lib.tgs_snapshot_wait(self._h, int(ticket), ctypes.byref(ptr), ctypes.byref(cyc)))
        n = self.pool_size(POOL_REG) + self._n_probes
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n else np.zeros(0, np.uint8), cyc.value

    def checkpoint(self, instance=0):
        """One instance's state (pools, cycle count, stop flag and cycle, stimulus
        generator) as bytes; restores into any context of the same design."""
        n = ctypes.c_uint64()
        self._check(self._lib.tgs_checkpoint_size(self._h, ctypes.byref(n)))
        buf = np.zeros(n.value, np.uint8)
        self._check(self._lib.tgs_checkpoint_save(self._h, int(instance), buf.ctypes.data_as(_u8p), buf.size))
        return buf.tobytes()

    def restore(self, image, instance=0):
        buf = np.frombuffer(bytes(image), np.uint8)
        self._check(self._lib.tgs_checkpoint_restore(self._h, int(instance), buf.ctypes.data_as(_u8p), buf.size))

    def cycles(self, instance=0):
        c = ctypes.c_uint64()
        self._check(self._lib.tgs_get_cycles(self._h, instance, ctypes.byref(c)))
        return c.value

    def print_log(self, instance=0):
        n = ctypes.c_uint64()
        self._check(self._lib.tgs_get_print_log(self._h, instance, None, None, 0, ctypes.byref(n)))
        k = n.value
        cyc = np.zeros(max(k, 1), dtype=np.uint64)
        msg = np.zeros(max(k, 1), dtype=np.uint32)
        self._check(self._lib.tgs_get_print_log(self._h, instance, cyc.ctypes.data_as(_u64p),
                                                msg.ctypes.data_as(_u32p), k, ctypes.byref(n)))
        return sorted(zip(cyc[:k].tolist(), msg[:k].tolist()))

    def clear_print_log(self, instance=0):
        """Drop the instance's held print events (consumed); returns how many
        events were lost since the last clear (0 unless a launch overflowed)."""
        d = ctypes.c_uint64()
        self._check(self._lib.tgs_clear_print_log(self._h, int(instance), ctypes.byref(d)))
        return d.value

    def set_part_profile(self, enable=True, start=100, count=100):
        """Per-partition useful/total clock ticks of launch-local cycles
        start .. start + count - 1 (ENABLE_SIM_PROFILE, src/SimEval.cu:894-983)."""
        self._check(self._lib.tgs_set_part_profile(self._h, 1 if enable else 0, int(start), int(count)))

    def part_profile(self):
        """(useful, total, cycles averaged): per-partition average ticks of the
        last launch that reached the profile window (copy_profile_data)."""
        n = sum(self.program_stats()["n_parts"])
        u = np.zeros(max(n, 1), np.int64)
        t = np.zeros(max(n, 1), np.int64)
        k = ctypes.c_uint32()
        self._check(self._lib.tgs_get_part_profile(self._h, u.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                                   t.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), n,
                                                   ctypes.byref(k)))
        return u[:n], t[:n], k.value

    def profile(self):
        p = _native.Profile()
        self._check(self._lib.tgs_get_profile(self._h, ctypes.byref(p)))
        return {n: getattr(p, n) for n, _ in _native.Profile._fields_}

    def reset_profile(self):
        self._check(self._lib.tgs_reset_profile(self._h))

    def program_stats(self):
        st = _native.ProgramStats()
        self._check(self._lib.tgs_get_program_stats(self._h, ctypes.byref(st)))
        return {n: (list(getattr(st, n)) if n == "n_parts" else getattr(st, n)) for n, _ in _native.ProgramStats._fields_}


def parse_symbols(blob: bytes) -> dict:
    """GPUSimIOSymbols.bin (src/Serializer.cpp:264-354) -> {'reg','signal','mem'} dicts."""
    pos = 0

    def u64():
        nonlocal pos
        v = int.from_bytes(blob[pos:pos + 8], "little")
        pos += 8
        return v

    def u32():
        nonlocal pos
        v = int.from_bytes(blob[pos:pos + 4], "little")
        pos += 4
        return v

    out = {}
    for key, k in (("reg", 2), ("signal", 3), ("mem", 3)):
        m = {}
        for _ in range(u64()):
            n = u64()
            name = blob[pos:pos + n].decode()
            pos += n
            m[name] = [tuple(u32() for _ in range(k)) for _ in range(u64())]
        out[key] = m
    if pos != len(blob):
        raise ValueError("trailing bytes in symbol file")
    return out


def run_batch_multi(sims, cycles):
    """tgs_run_batch_multi: run_batch of several batched simulators of one
    device in shared launches (small batches of different designs occupy the
    GPU together). cycles[k] holds one count per instance of sims[k].
    Returns each simulator's any-stopped flag."""
    n = len(sims)
    if n == 0:
        return []
    arrs = [np.ascontiguousarray(c, dtype=np.uint32) for c in cycles]
    if len(arrs) != n:
        raise ValueError("one cycle array per simulator")
    lib = sims[0]._lib
    hs = (ctypes.c_void_p * n)(*[s._h for s in sims])
    cs = (_u32p * n)(*[a.ctypes.data_as(_u32p) for a in arrs])
    st = (ctypes.c_int * n)()
    before = [s._err() for s in sims]
    rc = lib.tgs_run_batch_multi(hs, n, cs, st)
    if rc != TGS_OK:  # the context that failed keeps the error text
        after = [s._err() for s in sims]
        raise SimError(rc, "; ".join(a for a, b in zip(after, before) if a and a != b) or after[0])
    return [bool(v) for v in st]
