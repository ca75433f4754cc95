"""tgs_run_batch_noalloc: the allocation-free variant of tgs_run_batch (its
per-launch scratch in a static per-GPU array) gives bit-identical results to
tgs_run_batch, over two calls in a row on the same context, the second with
fewer instances and shorter runs, and matches the oracle on sampled
instances. The calls go through the C-ABI (ctypes)."""
import ctypes

import numpy as np
import pytest

from simharness import OracleRun, assert_pools_equal
from toucangpurt_amd import KERNEL_BATCH, _native, gen
from toucangpurt_amd._native import POOL_EXCHANGE, POOL_MEM, POOL_REG

pytestmark = pytest.mark.gpu

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)


def _ok(lib, h, rc):
    if rc != 0:
        buf = ctypes.create_string_buffer(512)
        lib.tgs_last_error(h, buf, len(buf))
        raise AssertionError(f"rc {rc}: {buf.value.decode()}")


def _context(lib, d, n, seed):
    h = ctypes.c_void_p()
    _ok(lib, h, lib.tgs_create(0, ctypes.byref(h)))
    _ok(lib, h, lib.tgs_load(h, _native.as_u8p(d.design), len(d.design), None, 0, 0, 0))
    _ok(lib, h, lib.tgs_set_kernel(h, KERNEL_BATCH))
    _ok(lib, h, lib.tgs_set_instances(h, n))
    ids = np.ascontiguousarray(d.inputs, np.uint32)
    _ok(lib, h, lib.tgs_set_stimulus(h, ids.ctypes.data_as(_u32p), ids.size, seed))
    return h


def _pools(lib, h, i):
    out = []
    for pool in (POOL_REG, POOL_MEM, POOL_EXCHANGE):
        n = ctypes.c_uint64()
        _ok(lib, h, lib.tgs_pool_size(h, pool, ctypes.byref(n)))
        a = np.zeros(n.value, np.uint8)
        _ok(lib, h, lib.tgs_read_pool(h, i, pool, 0, n.value, a.ctypes.data_as(_u8p)))
        out.append(a)
    return out


def _run(lib, h, fn, cycles):
    c = np.ascontiguousarray(cycles, np.uint32)
    st = ctypes.c_int(-1)
    _ok(lib, h, fn(h, c.ctypes.data_as(_u32p), c.size, ctypes.byref(st)))
    return st.value


def test_noalloc_equals_run_batch_twice(gpu_available):
    lib = _native.rt_lib()
    d = gen.generate(None, n_luts=6000, depth=8, n_regs=300, n_inputs=70, n_vecops=600, n_srams=3, sram_depth=1024,
                     sram_width=32, n_stops=2, part_luts=512, word_part_ops=48, seed=5)
    rng = np.random.default_rng(12)
    print("case noalloc: 200 instances, then 70 with shorter runs", flush=True)
    for n, lo, hi in ((200, 30, 160), (70, 5, 40)):
        cycles = rng.integers(lo, hi + 1, size=n).astype(np.uint32)
        a = _context(lib, d, n, 21)
        b = _context(lib, d, n, 21)
        sa = _run(lib, a, lib.tgs_run_batch, cycles)
        sb = _run(lib, b, lib.tgs_run_batch_noalloc, cycles)
        assert sa == sb
        # a second call on the same contexts continues from the first
        cycles2 = (cycles // 3).astype(np.uint32)
        assert _run(lib, a, lib.tgs_run_batch, cycles2) == _run(lib, b, lib.tgs_run_batch_noalloc, cycles2)
        for i in range(n):
            pa, pb = _pools(lib, a, i), _pools(lib, b, i)
            assert_pools_equal(pb, pa, f"instance {i} of {n}")
            ca, cb = ctypes.c_uint64(), ctypes.c_uint64()
            _ok(lib, a, lib.tgs_get_cycles(a, i, ctypes.byref(ca)))
            _ok(lib, b, lib.tgs_get_cycles(b, i, ctypes.byref(cb)))
            assert ca.value == cb.value == int(cycles[i]) + int(cycles2[i])
        for i in (0, n // 2, n - 1):
            orc = OracleRun(d.design, inputs=d.inputs, seed=21, instance=i)
            orc.run(int(cycles[i]) + int(cycles2[i]))
            assert_pools_equal(_pools(lib, b, i), orc.pools(), f"instance {i} against the oracle")
        lib.tgs_destroy(a)
        lib.tgs_destroy(b)


def test_noalloc_rejects_what_it_cannot_run(gpu_available):
    lib = _native.rt_lib()
    d = gen.generate("lut20k", seed=3, part_luts=512)
    h = ctypes.c_void_p()
    _ok(lib, h, lib.tgs_create(0, ctypes.byref(h)))
    _ok(lib, h, lib.tgs_load(h, _native.as_u8p(d.design), len(d.design), None, 0, 0, 0))
    one = np.ones(1, np.uint32)
    st = ctypes.c_int()
    # a single-instance (non-batched) context
    assert lib.tgs_run_batch_noalloc(h, one.ctypes.data_as(_u32p), 1, ctypes.byref(st)) == _native.TGS_E_STATE
    _ok(lib, h, lib.tgs_set_kernel(h, KERNEL_BATCH))
    _ok(lib, h, lib.tgs_set_instances(h, 4097))
    many = np.ones(4097, np.uint32)
    assert lib.tgs_run_batch_noalloc(h, many.ctypes.data_as(_u32p), 4097, ctypes.byref(st)) == _native.TGS_E_LIMIT
    lib.tgs_destroy(h)
