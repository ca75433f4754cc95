"""ctypes loaders for the in-tree native libraries.

The runtime (libtoucan_rt.so, HIP) and the design generator
(libtoucan_gen.so, host C++) are built in-tree by ``__graft_entry__.build()``
(or ``make -C toucangpurt_amd/csrc``). Nothing here falls back to another
implementation: a missing library raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))

TGS_OK = 0
TGS_E_ARG = -1
TGS_E_IO = -2
TGS_E_FORMAT = -3
TGS_E_LIMIT = -4
TGS_E_HIP = -5
TGS_E_NODEV = -6
TGS_E_BUDGET = -7
TGS_E_STATE = -8
TGS_E_INVALID = -9
TGS_E_UNSUPPORTED = -10

KERNEL_PLAIN = 0
KERNEL_WAVE64 = 1
KERNEL_BATCH = 2
KERNEL_STREAM = 3

POOL_REG = 0
POOL_MEM = 1
POOL_EXCHANGE = 2

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)


class GenParams(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("n_luts", ctypes.c_uint32),
        ("depth", ctypes.c_uint32),
        ("n_regs", ctypes.c_uint32),
        ("n_inputs", ctypes.c_uint32),
        ("n_vecops", ctypes.c_uint32),
        ("word_bits_mix", ctypes.c_uint32),
        ("n_srams", ctypes.c_uint32),
        ("sram_depth", ctypes.c_uint32),
        ("sram_width", ctypes.c_uint32),
        ("part_luts", ctypes.c_uint32),
        ("cluster_luts", ctypes.c_uint32),
        ("sources_a", ctypes.c_uint32),
        ("sources_b", ctypes.c_uint32),
        ("word_part_ops", ctypes.c_uint32),
        ("n_prints", ctypes.c_uint32),
        ("n_stops", ctypes.c_uint32),
        ("stop_rarity", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32 * 8),
    ]


class GenStats(ctypes.Structure):
    _fields_ = [
        ("n_parts", ctypes.c_uint64 * 2),
        ("n_luts", ctypes.c_uint64),
        ("n_vecops", ctypes.c_uint64),
        ("n_lut_ops", ctypes.c_uint64),
        ("n_microparts", ctypes.c_uint64),
        ("reg_pool", ctypes.c_uint64),
        ("mem_pool", ctypes.c_uint64),
        ("exchange_pool", ctypes.c_uint64),
        ("max_value_pool", ctypes.c_uint64),
        ("design_bytes", ctypes.c_uint64),
        ("symbol_bytes", ctypes.c_uint64),
        ("netlist_bytes", ctypes.c_uint64),
        ("n_inputs", ctypes.c_uint64),
        ("reserved", ctypes.c_uint64 * 8),
    ]


class ProgramStats(ctypes.Structure):
    _fields_ = [
        ("n_parts", ctypes.c_uint64 * 2),
        ("n_top", ctypes.c_uint64),
        ("n_mid", ctypes.c_uint64),
        ("n_last", ctypes.c_uint64),
        ("n_vecread", ctypes.c_uint64),
        ("n_vecop", ctypes.c_uint64),
        ("n_memread", ctypes.c_uint64),
        ("n_regread", ctypes.c_uint64),
        ("n_exgread", ctypes.c_uint64),
        ("n_regwrite", ctypes.c_uint64),
        ("n_exgwrite", ctypes.c_uint64),
        ("n_memwrite", ctypes.c_uint64),
        ("n_print", ctypes.c_uint64),
        ("n_stop", ctypes.c_uint64),
        ("n_microparts", ctypes.c_uint64),
        ("n_exec_levels", ctypes.c_uint64),
        ("b_prog", ctypes.c_uint64),
        ("b_state", ctypes.c_uint64),
        ("max_value_pool", ctypes.c_uint64),
        ("reg_pool", ctypes.c_uint64),
        ("mem_pool", ctypes.c_uint64),
        ("exchange_pool", ctypes.c_uint64),
    ]


class Profile(ctypes.Structure):
    _fields_ = [
        ("kernel_ms", ctypes.c_double),
        ("launches", ctypes.c_uint64),
        ("cycles", ctypes.c_uint64),
        ("grid", ctypes.c_uint32),
        ("block", ctypes.c_uint32),
        ("lds_bytes", ctypes.c_uint32),
        ("kernel", ctypes.c_uint32),
        ("resident", ctypes.c_uint32),
        ("prog_bytes", ctypes.c_uint64),
        ("barrier_ticks", ctypes.c_uint64),
        ("total_ticks", ctypes.c_uint64),
        ("batch_group", ctypes.c_uint32),
        ("batch_fast_tasks", ctypes.c_uint32),
        ("batch_nibble_rows", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


def lib_path(name):
    return os.path.join(_HERE, name)


_gen = None
_rt = None


def gen_lib():
    """The design generator library (host-only C++)."""
    global _gen
    if _gen is None:
        path = lib_path("libtoucan_gen.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(path)
        lib.tgs_gen_default_params.argtypes = [ctypes.POINTER(GenParams)]
        lib.tgs_gen_build.argtypes = [ctypes.POINTER(GenParams), ctypes.POINTER(ctypes.c_void_p)]
        lib.tgs_gen_get.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(_u8p),
                                    ctypes.POINTER(ctypes.c_size_t)]
        lib.tgs_gen_get_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(GenStats)]
        lib.tgs_gen_error.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
        lib.tgs_gen_free.argtypes = [ctypes.c_void_p]
        lib.tgs_gen_free.restype = None
        _gen = lib
    return _gen


# Exported symbols of libtoucan_rt.so, as declared in include/toucan_rt.h.
RT_SYMBOLS = [
    "tgs_program_stats_from_bytes", "tgs_initial_state", "tgs_create", "tgs_create_on_stream", "tgs_destroy",
    "tgs_last_error", "tgs_load",
    "tgs_load_files", "tgs_set_kernel", "tgs_set_launch", "tgs_set_print", "tgs_set_instances", "tgs_set_stimulus",
    "tgs_eval", "tgs_eval_free_running", "tgs_run_batch", "tgs_run_batch_multi", "tgs_device_bytes", "tgs_run_batch_noalloc", "tgs_read_regs", "tgs_write_regs", "tgs_step",
    "tgs_read_pool", "tgs_write_pool", "tgs_pool_size", "tgs_reset_instance", "tgs_snapshot_begin",
    "tgs_snapshot_wait", "tgs_set_probes", "tgs_read_probes", "tgs_checkpoint_size", "tgs_checkpoint_save", "tgs_checkpoint_restore",
    "tgs_get_stop", "tgs_get_cycles",
    "tgs_get_stop_cycle", "tgs_get_print_log", "tgs_clear_print_log", "tgs_get_profile", "tgs_reset_profile",
    "tgs_set_part_profile", "tgs_get_part_profile", "tgs_get_program_stats",
]


def rt_lib():
    """The HIP runtime library. Raises if it is not built."""
    global _rt
    if _rt is None:
        path = lib_path("libtoucan_rt.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        lib.tgs_program_stats_from_bytes.argtypes = [_u8p, ctypes.c_size_t, ctypes.POINTER(ProgramStats),
                                                     ctypes.c_char_p, ctypes.c_size_t]
        lib.tgs_initial_state.argtypes = [_u8p, ctypes.c_size_t, _u8p, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint32,
                                          _u8p, ctypes.c_size_t, _u8p, ctypes.c_size_t]
        lib.tgs_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        lib.tgs_create_on_stream.argtypes = [ctypes.c_int, vp, ctypes.POINTER(vp)]
        lib.tgs_destroy.argtypes = [vp]
        lib.tgs_destroy.restype = None
        lib.tgs_last_error.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
        lib.tgs_load.argtypes = [vp, _u8p, ctypes.c_size_t, _u8p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int]
        lib.tgs_load_files.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
        lib.tgs_set_kernel.argtypes = [vp, ctypes.c_int]
        lib.tgs_set_print.argtypes = [vp, ctypes.c_int]
        lib.tgs_set_launch.argtypes = [vp, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32]
        lib.tgs_set_instances.argtypes = [vp, ctypes.c_uint32]
        lib.tgs_set_stimulus.argtypes = [vp, _u32p, ctypes.c_uint32, ctypes.c_uint64]
        lib.tgs_eval.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        lib.tgs_eval_free_running.argtypes = [vp, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int)]
        lib.tgs_run_batch.argtypes = [vp, _u32p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int)]
        lib.tgs_run_batch_noalloc.argtypes = [vp, _u32p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int)]
        lib.tgs_run_batch_multi.argtypes = [ctypes.POINTER(vp), ctypes.c_uint32, ctypes.POINTER(_u32p),
                                            ctypes.POINTER(ctypes.c_int)]
        lib.tgs_device_bytes.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64)]
        lib.tgs_read_regs.argtypes = [vp, ctypes.c_uint32, _u32p, ctypes.c_uint32, _u8p]
        lib.tgs_write_regs.argtypes = [vp, ctypes.c_uint32, _u32p, ctypes.c_uint32, _u8p]
        lib.tgs_step.argtypes = [vp, _u32p, ctypes.c_uint32, _u8p, ctypes.c_uint32, _u32p, ctypes.c_uint32, _u8p,
                                 ctypes.POINTER(ctypes.c_int)]
        lib.tgs_read_pool.argtypes = [vp, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, _u8p]
        lib.tgs_write_pool.argtypes = [vp, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, _u8p]
        lib.tgs_pool_size.argtypes = [vp, ctypes.c_int, _u64p]
        lib.tgs_reset_instance.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32]
        lib.tgs_snapshot_begin.argtypes = [vp, ctypes.c_uint32, _u64p]
        lib.tgs_snapshot_wait.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(_u8p), _u64p]
        lib.tgs_set_probes.argtypes = [vp, _u32p, _u32p, ctypes.c_uint32]
        lib.tgs_read_probes.argtypes = [vp, ctypes.c_uint32, _u8p, ctypes.c_uint32]
        lib.tgs_checkpoint_size.argtypes = [vp, _u64p]
        lib.tgs_checkpoint_save.argtypes = [vp, ctypes.c_uint32, _u8p, ctypes.c_uint64]
        lib.tgs_checkpoint_restore.argtypes = [vp, ctypes.c_uint32, _u8p, ctypes.c_uint64]
        lib.tgs_get_stop.argtypes = [vp, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int)]
        lib.tgs_get_cycles.argtypes = [vp, ctypes.c_uint32, _u64p]
        lib.tgs_get_stop_cycle.argtypes = [vp, ctypes.c_uint32, _u64p]
        lib.tgs_get_print_log.argtypes = [vp, ctypes.c_uint32, _u64p, _u32p, ctypes.c_uint64, _u64p]
        lib.tgs_clear_print_log.argtypes = [vp, ctypes.c_uint32, _u64p]
        lib.tgs_set_part_profile.argtypes = [vp, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32]
        lib.tgs_get_part_profile.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                             ctypes.c_size_t, _u32p]
        lib.tgs_get_profile.argtypes = [vp, ctypes.POINTER(Profile)]
        lib.tgs_reset_profile.argtypes = [vp]
        lib.tgs_get_program_stats.argtypes = [vp, ctypes.POINTER(ProgramStats)]
        _rt = lib
    return _rt


def as_u8p(buf):
    """Pointer to the bytes of a bytes/bytearray/numpy buffer (kept alive by the caller)."""
    if isinstance(buf, bytes):
        return ctypes.cast(ctypes.c_char_p(buf), _u8p)
    import numpy as np
    arr = np.ascontiguousarray(buf, dtype=np.uint8)
    return arr.ctypes.data_as(_u8p)
