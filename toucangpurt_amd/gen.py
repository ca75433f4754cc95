"""Seeded synthetic designs (the workload side; the reference's compiler is absent).

``generate(**params)`` calls ``tgs_gen_build`` in libtoucan_gen.so and returns
the serialized design and symbol images in the reference's file format
(src/Serializer.cpp:199-300), the logical netlist the test simulator reads,
and the stimulus input register ids.

Presets name the five workloads by their sizes.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _native

# name -> generator parameters. Sizes follow the workload definitions:
#   lut20k       : 20,000 LUT4, depth 32, 256 regs, 64 inputs
#   lut1m        : 1,000,000 LUT4 + 20,000 word ops, depth 64, 50,000 regs,
#                  16 SRAMs of 4096 x 64 bit, 128 inputs
#   lut10m       : 10,000,000 LUT4, depth 128, 500,000 regs, 64 SRAMs (4 MiB), 256 inputs
#   lut200k      : 200,000 LUT4, depth 48, 8 SRAMs of 1024 x 32 bit (batched workload)
PRESETS = {
    "lut20k": dict(n_luts=20_000, depth=32, n_regs=256, n_inputs=64),
    "lut1m": dict(n_luts=1_000_000, depth=64, n_regs=50_000, n_inputs=128, n_vecops=20_000,
                  n_srams=16, sram_depth=4096, sram_width=64),
    "lut10m": dict(n_luts=10_000_000, depth=128, n_regs=500_000, n_inputs=256, n_srams=64,
                   sram_depth=8192, sram_width=64),
    "lut200k": dict(n_luts=200_000, depth=48, n_regs=4_096, n_inputs=64, n_vecops=2_000,
                    n_srams=8, sram_depth=1024, sram_width=32, word_part_ops=16),
}


@dataclass
class Design:
    design: bytes
    symbols: bytes
    netlist: bytes
    inputs: np.ndarray  # uint32 regPool ids of the primary inputs, stimulus order
    stats: dict
    params: dict


def default_params():
    p = _native.GenParams()
    _native.gen_lib().tgs_gen_default_params(ctypes.byref(p))
    return p


def generate(preset=None, **overrides) -> Design:
    lib = _native.gen_lib()
    p = default_params()
    fields = dict(PRESETS[preset]) if preset else {}
    fields.update(overrides)
    for k, v in fields.items():
        if not hasattr(p, k):
            raise KeyError(f"unknown generator parameter {k}")
        setattr(p, k, v)
    h = ctypes.c_void_p()
    rc = lib.tgs_gen_build(ctypes.byref(p), ctypes.byref(h))
    try:
        if rc != 0:
            buf = ctypes.create_string_buffer(512)
            if h:
                lib.tgs_gen_error(h, buf, len(buf))
            raise ValueError(f"generator failed ({rc}): {buf.value.decode()}")

        def get(which):
            ptr = _native._u8p()
            n = ctypes.c_size_t()
            if lib.tgs_gen_get(h, which, ctypes.byref(ptr), ctypes.byref(n)) != 0:
                raise RuntimeError("tgs_gen_get failed")
            return ctypes.string_at(ptr, n.value)

        st = _native.GenStats()
        lib.tgs_gen_get_stats(h, ctypes.byref(st))
        stats = {name: (list(getattr(st, name)) if name == "n_parts" else getattr(st, name))
                 for name, _ in _native.GenStats._fields_ if name != "reserved"}
        inputs = np.frombuffer(get(3), dtype=np.uint32).copy()
        params = {name: getattr(p, name) for name, _ in _native.GenParams._fields_ if name != "reserved"}
        return Design(get(0), get(1), get(2), inputs, stats, params)
    finally:
        if h:
            lib.tgs_gen_free(h)
