"""MI355X-native Toucan cycle-simulation runtime.

The product is the C-ABI library ``libtoucan_rt.so`` (include/toucan_rt.h)
with its gfx950 kernels; this package loads it in-tree and mirrors the
reference's ``toucanGPUSim::ToucanSimulator`` interface in Python
(``rt.Simulator``). ``gen`` builds seeded synthetic designs in the
reference's serialized program format.
"""
from ._native import (KERNEL_BATCH, KERNEL_PLAIN, KERNEL_STREAM, KERNEL_WAVE64, POOL_EXCHANGE, POOL_MEM, POOL_REG)  # noqa: F401

__all__ = ["rt", "gen", "KERNEL_PLAIN", "KERNEL_WAVE64", "KERNEL_BATCH", "KERNEL_STREAM", "POOL_REG", "POOL_MEM", "POOL_EXCHANGE"]
