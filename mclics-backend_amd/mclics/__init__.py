"""Python plumbing around the MI355X checkerboard Metropolis engine.

The engine itself is the HIP shared library with a C ABI (include/mclics_hip.h);
this package binds it with ctypes for tests and the benchmark.
"""
from .engine import (ANCHORING, Batch, Network, FREE, GHRL, KERNEL_AUTO, KERNEL_GENERIC, KERNEL_SLAB, KERNEL_SLAB_PERSISTENT, KERNEL_X4, LL, ORDER_CHECKER, ORDER_REF8, PEAR, PERIODIC, Config, EngineError, McParams,
                     Lattice, Surface, device_count, load_library)
from .ic import homogeneous_ic, random_ic

__all__ = [
    "ANCHORING", "Batch", "Network", "FREE", "GHRL", "KERNEL_AUTO", "KERNEL_GENERIC", "KERNEL_SLAB", "KERNEL_SLAB_PERSISTENT", "KERNEL_X4", "LL", "ORDER_CHECKER", "ORDER_REF8", "PEAR", "PERIODIC", "Config", "EngineError", "McParams",
    "Lattice", "Surface", "device_count", "load_library", "homogeneous_ic", "random_ic",
]
