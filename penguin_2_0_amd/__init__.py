"""MI355X-native PEnGUIn hydro step.

The product is libpenguin.so (C ABI in include/penguin.h, HIP kernels for gfx950); this
package holds its build recipe, the ctypes mirror of the ABI and the workload table.
"""
from .abi import (NPAD, PN_FP32, PN_FP64, Context, Disk, PnBody, PnError, PnGrid, PnParams, PnRunState,
                  PnSymdisk, init_disk, load_library, params_default, ring_dtype)

__all__ = ["NPAD", "PN_FP32", "PN_FP64", "Context", "Disk", "PnBody", "PnError", "PnGrid", "PnParams",
           "PnRunState", "PnSymdisk", "init_disk", "load_library", "params_default", "ring_dtype"]
