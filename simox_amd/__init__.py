"""simox_amd: MI355X-native SIMOX MD hot path (HIP kernels behind a C-ABI).

The compute path is libsimox_gpu.so (hand-written gfx950 HIP kernels); this
package is its host-side mirror: a ctypes binding whose methods carry the
reference routines' names (FORCE, PME, PME_SELF, CTIP_SOLVER, CELL_SORT,
VV_NVE1, VV_NVE2) and raise SimoxError with the reference's STOP text where
the reference would stop.  There is no CPU fallback: without the built
library every call fails loudly.
"""
from .api import (MODE_FAITHFUL, MODE_FP64, MODE_MIXED, Simox, SimoxError, batch_step,
                  load_library, potential_setup)

__all__ = ["Simox", "SimoxError", "batch_step", "load_library", "potential_setup",
           "MODE_FAITHFUL", "MODE_FP64", "MODE_MIXED"]
