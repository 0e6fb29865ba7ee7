"""amgt_amd: MI355X (gfx950) mBSR data path for HYPRE BoomerAMG.

Hand-written HIP kernels live in ``csrc/`` and build into ``libamgt_mbsr.so``
(C-ABI: include/amgt_mbsr.h).  This package is the Python host side over that
C-ABI; it has no CPU compute path.
"""
from ._lib import (AMGT_DOUBLE, AMGT_FLOAT, AMGT_HALF, AmgtError, AmgtLibraryMissing, LIB_PATH,
                   load)

__all__ = ["AMGT_DOUBLE", "AMGT_FLOAT", "AMGT_HALF", "AmgtError", "AmgtLibraryMissing",
           "LIB_PATH", "load"]
