"""hopefoam_amd -- MI355X (gfx950) implementation of HopeFOAM's 2D triangle DG hot path.

The compute path is libdghip.so (hand-written HIP kernels behind the C ABI of
include/dgh.h).  This package only locates, builds and binds that library; it has
no CPU fallback: every entry point fails loudly when the library or a GPU is missing.
"""
from .dgh import (DGH_BC_FIXED_VALUE, DGH_BC_WALL, DGH_BC_ZERO_GRADIENT, DGH_EQ_ADVECT, DGH_EQ_EULER,
                  DGH_FLUX_AVERAGE, DGH_FLUX_LF, DGH_FLUX_ROE, DGH_FLUX_UPWIND, DGH_PREC_BF16X3, DGH_PREC_FP32,
                  DGH_PREC_FP64, Context, DghError, RhsDesc, build, lib, lib_path, refelem_get)

__all__ = [
    "Context", "DghError", "RhsDesc", "build", "lib", "lib_path", "refelem_get",
    "DGH_EQ_ADVECT", "DGH_EQ_EULER", "DGH_FLUX_UPWIND", "DGH_FLUX_LF", "DGH_FLUX_ROE", "DGH_FLUX_AVERAGE",
    "DGH_PREC_FP64", "DGH_PREC_FP32", "DGH_PREC_BF16X3",
    "DGH_BC_FIXED_VALUE", "DGH_BC_ZERO_GRADIENT", "DGH_BC_WALL",
]
