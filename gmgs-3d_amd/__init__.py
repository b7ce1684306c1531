"""MI355X-native geometric-multigrid PCG for voxel hex8 elasticity.

The compute path is libgmgs.so (hand-written HIP for gfx950, C ABI in
include/gmgs.h); this package is its Python host side.
"""
from . import export, voxelize, workloads  # noqa: F401
from .gmgs import (GMGS3D, GMGSError, NotConverged, Solver, default_params,  # noqa: F401
                   batch_solve, copy_probe,
                   load_library, library_path, exported_symbols,
                   GMGS_PREC_FP64, GMGS_PREC_MIXED, FLAG_EXACT_GALERKIN, FLAG_NO_INDEX_TABLES,
                   FLAG_FR_BETA, FLAG_PR_BETA, FLAG_SIMPLE_KERNELS,
                   FLAG_AUTOTUNE, FLAG_NO_GRAPHS, FLAG_SYM_VCYCLE, FLAG_NO_JACOBI_STEP,
                   FLAG_CHEBYSHEV, FLAG_ASSEMBLED_COARSE, FLAG_SPLIT_COARSE,
                   FLAG_UNFUSED_PROLONG, FLAG_UNFUSED_RESIDUAL)
