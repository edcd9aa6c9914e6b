"""MI355X-native pressure-Poisson multigrid for MultiFlow3D.

The package drops in for the reference's FP64 pressure solver
(newsolv_mg / mgkcyc / coeff, newsolv_mg.for and mgsolver.for) through the
C ABI of include/mf3d_mg.h, implemented by hand-written gfx950 HIP kernels in
csrc/.  ``PressureMultigrid`` is the host-side mirror of the reference's
subroutine interface.
"""
from .topology import (CHANNEL_BC, NEUMANN, PERIODIC, Block, BlockGrid, block_grid,
                       max_levels, single_block)
from .solver import (Diverged, MGError, NotConverged, PressureMultigrid, SolveResult,
                     estimate_bytes)

__all__ = [
    "CHANNEL_BC", "NEUMANN", "PERIODIC", "Block", "BlockGrid", "block_grid", "max_levels",
    "single_block", "Diverged", "MGError", "NotConverged", "PressureMultigrid", "SolveResult",
    "estimate_bytes",
]
