"""Shared small cases for the parity tests (sizes the CPU oracle finishes in
well under a second)."""
from multiflow3d_amd import topology as T
from multiflow3d_amd import workloads as W

CH = T.CHANNEL_BC


def case(name):
    """(grid, levels, lsm) for a named case."""
    return {
        "single16_L4": (T.single_block(16, [1 / 16] * 3, CH), 4, False),
        "single16_L3": (T.single_block(16, [1 / 16] * 3, CH), 3, False),
        "aniso_L3": (T.single_block([16, 8, 32], [1 / 16, 1 / 8, 1 / 32], CH), 3, False),
        "multi_uneven_L3": (T.block_grid([8, 12], [16, 8], [12, 8], CH, [1 / 20, 1 / 24, 1 / 20]), 3, False),
        "multi_uneven_L2": (T.block_grid([8, 12], [16, 8], [12, 8], CH, [1 / 20, 1 / 24, 1 / 20]), 2, False),
        "lsm16_L4": (T.single_block(16, [1 / 16] * 3, CH), 4, True),
        "lsm_multi_L3": (T.block_grid([8, 12], [16, 8], [12, 8], CH, [1 / 20, 1 / 24, 1 / 20]), 3, True),
        "neumann_L4": (T.single_block(16, [1 / 16] * 3, [T.NEUMANN] * 6), 4, False),
        "periodic_multi_L3": (T.block_grid([8, 8], [8, 8], [8, 8], [T.PERIODIC] * 6, [1 / 16] * 3), 3, False),
        "mixed_bc_multi_L3": (T.block_grid([8, 16], [8], [12, 4, 8], [T.NEUMANN, T.NEUMANN, T.PERIODIC, T.PERIODIC, T.NEUMANN, T.NEUMANN], [1 / 24, 1 / 8, 1 / 24]), 3, False),
        "single32_L5": (T.single_block(32, [1 / 32] * 3, CH), 5, False),
        "single64_L5": (T.single_block(64, [1 / 64] * 3, CH), 5, False),
        "single1_L1": (T.single_block([2, 2, 2], [0.5] * 3, CH), 1, False),
    }[name]


SMALL = ["single16_L4", "single16_L3", "aniso_L3", "multi_uneven_L3", "multi_uneven_L2",
         "lsm16_L4", "lsm_multi_L3", "neumann_L4", "periodic_multi_L3", "mixed_bc_multi_L3",
         "single32_L5", "single1_L1"]


def inputs(grid, lsm, seed=W.BASE_SEED):
    f = W.split_blocks(grid, W.global_rhs(grid, seed, fv_volume=lsm))
    d = W.split_blocks(grid, W.sphere_density(grid)) if lsm else None
    return f, d
