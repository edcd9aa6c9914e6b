"""Write golden vectors with the compiled reference (oracle/_ref/mg_ref).

Runs in the build container only (the reference sources exist there):

    make -C oracle ref && python tests/golden/make_golden.py

Each case runs the reference's own coeff + mgkcyc + exchangepp in the
Poisson form of the newsolv_mg loop for a fixed number of cycles (tol = 0)
on seeded inputs, and stores inputs and outputs in <case>.npz (float64, no
pickles): f (RHS, blocks concatenated), dens (LSM only), rho and rmax per
cycle, p (interior, blocks concatenated) and pp1 (pp after cycle 1
including one ghost layer).  manifest.json records how each case was made.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from multiflow3d_amd import topology as T  # noqa: E402
from multiflow3d_amd import workloads as W  # noqa: E402
from oracle.oracle import run_ref  # noqa: E402

CH = T.CHANNEL_BC
# name: (ex, ey, ez, bc, h, levels, lsm, cycles, seed)
CASES = {
    "single16_channel_L4": ([16], [16], [16], CH, [1 / 16] * 3, 4, False, 8, W.BASE_SEED + 11),
    "aniso_channel_L3": ([16], [8], [32], CH, [1 / 16, 1 / 8, 1 / 32], 3, False, 6, W.BASE_SEED + 12),
    "multi_uneven_channel_L3": ([8, 12], [16, 8], [12, 8], CH, [1 / 20, 1 / 24, 1 / 20], 3, False, 6,
                                W.BASE_SEED + 13),
    "multi_periodic_L3": ([8, 8], [8, 8], [8, 8], [T.PERIODIC] * 6, [1 / 16] * 3, 3, False, 6,
                          W.BASE_SEED + 14),
    "multi_mixedbc_L3": ([8, 16], [8], [12, 4, 8],
                         [T.NEUMANN, T.NEUMANN, T.PERIODIC, T.PERIODIC, T.NEUMANN, T.NEUMANN],
                         [1 / 24, 1 / 8, 1 / 24], 3, False, 6, W.BASE_SEED + 15),
    "neumann16_L4": ([16], [16], [16], [T.NEUMANN] * 6, [1 / 16] * 3, 4, False, 6, W.BASE_SEED + 16),
    "lsm16_sphere_L4": ([16], [16], [16], CH, [1 / 16] * 3, 4, True, 6, W.BASE_SEED + 17),
    "lsm_multi_sphere_L3": ([8, 12], [16, 8], [12, 8], CH, [1 / 20, 1 / 24, 1 / 20], 3, True, 6,
                            W.BASE_SEED + 18),
}
# full-size config 1 (64^3, 5 levels, to 1e-8): history and a few planes only
CONFIG1 = ("config1_64_L5", [64], [64], [64], CH, [1 / 64] * 3, 5, False, 40, W.BASE_SEED + 1)


def grid_of(spec):
    ex, ey, ez, bc, h = spec[:5]
    return T.block_grid(ex, ey, ez, bc, h)


def main():
    manifest = {}
    for name, spec in CASES.items():
        ex, ey, ez, bc, h, L, lsm, cyc, seed = spec
        g = grid_of(spec)
        f = W.split_blocks(g, W.global_rhs(g, seed, fv_volume=lsm))
        d = W.split_blocks(g, W.sphere_density(g)) if lsm else None
        p, rho, rmax, k, pp1 = run_ref(g, L, f, 0.0, cyc, lsm=lsm, dens_blocks=d)
        arrays = dict(f=np.concatenate([a.ravel() for a in f]), rho=rho, rmax=rmax,
                      p=np.concatenate([a.ravel() for a in p]),
                      pp1=np.concatenate([a.ravel() for a in pp1]))
        if lsm:
            arrays["dens"] = np.concatenate([a.ravel() for a in d])
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        manifest[name] = dict(ex=ex, ey=ey, ez=ez, bc=bc, h=h, levels=L, lsm=lsm, cycles=cyc,
                              seed=seed, tol=0.0)
        print(name, "cycles", k, "rho_last", rho[-1], flush=True)

    name, ex, ey, ez, bc, h, L, lsm, cyc, seed = CONFIG1
    g = T.block_grid(ex, ey, ez, bc, h)
    f = W.split_blocks(g, W.global_rhs(g, seed))
    p, rho, rmax, k, _ = run_ref(g, L, f, 1e-8, cyc)
    p0 = p[0]
    np.savez_compressed(os.path.join(HERE, name + ".npz"), rho=rho, rmax=rmax,
                        plane_k0=p0[0].copy(), plane_k32=p0[32].copy(),
                        p_mean=np.array([p0.mean()]), p_l2_mean_removed=np.array([np.linalg.norm(p0 - p0.mean())]))
    manifest[name] = dict(ex=ex, ey=ey, ez=ez, bc=bc, h=h, levels=L, lsm=lsm, maxcy=cyc, seed=seed,
                          tol=1e-8, cycles=k)
    print(name, "cycles", k, "rho_last", rho[-1], flush=True)
    with open(os.path.join(HERE, "manifest.json"), "w") as fh:
        json.dump(manifest, fh, indent=1)


if __name__ == "__main__":
    main()
