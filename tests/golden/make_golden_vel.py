"""Write golden vectors of the velocity form of newsolv_mg with the compiled
reference (oracle/_ref/mg_ref_vel, built from the reference's own press.for,
exchange*.for, mgsolver.for and exchangepp.for by oracle/Makefile).

Runs in the build container only (the reference sources exist there):

    make -C oracle ref && python tests/golden/make_golden_vel.py [pl]

(`pl`: only the wide-shell cases of CASES_PL, manifest_vel_pl.json.)

Each case runs the reference's coeff, calvel, calmas, then the loop
mgkcyc / p += pp / exchange(4) / pbound / calvel / calmas of
newsolv_mg.for:12-58 until rmax < eps (after at least two cycles) or maxcy.
Inputs are seeded u*, v*, w* (whole ghosted arrays), a seeded starting p
including its ghosts, and for LSM the two-phase sphere density with
ghosts.  vel_<case>.npz holds (float64, no pickles, blocks concatenated,
each block's ghosted array i fastest): ustar, vstar, wstar, p0, dens (LSM
only) and the reference's outputs iters, rmax (per cycle, index 0 =
before the first cycle), p, u, v, w (ghosted) and su (sup under LSM,
interior).  manifest_vel.json records how each case was made.
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
from oracle.oracle import run_ref_vel  # noqa: E402

CH = T.CHANNEL_BC
P = T.PERIODIC
N = T.NEUMANN
# name: (ex, ey, ez, bc, h, levels, lsm, maxcy, eps, dt, alfapr, seed)
CASES = {
    "single12_channel_L3": ([12], [12], [12], CH, [1 / 12] * 3, 3, False, 6, 0.0, 0.01, 1.0,
                            W.BASE_SEED + 31),
    "single12_channel_L3_eps": ([12], [12], [12], CH, [1 / 12] * 3, 3, False, 40, 1e-9, 0.01, 1.0,
                                W.BASE_SEED + 32),
    "multi222_channel_L2": ([8, 8], [8, 8], [8, 8], CH, [1 / 16] * 3, 2, False, 4, 0.0, 0.02, 0.8,
                            W.BASE_SEED + 33),
    "uneven_channel_L2": ([4, 8], [8, 4], [4, 8], CH, [1 / 12, 1 / 12, 1 / 12], 2, False, 4, 0.0, 0.01, 1.0,
                          W.BASE_SEED + 34),
    "mixedbc_L3": ([8, 8], [8], [8, 8], [N, N, P, P, N, N], [1 / 16] * 3, 3, False, 4, 0.0, 0.01, 1.0,
                   W.BASE_SEED + 35),
    "allper312_L2": ([4, 4, 4], [8], [4, 4], [P] * 6, [1 / 12, 1 / 8, 1 / 8], 2, False, 4, 0.0, 0.01, 1.0,
                     W.BASE_SEED + 36),
    "walls221_L2": ([4, 4], [4, 4], [8], [N] * 6, [1 / 8] * 3, 2, False, 4, 0.0, 0.01, 1.0,
                    W.BASE_SEED + 37),
    "lsm222_sphere_L2": ([8, 8], [8, 8], [8, 8], CH, [1 / 16] * 3, 2, True, 3, 0.0, 0.01, 1.0,
                         W.BASE_SEED + 38),
}
# wider ghost shells: name: (ex, ey, ez, bc, h, levels, maxcy, dt, alfapr, seed, pl_ex, LMR,
# differencing); pl = LMR + pl_ex ghost layers (localparameters.for:17-24).  The first is the
# shipped control.cin's LMR 2, pl_ex 2 on a periodic channel of 2x2x2 blocks.
CASES_PL = {
    "deck_channel222_pl4": ([8, 8], [8, 8], [8, 8], CH, [1 / 16] * 3, 2, 3, 0.01, 1.0, W.BASE_SEED + 41, 2, 2, 1),
    "allper312_pl2": ([4, 4, 4], [8], [4, 4], [P] * 6, [1 / 12, 1 / 8, 1 / 8], 2, 3, 0.01, 1.0,
                      W.BASE_SEED + 42, 1, 1, 1),
    "d4_mixedbc_pl3": ([8, 8], [8], [8, 8], [N, N, P, P, N, N], [1 / 16] * 3, 3, 3, 0.01, 1.0,
                       W.BASE_SEED + 43, 1, 2, 2),
    "d4_single12_pl2": ([12], [12], [12], CH, [1 / 12] * 3, 3, 4, 0.01, 1.0, W.BASE_SEED + 44, 0, 2, 2),
}


def grid_of(spec):
    ex, ey, ez, bc, h = spec[:5]
    return T.block_grid(ex, ey, ez, bc, h)


def inputs(g, seed, lsm):
    """Seeded inputs of one case (also used by the tests to re-derive them)."""
    us, vs, ws = W.velocity_inputs(g, seed)
    rng_p = W.uniform_pm1(seed + 1, 0, sum((b.n[0] + 2) * (b.n[1] + 2) * (b.n[2] + 2) for b in g.blocks))
    p0, o = [], 0
    for b in g.blocks:
        shp = (b.n[2] + 2, b.n[1] + 2, b.n[0] + 2)
        cnt = shp[0] * shp[1] * shp[2]
        p0.append(0.1 * rng_p[o:o + cnt].reshape(shp))
        o += cnt
    dens = None
    if lsm:
        d = W.sphere_density(g)  # interior (z, y, x) of the whole grid
        dens = []
        for blk in g.blocks:
            ox, oy, oz = blk.origin
            nx, ny, nz = blk.n
            # ghosted block density: clamp to the grid (walls) or wrap (periodic)
            zi = np.arange(oz - 1, oz + nz + 1) % d.shape[0]
            yi = np.clip(np.arange(oy - 1, oy + ny + 1), 0, d.shape[1] - 1)
            xi = np.arange(ox - 1, ox + nx + 1) % d.shape[2]
            dens.append(np.ascontiguousarray(d[np.ix_(zi, yi, xi)]))
    return us, vs, ws, p0, dens


def inputs_pl(g, seed, G):
    """Seeded inputs of a wide-shell case: u*, v*, w* and p0 over every
    block's whole (z+2G, y+2G, x+2G) array, drawn block by block, u*, v*,
    w*, then p (scaled by 0.1), i fastest; wall faces (coordinates 0 and n
    on a physical non-periodic face) zero."""
    us, vs, ws, p0 = [], [], [], []
    pos = 0
    for blk in g.blocks:
        nx, ny, nz = blk.n
        shp = (nz + 2 * G, ny + 2 * G, nx + 2 * G)
        cnt = shp[0] * shp[1] * shp[2]
        for c, out in enumerate((us, vs, ws)):
            a = W.uniform_pm1(seed, pos, cnt).reshape(shp)
            pos += cnt
            n = blk.n[c]
            for side, coord in ((0, 0), (1, n)):
                if blk.nbr[2 * c + side] < 0 and blk.bc[2 * c + side] != P:
                    ix = coord + G - 1
                    if c == 0:
                        a[:, :, ix] = 0.0
                    elif c == 1:
                        a[:, ix, :] = 0.0
                    else:
                        a[ix] = 0.0
            out.append(a)
        p0.append(0.1 * W.uniform_pm1(seed, pos, cnt).reshape(shp))
        pos += cnt
    return us, vs, ws, p0


def cat(parts):
    return np.concatenate([a.ravel() for a in parts])


def main_pl():
    """The wide-shell cases only (manifest_vel_pl.json, vel_<case>.npz; su
    over the interior)."""
    manifest = {}
    for name, spec in CASES_PL.items():
        ex, ey, ez, bc, h, L, maxcy, dt, alfapr, seed, pl_ex, lmr, diff = spec
        g = T.block_grid(ex, ey, ez, bc, h)
        G = pl_ex + lmr
        us, vs, ws, p0 = inputs_pl(g, seed, G)
        k, rh, p, u, v, w, su = run_ref_vel(g, L, us, vs, ws, p0, 0.0, maxcy, dt, alfapr, pl_ex=pl_ex, lmr=lmr,
                                            differencing=diff)
        arrs = dict(iters=np.array([float(k)]), rmax=np.asarray(rh[:k + 1], dtype=np.float64), p=cat(p),
                    u=cat(u), v=cat(v), w=cat(w), su=cat([a[G:-G, G:-G, G:-G] for a in su]))
        np.savez_compressed(os.path.join(HERE, f"vel_{name}.npz"), **arrs)
        manifest[name] = dict(ex=ex, ey=ey, ez=ez, bc=bc, h=h, levels=L, maxcy=maxcy, eps=0.0, dt=dt,
                              alfapr=alfapr, seed=seed, pl_ex=pl_ex, lmr=lmr, differencing=diff, ghost=G,
                              iters=int(k), inputs="inputs_pl(grid, seed, ghost) of this script",
                              made_by="oracle/_ref/mg_ref_vel (reference press.for, exchange*.for, "
                                      "mgsolver.for, exchangepp.for)")
        print(name, "iters", k, "rmax", rh[:k + 1][-1], flush=True)
    with open(os.path.join(HERE, "manifest_vel_pl.json"), "w") as fh:
        json.dump(manifest, fh, indent=1)


def main():
    manifest = {}
    for name, spec in CASES.items():
        ex, ey, ez, bc, h, L, lsm, maxcy, eps, dt, alfapr, seed = spec
        g = grid_of(spec)
        us, vs, ws, p0, dens = inputs(g, seed, lsm)
        k, rh, p, u, v, w, su = run_ref_vel(g, L, us, vs, ws, p0, eps, maxcy, dt, alfapr, lsm=lsm, dens=dens)
        arrs = dict(ustar=cat(us), vstar=cat(vs), wstar=cat(ws), p0=cat(p0), iters=np.array([float(k)]),
                    rmax=np.asarray(rh[:k + 1], dtype=np.float64), p=cat(p), u=cat(u), v=cat(v), w=cat(w),
                    su=cat([a[1:-1, 1:-1, 1:-1] for a in su]))
        if lsm:
            arrs["dens"] = cat(dens)
        np.savez_compressed(os.path.join(HERE, f"vel_{name}.npz"), **arrs)
        manifest[name] = dict(ex=ex, ey=ey, ez=ez, bc=bc, h=h, levels=L, lsm=lsm, maxcy=maxcy, eps=eps,
                              dt=dt, alfapr=alfapr, seed=seed, iters=int(k),
                              made_by="oracle/_ref/mg_ref_vel (reference press.for, exchange*.for, "
                                      "mgsolver.for, exchangepp.for)")
        print(name, "iters", k, "rmax", rh[:k + 1][-1], flush=True)
    with open(os.path.join(HERE, "manifest_vel.json"), "w") as fh:
        json.dump(manifest, fh, indent=1)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "pl":
        main_pl()
    else:
        main()
        main_pl()
