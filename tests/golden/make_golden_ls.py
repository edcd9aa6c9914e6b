"""Write golden vectors of the level-set density with the compiled reference
(oracle/_ref/mg_ref_vel, which links the reference's own lsm.for and
bounds_lsm.for besides the velocity-form sources; oracle/Makefile).

Runs in the build container only (the reference sources exist there):

    make -C oracle ref && python tests/golden/make_golden_ls.py

Two kinds of case:

- ls_<case>.npz (grids without periodic faces, where the reference's
  exchange(18) is defined): a level set phi whose ghost layer the
  reference's bound_lsm(19) filled (the copy sequence bound_lsm(14) applies
  to phi), then the reference's heaviside (lsm.for:1456-1509, with
  bound_lsm(18/19)), coeff and the velocity-form newsolv_mg loop for maxcy
  cycles (eps = 0).  Holds phi, ustar, vstar, wstar, p0, densl/densg/mul/mug
  (props), and the reference's dens, mu, iters, rmax, p, u, v, w (ghosted)
  and sup (interior).
- lsb_<case>.npz (periodic faces allowed): a seeded cell field x and the
  reference's bound_lsm(19) of it (out), the ghost sequence this library
  applies to the density and the viscosity.

All arrays float64, no pickles, blocks concatenated, each block's ghosted
array i fastest.  manifest_ls.json records how each case was made.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from make_golden_vel import inputs as vel_inputs  # noqa: E402
from multiflow3d_amd import topology as T  # noqa: E402
from multiflow3d_amd import workloads as W  # noqa: E402
from oracle.oracle import run_ref_bound_lsm, run_ref_vel  # noqa: E402

P = T.PERIODIC
N = T.NEUMANN
CH = T.CHANNEL_BC
WATER_AIR = (1000.0, 1.2, 1.0e-3, 1.8e-5)
MILD = (1.0, 0.25, 2.0, 1.0)
# name: (ex, ey, ez, bc, h, levels, maxcy, dt, alfapr, phi kind, props, seed)
LS_CASES = {
    "walls221_band_L2": ([6, 4], [6, 8], [6], [N] * 6, [1 / 12] * 3, 2, 3, 0.01, 1.0, "band", WATER_AIR,
                         W.BASE_SEED + 61),
    "single16_sphere_L3": ([16], [16], [16], [N] * 6, [1 / 16] * 3, 3, 3, 0.01, 1.0, "sphere", WATER_AIR,
                           W.BASE_SEED + 62),
    "walls312_band_L2": ([4, 4, 4], [8], [4, 4], [N] * 6, [1 / 12, 1 / 8, 1 / 8], 2, 3, 0.02, 0.8, "band", MILD,
                         W.BASE_SEED + 63),
    "walls222_sphere_L2": ([8, 8], [8, 8], [8, 8], [N] * 6, [1 / 16] * 3, 2, 4, 0.01, 1.0, "sphere", MILD,
                           W.BASE_SEED + 64),
}
# name: (ex, ey, ez, bc, seed)
LSB_CASES = {
    "single_channel": ([6], [5], [4], CH, W.BASE_SEED + 71),
    "channel222": ([8, 8], [8, 8], [8, 8], CH, W.BASE_SEED + 72),
    "allper312": ([4, 4, 4], [8], [4, 4], [P] * 6, W.BASE_SEED + 73),
    "mixed221": ([6, 4], [5, 7], [6], [N, N, P, P, N, N], W.BASE_SEED + 74),
    "walls221": ([6, 4], [5, 7], [6], [N] * 6, W.BASE_SEED + 75),
    "zper_single": ([5], [6], [7], [N, N, N, N, P, P], W.BASE_SEED + 76),
}


def cat(parts):
    return np.concatenate([a.ravel() for a in parts])


def seeded_boxes(g, seed, scale=1.0):
    """Seeded uniform(-1, 1) values over every block's ghosted box."""
    tot = sum((b.n[0] + 2) * (b.n[1] + 2) * (b.n[2] + 2) for b in g.blocks)
    r = W.uniform_pm1(seed, 0, tot)
    out, o = [], 0
    for b in g.blocks:
        shp = (b.n[2] + 2, b.n[1] + 2, b.n[0] + 2)
        m = shp[0] * shp[1] * shp[2]
        out.append(scale * r[o:o + m].reshape(shp))
        o += m
    return out


def raw_phi(g, kind, seed):
    """A level set before its ghosts are filled: 'band' = seeded values in
    [-3 epsl, 3 epsl] (a third of them in the transition band); 'sphere' =
    the signed distance to a sphere of radius 0.3 at the domain centre plus
    a seeded perturbation of 0.1 epsl."""
    epsl = 2 * max(g.h)
    if kind == "band":
        return seeded_boxes(g, seed, 3 * epsl)
    noise = seeded_boxes(g, seed, 0.1 * epsl)
    ext = [sum(g.ex) * g.h[0], sum(g.ey) * g.h[1], sum(g.ez) * g.h[2]]
    out = []
    for blk, nz in zip(g.blocks, noise):
        ox, oy, oz = blk.origin
        nx, ny, nzz = blk.n
        x = (ox + np.arange(nx + 2) - 0.5) * g.h[0] - 0.5 * ext[0]
        y = (oy + np.arange(ny + 2) - 0.5) * g.h[1] - 0.5 * ext[1]
        z = (oz + np.arange(nzz + 2) - 0.5) * g.h[2] - 0.5 * ext[2]
        r = np.sqrt(z[:, None, None] ** 2 + y[None, :, None] ** 2 + x[None, None, :] ** 2)
        out.append((0.3 - r) + nz)
    return out


def main():
    manifest = {"ls": {}, "lsb": {}}
    for name, spec in LS_CASES.items():
        ex, ey, ez, bc, h, L, maxcy, dt, alfapr, kind, props, seed = spec
        g = T.block_grid(ex, ey, ez, bc, h)
        phi = run_ref_bound_lsm(g, raw_phi(g, kind, seed))
        us, vs, ws, p0, _ = vel_inputs(g, seed + 1000, False)
        k, rh, p, u, v, w, su, dens, mu = run_ref_vel(g, L, us, vs, ws, p0, 0.0, maxcy, dt, alfapr,
                                                      phi=phi, props=props)
        np.savez_compressed(
            os.path.join(HERE, f"ls_{name}.npz"), phi=cat(phi), ustar=cat(us), vstar=cat(vs), wstar=cat(ws),
            p0=cat(p0), props=np.array(props, dtype=np.float64), dens=cat(dens), mu=cat(mu),
            iters=np.array([float(k)]), rmax=np.asarray(rh[:k + 1], dtype=np.float64), p=cat(p), u=cat(u),
            v=cat(v), w=cat(w), su=cat([a[1:-1, 1:-1, 1:-1] for a in su]))
        epsl = 2 * max(h)
        band = int(sum(int((np.abs(a) <= epsl).sum()) for a in phi))
        manifest["ls"][name] = dict(ex=ex, ey=ey, ez=ez, bc=bc, h=h, levels=L, maxcy=maxcy, dt=dt, alfapr=alfapr,
                                    phi=kind, props=list(props), seed=seed, iters=int(k), band_cells=band,
                                    made_by="oracle/_ref/mg_ref_vel mode 2 (reference lsm.for heaviside, "
                                            "bounds_lsm.for, press.for, exchange*.for, mgsolver.for)")
        print("ls", name, "iters", k, "band cells", band, flush=True)
    for name, (ex, ey, ez, bc, seed) in LSB_CASES.items():
        g = T.block_grid(ex, ey, ez, bc, [0.1] * 3)
        x = seeded_boxes(g, seed)
        out = run_ref_bound_lsm(g, x)
        np.savez_compressed(os.path.join(HERE, f"lsb_{name}.npz"), x=cat(x), out=cat(out))
        manifest["lsb"][name] = dict(ex=ex, ey=ey, ez=ez, bc=bc, seed=seed,
                                     made_by="oracle/_ref/mg_ref_vel mode 3 (reference bounds_lsm.for "
                                             "bound_lsm(19), exchange*.for)")
        print("lsb", name, flush=True)
    with open(os.path.join(HERE, "manifest_ls.json"), "w") as fh:
        json.dump(manifest, fh, indent=1)


if __name__ == "__main__":
    main()
