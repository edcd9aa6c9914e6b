"""Seeded synthetic inputs for the pressure-Poisson workloads.

PRNG: SplitMix64 (Steele, Lea & Flood, "Fast splittable pseudorandom number
generators", OOPSLA 2014).  The k-th draw of a stream seeded with s is
mix(s + (k+1)*0x9E3779B97F4A7C15); uniform(-1, 1) = 2*((x >> 11) * 2**-53) - 1.

Right-hand side: f = div(u)/dt with dt = 1 from seeded face velocities (the
2nd-order divergence of press.for:46-57), drawn u, then v, then w, each in
i-fastest order, then the mean removed so the periodic/Neumann problem is
consistent.  Periodic axes carry n faces (the face left of cell 0 is the
last face), wall axes carry n+1 faces including both walls.  The
variable-coefficient form multiplies by the cell volume (press.for:51-53).

Two-phase density: a sphere of radius 0.2 at the box centre, level set
phi = |x - c| - 0.2, rho = rho_g + (rho_l - rho_g) * H(phi) with the smoothed
Heaviside of lsm.for:1480-1494 (including its pi = 3.14159265359) and a
half-width of 1.5 dx.
"""
from __future__ import annotations

import numpy as np

from .topology import PERIODIC, BlockGrid

BASE_SEED = 20261202
_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, start: int, count: int) -> np.ndarray:
    """Draws start .. start+count-1 of the SplitMix64 stream seeded with ``seed``."""
    k = np.arange(start + 1, start + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform_pm1(seed: int, start: int, count: int) -> np.ndarray:
    x = splitmix64(seed, start, count)
    return 2.0 * ((x >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)) - 1.0


def face_velocities(seed: int, gn, periodic):
    """u, v, w face arrays shaped (z, y, x), drawn u then v then w."""
    nx, ny, nz = gn
    fx = nx if periodic[0] else nx + 1
    fy = ny if periodic[1] else ny + 1
    fz = nz if periodic[2] else nz + 1
    shapes = [(nz, ny, fx), (nz, fy, nx), (fz, ny, nx)]
    out, pos = [], 0
    for shp in shapes:
        cnt = shp[0] * shp[1] * shp[2]
        out.append(uniform_pm1(seed, pos, cnt).reshape(shp))
        pos += cnt
    return out


def divergence(u, v, w, h, periodic):
    """(u_e - u_w)/dx + (v_n - v_s)/dy + (w_t - w_b)/dz on the cell grid (z, y, x)."""
    dx, dy, dz = h
    if periodic[0]:
        ue, uw = u, np.roll(u, 1, axis=2)
    else:
        ue, uw = u[:, :, 1:], u[:, :, :-1]
    if periodic[1]:
        vn, vs = v, np.roll(v, 1, axis=1)
    else:
        vn, vs = v[:, 1:, :], v[:, :-1, :]
    if periodic[2]:
        wt, wb = w, np.roll(w, 1, axis=0)
    else:
        wt, wb = w[1:, :, :], w[:-1, :, :]
    return (ue - uw) / dx + (vn - vs) / dy + (wt - wb) / dz


def global_rhs(grid: BlockGrid, seed: int, fv_volume: bool = False, dt: float = 1.0) -> np.ndarray:
    """Mean-free RHS on the whole grid, shaped (z, y, x)."""
    periodic = [grid.bc[0] == PERIODIC, grid.bc[2] == PERIODIC, grid.bc[4] == PERIODIC]
    u, v, w = face_velocities(seed, grid.global_n, periodic)
    fact = divergence(u, v, w, grid.h, periodic)
    if fv_volume:
        f = fact * grid.h[0] * grid.h[1] * grid.h[2] / dt
    else:
        f = fact / dt
    f -= f.mean()
    return f


def block_rhs(grid: BlockGrid, seed: int):
    """Per-block RHS of a grid whose blocks need not share a resolution
    (refined blocks): uniform(-1, 1) SplitMix64 draws of ``seed``, block by
    block in block order (z, y, x within a block), less the mean over all
    cells."""
    out, o = [], 0
    for blk in grid.blocks:
        m = blk.cells
        out.append(uniform_pm1(seed, o, m).reshape(blk.n[2], blk.n[1], blk.n[0]))
        o += m
    mean = sum(float(a.sum()) for a in out) / o
    return [np.ascontiguousarray(a - mean) for a in out]


def split_blocks(grid: BlockGrid, field: np.ndarray):
    """Cut a global (z, y, x) field into per-block contiguous (z, y, x) arrays."""
    out = []
    for blk in grid.blocks:
        ox, oy, oz = blk.origin
        nx, ny, nz = blk.n
        out.append(np.ascontiguousarray(field[oz:oz + nz, oy:oy + ny, ox:ox + nx]))
    return out


def join_blocks(grid: BlockGrid, parts) -> np.ndarray:
    gx, gy, gz = grid.global_n
    out = np.empty((gz, gy, gx), dtype=parts[0].dtype)
    for blk, a in zip(grid.blocks, parts):
        ox, oy, oz = blk.origin
        nx, ny, nz = blk.n
        out[oz:oz + nz, oy:oy + ny, ox:ox + nx] = a.reshape(nz, ny, nx)
    return out


UNEVEN_EXTENTS = (96, 128, 160, 192)


def uneven_extents(seed: int, nblocks=(2, 2, 2), choices=UNEVEN_EXTENTS):
    """Per-axis block extents of an uneven block grid: successive SplitMix64
    draws x of ``seed`` (the x-axis blocks first, then y, then z) pick
    ``choices[x % len(choices)]``."""
    x = splitmix64(seed, 0, sum(nblocks))
    pick = [int(choices[int(v % np.uint64(len(choices)))]) for v in x]
    out, o = [], 0
    for k in nblocks:
        out.append(pick[o:o + k])
        o += k
    return out


LSM_PI = 3.14159265359   # lsm.for:1466


def heaviside(phi: np.ndarray, eps: float) -> np.ndarray:
    """Smoothed Heaviside of lsm.for:1480-1494."""
    hh = np.empty_like(phi)
    hh[phi < -eps] = 0.0
    hh[phi > eps] = 1.0
    mid = np.abs(phi) <= eps
    p = phi[mid]
    hh[mid] = 0.5 * (1.0 + p / eps + 1.0 / LSM_PI * np.sin(LSM_PI * p / eps))
    return hh


def sphere_density(grid: BlockGrid, radius: float = 0.2, rho_l: float = 1000.0,
                   rho_g: float = 1.0, eps_cells: float = 1.5) -> np.ndarray:
    """Two-phase cell density on the global grid (z, y, x); gas inside the sphere."""
    gx, gy, gz = grid.global_n
    dx, dy, dz = grid.h
    x = (np.arange(gx) + 0.5) * dx
    y = (np.arange(gy) + 0.5) * dy
    z = (np.arange(gz) + 0.5) * dz
    cx, cy, cz = 0.5 * gx * dx, 0.5 * gy * dy, 0.5 * gz * dz
    phi = np.sqrt((z[:, None, None] - cz) ** 2 + (y[None, :, None] - cy) ** 2
                  + (x[None, None, :] - cx) ** 2) - radius
    eps = eps_cells * max(dx, dy, dz)
    return rho_g + (rho_l - rho_g) * heaviside(phi, eps)


def velocity_inputs(grid: BlockGrid, seed: int, wall_faces_zero: bool = True):
    """Predicted face velocities u*, v*, w* for the velocity form of
    newsolv_mg: per block a ghosted (z+2, y+2, x+2) array per component
    (index i holds the face east of cell i, index 0 the face below the
    first cell), drawn block by block, u* then v* then w*, i fastest, over
    the whole ghosted array.  Faces on a physical non-periodic wall are set
    to zero (no flow through walls) unless wall_faces_zero is False."""
    out = ([], [], [])
    pos = 0
    for blk in grid.blocks:
        nx, ny, nz = blk.n
        shp = (nz + 2, ny + 2, nx + 2)
        cnt = shp[0] * shp[1] * shp[2]
        for c in range(3):
            a = uniform_pm1(seed, pos, cnt).reshape(shp)
            pos += cnt
            if wall_faces_zero:
                lo, hi = 2 * c, 2 * c + 1
                n = blk.n[c]
                if blk.nbr[lo] < 0 and blk.bc[lo] != PERIODIC:
                    _axis_plane(a, c, 0)[...] = 0.0
                if blk.nbr[hi] < 0 and blk.bc[hi] != PERIODIC:
                    _axis_plane(a, c, n)[...] = 0.0
            out[c].append(a)
    return out


def _axis_plane(a, axis, idx):
    return a[:, :, idx] if axis == 0 else (a[:, idx, :] if axis == 1 else a[idx, :, :])
