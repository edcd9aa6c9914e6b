"""CPU restatement of the reference's level-set density (TEST INFRASTRUCTURE
ONLY: what the tests compare the library's mf3d_mg_heaviside entry points
against; never imported by the product).

Fields are lists of ghosted block arrays (one ghost layer), shaped
(z+2, y+2, x+2), as in oracle.velocity.

- heaviside (lsm.for:1456-1509): per block epsl = 2*max(dx, dy, dz); over
  every ghosted cell h = 0 where phi < -epsl, 1 where phi > epsl, and
  0.5*(1 + phi/epsl + (1/pi)*sin(pi*phi/epsl)) where |phi| <= epsl, with
  pi = 3.14159265359 (a default real: FP64 under -fdefault-real-8);
  dens = densg + (densl-densg)*h, mu = mug + (mul-mug)*h.  sin is the
  host libm's (Python's math.sin), as the compiled reference calls it.
- bound_lsm(18/19) (bounds_lsm.for:2-104): on each physical non-periodic
  face, in the order E, W, B, T, S, N, the whole ghosted ghost plane takes
  the plane inside it.  The north loop starts one plane early
  (bounds_lsm.for:91, ipl = 0..pl), so the last interior row j = n first
  takes row n-1 and the ghost row then takes that value.
- then exchange(18/19) (exchange.for:2-341): the same exchange_bc +
  exchange_smlvl sequence as any cell field, restated in oracle.velocity.
  For the density the reference's exchange_bc selects no array in its two
  pack loops (exchange_bc.for:105,422: "case (10) !used to be 18"), so with
  a periodic face its result is not defined; the viscosity (op 19) takes
  the defined path, and that is the one restated here for both fields.
"""
from __future__ import annotations

import math

import numpy as np

from .velocity import _plane, exchange

PI = 3.14159265359  # lsm.for:1466
N_EPSL = 2          # lsm.for:1471


def epsl_of(blk):
    """lsm.for:1475: n_epsl*max(dx, dy, dz)."""
    return N_EPSL * max(blk.h)


def h_phi(phi, epsl):
    """The smoothed Heaviside of lsm.for:1480-1492, elementwise, in the
    reference's operand order: 0.5*((1 + phi/epsl) + (1/pi)*sin((pi*phi)/epsl))."""
    phi = np.asarray(phi, dtype=np.float64)
    h = np.empty_like(phi)
    h[phi < -1.0 * epsl] = 0.0
    h[phi > epsl] = 1.0
    band = np.abs(phi) <= epsl
    x = phi[band]
    s = np.array([math.sin(v) for v in ((PI * x) / epsl).tolist()], dtype=np.float64)
    h[band] = 0.5 * ((1.0 + x / epsl) + (1.0 / PI) * s)
    return h


def bound_lsm(grid, F):
    """bounds_lsm.for:30-97 with one ghost layer: zero-gradient ghosts on
    the physical non-periodic faces, E, W, B, T, S, N."""
    for b, blk in enumerate(grid.blocks):
        A = F[b]
        nx, ny, nz = blk.n

        def wall(f):
            return blk.nbr[f] < 0 and blk.bc[f] != 5
        if wall(1):
            _plane(A, 0, nx + 1)[...] = _plane(A, 0, nx)
        if wall(0):
            _plane(A, 0, 0)[...] = _plane(A, 0, 1)
        if wall(4):
            _plane(A, 2, 0)[...] = _plane(A, 2, 1)
        if wall(5):
            _plane(A, 2, nz + 1)[...] = _plane(A, 2, nz)
        if wall(2):
            _plane(A, 1, 0)[...] = _plane(A, 1, 1)
        if wall(3):
            _plane(A, 1, ny)[...] = _plane(A, 1, ny - 1)
            _plane(A, 1, ny + 1)[...] = _plane(A, 1, ny)
    return F


def bound_and_exchange(grid, F):
    """bound_lsm(op) for a cell field: the zero-gradient faces, then
    exchange(op) (bounds_lsm.for:99)."""
    return exchange(grid, bound_lsm(grid, F))


def heaviside(grid, phi, densl, densg, mul, mug):
    """Returns (dens, mu) as lists of ghosted block arrays after heaviside
    and bound_lsm(18), bound_lsm(19)."""
    dens, mu = [], []
    for b, blk in enumerate(grid.blocks):
        h = h_phi(phi[b], epsl_of(blk))
        dens.append(densg + (densl - densg) * h)
        mu.append(mug + (mul - mug) * h)
    return bound_and_exchange(grid, dens), bound_and_exchange(grid, mu)
