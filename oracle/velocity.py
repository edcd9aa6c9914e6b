"""CPU restatement of the velocity form of newsolv_mg (TEST INFRASTRUCTURE
ONLY: the checker for the library's mf3d_mg_newsolv; never imported by the
product).

Fields are lists of ghosted block arrays (one ghost layer), shaped
(z+2, y+2, x+2); index 0 and n+1 are the ghost layer.  A face field u at
index i is the face east of cell i (u(i,j,k) between cells i and i+1),
likewise v (north) and w (top).  All arithmetic is FP64 in the reference's
operand order (numpy never contracts to FMA).

- calvel (press.for:139-200): u = ustar - dt*alfapr*(p(i+1)-p(i))/dx, divided
  by the face density 0.5*(dens(i)+dens(i+1)) under LSM, on the faces
  isu..ieu (1..n, minus the face on a physical non-periodic east wall,
  localparameters.for:185-190), then exchange(1..3).
- exchange(1..3) (exchange.for:2-341, exchange_bc.for:2-841): restated
  step by step, because its edges and corners keep values of earlier
  ghosts (periodic planes are copied whole, ghost rows included).
- calmas (press.for:38-72), 2nd order: fact = (u(i)-u(i-1))/dx +
  (v(j)-v(j-1))/dy + (w(k)-w(k-1))/dz; su = fact/dt (LSM: sup =
  fact*dx*dy*dz/dt); rmax = max |fact*dx*dy*dz|.
- The pressure update of newsolv_mg.for:38-51: p += pp over every ghosted
  cell, then exchange(4) + pbound (press.for:77-137).
"""
from __future__ import annotations

import numpy as np

from .oracle import OracleMG


def _faces(blk, axis):
    """Interior face range 1..n on an axis, minus the last face on a
    physical non-periodic wall without a neighbour (the wall face)."""
    n = blk.n[axis]
    hi = n
    f = 2 * axis + 1
    if blk.nbr[f] < 0 and blk.bc[f] != 5:
        hi = n - 1
    return 1, hi


def calvel(grid, ustar, vstar, wstar, u, v, w, p, dt, alfapr, dens=None):
    dta = dt * alfapr
    for b, blk in enumerate(grid.blocks):
        dx, dy, dz = blk.h
        nx, ny, nz = blk.n
        P = p[b]
        lo, hi = _faces(blk, 0)
        i = slice(lo, hi + 1)
        ip = slice(lo + 1, hi + 2)
        J, K = slice(1, ny + 1), slice(1, nz + 1)
        g = dta * (P[K, J, ip] - P[K, J, i]) / dx
        if dens is not None:
            g = g / (0.5 * (dens[b][K, J, i] + dens[b][K, J, ip]))
        u[b][K, J, i] = ustar[b][K, J, i] - g
        lo, hi = _faces(blk, 1)
        j, jp = slice(lo, hi + 1), slice(lo + 1, hi + 2)
        I = slice(1, nx + 1)
        g = dta * (P[K, jp, I] - P[K, j, I]) / dy
        if dens is not None:
            g = g / (0.5 * (dens[b][K, j, I] + dens[b][K, jp, I]))
        v[b][K, j, I] = vstar[b][K, j, I] - g
        lo, hi = _faces(blk, 2)
        k, kp = slice(lo, hi + 1), slice(lo + 1, hi + 2)
        g = dta * (P[kp, J, I] - P[k, J, I]) / dz
        if dens is not None:
            g = g / (0.5 * (dens[b][k, J, I] + dens[b][kp, J, I]))
        w[b][k, J, I] = wstar[b][k, J, I] - g
    exchange(grid, u, 0)
    exchange(grid, v, 1)
    exchange(grid, w, 2)


def _diag(grid, b, t):
    """Block in direction t (components -1, 0, +1) through chained face
    neighbours, or -1.  On a conforming block grid this is the edge or
    corner neighbour that alloc_dom.for:617-1163 finds by matching
    coordinates (no periodic wrap)."""
    c = b
    for a in range(3):
        if t[a] == 0:
            continue
        c = grid.blocks[c].nbr[2 * a + (1 if t[a] > 0 else 0)]
        if c < 0:
            return -1
    return c


def _plane(a, axis, idx):
    """View of the full ghosted plane idx of axis (x = last array axis)."""
    if axis == 0:
        return a[:, :, idx]
    if axis == 1:
        return a[:, idx, :]
    return a[idx, :, :]


def exchange_bc(grid, F, ie_of):
    """exchange_bc.for:14-841 with one ghost layer on one rank: for every
    periodic face without a neighbour, in block order and face order W, E,
    S, N, B, T, either copy the block's own opposite plane (a block alone
    along that axis) or hand the partner its boundary plane; the full
    ghosted plane each time.  The send loop runs twice (:16-330 posts,
    :332-640 sends again), so a plane handed over is taken after the first
    loop's self copies.  The unpack (:642-838) then writes every received
    plane."""
    bufs = {}
    for _ in range(2):
        for b, blk in enumerate(grid.blocks):
            A = F[b]
            for axis in range(3):
                for side in range(2):
                    f = 2 * axis + side
                    if not (blk.bc[f] == 5 and blk.nbr[f] < 0):
                        continue
                    ie = ie_of(b, axis)
                    if blk.nbr[2 * axis + 1 - side] < 0:
                        if side == 0:
                            _plane(A, axis, 0)[...] = _plane(A, axis, ie)
                        else:
                            _plane(A, axis, ie + 1)[...] = _plane(A, axis, 1)
                    else:
                        part = blk.per[f]
                        src = 1 if side == 0 else ie
                        bufs[(part, 2 * axis + 1 - side)] = _plane(A, axis, src).copy()
    for b, blk in enumerate(grid.blocks):
        A = F[b]
        for axis in range(3):
            for side in range(2):
                f = 2 * axis + side
                if blk.bc[f] == 5 and blk.nbr[f] < 0 and blk.nbr[2 * axis + 1 - side] >= 0:
                    ie = ie_of(b, axis)
                    _plane(A, axis, 0 if side == 0 else ie + 1)[...] = bufs[(b, f)]


# corner and edge directions in the reference's unpack order
# (exchange.for:2118-2389): corprev1..4, cornext1..4, edgprev1..6, edgnext1..6
_CORNERS = [(-1, -1, -1), (-1, 1, -1), (1, 1, -1), (1, -1, -1),
            (1, 1, 1), (1, -1, 1), (-1, -1, 1), (-1, 1, 1)]
_EDGES = [(-1, 0, -1), (0, 1, -1), (1, 0, -1), (0, -1, -1), (-1, -1, 0), (-1, 1, 0),
          (1, 0, 1), (0, -1, 1), (-1, 0, 1), (0, 1, 1), (1, 1, 0), (1, -1, 0)]


def exchange_smlvl(grid, F):
    """exchange.for:342-2499, same resolution, one ghost layer, one rank.
    Every block packs before any block unpacks, so all values sent are the
    ones after exchange_bc.  Per receiving block the unpack order is the
    8 corners, the 12 edges, then the 6 faces.  Along an axis where the
    sender lies below (t = -1) the receiver's index 1-q takes the sender's
    n+1-q, where it lies above (t = +1) the receiver's n+q takes the
    sender's q, q = 0, 1 (exchange.for:664-1210 pack, 2118-2389 unpack);
    an entry is written when some q = 1, i.e. when it lies in the
    receiver's ghost layer.  Edges run over the interior cells of their
    free axis; faces write the interior rows of the ghost plane from the
    sender's boundary plane (:2394-2484)."""
    pre = [a.copy() for a in F]
    for d, blk in enumerate(grid.blocks):
        A = F[d]
        n = blk.n
        for t in _CORNERS + _EDGES:
            s = _diag(grid, d, t)
            if s < 0:
                continue
            ns = grid.blocks[s].n
            axes = [a for a in range(3) if t[a] != 0]
            free = [a for a in range(3) if t[a] == 0]
            for qs in np.ndindex(*([2] * len(axes))):
                if not any(qs):
                    continue
                di = [None] * 3
                si = [None] * 3
                for a, q in zip(axes, qs):
                    if t[a] < 0:
                        di[a], si[a] = 1 - q, ns[a] + 1 - q
                    else:
                        di[a], si[a] = n[a] + q, q
                for a in free:
                    di[a] = si[a] = slice(1, n[a] + 1)
                A[di[2], di[1], di[0]] = pre[s][si[2], si[1], si[0]]
        for axis in range(3):
            for side in range(2):
                s = blk.nbr[2 * axis + side]
                if s < 0:
                    continue
                ns = grid.blocks[s].n
                di = [slice(1, n[a] + 1) for a in range(3)]
                si = list(di)
                di[axis] = 0 if side == 0 else n[axis] + 1
                si[axis] = ns[axis] if side == 0 else 1
                A[di[2], di[1], di[0]] = pre[s][si[2], si[1], si[0]]


def exchange(grid, F, comp=None):
    """exchange(op) (exchange.for:2-341) for u (comp 0), v (1), w (2) or a
    cell field (None): exchange_bc if any face is periodic, then
    exchange_smlvl.  The LMR = 2 branch and the rdiv > 1 paths are out of
    scope."""
    def ie_of(b, axis):
        if comp is not None and axis == comp:
            return _faces(grid.blocks[b], axis)[1]
        return grid.blocks[b].n[axis]
    if any(5 in blk.bc for blk in grid.blocks):
        exchange_bc(grid, F, ie_of)
    exchange_smlvl(grid, F)
    return F


def pbound(grid, P):
    """press.for:77-137: on each physical non-periodic face, in the order
    W, E, S, N, B, T, the full ghosted plane takes the adjacent interior
    plane."""
    for b, blk in enumerate(grid.blocks):
        A = P[b]
        for axis in range(3):
            n = blk.n[axis]
            if blk.nbr[2 * axis] < 0 and blk.bc[2 * axis] != 5:
                _plane(A, axis, 0)[...] = _plane(A, axis, 1)
            if blk.nbr[2 * axis + 1] < 0 and blk.bc[2 * axis + 1] != 5:
                _plane(A, axis, n + 1)[...] = _plane(A, axis, n)
    return P


def calmas(grid, u, v, w, dt, lsm=False):
    """Returns (su interior blocks, rmax)."""
    out, rmax = [], 0.0
    for b, blk in enumerate(grid.blocks):
        dx, dy, dz = blk.h
        nx, ny, nz = blk.n
        I, J, K = slice(1, nx + 1), slice(1, ny + 1), slice(1, nz + 1)
        Im, Jm, Km = slice(0, nx), slice(0, ny), slice(0, nz)
        fact = ((u[b][K, J, I] - u[b][K, J, Im]) / dx + (v[b][K, J, I] - v[b][K, Jm, I]) / dy) \
            + (w[b][K, J, I] - w[b][Km, J, I]) / dz
        vol = fact * dx * dy * dz
        out.append(vol / dt if lsm else fact / dt)
        rmax = max(rmax, float(np.abs(vol).max()))
    return out, rmax


def newsolv_velocity(grid, levels, ustar, vstar, wstar, p0, eps, maxcy, dt, alfapr, lsm=False,
                     dens=None, coarse_op="reference", smoother=3, nu_pre=2, nu_post=2):
    """newsolv_mg.for:12-58 in velocity form (mg_itrsch = smoother, irestr
    = nu_pre, iproln = nu_post).  Returns (iterations, rmax history, p, u,
    v, w, su) with ghosted block arrays."""
    o = OracleMG(grid.blocks, levels, lsm=lsm, coarse_op=coarse_op, smoother=smoother, nu_pre=nu_pre,
                 nu_post=nu_post)
    if lsm:
        o.set_density([d[1:-1, 1:-1, 1:-1] for d in dens])
    o.coeff()
    p = [a.copy() for a in p0]
    u = [a.copy() for a in ustar]
    v = [a.copy() for a in vstar]
    w = [a.copy() for a in wstar]
    dn = dens if lsm else None
    calvel(grid, ustar, vstar, wstar, u, v, w, p, dt, alfapr, dn)
    su, rm = calmas(grid, u, v, w, dt, lsm)
    hist = [rm]
    it = maxcy
    for k in range(1, maxcy + 1):
        pp = o.vcycle(su)
        p = [a + (q - 0.0) for a, q in zip(p, pp)]
        pbound(grid, exchange(grid, p))
        calvel(grid, ustar, vstar, wstar, u, v, w, p, dt, alfapr, dn)
        su, rm = calmas(grid, u, v, w, dt, lsm)
        hist.append(rm)
        if rm < eps and k > 1:
            it = k
            break
    o.close()
    return it, np.array(hist), p, u, v, w, su
