"""CPU restatement of the velocity form of newsolv_mg (TEST INFRASTRUCTURE
ONLY: what the tests compare the library's mf3d_mg_newsolv against; never
imported by the product).

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
- calmas, 4th order (differencing = 2, press.for:13-36): fact = sum over
  the axes of (-u(i+1) + 27 u(i) - 27 u(i-1) + u(i-2)) / (24 dx); su =
  fact/dt (su even under LSM, which the reference only runs with WENO).
- The pressure update of newsolv_mg.for:38-51: p += pp over the one-cell
  ghosted box, then exchange(4) + pbound (press.for:77-137).

Ghost width: the arrays may carry G >= 1 ghost layers (the reference's pl;
G is read off the array shape), index 0 / n+1 being the first layer as
above and -(G-1) / n+G the outermost.  exchange(1..3) then refreshes
ly = 0..pl_ex layers of u, v, w (exchange.for:14-23), exchange(4) one
(nly = 0); the periodic planes span the whole ghosted array, the corner
and edge cubes of exchange_smlvl are (pl+1) deep (exchange.for:664-1210,
2118-2389, at ly = 0 only), the faces write interior rows
(exchange.for:2394-2484) -- so even with pl_ex = 0 entries of the outer
layers change.  pbound and the pressure update touch the one-cell box.
"""
from __future__ import annotations

import numpy as np

from .oracle import OracleMG


def _G(blk, A):
    """Ghost layers of array A of block blk (shape (nz+2G, ny+2G, nx+2G))."""
    return (A.shape[2] - blk.n[0]) // 2


def _ix(G, c):
    """Array index of coordinate c (interior 1..n, ghosts 0, -1, ... and
    n+1, n+2, ...) in an array with G ghost layers."""
    return c + G - 1


def _sl(G, lo, hi):
    """Slice of the coordinates lo..hi (inclusive)."""
    return slice(lo + G - 1, hi + G)


def _faces(blk, axis):
    """Interior face range 1..n on an axis, minus the last face on a
    physical non-periodic wall without a neighbour (the wall face)."""
    n = blk.n[axis]
    hi = n
    f = 2 * axis + 1
    if blk.nbr[f] < 0 and blk.bc[f] != 5:
        hi = n - 1
    return 1, hi


def calvel(grid, ustar, vstar, wstar, u, v, w, p, dt, alfapr, dens=None, pl_ex=0):
    """dens (LSM): one-ghost-layer cell densities; the fields may carry G
    layers.  The exchange then refreshes pl_ex extra layers of u, v, w."""
    dta = dt * alfapr
    for b, blk in enumerate(grid.blocks):
        dx, dy, dz = blk.h
        nx, ny, nz = blk.n
        P = p[b]
        G = _G(blk, P)
        D = dens[b] if dens is not None else None
        lo, hi = _faces(blk, 0)
        i, ip = _sl(G, lo, hi), _sl(G, lo + 1, hi + 1)
        J, K = _sl(G, 1, ny), _sl(G, 1, nz)
        di, dip, dJ, dK = _sl(1, lo, hi), _sl(1, lo + 1, hi + 1), _sl(1, 1, ny), _sl(1, 1, nz)
        g = dta * (P[K, J, ip] - P[K, J, i]) / dx
        if D is not None:
            g = g / (0.5 * (D[dK, dJ, di] + D[dK, dJ, dip]))
        u[b][K, J, i] = ustar[b][K, J, i] - g
        lo, hi = _faces(blk, 1)
        j, jp = _sl(G, lo, hi), _sl(G, lo + 1, hi + 1)
        I = _sl(G, 1, nx)
        dj, djp, dI = _sl(1, lo, hi), _sl(1, lo + 1, hi + 1), _sl(1, 1, nx)
        g = dta * (P[K, jp, I] - P[K, j, I]) / dy
        if D is not None:
            g = g / (0.5 * (D[dK, dj, dI] + D[dK, djp, dI]))
        v[b][K, j, I] = vstar[b][K, j, I] - g
        lo, hi = _faces(blk, 2)
        k, kp = _sl(G, lo, hi), _sl(G, lo + 1, hi + 1)
        dk, dkp = _sl(1, lo, hi), _sl(1, lo + 1, hi + 1)
        g = dta * (P[kp, J, I] - P[k, J, I]) / dz
        if D is not None:
            g = g / (0.5 * (D[dk, dJ, dI] + D[dkp, dJ, dI]))
        w[b][k, J, I] = wstar[b][k, J, I] - g
    exchange(grid, u, 0, pl_ex)
    exchange(grid, v, 1, pl_ex)
    exchange(grid, w, 2, pl_ex)


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


def _plane(a, axis, idx, G=1):
    """View of the full ghosted plane at coordinate idx of axis (x = last
    array axis), in an array with G ghost layers."""
    idx = _ix(G, idx)
    if axis == 0:
        return a[:, :, idx]
    if axis == 1:
        return a[:, idx, :]
    return a[idx, :, :]


def exchange_bc(grid, F, ie_of, nly=0):
    """exchange_bc.for:14-841 on one rank, layer by layer ly = 0..nly
    (:16): for every periodic face without a neighbour, in block order and
    face order W, E, S, N, B, T, either copy the block's own opposite plane
    (a block alone along that axis: ghost plane 0-ly <- plane ie-ly, plane
    ie+1+ly <- plane 1+ly) or hand the partner its plane 1+ly (ie-ly); the
    full ghosted plane each time.  The send loop runs twice (:16-330 posts,
    :332-640 sends again), so a plane handed over is taken after the first
    loop's self copies.  The unpack (:642-838) then writes every received
    plane."""
    for ly in range(nly + 1):
        bufs = {}
        for _ in range(2):
            for b, blk in enumerate(grid.blocks):
                A = F[b]
                G = _G(blk, A)
                for axis in range(3):
                    for side in range(2):
                        f = 2 * axis + side
                        if not (blk.bc[f] == 5 and blk.nbr[f] < 0):
                            continue
                        ie = ie_of(b, axis)
                        if blk.nbr[2 * axis + 1 - side] < 0:
                            if side == 0:
                                _plane(A, axis, 0 - ly, G)[...] = _plane(A, axis, ie - ly, G)
                            else:
                                _plane(A, axis, ie + 1 + ly, G)[...] = _plane(A, axis, 1 + ly, G)
                        else:
                            part = blk.per[f]
                            src = 1 + ly if side == 0 else ie - ly
                            bufs[(part, 2 * axis + 1 - side)] = _plane(A, axis, src, G).copy()
        for b, blk in enumerate(grid.blocks):
            A = F[b]
            G = _G(blk, A)
            for axis in range(3):
                for side in range(2):
                    f = 2 * axis + side
                    if blk.bc[f] == 5 and blk.nbr[f] < 0 and blk.nbr[2 * axis + 1 - side] >= 0:
                        ie = ie_of(b, axis)
                        _plane(A, axis, 0 - ly if side == 0 else ie + 1 + ly, G)[...] = bufs[(b, f)]


# corner and edge directions in the reference's unpack order
# (exchange.for:2118-2389): corprev1..4, cornext1..4, edgprev1..6, edgnext1..6
_CORNERS = [(-1, -1, -1), (-1, 1, -1), (1, 1, -1), (1, -1, -1),
            (1, 1, 1), (1, -1, 1), (-1, -1, 1), (-1, 1, 1)]
_EDGES = [(-1, 0, -1), (0, 1, -1), (1, 0, -1), (0, -1, -1), (-1, -1, 0), (-1, 1, 0),
          (1, 0, 1), (0, -1, 1), (-1, 0, 1), (0, 1, 1), (1, 1, 0), (1, -1, 0)]


def exchange_smlvl(grid, F, ie_of=None, nly=0):
    """exchange.for:342-2499, same resolution, one rank, layer by layer
    ly = 0..nly (:362).  Every block packs before any block unpacks, so
    all values sent are the ones after the previous layer (and
    exchange_bc).  Per receiving block the unpack order is the 8 corners
    and the 12 edges (ly = 0 only), then the 6 faces.  Corners and edges
    carry (pl+1)-deep cubes (pl = G): along an axis where the sender lies
    below (t = -1) the receiver's index 1-q takes the sender's n+1-q, where
    it lies above (t = +1) the receiver's n+q takes the sender's q,
    q = 0..pl (exchange.for:664-1210 pack, 2118-2389 unpack); an entry is
    written when some q >= 1, i.e. when it lies outside the receiver's
    interior.  Edges run over the interior cells of their free axis.
    Faces write the interior rows of ghost plane 0-ly (n+1+ly) from the
    sender's plane ie-ly (1+ly) (:2394-2484), ie the field's last index on
    that axis."""
    if ie_of is None:
        def ie_of(b, axis):
            return grid.blocks[b].n[axis]
    for ly in range(nly + 1):
        pre = [a.copy() for a in F]
        for d, blk in enumerate(grid.blocks):
            A = F[d]
            G = _G(blk, A)
            n = blk.n
            if ly == 0:
                for t in _CORNERS + _EDGES:
                    s = _diag(grid, d, t)
                    if s < 0:
                        continue
                    ns = grid.blocks[s].n
                    Gs = _G(grid.blocks[s], pre[s])
                    axes = [a for a in range(3) if t[a] != 0]
                    free = [a for a in range(3) if t[a] == 0]
                    for qs in np.ndindex(*([G + 1] * len(axes))):
                        if not any(qs):
                            continue
                        di = [None] * 3
                        si = [None] * 3
                        for a, q in zip(axes, qs):
                            if t[a] < 0:
                                di[a], si[a] = _ix(G, 1 - q), _ix(Gs, ns[a] + 1 - q)
                            else:
                                di[a], si[a] = _ix(G, n[a] + q), _ix(Gs, q)
                        for a in free:
                            di[a], si[a] = _sl(G, 1, n[a]), _sl(Gs, 1, n[a])
                        A[di[2], di[1], di[0]] = pre[s][si[2], si[1], si[0]]
            for axis in range(3):
                for side in range(2):
                    s = blk.nbr[2 * axis + side]
                    if s < 0:
                        continue
                    Gs = _G(grid.blocks[s], pre[s])
                    di = [_sl(G, 1, n[a]) for a in range(3)]
                    si = [_sl(Gs, 1, n[a]) for a in range(3)]
                    di[axis] = _ix(G, 0 - ly if side == 0 else ie_of(d, axis) + 1 + ly)
                    si[axis] = _ix(Gs, ie_of(s, axis) - ly if side == 0 else 1 + ly)
                    A[di[2], di[1], di[0]] = pre[s][si[2], si[1], si[0]]


def exchange(grid, F, comp=None, nly=0):
    """exchange(op) (exchange.for:2-341) for u (comp 0), v (1), w (2) or a
    cell field (None): exchange_bc if any face is periodic, then
    exchange_smlvl, over the layers ly = 0..nly (pl_ex for u, v, w; 0 for
    p, exchange.for:14-23).  The rdiv > 1 paths are out of scope."""
    def ie_of(b, axis):
        if comp is not None and axis == comp:
            return _faces(grid.blocks[b], axis)[1]
        return grid.blocks[b].n[axis]
    if any(5 in blk.bc for blk in grid.blocks):
        exchange_bc(grid, F, ie_of, nly)
    exchange_smlvl(grid, F, ie_of, nly)
    return F


def pbound(grid, P):
    """press.for:77-137: on each physical non-periodic face, in the order
    W, E, S, N, B, T, the ghost plane of the one-cell box (isp-1..iep+1 on
    the other axes) takes the adjacent interior plane."""
    for b, blk in enumerate(grid.blocks):
        A = P[b]
        G = _G(blk, A)
        for axis in range(3):
            n = blk.n[axis]
            box = [_sl(G, 0, blk.n[a] + 1) for a in range(3)]

            def pl_(c):
                ix = list(box)
                ix[axis] = _ix(G, c)
                return (ix[2], ix[1], ix[0])
            if blk.nbr[2 * axis] < 0 and blk.bc[2 * axis] != 5:
                A[pl_(0)] = A[pl_(1)]
            if blk.nbr[2 * axis + 1] < 0 and blk.bc[2 * axis + 1] != 5:
                A[pl_(n + 1)] = A[pl_(n)]
    return P


def calmas(grid, u, v, w, dt, lsm=False, differencing=1):
    """Returns (su interior blocks, rmax).  differencing = 2: the 4th-order
    branch (press.for:13-36, needs two ghost layers), which writes su (not
    sup) and keeps the literal 27.0 and 24.0 in double precision."""
    out, rmax = [], 0.0
    for b, blk in enumerate(grid.blocks):
        dx, dy, dz = blk.h
        nx, ny, nz = blk.n
        G = _G(blk, u[b])
        I, J, K = _sl(G, 1, nx), _sl(G, 1, ny), _sl(G, 1, nz)
        if differencing == 2:
            assert G >= 2, "4th-order calmas reads two ghost layers"
            U, V_, W_ = u[b], v[b], w[b]

            def sh(a, lo):  # coordinates lo .. lo+n-1 on one axis
                return _sl(G, lo, lo + [nx, ny, nz][a] - 1)
            fact = ((((-U[K, J, sh(0, 2)] + 27.0 * U[K, J, I]) - 27.0 * U[K, J, sh(0, 0)]) + U[K, J, sh(0, -1)])
                    / (24.0 * dx)
                    + (((-V_[K, sh(1, 2), I] + 27.0 * V_[K, J, I]) - 27.0 * V_[K, sh(1, 0), I]) + V_[K, sh(1, -1), I])
                    / (24.0 * dy)) \
                + (((-W_[sh(2, 2), J, I] + 27.0 * W_[K, J, I]) - 27.0 * W_[sh(2, 0), J, I]) + W_[sh(2, -1), J, I]) \
                / (24.0 * dz)
            vol = fact * dx * dy * dz
            out.append(fact / dt)
            rmax = max(rmax, float(np.abs(vol).max()))
            continue
        Im, Jm, Km = _sl(G, 0, nx - 1), _sl(G, 0, ny - 1), _sl(G, 0, nz - 1)
        fact = ((u[b][K, J, I] - u[b][K, J, Im]) / dx + (v[b][K, J, I] - v[b][K, Jm, I]) / dy) \
            + (w[b][K, J, I] - w[b][Km, J, I]) / dz
        vol = fact * dx * dy * dz
        out.append(vol / dt if lsm else fact / dt)
        rmax = max(rmax, float(np.abs(vol).max()))
    return out, rmax


def newsolv_velocity(grid, levels, ustar, vstar, wstar, p0, eps, maxcy, dt, alfapr, lsm=False,
                     dens=None, coarse_op="reference", smoother=3, nu_pre=2, nu_post=2, pl_ex=0,
                     differencing=1):
    """newsolv_mg.for:12-58 in velocity form (mg_itrsch = smoother, irestr
    = nu_pre, iproln = nu_post).  Returns (iterations, rmax history, p, u,
    v, w, su) with ghosted block arrays (G layers, as given; dens: one
    layer)."""
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
    calvel(grid, ustar, vstar, wstar, u, v, w, p, dt, alfapr, dn, pl_ex)
    su, rm = calmas(grid, u, v, w, dt, lsm, differencing)
    hist = [rm]
    it = maxcy
    for k in range(1, maxcy + 1):
        pp = o.vcycle(su)
        for blk, a, q in zip(grid.blocks, p, pp):  # the one-cell box (newsolv_mg.for:38-51)
            G = _G(blk, a)
            box = tuple(_sl(G, 0, blk.n[ax] + 1) for ax in (2, 1, 0))
            a[box] = a[box] + (q - 0.0)
        pbound(grid, exchange(grid, p))
        calvel(grid, ustar, vstar, wstar, u, v, w, p, dt, alfapr, dn, pl_ex)
        su, rm = calmas(grid, u, v, w, dt, lsm, differencing)
        hist.append(rm)
        if rm < eps and k > 1:
            it = k
            break
    o.close()
    return it, np.array(hist), p, u, v, w, su
