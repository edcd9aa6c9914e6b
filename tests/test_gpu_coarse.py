"""The FP64 direct solve of the coarsest level (mf3d_mg_cfg.coarse_solve).

- Two levels, no smoothing sweeps: one V-cycle is then restriction, the
  coarse solve and prolongation, so the GPU's pp must equal the oracle's
  restriction and prolongation around a dense numpy solve of the coarse
  operator -- built column by column from the oracle's own residual on
  unit vectors (ghosts by its exchange), one cell pinned to zero as the
  library pins it.  Single block, multi-block with mixed walls, two-phase
  density.
- Whole solves: the direct form reaches the tolerance in fewer cycles than
  the reference's smoothing-only coarsest level, and its pressure equals
  the FP64 oracle's converged pressure (mean removed) to 1e-6 -- both stop
  at a 1e-10 relative residual, and the operator's condition number (about
  1e3 at 64^3) lets their smooth error components differ by that much.
- The mixed precision's default keeps the reference's sweeps on the
  coarsest level, computed in FP64: its pressure stays within 1e-8 of the
  FP64 form's at the same tolerance (the same path); "inner" sweeps in FP32
  and also stays on that path; "direct" takes fewer cycles.
- A coarsest level beyond MF3D_COARSEST_MAX_CELLS is refused when the
  direct form is asked for.
"""
import warnings

import numpy as np
import pytest

from multiflow3d_amd import MGError, NotConverged, PressureMultigrid
from multiflow3d_amd import topology as T
from oracle.oracle import OracleMG

from .cases import case, inputs

pytestmark = pytest.mark.gpu


def _dense_coarse(o, grid):
    """The oracle's level-1 operator as a dense matrix (cells block by
    block, i fastest): column q = -(residual of unit vector q with a zero
    right-hand side), ghosts by mgbound + exchange."""
    dims = [o.dims(b, 1) for b in range(len(grid.blocks))]
    sizes = [int(np.prod(d)) for d in dims]
    nc = sum(sizes)
    A = np.zeros((nc, nc))
    for q in range(nc):
        col = 0
        for b, (nx, ny, nz) in enumerate(dims):
            ph = o.phi(b, 1)
            ph[...] = 0.0
            o.rhs(b, 1)[...] = 0.0
        b, r = 0, q
        while r >= sizes[b]:
            r -= sizes[b]
            b += 1
        nx, ny, nz = dims[b]
        k, rem = divmod(r, nx * ny)
        j, i = divmod(rem, nx)
        o.phi(b, 1)[k + 1, j + 1, i + 1] = 1.0
        o.bound(1)
        o.exchange(1)
        res = np.concatenate([x.ravel() for x in o.residual(1)])
        A[:, q] = -res
        col += 1
    return A, dims, sizes


def _two_level_reference(grid, lsm, dens, rhs):
    """pp of one V-cycle with no sweeps and an exact coarse solve, from the
    oracle's restriction and prolongation."""
    o = OracleMG(grid.blocks, 2, lsm=lsm, nu_pre=0, nu_post=0)
    if lsm:
        o.set_density(dens)
    o.coeff()
    A, dims, sizes = _dense_coarse(o, grid)
    for b in range(len(grid.blocks)):
        o.phi(b, 0)[...] = 0.0
        o.rhs(b, 0)[...] = rhs[b]
    o.restrict(1)
    bvec = np.concatenate([o.rhs(b, 1).ravel().copy() for b in range(len(grid.blocks))])
    nc = len(bvec)
    x = np.zeros(nc)
    x[:-1] = np.linalg.solve(A[:-1, :-1], bvec[:-1])  # the last cell pinned (one connected group)
    o2 = 0
    for b, (nx, ny, nz) in enumerate(dims):
        ph = o.phi(b, 1)
        ph[...] = 0.0
        ph[1:-1, 1:-1, 1:-1] = x[o2:o2 + sizes[b]].reshape(nz, ny, nx)
        o2 += sizes[b]
    o.bound(1)
    o.exchange(1)
    o.correct(0)
    out = [o.phi(b, 0).copy() for b in range(len(grid.blocks))]
    o.close()
    return out


@pytest.mark.parametrize("name", ["single16_L3", "mixed_bc_multi_L3", "lsm_multi_L3"])
def test_two_level_cycle_is_restrict_solve_prolong(name):
    grid, _, lsm = case(name)
    f, dens = inputs(grid, lsm, seed=21)
    print(f"[coarse direct two-level] {name}", flush=True)
    exp = _two_level_reference(grid, lsm, dens, f)
    with PressureMultigrid(grid, 2, nu_pre=0, nu_post=0, variable_coef=lsm, coarse_solve="direct") as mg:
        if lsm:
            mg.coeff(dens)
        pp = [np.zeros_like(e) for e in exp]
        mg.mgkcyc(f, pp)
    num = np.sqrt(sum(float(np.sum((a - e) ** 2)) for a, e in zip(pp, exp)))
    den = np.sqrt(sum(float(np.sum(e ** 2)) for e in exp))
    assert den > 0 and num / den < 1e-11, num / den


def _solve(grid, L, lsm, dens, f, **kw):
    with PressureMultigrid(grid, L, variable_coef=lsm, maxcy=40, **kw) as mg:
        if lsm:
            mg.coeff(dens)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", NotConverged)
            return mg.newsolv_mg(f)


def _mean_free(parts):
    a = np.concatenate([x.ravel() for x in parts])
    return a - a.mean()


@pytest.mark.parametrize("name", ["single64_L5", "multi_uneven_L3"])
def test_direct_coarse_converges_faster_to_the_same_pressure(name):
    grid, L, lsm = case(name)
    f, dens = inputs(grid, lsm, seed=5)
    print(f"[coarse direct solve] {name}", flush=True)
    o = OracleMG(grid.blocks, L)
    o.coeff()
    p_o, rho_o, _, k_o = o.solve(f, 1e-11, 60)
    o.close()
    p_s, r_s = _solve(grid, L, lsm, dens, f, eps=1e-10)
    p_d, r_d = _solve(grid, L, lsm, dens, f, eps=1e-10, coarse_solve="direct")
    assert r_s.status == 0 and r_d.status == 0
    assert r_d.iters < r_s.iters, (r_d.iters, r_s.iters)
    ref = _mean_free(p_o)
    for got in (p_s, p_d):
        g = _mean_free(got)
        assert np.linalg.norm(g - ref) / np.linalg.norm(ref) < 1e-6


@pytest.mark.parametrize("name", ["single64_L5", "multi_uneven_L3"])
def test_mixed_coarsest_forms(name):
    """Mixed precision: the default (FP64 sweeps on the coarsest level, a
    tail kernel on lone blocks, per-level launches on block grids) and the
    FP32 sweeps both stay on the FP64 form's path; the direct solve takes
    no more cycles than the FP64 form."""
    grid, L, lsm = case(name)
    f, dens = inputs(grid, lsm, seed=6)
    print(f"[coarsest forms mixed] {name}", flush=True)
    kw = {}
    p64, r64 = _solve(grid, L, lsm, dens, f, eps=1e-9, **kw)
    pdf, rdf = _solve(grid, L, lsm, dens, f, eps=1e-9, precision="mixed", **kw)
    pin, rin = _solve(grid, L, lsm, dens, f, eps=1e-9, precision="mixed", coarse_solve="inner", **kw)
    pdi, rdi = _solve(grid, L, lsm, dens, f, eps=1e-9, precision="mixed", coarse_solve="direct", **kw)
    assert r64.status == rdf.status == rin.status == rdi.status == 0
    ref = _mean_free(p64)
    for got in (pdf, pin):
        assert np.linalg.norm(_mean_free(got) - ref) / np.linalg.norm(ref) < 1e-8
    assert rdf.iters == r64.iters
    assert rdi.iters <= r64.iters


def test_direct_refused_beyond_its_size():
    g = T.single_block([128, 128, 128], [1 / 128] * 3, T.CHANNEL_BC)  # L = 3: 32^3 coarsest cells
    with pytest.raises(MGError, match="direct solve"):
        PressureMultigrid(g, 3, coarse_solve="direct")
    with PressureMultigrid(g, 3, precision="mixed") as mg:  # the default falls back to smoothing
        assert mg.cells == 128 ** 3
