"""Refined block grids (rdiv > 1) in the CPU oracle.

- Against golden vectors of the compiled reference (make_golden_refined.py):
  rows of 2-3 blocks along x, y or z, rdiv 1/2 and 2/4 side by side,
  walls or a block's own periodic wrap on the other axes, with the
  red-black, lexicographic and ADI smoothers: p and pp1 bitwise, residual
  histories to 1e-14.
- The results do not depend on the 2:1 ghost rows no sender fills (the
  reference rewrites them from stale receive-buffer entries): with those
  rows poisoned with NaN the fixtures' cycles come out unchanged.
- What the mathematics fixes, independently of the reference: across a 2:1
  face the exchange reproduces a field linear along every axis exactly --
  the coarse ghost is the 8-cell average of the first two fine planes, the
  fine ghost (2 phi_c + 3 phi_1 - phi_2) / 4 -- on dyadic values that FP64
  holds exactly; a repeated coarsest grid restricts by injection (coarse
  RHS = fine residual) and prolongs by adding the coarse correction.
- Where oracle/_ref is built: randomised refined rows against the compiled
  reference.
- Block sets outside the covered scope are refused.
"""
import json
import os

import numpy as np
import pytest

from multiflow3d_amd import topology as T
from multiflow3d_amd import workloads as W
from oracle.oracle import OracleMG, have_ref, run_ref

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MAN = json.load(open(os.path.join(GOLDEN, "manifest_refined.json")))
N, P = T.NEUMANN, T.PERIODIC


def _grid(m):
    return T.refined_grid(m["ex"], m["ey"], m["ez"], m["bc"], m["h"], m["rdiv"])


def _split(grid, flat, ghost=0):
    out, o = [], 0
    for b in grid.blocks:
        nx, ny, nz = (x + 2 * ghost for x in b.n)
        m = nx * ny * nz
        out.append(flat[o:o + m].reshape(nz, ny, nx))
        o += m
    return out


def _cat(parts):
    return np.concatenate([a.ravel() for a in parts])


@pytest.mark.parametrize("name", list(MAN))
def test_oracle_reproduces_refined_fixture(name):
    m = MAN[name]
    g = _grid(m)
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    f = _split(g, z["f"])
    assert np.array_equal(_cat(W.block_rhs(g, m["seed"])), z["f"])
    o = OracleMG(g.blocks, m["levels"], smoother=m.get("smoother", 3))
    o.coeff()
    p, rho, rmax, k = o.solve(f, 0.0, m["cycles"])
    assert k == m["cycles"]
    assert np.array_equal(_cat(p), z["p"])
    np.testing.assert_allclose(rho, z["rho"], rtol=1e-14, atol=0)
    np.testing.assert_allclose(rmax, z["rmax"], rtol=1e-14, atol=0)
    assert np.array_equal(_cat(o.vcycle(f)), z["pp1"])
    o.close()


@pytest.mark.parametrize("name", list(MAN))
def test_fixture_independent_of_unfilled_rows(name):
    m = MAN[name]
    g = _grid(m)
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    o = OracleMG(g.blocks, m["levels"], smoother=m.get("smoother", 3))
    o.coeff()
    o.set_poison(True)
    p, rho, _, _ = o.solve(_split(g, z["f"]), 0.0, m["cycles"])
    assert np.array_equal(_cat(p), z["p"])
    o.close()


def _centres(blk, axis):
    """Cell-centre coordinates (ghosts included) along axis, in quarters of
    an unrefined cell from the grid's origin (integers for rdiv 1 and 2)."""
    r = blk.rdiv
    return np.array([4 * blk.origin[axis] + (4 * c - 2) // r for c in range(blk.n[axis] + 2)], dtype=np.float64)


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("rdiv", [(1, 2), (2, 1)])
@pytest.mark.parametrize("along", ["all", "normal"])
def test_exchange_reproduces_linear_fields(axis, rdiv, along):
    """A field linear along every axis reaches the coarse side's interface
    ghosts exactly (the 8-cell average sits at the coarse ghost's centre);
    the fine side gets the coarse cell, not a tangential interpolation, so
    only a field linear along the face normal reaches it exactly."""
    ext = [[4], [4], [4]]
    ext[axis] = [4, 4]
    g = T.refined_grid(*ext, [N] * 6, [1 / 8] * 3, list(rdiv))
    o = OracleMG(g.blocks, 2)  # ngrd_gl 2: one level of the unrefined block
    o.coeff()
    coef = [3.0, -5.0, 7.0] if along == "all" else [3.0 if a == axis else 0.0 for a in range(3)]

    def field(blk):
        xs = [_centres(blk, a) for a in range(3)]
        return (coef[0] * xs[0][None, None, :] + coef[1] * xs[1][None, :, None]
                + coef[2] * xs[2][:, None, None])

    for b, blk in enumerate(g.blocks):
        ph = o.phi(b, 0)
        ph[...] = np.nan
        ph[1:-1, 1:-1, 1:-1] = field(blk)[1:-1, 1:-1, 1:-1]
    o.exchange(0)
    for b, blk in enumerate(g.blocks):
        fine = blk.rdiv > g.blocks[1 - b].rdiv
        if along == "all" and fine:
            continue
        got, exp = o.phi(b, 0), field(blk)
        sl = [slice(1, -1)] * 3  # the interface ghost plane, interior rows (what a 2:1 face sets)
        sl[2 - axis] = -1 if b == 0 else 0
        assert np.array_equal(got[tuple(sl)], exp[tuple(sl)]), (b, blk.rdiv)
    o.close()


def test_repeated_grid_restricts_and_prolongs_by_injection():
    g = T.refined_grid([4, 4], [4], [4], [N] * 6, [1 / 8] * 3, [1, 2])
    o = OracleMG(g.blocks, 4)  # the unrefined block has 3 levels; level 3 repeats its level 2
    o.coeff()
    assert o.dims(0, 3) == o.dims(0, 2) == (1, 1, 1) and o.dims(1, 3) == (1, 1, 1)
    rng = np.random.default_rng(5)
    for b in range(2):
        o.phi(b, 2)[...] = rng.standard_normal(o.phi(b, 2).shape)
        o.rhs(b, 2)[...] = rng.standard_normal(o.rhs(b, 2).shape)
    r2 = o.residual(2)
    o.restrict(3)
    assert np.array_equal(o.rhs(0, 3).ravel(), r2[0].ravel())
    before = o.phi(0, 2).copy()
    o.phi(0, 3)[...] = rng.standard_normal(o.phi(0, 3).shape)
    corr = o.phi(0, 3).copy()
    o.correct(2)
    summed = before + corr
    got = o.phi(0, 2)
    # interior: fine + coarse; Neumann ghosts then copy the interior (mgbound)
    assert got[1, 1, 1] == summed[1, 1, 1]
    o.close()


@pytest.mark.parametrize("blocks", [
    dict(ex=[8, 8], ey=[8], ez=[8], bc=[P] * 6, rdiv=[1, 2]),          # periodic partner blocks
    dict(ex=[8, 8], ey=[8, 8], ez=[8], bc=[N] * 6, rdiv=[1, 2, 1, 1]),  # edge neighbours
    dict(ex=[8, 8], ey=[8], ez=[8], bc=[N] * 6, rdiv=[1, 4]),          # 4:1 across a face
])
def test_out_of_scope_refined_sets_refused(blocks):
    g = T.refined_grid(blocks["ex"], blocks["ey"], blocks["ez"], blocks["bc"], [1 / 16] * 3, blocks["rdiv"])
    with pytest.raises(ValueError):
        OracleMG(g.blocks, 3)


@pytest.mark.skipif(not have_ref(), reason="compiled reference (oracle/_ref) not built here")
@pytest.mark.parametrize("seed", range(6))
def test_refined_oracle_vs_compiled_reference_random(seed):
    rng = np.random.default_rng(2000 + seed)
    axis = int(rng.integers(0, 3))
    nb = int(rng.integers(2, 4))
    lin = int(rng.integers(1, 3))
    q = 1 << lin
    ext = [[q * int(rng.integers(1, 3))] for _ in range(3)]
    ext[axis] = [q * int(rng.integers(1, 3)) for _ in range(nb)]
    rdiv = [1] * nb
    while all(r == rdiv[0] for r in rdiv):
        rdiv = [int(rng.choice([1, 2])) for _ in range(nb)]
    bc = []
    for d in range(3):
        c = N if d == axis else int(rng.choice([P, N]))
        bc += [c, c]
    L = lin + 1
    g = T.refined_grid(*ext, bc, [1.0 / (8 * sum(e)) for e in ext], rdiv)
    f = W.block_rhs(g, 77 + seed)
    p_r, rho_r, _, k, pp1_r = run_ref(g, L, f, 0.0, 3)
    o = OracleMG(g.blocks, L)
    o.coeff()
    p, rho, _, _ = o.solve(f, 0.0, 3)
    assert all(np.array_equal(a, b) for a, b in zip(p, p_r)), (ext, rdiv, bc)
    np.testing.assert_allclose(rho, rho_r, rtol=1e-14, atol=0)
    o.close()
