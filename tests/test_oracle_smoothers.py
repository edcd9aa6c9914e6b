"""The other two smoothers of mgrelax in the CPU oracle: lexicographic
Gauss-Seidel (mg_itrsch = 1) and ADI line relaxation with TDMA
(mg_itrsch = 2), mgsolver.for:403-545.  Pinned by golden vectors of the
compiled reference (tests/golden/make_golden_sm.py), live randomised
comparisons with it, and what the methods fix without the reference: the
last line (cell) a sweep solves satisfies its equations exactly, and the
exact solution is a fixed point of a sweep."""
import json
import os

import numpy as np
import pytest

from multiflow3d_amd import topology as T
from multiflow3d_amd import workloads as W
from oracle.oracle import OracleMG, have_ref, run_ref

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MAN = json.load(open(os.path.join(GOLDEN, "manifest_sm.json")))


def _split(grid, flat, ghost=0):
    out, o = [], 0
    for b in grid.blocks:
        nx, ny, nz = (x + 2 * ghost for x in b.n)
        m = nx * ny * nz
        out.append(flat[o:o + m].reshape(nz, ny, nx).copy())
        o += m
    return out


def load_sm(name):
    """(grid, manifest entry, arrays) of a smoother case."""
    m = MAN[name]
    g = T.block_grid(m["ex"], m["ey"], m["ez"], m["bc"], m["h"])
    z = np.load(os.path.join(GOLDEN, f"sm_{name}.npz"))
    a = dict(f=_split(g, z["f"]), p=_split(g, z["p"]), pp1=_split(g, z["pp1"], ghost=1), rho=z["rho"],
             rmax=z["rmax"], dens=_split(g, z["dens"]) if m["lsm"] else None)
    return g, m, a


def _oracle(g, m, a):
    o = OracleMG(g.blocks, m["levels"], lsm=m["lsm"], nu_pre=m["nu_pre"], nu_post=m["nu_post"],
                 smoother=m["smoother"])
    if m["lsm"]:
        o.set_density(a["dens"])
    o.coeff()
    return o


@pytest.mark.parametrize("name", sorted(MAN))
def test_smoother_oracle_reproduces_reference_fixture(name):
    g, m, a = load_sm(name)
    o = _oracle(g, m, a)
    pp1 = o.vcycle(a["f"])
    for x, e in zip(pp1, a["pp1"]):
        assert np.array_equal(x, e)
    o2 = _oracle(g, m, a)
    p, rho, rmax, k = o2.solve(a["f"], 0.0, m["cycles"])
    assert k == m["cycles"]
    assert np.array_equal(rho, a["rho"]) and np.array_equal(rmax, a["rmax"])
    for x, e in zip(p, a["p"]):
        assert np.array_equal(x, e)


def test_smoothers_differ_from_each_other():
    """The three smoothers give three different first cycles on the same
    input (the fixtures do not pin one smoother through another)."""
    g, m, a = load_sm("sm1_single16_channel_L3")
    outs = []
    for sm in (1, 2, 3):
        o = OracleMG(g.blocks, 3, smoother=sm)
        o.coeff()
        outs.append(o.vcycle(a["f"])[0])
    assert not np.array_equal(outs[0], outs[1])
    assert not np.array_equal(outs[0], outs[2])
    assert not np.array_equal(outs[1], outs[2])


def _walls_problem(n=(6, 5, 7), seed=3):
    g = T.block_grid([n[0]], [n[1]], [n[2]], [T.NEUMANN] * 6, [0.1, 0.125, 0.08])
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((n[2], n[1], n[0]))
    f -= f.mean()
    return g, f


def _one_sweep(g, f, sm, p0=None):
    o = OracleMG(g.blocks, 1, nu_pre=0, nu_post=0, smoother=sm)
    o.coeff()
    o.rhs(0, 0)[...] = f
    if p0 is not None:
        o.phi(0, 0)[...] = p0
    o.relax(0, 1)
    return o


def test_lexicographic_last_cell_is_exact():
    """Within one sweep the last cell updated, (n0, n1, n2), sees only
    final neighbours (W, S, B updated before it; E, N, T are Neumann
    ghosts with zero coefficient), so its own equation holds to rounding."""
    g, f = _walls_problem()
    o = _one_sweep(g, f, 1)
    r = o.residual(0)[0]
    assert abs(r[-1, -1, -1]) < 1e-12 * np.abs(f).max()
    assert np.abs(r).max() > 1e-3 * np.abs(f).max()  # the sweep is not a solve


def test_adi_last_line_is_exact():
    """The k-line pass solves lines in the order (i, j); the last line
    (n0, n1, :) is solved with final W and S neighbours and wall ghosts, so
    its tridiagonal equations -- the 3-D residual on that line -- hold to
    rounding after the sweep."""
    g, f = _walls_problem()
    o = _one_sweep(g, f, 2)
    r = o.residual(0)[0]
    assert np.abs(r[:, -1, -1]).max() < 1e-11 * np.abs(f).max()
    assert np.abs(r).max() > 1e-3 * np.abs(f).max()


@pytest.mark.parametrize("sm", [1, 2])
def test_exact_solution_is_a_fixed_point(sm):
    """A consistent smoother leaves the solution of A p = f unchanged: solve
    to rounding with many red-black cycles, then one sweep of the other
    smoother moves p by rounding only."""
    g, f = _walls_problem((8, 8, 8), seed=7)
    ref = OracleMG(g.blocks, 2, smoother=3)
    ref.coeff()
    p, rho, _, _ = ref.solve([f], 1e-14, 200)
    assert rho[-1] < 1e-13
    pg = np.zeros((10, 10, 10))
    pg[1:-1, 1:-1, 1:-1] = p[0]
    pg[0], pg[-1] = pg[1], pg[-2]
    pg[:, 0], pg[:, -1] = pg[:, 1], pg[:, -2]
    pg[:, :, 0], pg[:, :, -1] = pg[:, :, 1], pg[:, :, -2]
    o = _one_sweep(g, f, sm, pg)
    d = o.phi(0, 0)[1:-1, 1:-1, 1:-1] - p[0]
    assert np.abs(d).max() < 1e-10 * np.abs(p[0]).max()


@pytest.mark.skipif(not have_ref(), reason="compiled reference not built (make -C oracle ref)")
@pytest.mark.parametrize("seed", range(4))
def test_smoother_oracle_matches_compiled_reference_randomised(seed):
    rng = np.random.default_rng(700 + seed)
    L = int(rng.integers(1, 3))
    div = 1 << (L - 1)
    ex = (rng.integers(2, 6, int(rng.integers(1, 3))) * div).tolist()
    ey = (rng.integers(2, 6, int(rng.integers(1, 3))) * div).tolist()
    ez = (rng.integers(2, 6, int(rng.integers(1, 3))) * div).tolist()
    kinds = [[T.PERIODIC] * 2, [T.NEUMANN] * 2, [T.NEUMANN, 3]]
    bc = sum((kinds[int(rng.integers(0, 3))] for _ in range(3)), [])
    g = T.block_grid(ex, ey, ez, bc, [0.1, 0.12, 0.09])
    lsm = bool(rng.integers(0, 2))
    sm = int(rng.integers(1, 3))
    f = W.split_blocks(g, W.global_rhs(g, W.BASE_SEED + 700 + seed, fv_volume=lsm))
    d = [rng.uniform(1, 50, x.shape) for x in f] if lsm else None
    pr, rho_r, rmax_r, k_r, pp1_r = run_ref(g, L, f, 0.0, 3, lsm=lsm, dens_blocks=d, smoother=sm)
    o = OracleMG(g.blocks, L, lsm=lsm, smoother=sm)
    if lsm:
        o.set_density(d)
    o.coeff()
    p, rho, rmax, k = o.solve(f, 0.0, 3)
    assert k == k_r and np.array_equal(rho, rho_r)
    for x, e in zip(p, pr):
        assert np.array_equal(x, e)
