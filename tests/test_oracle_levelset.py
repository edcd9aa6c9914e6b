"""The level-set density in the CPU oracle (oracle/levelset.py): against
golden vectors written by the compiled reference (heaviside, bound_lsm and
the chain into newsolv_mg; tests/golden/make_golden_ls.py), and against what
the formula fixes without the reference (closed-form values, symmetry,
monotonicity, slope at the interface)."""
import json
import math
import os

import numpy as np
import pytest

from multiflow3d_amd import topology as T
from oracle import levelset as LS
from oracle import velocity as V
from oracle.oracle import have_ref_vel, run_ref_bound_lsm, run_ref_vel

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MAN = json.load(open(os.path.join(GOLDEN, "manifest_ls.json")))


def _split(grid, flat, ghost=1):
    out, o = [], 0
    for b in grid.blocks:
        nx, ny, nz = (x + 2 * ghost for x in b.n)
        m = nx * ny * nz
        out.append(flat[o:o + m].reshape(nz, ny, nx).copy())
        o += m
    return out


def load_ls(name):
    """(grid, manifest entry, arrays) of a heaviside chain case; the
    arrays' fields are split into per-block ghosted boxes."""
    m = MAN["ls"][name]
    g = T.block_grid(m["ex"], m["ey"], m["ez"], m["bc"], m["h"])
    z = np.load(os.path.join(GOLDEN, f"ls_{name}.npz"))
    a = {k: _split(g, z[k]) for k in ("phi", "ustar", "vstar", "wstar", "p0", "dens", "mu", "p", "u", "v", "w")}
    a["su"] = _split(g, z["su"], ghost=0)
    a["props"] = tuple(float(x) for x in z["props"])
    a["iters"] = int(z["iters"][0])
    a["rmax"] = z["rmax"]
    return g, m, a


def load_lsb(name):
    m = MAN["lsb"][name]
    g = T.block_grid(m["ex"], m["ey"], m["ez"], m["bc"], [0.1] * 3)
    z = np.load(os.path.join(GOLDEN, f"lsb_{name}.npz"))
    return g, _split(g, z["x"]), _split(g, z["out"])


@pytest.mark.parametrize("name", sorted(MAN["ls"]))
def test_heaviside_oracle_reproduces_reference_fixture(name):
    g, m, a = load_ls(name)
    dens, mu = LS.heaviside(g, a["phi"], *a["props"])
    for got, exp in ((dens, a["dens"]), (mu, a["mu"])):
        for x, y in zip(got, exp):
            assert np.array_equal(x, y), np.argwhere(x != y)[:5]
    assert m["band_cells"] > 100  # the sin branch is exercised


@pytest.mark.parametrize("name", sorted(MAN["ls"]))
def test_heaviside_chain_into_newsolv_matches_reference_fixture(name):
    """heaviside -> coeff -> the velocity form, the order of the
    reference's LSM step (flosol.for:114-117, newsolv_mg.for:12-58)."""
    g, m, a = load_ls(name)
    dens, _ = LS.heaviside(g, a["phi"], *a["props"])
    k, hist, p, u, v, w, su = V.newsolv_velocity(
        g, m["levels"], a["ustar"], a["vstar"], a["wstar"], a["p0"], 0.0, m["maxcy"], m["dt"], m["alfapr"],
        lsm=True, dens=dens)
    assert k == a["iters"]
    assert np.array_equal(hist[:k + 1], a["rmax"])
    for key, got in (("p", p), ("u", u), ("v", v), ("w", w), ("su", su)):
        for x, y in zip(got, a[key]):
            assert np.array_equal(x, y), key


@pytest.mark.parametrize("name", sorted(MAN["lsb"]))
def test_bound_lsm_oracle_reproduces_reference_fixture(name):
    g, x, exp = load_lsb(name)
    got = LS.bound_and_exchange(g, [b.copy() for b in x])
    for b, (y, e) in enumerate(zip(got, exp)):
        assert np.array_equal(y, e), (b, np.argwhere(y != e)[:5])


def test_north_wall_rewrites_the_last_interior_row():
    """bounds_lsm.for:91 runs ipl = 0..pl: at a north wall the last interior
    row takes the row below it; no other interior entry changes."""
    g = T.block_grid([5], [6], [4], [T.NEUMANN] * 6, [0.1] * 3)
    x = [np.arange(7 * 8 * 6, dtype=np.float64).reshape(6, 8, 7)]
    y = LS.bound_lsm(g, [x[0].copy()])[0]
    ny = 6
    assert np.array_equal(y[1:-1, ny, 1:-1], x[0][1:-1, ny - 1, 1:-1])
    assert np.array_equal(y[1:-1, ny + 1, 1:-1], x[0][1:-1, ny - 1, 1:-1])
    keep = np.ones(y.shape, bool)
    keep[:, ny:, :] = False
    keep[0], keep[-1], keep[:, 0], keep[:, :, 0], keep[:, :, -1] = False, False, False, False, False
    assert np.array_equal(y[keep], x[0][keep])
    # a periodic north face is left to the exchange
    gp = T.block_grid([5], [6], [4], [T.NEUMANN] * 2 + [T.PERIODIC] * 2 + [T.NEUMANN] * 2, [0.1] * 3)
    yp = LS.bound_lsm(gp, [x[0].copy()])[0]
    assert np.array_equal(yp[1:-1, 1:-1, 1:-1], x[0][1:-1, 1:-1, 1:-1])


def test_h_phi_closed_forms():
    """What the smoothed Heaviside fixes without the reference: 0 and 1
    beyond the band, 1/2 at the interface, h(x) + h(-x) = 1, monotone
    increase across the band, slope 1/epsl at the interface, and the band
    edges within |sin(pi_ref) / pi| of 0 and 1 (pi_ref = 3.14159265359 is
    not pi)."""
    epsl = 0.125
    h = LS.h_phi
    assert h(np.array([-1.0, -epsl * 1.0000001, epsl * 1.0000001, 7.0]), epsl).tolist() == [0, 0, 1, 1]
    assert h(np.array([0.0]), epsl)[0] == 0.5
    edge = abs(math.sin(LS.PI)) / LS.PI
    assert abs(h(np.array([epsl]), epsl)[0] - 1.0) <= edge + 1e-16
    assert abs(h(np.array([-epsl]), epsl)[0]) <= edge + 1e-16
    x = np.linspace(-epsl, epsl, 4097)
    hx = h(x, epsl)
    assert np.all(np.diff(hx) >= 0)
    assert np.max(np.abs(hx + hx[::-1] - 1.0)) < 4e-16
    d = 1e-6 * epsl
    slope = (h(np.array([d]), epsl)[0] - h(np.array([-d]), epsl)[0]) / (2 * d)
    assert abs(slope * epsl - 1.0) < 1e-6
    # the formula against a direct evaluation with the textbook form at a few points
    for v in (-0.1, -0.03, 0.01, 0.07, 0.124):
        ref = 0.5 * (1 + v / epsl + math.sin(math.pi * v / epsl) / math.pi)
        assert abs(h(np.array([v]), epsl)[0] - ref) < 1e-12


def test_heaviside_density_limits():
    """dens = densg where phi < -epsl (gas), densl where phi > epsl (liquid),
    the mean at the interface; mu likewise."""
    g = T.block_grid([4], [4], [4], [T.NEUMANN] * 6, [0.25] * 3)
    epsl = LS.epsl_of(g.blocks[0])
    assert epsl == 0.5
    phi = np.full((6, 6, 6), -1.0)
    phi[:, :, 3:] = 1.0
    phi[1, 1, 1] = 0.0
    dens, mu = LS.heaviside(g, [phi], 1000.0, 1.2, 1e-3, 1.8e-5)
    d = dens[0]
    assert d[2, 2, 1] == 1.2 and d[2, 2, 4] == 1000.0 and mu[0][2, 2, 4] == 1e-3
    assert d[1, 1, 1] == 1.2 + (1000.0 - 1.2) * 0.5


@pytest.mark.skipif(not have_ref_vel(), reason="compiled reference not built (make -C oracle ref)")
@pytest.mark.parametrize("seed", range(4))
def test_levelset_oracle_matches_compiled_reference_randomised(seed):
    rng = np.random.default_rng(900 + seed)
    nbx, nby, nbz = rng.integers(1, 3, 3)
    ex = rng.integers(3, 7, nbx).tolist()
    ey = rng.integers(3, 7, nby).tolist()
    ez = rng.integers(3, 7, nbz).tolist()
    h = [0.1, 0.125, 0.1]
    g = T.block_grid(ex, ey, ez, [T.NEUMANN] * 6, h)
    epsl = 2 * max(h)
    phi = [rng.uniform(-2.5 * epsl, 2.5 * epsl, (b.n[2] + 2, b.n[1] + 2, b.n[0] + 2)) for b in g.blocks]
    z = [np.zeros_like(p) for p in phi]
    props = (float(rng.uniform(1, 1000)), float(rng.uniform(0.5, 2)), 1e-3, 2e-5)
    res = run_ref_vel(g, 1, z, z, z, z, 0.0, 0, 0.01, 1.0, phi=phi, props=props)
    dens, mu = LS.heaviside(g, phi, *props)
    for x, y in zip(dens + mu, res[7] + res[8]):
        assert np.array_equal(x, y)
    bc = [int(c) for c in rng.choice([T.NEUMANN, T.PERIODIC], 3)]
    gp = T.block_grid(ex, ey, ez, [bc[0], bc[0], bc[1], bc[1], bc[2], bc[2]], h)
    x = [rng.standard_normal(p.shape) for p in phi]
    for y, e in zip(LS.bound_and_exchange(gp, [a.copy() for a in x]), run_ref_bound_lsm(gp, x)):
        assert np.array_equal(y, e)
