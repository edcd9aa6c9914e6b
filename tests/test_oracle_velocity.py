"""The velocity form of newsolv_mg in the CPU oracle (oracle/velocity.py):
against golden vectors written by the compiled reference
(tests/golden/make_golden_vel.py), and against what the discretisation
fixes without the reference (projection = -alpha * the multigrid operator,
ghost maps as brute-force permutations, untouched wall faces)."""
import json
import os

import numpy as np
import pytest

from multiflow3d_amd import topology as T
from oracle import velocity as V
from oracle.oracle import OracleMG, have_ref_vel, run_ref_vel

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MAN = json.load(open(os.path.join(GOLDEN, "manifest_vel.json")))


def _split(grid, flat, ghost=1):
    out, o = [], 0
    for b in grid.blocks:
        nx, ny, nz = (x + 2 * ghost for x in b.n)
        m = nx * ny * nz
        out.append(flat[o:o + m].reshape(nz, ny, nx).copy())
        o += m
    return out


def _grid(m):
    return T.block_grid(m["ex"], m["ey"], m["ez"], m["bc"], m["h"])


def load_case(name):
    """(grid, manifest entry, inputs dict, expected dict) of a golden case."""
    m = MAN[name]
    g = _grid(m)
    z = np.load(os.path.join(GOLDEN, f"vel_{name}.npz"))
    inp = {k: _split(g, z[k]) for k in ("ustar", "vstar", "wstar", "p0")}
    inp["dens"] = _split(g, z["dens"]) if m["lsm"] else None
    exp = {k: _split(g, z[k]) for k in ("p", "u", "v", "w")}
    exp["su"] = _split(g, z["su"], ghost=0)
    exp["iters"] = int(z["iters"][0])
    exp["rmax"] = z["rmax"]
    return g, m, inp, exp


@pytest.mark.parametrize("name", sorted(MAN))
def test_velocity_oracle_reproduces_reference_fixture(name):
    g, m, inp, exp = load_case(name)
    # the seeded recipe regenerates the fixture's inputs exactly
    import importlib.util
    spec = importlib.util.spec_from_file_location("mkv", os.path.join(GOLDEN, "make_golden_vel.py"))
    mkv = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mkv)
    us, vs, ws, p0, dens = mkv.inputs(g, m["seed"], m["lsm"])
    for a, b in zip(us + vs + ws + p0, inp["ustar"] + inp["vstar"] + inp["wstar"] + inp["p0"]):
        assert np.array_equal(a, b)
    k, hist, p, u, v, w, su = V.newsolv_velocity(
        g, m["levels"], inp["ustar"], inp["vstar"], inp["wstar"], inp["p0"], m["eps"], m["maxcy"], m["dt"],
        m["alfapr"], lsm=m["lsm"], dens=inp["dens"])
    assert k == exp["iters"]
    assert np.array_equal(hist[:k + 1], exp["rmax"])
    for key, got in (("p", p), ("u", u), ("v", v), ("w", w), ("su", su)):
        for a, b in zip(got, exp[key]):
            assert np.array_equal(a, b), key


def test_eps_case_stops_on_rmax():
    """newsolv_mg.for:55: the loop ends once rmax < eps after at least two
    cycles; the recorded history shows the stop cycle is the first below."""
    _, m, _, exp = load_case("single12_channel_L3_eps")
    r = exp["rmax"]
    k = exp["iters"]
    assert k < m["maxcy"] and r[k] < m["eps"]
    assert all(r[j] >= m["eps"] for j in range(2, k))


def _mapped_p(g, p):
    return V.pbound(g, V.exchange(g, p))


@pytest.mark.parametrize("bc", [T.CHANNEL_BC, [T.PERIODIC] * 6, [T.NEUMANN] * 6])
def test_projection_is_minus_alpha_times_multigrid_operator(bc):
    """With u* = 0 and no flow through walls, calmas(calvel(p)) / alpha is
    -(A p) for the multigrid's level-1 operator A (coefficients 1/h^2, wall
    coefficients zero, mgsolver.for:184-208): the pressure correction and
    the Poisson operator are the same discretisation."""
    g = T.block_grid([8, 8], [8, 4], [8], bc, [1 / 16, 1 / 12, 1 / 8])
    rng = np.random.default_rng(3)
    p = _mapped_p(g, [rng.standard_normal((b.n[2] + 2, b.n[1] + 2, b.n[0] + 2)) for b in g.blocks])
    zero = [np.zeros_like(a) for a in p]
    u, v, w = [a.copy() for a in zero], [a.copy() for a in zero], [a.copy() for a in zero]
    dt, alfapr = 0.01, 0.7
    V.calvel(g, zero, zero, zero, u, v, w, p, dt, alfapr)
    su, _ = V.calmas(g, u, v, w, dt)
    o = OracleMG(g.blocks, 1)
    o.coeff()
    for b in range(len(g.blocks)):
        o.phi(b, 0)[...] = p[b]
        o.rhs(b, 0)[...] = 0.0
    r = o.residual(0)  # 0 - A p
    o.close()
    for s, rr in zip(su, r):
        np.testing.assert_allclose(s, alfapr * rr, rtol=1e-11, atol=1e-9 * np.abs(rr).max())


def _brute_source(g, b, idx):
    """Block and index a ghost entry maps to through face neighbours and
    periodic partners (None through a wall)."""
    t = list(idx)
    for a in range(3):
        n = g.blocks[b].n[a]
        if t[a] == 0 or t[a] == n + 1:
            side = 0 if t[a] == 0 else 1
            blk = g.blocks[b]
            f = 2 * a + side
            s = blk.nbr[f] if blk.nbr[f] >= 0 else (blk.per[f] if blk.per[f] >= 0 else (b if blk.bc[f] == 5 else None))
            if s is None:
                return None
            t[a] = g.blocks[s].n[a] if side == 0 else 1
            b = s
    return b, tuple(t)


@pytest.mark.parametrize("shape", [([8, 8], [8, 8], [8, 8]), ([4, 8, 4], [8], [4, 4]), ([8], [8], [8])])
def test_exchange_face_ghosts_hold_mapped_interior(shape):
    """On an all-periodic grid every face ghost (one index out of range)
    holds the interior value it maps to, wherever the block's neighbours
    and partners lie (brute-force index map)."""
    g = T.block_grid(*shape, [T.PERIODIC] * 6, [1 / 16] * 3)
    rng = np.random.default_rng(11)
    F = [rng.standard_normal((b.n[2] + 2, b.n[1] + 2, b.n[0] + 2)) for b in g.blocks]
    pre = [a.copy() for a in F]
    V.exchange(g, F)
    for b, blk in enumerate(g.blocks):
        n = blk.n
        for k in range(n[2] + 2):
            for j in range(n[1] + 2):
                for i in range(n[0] + 2):
                    idx = (i, j, k)
                    out = sum(1 for a in range(3) if idx[a] in (0, n[a] + 1))
                    if out != 1:
                        continue
                    s, t = _brute_source(g, b, idx)
                    assert F[b][k, j, i] == pre[s][t[2], t[1], t[0]]
        # the interior is never written
        assert np.array_equal(F[b][1:-1, 1:-1, 1:-1], pre[b][1:-1, 1:-1, 1:-1])


def test_walls_keep_their_faces_and_pbound_copies_interior():
    """calvel leaves the face on a non-periodic wall to the caller
    (localparameters.for:185-190) and exchange never writes a ghost behind
    a wall; pbound then copies the adjacent interior plane into it."""
    g = T.block_grid([8], [8], [8], [T.NEUMANN] * 6, [1 / 8] * 3)
    rng = np.random.default_rng(5)
    sh = (10, 10, 10)
    us, vs, ws = (rng.standard_normal(sh) for _ in range(3))
    u, v, w = us.copy() + 1.0, vs.copy() + 1.0, ws.copy() + 1.0
    u0, v0, w0 = u.copy(), v.copy(), w.copy()
    p = [rng.standard_normal(sh)]
    V.calvel(g, [us], [vs], [ws], [u], [v], [w], p, 0.01, 1.0)
    assert np.array_equal(u[:, :, 8], u0[:, :, 8]) and np.array_equal(u[:, :, 0], u0[:, :, 0])
    assert np.array_equal(v[:, 8, :], v0[:, 8, :]) and np.array_equal(w[8], w0[8])
    assert not np.array_equal(u[1:-1, 1:-1, 1:8], u0[1:-1, 1:-1, 1:8])
    P = V.pbound(g, [p[0].copy()])[0]
    assert np.array_equal(P[:, :, 0], P[:, :, 1]) and np.array_equal(P[9], P[8])


@pytest.mark.skipif(not have_ref_vel(), reason="oracle/_ref/mg_ref_vel not built (build container only)")
def test_velocity_oracle_matches_compiled_reference_randomised():
    """Randomised block layouts and boundary codes against the compiled
    reference: full ghosted p, u, v, w and su bitwise."""
    rng = np.random.default_rng(2024)
    for trial in range(4):
        nb = [int(x) for x in rng.integers(1, 3, size=3)]
        ext = [[int(rng.choice([4, 8])) for _ in range(n)] for n in nb]
        bc = []
        for a in range(3):
            c = int(rng.choice([T.PERIODIC, T.NEUMANN]))
            bc += [c, c]
        g = T.block_grid(*ext, bc, [1 / 8] * 3)
        shp = [(b.n[2] + 2, b.n[1] + 2, b.n[0] + 2) for b in g.blocks]
        us = [rng.uniform(-1, 1, s) for s in shp]
        vs = [rng.uniform(-1, 1, s) for s in shp]
        ws = [rng.uniform(-1, 1, s) for s in shp]
        p0 = [rng.uniform(-1, 1, s) for s in shp]
        r = run_ref_vel(g, 2, us, vs, ws, p0, 0.0, 3, 0.01, 1.0)
        o = V.newsolv_velocity(g, 2, us, vs, ws, p0, 0.0, 3, 0.01, 1.0)
        assert r[0] == o[0]
        assert np.array_equal(r[1], o[1][:len(r[1])])
        for i in (2, 3, 4, 5):
            for a, b in zip(r[i], o[i]):
                assert np.array_equal(a, b), (trial, i)
        for a, b in zip(r[6], o[6]):
            assert np.array_equal(a[1:-1, 1:-1, 1:-1], b)


MAN_PL = json.load(open(os.path.join(GOLDEN, "manifest_vel_pl.json")))


def load_case_pl(name):
    """(grid, manifest entry, inputs, expected) of a wide-shell golden case
    (pl = LMR + pl_ex ghost layers)."""
    import importlib.util
    m = MAN_PL[name]
    g = _grid(m)
    spec = importlib.util.spec_from_file_location("mkv", os.path.join(GOLDEN, "make_golden_vel.py"))
    mkv = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mkv)
    G = m["ghost"]
    us, vs, ws, p0 = mkv.inputs_pl(g, m["seed"], G)
    z = np.load(os.path.join(GOLDEN, f"vel_{name}.npz"))
    exp = {k: _split(g, z[k], ghost=G) for k in ("p", "u", "v", "w")}
    exp["su"] = _split(g, z["su"], ghost=0)
    exp["iters"] = int(z["iters"][0])
    exp["rmax"] = z["rmax"]
    return g, m, dict(ustar=us, vstar=vs, wstar=ws, p0=p0), exp


@pytest.mark.parametrize("name", sorted(MAN_PL))
def test_wide_shell_oracle_reproduces_reference_fixture(name):
    """pl_ex extra layers refreshed by exchange(1..3), whole-plane periodic
    copies and (pl+1)-deep corner/edge cubes reaching the outer layers, and
    the 4th-order calmas: the oracle's whole ghosted p, u, v, w, su and
    rmax history equal the compiled reference's bitwise."""
    g, m, inp, exp = load_case_pl(name)
    k, hist, p, u, v, w, su = V.newsolv_velocity(
        g, m["levels"], inp["ustar"], inp["vstar"], inp["wstar"], inp["p0"], m["eps"], m["maxcy"], m["dt"],
        m["alfapr"], pl_ex=m["pl_ex"], differencing=m["differencing"])
    assert k == exp["iters"]
    assert np.array_equal(hist, exp["rmax"])
    for key, got in (("p", p), ("u", u), ("v", v), ("w", w), ("su", su)):
        for a, b in zip(got, exp[key]):
            assert np.array_equal(a, b), (name, key)


def test_wide_shell_exchange_depths():
    """On the shipped deck's shell (pl = 4, pl_ex = 2): entries beyond the
    first ghost layer change (whole-plane periodic copies, 5-deep corner
    cubes), so a one-layer refresh fails; the x-face ghost planes of u at
    depth pl_ex + 2 keep their input values on the interior rows, so a
    refresh of all pl layers fails too (rows 2..n-1: the corner and edge
    cubes also reach rows 1 and n)."""
    g, m, inp, exp = load_case_pl("deck_channel222_pl4")
    G = m["ghost"]
    for key, src in (("p", inp["p0"]), ("u", inp["ustar"])):
        beyond = np.ones(exp[key][0].shape, dtype=bool)
        beyond[G - 1:-(G - 1), G - 1:-(G - 1), G - 1:-(G - 1)] = False
        changed = sum(int(np.count_nonzero((a != b)[beyond])) for a, b in zip(exp[key], src))
        assert changed > 0, key
    d = m["pl_ex"] + 2  # first layer the exchange does not refresh
    for a, b, blk in zip(exp["u"], inp["ustar"], g.blocks):
        nx, ny, nz = blk.n
        rows = (slice(G + 1, G + nz - 1), slice(G + 1, G + ny - 1))
        assert np.array_equal(a[rows + (G - d,)], b[rows + (G - d,)])
        assert np.array_equal(a[rows + (G + nx - 1 + d,)], b[rows + (G + nx - 1 + d,)])


def test_fourth_order_calmas_exact_on_cubic_faces():
    """The 4th-order staggered divergence is exact for face velocities that
    are cubic along their own axis: u(x) = x^3 at faces x_i = i dx gives
    du/dx = 3 x^2 at the cell centres (i - 1/2) dx, and likewise for v, w."""
    n, G = 6, 2
    g = T.block_grid([n], [n], [n], [T.PERIODIC] * 6, [0.1, 0.2, 0.25])
    c = np.arange(1 - G, n + G + 1, dtype=np.float64)  # coordinates of the array indices
    shp = (n + 2 * G,) * 3
    dx, dy, dz = g.blocks[0].h
    u = np.broadcast_to((c * dx) ** 3, shp).copy()
    v = np.broadcast_to(((c * dy) ** 3)[None, :, None], shp).copy()
    w = np.broadcast_to(((c * dz) ** 3)[:, None, None], shp).copy()
    su, _ = V.calmas(g, [u], [v], [w], dt=1.0, differencing=2)
    ctr = c[G:G + n] - 0.5
    expect = 3 * (ctr[None, None, :] * dx) ** 2 + 3 * (ctr[None, :, None] * dy) ** 2 + 3 * (ctr[:, None, None] * dz) ** 2
    np.testing.assert_allclose(su[0], expect, rtol=1e-12, atol=1e-13)
    su2, _ = V.calmas(g, [u], [v], [w], dt=1.0)  # the 2nd-order branch is not exact on cubics
    assert not np.allclose(su2[0], expect, rtol=1e-6)


@pytest.mark.skipif(not have_ref_vel(), reason="oracle/_ref/mg_ref_vel not built (build container only)")
def test_wide_shell_oracle_matches_compiled_reference_randomised():
    """Randomised layouts, boundary codes, pl_ex in 0..2, LMR 1 or 2 and
    the 4th-order calmas where two layers exist, run live against the
    compiled reference: whole ghosted fields bitwise."""
    rng = np.random.default_rng(77)
    for trial in range(6):
        pl_ex, lmr = int(rng.integers(0, 3)), int(rng.integers(1, 3))
        G = pl_ex + lmr
        diff = 2 if G >= 2 and trial % 2 else 1
        nb = [int(x) for x in rng.integers(1, 3, size=3)]
        ext = [[int(rng.choice([4, 8])) for _ in range(n)] for n in nb]
        bc = []
        for a in range(3):
            cc = int(rng.choice([T.PERIODIC, T.NEUMANN]))
            bc += [cc, cc]
        g = T.block_grid(*ext, bc, [1 / 8] * 3)
        shp = [(b.n[2] + 2 * G, b.n[1] + 2 * G, b.n[0] + 2 * G) for b in g.blocks]
        us, vs, ws, p0 = ([rng.uniform(-1, 1, s) for s in shp] for _ in range(4))
        r = run_ref_vel(g, 2, us, vs, ws, p0, 0.0, 3, 0.01, 1.0, pl_ex=pl_ex, lmr=lmr, differencing=diff)
        o = V.newsolv_velocity(g, 2, us, vs, ws, p0, 0.0, 3, 0.01, 1.0, pl_ex=pl_ex, differencing=diff)
        assert r[0] == o[0] and np.array_equal(r[1], o[1][:len(r[1])])
        for i in (2, 3, 4, 5):
            for a, b in zip(r[i], o[i]):
                assert np.array_equal(a, b), (trial, i)
        for a, b in zip(r[6], o[6]):
            assert np.array_equal(a[G:-G, G:-G, G:-G], b)
