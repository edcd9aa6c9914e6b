"""GPU parity of the level-set density (mf3d_mg_heaviside and
mf3d_mg_heaviside_device) against the compiled reference's golden vectors
and the CPU oracle.

Tolerance: sin is the device library's, the reference's the host libm's,
so inside the transition band |phi| <= epsl a density may differ in the
last place.  The bar: every entry whose reference value is densg or densl
(h = 0 or 1, outside the band) bitwise; every other entry within
4 * 2^-52 * max(|densl|, |densg|) (mu likewise).  The chain heaviside ->
coeff -> velocity-form newsolv_mg is then bitwise where the densities are,
and within 1e-10 relative L2 per field otherwise.
"""
import warnings

import numpy as np
import pytest

from multiflow3d_amd import NotConverged, PressureMultigrid
from multiflow3d_amd import hip
from multiflow3d_amd import topology as T
from oracle import levelset as LS

from .test_gpu_velocity import _dev, _undev
from .test_oracle_levelset import MAN, load_ls

pytestmark = pytest.mark.gpu

ULP = 2.0 ** -52


def _check_dens(got, exp, lo, hi):
    """The bar of the module docstring; returns the count of entries that
    differ (all inside the band)."""
    tol = 4 * ULP * max(abs(lo), abs(hi))
    nd = 0
    for b, (x, e) in enumerate(zip(got, exp)):
        pure = (e == lo) | (e == hi)
        assert np.array_equal(x[pure], e[pure]), (b, np.argwhere((x != e) & pure)[:5])
        err = np.abs(x - e)
        assert err.max() <= tol, (b, float(err.max()), tol)
        nd += int((x != e).sum())
    return nd


def _rel(a_parts, b_parts):
    a = np.concatenate([x.ravel() for x in a_parts])
    b = np.concatenate([x.ravel() for x in b_parts])
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _newsolv(mg, a, m):
    us, vs, ws = a["ustar"], a["vstar"], a["wstar"]
    u, v, w, p = ([x.copy() for x in us], [x.copy() for x in vs], [x.copy() for x in ws],
                  [x.copy() for x in a["p0"]])
    su = [np.zeros_like(x) for x in p]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", NotConverged)
        res = mg.newsolv(us, vs, ws, u, v, w, p, m["dt"], m["alfapr"], 0.0, su=su)
    return res, dict(p=p, u=u, v=v, w=w, su=[x[1:-1, 1:-1, 1:-1] for x in su])


@pytest.mark.parametrize("name", sorted(MAN["ls"]))
def test_heaviside_chain_matches_reference_fixture(name):
    g, m, a = load_ls(name)
    densl, densg, mul, mug = a["props"]
    print(f"[heaviside host] {name}", flush=True)
    with PressureMultigrid(g, m["levels"], maxcy=m["maxcy"], variable_coef=True) as mg:
        dens, mu = mg.heaviside(a["phi"], densl, densg, mul, mug)
        nd = _check_dens(dens, a["dens"], densg, densl)
        nm = _check_dens(mu, a["mu"], mug, mul)
        print(f"  entries differing in the last place: dens {nd}, mu {nm}", flush=True)
        res, out = _newsolv(mg, a, m)
    assert res.iters == a["iters"]
    if nd == 0:
        assert np.array_equal(res.hist, a["rmax"])
        for key in ("p", "u", "v", "w", "su"):
            for x, e in zip(out[key], a[key]):
                assert np.array_equal(x, e), key
    else:
        assert np.allclose(res.hist, a["rmax"], rtol=1e-9, atol=0)
        for key in ("p", "u", "v", "w", "su"):
            assert _rel(out[key], a[key]) < 1e-10, key


@pytest.mark.parametrize("name", ["walls221_band_L2", "walls222_sphere_L2"])
def test_heaviside_device_entry_matches_host_entry(name):
    """Device phi in, device dens and mu out, coefficients rebuilt, then the
    device-resident velocity form: bitwise the host entries' results."""
    g, m, a = load_ls(name)
    props = a["props"]
    print(f"[heaviside device] {name}", flush=True)
    with PressureMultigrid(g, m["levels"], maxcy=m["maxcy"], variable_coef=True) as mg:
        dens_h, mu_h = mg.heaviside(a["phi"], *props)
        _, out_h = _newsolv(mg, a, m)
    phi = _dev(a["phi"])
    dens = hip.DeviceArray(phi.n)
    mu = hip.DeviceArray(phi.n)
    bufs = [_dev(a[k]) for k in ("ustar", "vstar", "wstar")] + \
           [_dev(a[k]) for k in ("ustar", "vstar", "wstar", "p0")]
    try:
        with PressureMultigrid(g, m["levels"], maxcy=m["maxcy"], variable_coef=True) as mg:
            mg.heaviside_device(phi.ptr, *props, dens_ptr=dens.ptr, mu_ptr=mu.ptr)
            for x, e in zip(_undev(g, dens), dens_h):
                assert np.array_equal(x, e)
            for x, e in zip(_undev(g, mu), mu_h):
                assert np.array_equal(x, e)
            with warnings.catch_warnings():
                warnings.simplefilter("ignore", NotConverged)
                mg.newsolv_device([b.ptr for b in bufs], m["dt"], m["alfapr"], 0.0)
            for key, d in (("u", bufs[3]), ("v", bufs[4]), ("w", bufs[5]), ("p", bufs[6])):
                for x, e in zip(_undev(g, d), out_h[key]):
                    assert np.array_equal(x, e), key
            # without outputs, from the caller's density: the same coefficients
            mg.heaviside_device(phi.ptr, *props)
            mg.coeff_device(dens.ptr)
    finally:
        for b in bufs + [phi, dens, mu]:
            b.free()


P, N = T.PERIODIC, T.NEUMANN
# grids with periodic faces: (ex, ey, ez, bc, levels)
PERIODIC_GRIDS = {
    "single_channel": ([8], [6], [4], T.CHANNEL_BC, 2),
    "channel222": ([8, 8], [8, 8], [8, 8], T.CHANNEL_BC, 2),
    "allper312": ([4, 4, 4], [8], [4, 4], [P] * 6, 2),
    "mixed221": ([6, 4], [6, 8], [6], [N, N, P, P, N, N], 2),
    "zper_single": ([4], [6], [8], [N, N, N, N, P, P], 2),
}


@pytest.mark.parametrize("name", sorted(PERIODIC_GRIDS))
def test_heaviside_periodic_grids_match_oracle(name):
    """Periodic faces: the ghost sequence of bound_lsm(19) on the density
    (the oracle's, pinned by the reference's lsb fixtures), seeded phi with
    a third of its cells in the band."""
    ex, ey, ez, bc, L = PERIODIC_GRIDS[name]
    g = T.block_grid(ex, ey, ez, bc, [0.1] * 3)
    epsl = 2 * max(g.h)
    rng = np.random.default_rng(sum(map(ord, name)))
    phi = [rng.uniform(-3 * epsl, 3 * epsl, (b.n[2] + 2, b.n[1] + 2, b.n[0] + 2)) for b in g.blocks]
    props = (1000.0, 1.2, 1e-3, 1.8e-5)
    print(f"[heaviside periodic] {name}", flush=True)
    with PressureMultigrid(g, L, maxcy=1, variable_coef=True) as mg:
        dens, mu = mg.heaviside(phi, *props)
    od, om = LS.heaviside(g, phi, *props)
    _check_dens(dens, od, 1.2, 1000.0)
    _check_dens(mu, om, 1.8e-5, 1e-3)


@pytest.mark.slow
def test_heaviside_256cubed_matches_oracle():
    """Full size: 256^3 single block with walls, 7 levels, a level set with
    a third of its cells in the band; device entry against the oracle."""
    g = T.single_block([256, 256, 256], [1 / 256] * 3, [T.NEUMANN] * 6)
    epsl = 2 / 256
    rng = np.random.default_rng(1234)
    phi = [rng.uniform(-3 * epsl, 3 * epsl, (258, 258, 258))]
    props = (1000.0, 1.2, 1e-3, 1.8e-5)
    print("[heaviside 256^3] gpu", flush=True)
    d_phi = _dev(phi)
    d_dens = hip.DeviceArray(d_phi.n)
    try:
        with PressureMultigrid(g, 7, maxcy=2, variable_coef=True) as mg:
            mg.heaviside_device(d_phi.ptr, *props, dens_ptr=d_dens.ptr)
            got = _undev(g, d_dens)
    finally:
        d_phi.free()
        d_dens.free()
    print("[heaviside 256^3] oracle", flush=True)
    od, _ = LS.heaviside(g, phi, *props)
    nd = _check_dens(got, od, 1.2, 1000.0)
    print(f"  entries differing in the last place: {nd} of {got[0].size}", flush=True)
