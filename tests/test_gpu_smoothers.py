"""GPU parity of the lexicographic Gauss-Seidel (mg_itrsch = 1) and ADI-TDMA
(mg_itrsch = 2) smoothers, through the C ABI, against the compiled
reference's golden vectors and the CPU oracle.

- Every sm_ fixture (single block and multi-block, mixed walls, LSM, and a
  slice of the reference's shipped deck: one level, ADI, irestr 2, iproln
  1): FP64 pp after one V-cycle with its ghost layer, and p after the
  fixed cycles, bitwise the reference's; the residual history to 1e-12
  (the GPU sums the norm in a different order).
- The velocity form with each smoother: bitwise the oracle.
- Mixed precision (FP32 inner cycle) converges like FP64.
"""
import warnings

import numpy as np
import pytest

from multiflow3d_amd import NotConverged, PressureMultigrid
from multiflow3d_amd import topology as T
from multiflow3d_amd import workloads as W
from oracle import velocity as V

from .test_oracle_smoothers import MAN, load_sm

pytestmark = pytest.mark.gpu


def _mg(g, m, a, **kw):
    mg = PressureMultigrid(g, m["levels"], nu_pre=m["nu_pre"], nu_post=m["nu_post"], smoother=m["smoother"],
                           variable_coef=m["lsm"], **kw)
    if m["lsm"]:
        mg.coeff(a["dens"])
    return mg


@pytest.mark.parametrize("name", sorted(MAN))
def test_smoother_fp64_matches_reference_fixture(name):
    g, m, a = load_sm(name)
    print(f"[smoother fp64] {name}", flush=True)
    with _mg(g, m, a) as mg:
        pp = [np.zeros((b.n[2] + 2, b.n[1] + 2, b.n[0] + 2)) for b in g.blocks]
        mg.mgkcyc(a["f"], pp)
    for x, e in zip(pp, a["pp1"]):
        assert np.array_equal(x, e), np.abs(x - e).max()
    with _mg(g, m, a, maxcy=m["cycles"], eps=0.0) as mg:
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", NotConverged)
            p, res = mg.newsolv_mg(a["f"])
    assert res.iters == m["cycles"]
    np.testing.assert_allclose(res.hist, a["rho"], rtol=1e-12, atol=0)
    for x, e in zip(p, a["p"]):
        assert np.array_equal(x, e), np.abs(x - e).max()


@pytest.mark.parametrize("sm", [1, 2])
def test_velocity_form_with_smoother_matches_oracle(sm):
    """The shipped deck's layout in miniature (blocks in x, one level, irestr
    2, iproln 1) through mf3d_mg_newsolv, against the oracle's loop."""
    g = T.block_grid([40, 40], [10], [25], [T.PERIODIC, T.PERIODIC, T.NEUMANN, T.NEUMANN, T.NEUMANN, 3],
                     [0.005] * 3)
    us, vs, ws = W.velocity_inputs(g, W.BASE_SEED + 90 + sm)
    p0 = [np.zeros_like(x) for x in us]
    print(f"[velocity smoother {sm}]", flush=True)
    k, hist, po, uo, vo, wo, suo = V.newsolv_velocity(g, 1, us, vs, ws, p0, 0.0, 4, 0.0025, 1.0, smoother=sm,
                                                      nu_pre=2, nu_post=1)
    u, v, w, p = [x.copy() for x in us], [x.copy() for x in vs], [x.copy() for x in ws], [x.copy() for x in p0]
    with PressureMultigrid(g, 1, nu_pre=2, nu_post=1, smoother=sm, maxcy=4) as mg:
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", NotConverged)
            res = mg.newsolv(us, vs, ws, u, v, w, p, 0.0025, 1.0, 0.0)
    assert res.iters == k == 4
    assert np.array_equal(res.hist, hist)
    for got, exp in ((p, po), (u, uo), (v, vo), (w, wo)):
        for x, e in zip(got, exp):
            assert np.array_equal(x, e)


@pytest.mark.parametrize("name", ["sm1_single16_channel_L3", "sm2_single16_channel_L3"])
def test_smoother_mixed_converges_like_fp64(name):
    g, m, a = load_sm(name)
    print(f"[smoother mixed] {name}", flush=True)
    out = {}
    for prec in ("fp64", "mixed"):
        with _mg(g, m, a, maxcy=200, eps=1e-8, precision=prec) as mg:
            p, res = mg.newsolv_mg(a["f"])
        assert res.status == 0 and res.hist[-1] <= 1e-8
        out[prec] = (np.concatenate([x.ravel() for x in p]), res.iters)
    pa, pb = out["mixed"][0], out["fp64"][0]
    pa, pb = pa - pa.mean(), pb - pb.mean()
    assert np.linalg.norm(pa - pb) / np.linalg.norm(pb) < 1e-6
    assert out["mixed"][1] <= out["fp64"][1] + 2


def test_shipped_slice_mixed_tracks_fp64():
    """One level of ADI (the shipped deck's setting) converges slowly; over
    the deck's 30 cycles the mixed form's residual history stays within
    1e-3 of FP64's and its pressure within 1e-5 relative L2."""
    g, m, a = load_sm("sm2_shipped_slice_L1")
    print("[smoother mixed] shipped slice, 30 cycles", flush=True)
    out = {}
    for prec in ("fp64", "mixed"):
        with _mg(g, m, a, maxcy=30, eps=0.0, precision=prec) as mg:
            with warnings.catch_warnings():
                warnings.simplefilter("ignore", NotConverged)
                p, res = mg.newsolv_mg(a["f"])
        out[prec] = (np.concatenate([x.ravel() for x in p]), res.hist)
    np.testing.assert_allclose(out["mixed"][1], out["fp64"][1], rtol=1e-3, atol=0)
    pa, pb = out["mixed"][0], out["fp64"][0]
    assert np.linalg.norm(pa - pb) / np.linalg.norm(pb) < 1e-5


@pytest.mark.parametrize("name", ["sm2_multi_uneven_L2", "sm2_lsm_multi_L3", "sm2_shipped_slice_L1"])
def test_adi_global_memory_path_matches_reference_fixture(name, monkeypatch):
    """MF3D_ADI_LDS=0 forces the per-pass kernels that work in global
    memory (the path of blocks too large for LDS): the same bitwise
    result."""
    monkeypatch.setenv("MF3D_ADI_LDS", "0")
    g, m, a = load_sm(name)
    print(f"[adi global] {name}", flush=True)
    with _mg(g, m, a, maxcy=m["cycles"], eps=0.0) as mg:
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", NotConverged)
            p, res = mg.newsolv_mg(a["f"])
    np.testing.assert_allclose(res.hist, a["rho"], rtol=1e-12, atol=0)
    for x, e in zip(p, a["p"]):
        assert np.array_equal(x, e), np.abs(x - e).max()
