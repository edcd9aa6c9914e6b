"""GPU parity: the HIP library through its C ABI against the CPU oracle.

FP64 results must be bitwise identical to the oracle (same operations in
the same order, no FMA contraction); residual histories may differ only by
the summation order of the norms (relative 1e-12 here, the bar is 1e-10).
"""
import warnings

import numpy as np
import pytest

from multiflow3d_amd import PressureMultigrid, NotConverged
from oracle.oracle import OracleMG

from .cases import SMALL, case, inputs

pytestmark = pytest.mark.gpu


def _oracle(grid, L, lsm, dens):
    o = OracleMG(grid.blocks, L, lsm=lsm)
    if lsm:
        o.set_density(dens)
    o.coeff()
    return o


def _gpu(grid, L, lsm, dens, **kw):
    mg = PressureMultigrid(grid, L, variable_coef=lsm, **kw)
    if lsm:
        mg.coeff(dens)
    return mg


@pytest.mark.parametrize("name", SMALL)
def test_fixed_cycles_bitwise(name):
    grid, L, lsm = case(name)
    f, dens = inputs(grid, lsm)
    ncyc = 5
    p_o, rho_o, rmax_o, k_o = _oracle(grid, L, lsm, dens).solve(f, 0.0, ncyc)
    with _gpu(grid, L, lsm, dens, maxcy=ncyc, eps=0.0) as mg:
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", NotConverged)
            p_g, res = mg.newsolv_mg(f)
    assert res.iters == k_o == ncyc
    np.testing.assert_allclose(res.hist, rho_o, rtol=1e-12, atol=0)
    np.testing.assert_allclose(res.rmax, rmax_o[-1], rtol=1e-12, atol=0)
    for a, b in zip(p_g, p_o):
        assert np.array_equal(a, b), f"{name}: p differs, max {np.max(np.abs(a - b))}"


@pytest.mark.parametrize("name", ["single16_L4", "multi_uneven_L3", "lsm_multi_L3",
                                  "periodic_multi_L3", "mixed_bc_multi_L3"])
def test_vcycle_bitwise_with_ghosts(name):
    """mgkcyc alone: pp after one V-cycle, ghost layer included."""
    grid, L, lsm = case(name)
    f, dens = inputs(grid, lsm, seed=7)
    pp_o = _oracle(grid, L, lsm, dens).vcycle(f)
    with _gpu(grid, L, lsm, dens) as mg:
        pp = [np.zeros((b.n[2] + 2, b.n[1] + 2, b.n[0] + 2)) for b in grid.blocks]
        mg.mgkcyc(f, pp)
    for a, b in zip(pp, pp_o):
        assert np.array_equal(a, b), f"{name}: pp differs, max {np.max(np.abs(a - b))}"


@pytest.mark.parametrize("name", ["single16_L4", "multi_uneven_L3", "lsm16_L4"])
def test_converged_solve_matches(name):
    grid, L, lsm = case(name)
    f, dens = inputs(grid, lsm)
    p_o, rho_o, _, k_o = _oracle(grid, L, lsm, dens).solve(f, 1e-8, 200)
    with _gpu(grid, L, lsm, dens, maxcy=200, eps=1e-8) as mg:
        p_g, res = mg.newsolv_mg(f)
    assert res.status == 0 and res.iters == k_o
    assert res.hist[-1] <= 1e-8
    for a, b in zip(p_g, p_o):
        assert np.array_equal(a, b)


def test_min_cycles_and_status():
    grid, L, lsm = case("single16_L4")
    f, _ = inputs(grid, False)
    with _gpu(grid, L, False, None, maxcy=3, eps=1e-30) as mg:
        with pytest.warns(NotConverged):
            _, res = mg.newsolv_mg(f)
        assert res.status == 1 and res.iters == 3
    with _gpu(grid, L, False, None, maxcy=10, eps=1.0, min_cycles=2) as mg:
        _, res = mg.newsolv_mg(f)
        assert res.status == 0 and res.iters == 2   # newsolv_mg.for:56: iter > 1


def test_deterministic_repeat():
    grid, L, lsm = case("multi_uneven_L3")
    f, _ = inputs(grid, False)
    # a fixed 12 cycles (eps 0: no tolerance ends the solve early)
    with _gpu(grid, L, False, None, maxcy=12, eps=0.0) as mg, warnings.catch_warnings():
        warnings.simplefilter("ignore", NotConverged)
        p1, r1 = mg.newsolv_mg(f)
        p2, r2 = mg.newsolv_mg(f)
    assert r1.iters == r2.iters == 12
    assert np.array_equal(r1.hist, r2.hist)
    for a, b in zip(p1, p2):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("name", ["single16_L4", "multi_uneven_L3", "lsm_multi_L3", "periodic_multi_L3",
                                  "mixed_bc_multi_L3"])
def test_both_sweep_kernels_bitwise(name, monkeypatch):
    """The one-pass 3-D tile sweep (k_rb3d) and the z-march (k_rbs), each
    forced on every level through MF3D_SWEEP_3D (the tuner otherwise picks
    per level by timing): in FP64 both give the oracle's pressure bitwise;
    in mixed precision they give each other's bitwise (same FP32 update
    expression, same operand order)."""
    grid, L, lsm = case(name)
    f, dens = inputs(grid, lsm, seed=11)
    ncyc = 3
    p_o, _, _, _ = _oracle(grid, L, lsm, dens).solve(f, 0.0, ncyc)
    got = {}
    for tiles in ("3d", "march"):
        monkeypatch.setenv("MF3D_SWEEP_3D", "1" if tiles == "3d" else "0")
        for precision in ("fp64", "mixed"):
            with _gpu(grid, L, lsm, dens, maxcy=ncyc, eps=0.0, precision=precision) as mg:
                with warnings.catch_warnings():
                    warnings.simplefilter("ignore", NotConverged)
                    p_g, res = mg.newsolv_mg(f)
            assert res.iters == ncyc
            got[tiles, precision] = p_g
    for tiles in ("3d", "march"):
        for a, b in zip(got[tiles, "fp64"], p_o):
            assert np.array_equal(a, b), f"{name}/{tiles}: p differs, max {np.max(np.abs(a - b))}"
    for a, b in zip(got["3d", "mixed"], got["march", "mixed"]):
        assert np.array_equal(a, b), f"{name}: mixed 3-D and march differ, max {np.max(np.abs(a - b))}"


SWITCHES = {
    "no_tail": {"MF3D_TAIL": "0"},
    "gather_ghosts": {"MF3D_OWN_GHOSTS": "0"},
    "spatial_tiles": {"MF3D_XCD_ORDER": "0"},
    "default_depths": {"MF3D_AUTOTUNE": "0"},
    "pinned_depth": {"MF3D_SWEEP_TZ": "4"},
    "zero_start_by_full_sweep": {"MF3D_ZERO_SWEEP": "0"},
    "prolongation_apart": {"MF3D_FUSE_PROLONG": "0"},
    "all_off": {"MF3D_TAIL": "0", "MF3D_OWN_GHOSTS": "0", "MF3D_XCD_ORDER": "0", "MF3D_AUTOTUNE": "0"},
}


@pytest.mark.parametrize("switch", sorted(SWITCHES))
@pytest.mark.parametrize("name", ["single32_L5", "neumann_L4", "lsm16_L4", "mixed_bc_multi_L3"])
def test_layout_switches_keep_results_bitwise(name, switch, monkeypatch):
    """Each environment switch of the library changes only how the work is
    laid out (tail kernel, who writes a lone block's ghosts, tile order,
    chunk depths, which kernel runs the zero-start sweep, whether a 3-D
    tile level's first post-sweep forms the prolongation itself): FP64 stays
    bitwise the oracle's."""
    for k, v in SWITCHES[switch].items():
        monkeypatch.setenv(k, v)
    grid, L, lsm = case(name)
    f, dens = inputs(grid, lsm, seed=13)
    ncyc = 3
    p_o, rho_o, _, _ = _oracle(grid, L, lsm, dens).solve(f, 0.0, ncyc)
    with _gpu(grid, L, lsm, dens, maxcy=ncyc, eps=0.0) as mg:
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", NotConverged)
            p_g, res = mg.newsolv_mg(f)
    assert res.iters == ncyc
    np.testing.assert_allclose(res.hist, rho_o, rtol=1e-12, atol=0)
    for a, b in zip(p_g, p_o):
        assert np.array_equal(a, b), f"{name}/{switch}: p differs, max {np.max(np.abs(a - b))}"


def test_invalid_pinned_depth_is_rejected(monkeypatch):
    """MF3D_SWEEP_TZ must be an even depth: an odd one fails the handle's
    creation with a message, not a silent fallback."""
    monkeypatch.setenv("MF3D_SWEEP_TZ", "3")
    grid, L, lsm = case("multi_uneven_L3")
    with pytest.raises(Exception, match="MF3D_SWEEP_TZ"):
        PressureMultigrid(grid, L)


@pytest.mark.parametrize("nu", [(2, 2), (1, 1), (2, 1), (1, 0), (3, 2)])
@pytest.mark.parametrize("name", ["single32_L5", "multi_uneven_L3", "periodic_multi_L3", "mixed_bc_multi_L3"])
def test_cycle_shapes_bitwise(name, nu):
    """V(nu_pre, nu_post) cycles other than the default V(2,2), including
    ones that toggle the level-0 potential an even number of times (nu_pre
    + 1 + nu_post): FP64 bitwise the oracle's, through the cycle graphs
    and through the timing hook's launches (timed graphs on the first
    solve, eager launches where graphs cannot record events)."""
    from multiflow3d_amd import _lib

    grid, L, _ = case(name)
    f, _ = inputs(grid, False, seed=17)
    ncyc = 5
    o = OracleMG(grid.blocks, L, nu_pre=nu[0], nu_post=nu[1])
    o.coeff()
    p_o, rho_o, _, _ = o.solve(f, 0.0, ncyc)
    for hook in (False, True):
        with _gpu(grid, L, False, None, maxcy=ncyc, eps=0.0, nu_pre=nu[0], nu_post=nu[1]) as mg:
            if hook:
                mg.time_kernel(_lib.MF3D_K_OUTER)
            for run in range(2 if hook else 1):
                with warnings.catch_warnings():
                    warnings.simplefilter("ignore", NotConverged)
                    p_g, res = mg.newsolv_mg(f)
                assert res.iters == ncyc
                np.testing.assert_allclose(res.hist, rho_o, rtol=1e-12, atol=0)
                for a, b in zip(p_g, p_o):
                    assert np.array_equal(a, b), f"{name}/{nu}/hook={hook}/{run}: max {np.max(np.abs(a - b))}"
            if hook:
                assert mg.kernel_time()[0] >= 1


@pytest.mark.parametrize("zero", ["1", "0"], ids=["zero_march", "zero_full"])
@pytest.mark.parametrize("tz", ["2", "4", "6", "16"])
@pytest.mark.parametrize("name", ["single64_L5", "multi_uneven_L3", "mixed_bc_multi_L3"])
def test_mixed_march_depths_bitwise(name, tz, zero, monkeypatch):
    """The mixed form's march at pinned chunk depths (the finest level's
    FP32 march trails the second colour by two planes, coarser levels by
    one; each chunk pads its steps to whole turns of the LDS ring), its
    zero-start sweep by the march over the RHS or by the full sweep on
    zeroed planes, against the one-pass 3-D tiles: bitwise the same
    pressure."""
    monkeypatch.setenv("MF3D_ZERO_SWEEP", zero)
    grid, L, _ = case(name)
    f, _ = inputs(grid, False, seed=29)
    got = {}
    for tiles in ("3d", "march"):
        monkeypatch.setenv("MF3D_SWEEP_3D", "1" if tiles == "3d" else "0")
        if tiles == "march":
            monkeypatch.setenv("MF3D_SWEEP_TZ", tz)
        with _gpu(grid, L, False, None, maxcy=3, eps=0.0, precision="mixed") as mg:
            with warnings.catch_warnings():
                warnings.simplefilter("ignore", NotConverged)
                got[tiles] = mg.newsolv_mg(f)
    for a, b in zip(got["3d"][0], got["march"][0]):
        assert np.array_equal(a, b), f"{name}/tz={tz}: max {np.max(np.abs(a - b))}"


@pytest.mark.parametrize("name", ["single32_L5", "single64_L5", "multi_uneven_L3", "periodic_multi_L3",
                                  "mixed_bc_multi_L3", "neumann_L4"])
def test_fused_prolongation_bitwise(name, monkeypatch):
    """The first post-sweep of a level on 3-D tiles forms the corrected
    potential itself (prolongation fused into k_rb3d): mixed bitwise the
    same as with k_prolong apart (MF3D_FUSE_PROLONG=0), with the 3-D tiles
    forced on every level so that level 0 takes the fused path too."""
    monkeypatch.setenv("MF3D_SWEEP_3D", "1")
    grid, L, _ = case(name)
    f, _ = inputs(grid, False, seed=31)
    got = {}
    for fuse in ("1", "0"):
        monkeypatch.setenv("MF3D_FUSE_PROLONG", fuse)
        with _gpu(grid, L, False, None, maxcy=4, eps=0.0, precision="mixed") as mg:
            with warnings.catch_warnings():
                warnings.simplefilter("ignore", NotConverged)
                got[fuse] = mg.newsolv_mg(f)
    assert np.array_equal(got["1"][1].hist, got["0"][1].hist)
    for a, b in zip(got["1"][0], got["0"][0]):
        assert np.array_equal(a, b), f"{name}: fused and apart differ, max {np.max(np.abs(a - b))}"


@pytest.mark.parametrize("var,val", [("MF3D_FUSE_PROLONG", "2"), ("MF3D_ZERO_SWEEP", "2")])
def test_invalid_switch_is_rejected(var, val, monkeypatch):
    monkeypatch.setenv(var, val)
    grid, L, _ = case("multi_uneven_L3")
    with pytest.raises(Exception, match=var):
        PressureMultigrid(grid, L)
