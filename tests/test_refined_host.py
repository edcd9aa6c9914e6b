"""The library's ghost fill on the host: mf3d_mg_ghost_table (the table of
ghost expressions a level's exchange becomes -- applied by k_gather_expr on
refined grids) against the CPU oracle's exchangepp + mgbound + exchbc_mgpp,
bitwise, on every level of the refined fixtures' grids and of conforming
block grids; and the refined block sets the library refuses (create returns
MF3D_ENOTSUP before any device work)."""
import ctypes
import json
import os

import numpy as np
import pytest

from multiflow3d_amd import MGError, PressureMultigrid
from multiflow3d_amd import _lib
from multiflow3d_amd import topology as T
from multiflow3d_amd.solver import _descs, make_cfg
from oracle.oracle import OracleMG

from .cases import SMALL, case

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MAN = json.load(open(os.path.join(GOLDEN, "manifest_refined.json")))
N, P = T.NEUMANN, T.PERIODIC


def _table(grid, L, level):
    cfg = make_cfg(len(grid.blocks), L)
    blocks = _descs(grid.blocks, 1)
    cnt = ctypes.c_int64(0)
    err = ctypes.create_string_buffer(512)
    lib = _lib.lib()
    rc = lib.mf3d_mg_ghost_table(ctypes.byref(cfg), blocks, level, None, 0, ctypes.byref(cnt), err, 512)
    if rc != 0:
        return rc, err.value.decode()
    out = np.zeros(10 * cnt.value, dtype=np.int32)
    rc = lib.mf3d_mg_ghost_table(ctypes.byref(cfg), blocks, level,
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), cnt.value,
                                 ctypes.byref(cnt), err, 512)
    assert rc == 0
    return 0, out.reshape(-1, 10)


def _apply(tab, a):
    dst, kind, src = tab[:, 0], tab[:, 1], tab[:, 2:]
    v = np.empty(len(tab))
    k0, k1, k2 = kind == 0, kind == 1, kind == 2
    v[k0] = a[src[k0, 0]]
    s = a[src[k1, 0]]
    for q in range(1, 8):
        s = s + a[src[k1, q]]
    v[k1] = 0.125 * s
    v[k2] = ((2.0 * a[src[k2, 0]] + 3.0 * a[src[k2, 1]]) - a[src[k2, 2]]) / 4.0
    out = a.copy()
    out[dst] = v
    return out


def _check_levels(grid, L, seed):
    rng = np.random.default_rng(seed)
    o = OracleMG(grid.blocks, L)
    o.coeff()
    for level in range(L):
        rc, tab = _table(grid, L, level)
        assert rc == 0, tab
        parts = []
        for b in range(len(grid.blocks)):
            ph = o.phi(b, level)
            ph[...] = np.nan
            ph[1:-1, 1:-1, 1:-1] = rng.standard_normal(ph[1:-1, 1:-1, 1:-1].shape)
            parts.append(ph.copy().ravel())
        got = _apply(tab, np.concatenate(parts))
        o.exchange(level)
        exp = np.concatenate([o.phi(b, level).ravel() for b in range(len(grid.blocks))])
        assert not np.isnan(got).any()
        assert np.array_equal(got, exp), level
        # every ghost is in the table exactly once
        assert len(np.unique(tab[:, 0])) == len(tab) == sum(
            (x[0] + 2) * (x[1] + 2) * (x[2] + 2) - x[0] * x[1] * x[2]
            for x in (o.dims(b, level) for b in range(len(grid.blocks))))
    o.close()


@pytest.mark.parametrize("name", list(MAN))
def test_ghost_table_matches_oracle_exchange_refined(name):
    m = MAN[name]
    g = T.refined_grid(m["ex"], m["ey"], m["ez"], m["bc"], m["h"], m["rdiv"])
    _check_levels(g, m["levels"], 31)


@pytest.mark.parametrize("name", [c for c in SMALL if c != "single1_L1"])
def test_ghost_table_matches_oracle_exchange_conforming(name):
    grid, L, _ = case(name)
    _check_levels(grid, L, 32)


@pytest.mark.parametrize("spec", [
    dict(ex=[8, 8], ey=[8], ez=[8], bc=[P] * 6, rdiv=[1, 2]),           # periodic partner blocks
    dict(ex=[8, 8], ey=[8, 8], ez=[8], bc=[N] * 6, rdiv=[1, 2, 1, 1]),   # blocks touching along an edge
    dict(ex=[8, 8], ey=[8], ez=[8], bc=[N] * 6, rdiv=[1, 4]),           # 4:1 across a face
])
def test_out_of_scope_refined_sets_refused(spec):
    g = T.refined_grid(spec["ex"], spec["ey"], spec["ez"], spec["bc"], [1 / 16] * 3, spec["rdiv"])
    rc, why = _table(g, 3, 0)
    assert rc == _lib.MF3D_ENOTSUP and "refined" in why
    with pytest.raises(MGError, match="refined"):
        PressureMultigrid(g, 3)


@pytest.mark.parametrize("kw", [dict(variable_coef=True)])
def test_refined_options_refused(kw):
    g = T.refined_grid([8, 8], [8], [8], [N] * 6, [1 / 16] * 3, [1, 2])
    with pytest.raises(MGError, match="refined"):
        PressureMultigrid(g, 3, **kw)


def test_refined_extents_must_conform():
    g = T.refined_grid([8, 8], [8], [8], [N] * 6, [1 / 16] * 3, [1, 2])
    g.blocks[1].n[1] = 8  # tangential extent of the refined block no longer 2 x
    with pytest.raises(MGError, match="conform"):
        PressureMultigrid(g, 3)
