"""The allocation-free solve (mf3d_mg_solve_device_noalloc) against the
handle-based mf3d_mg_solve_device on the same inputs: two calls in a row
through the C ABI, the second on a smaller problem, each bitwise equal to a
handle's solve (pressure, residual history, rmax, cycle count)."""
import ctypes

import numpy as np
import pytest

from multiflow3d_amd import _lib, hip
from multiflow3d_amd import topology as T
from multiflow3d_amd import workloads as W
from multiflow3d_amd.solver import PressureMultigrid, _descs, make_cfg

pytestmark = pytest.mark.gpu


def _noalloc(grid, levels, f, **kw):
    L = _lib.lib()
    cfg = make_cfg(len(grid.blocks), levels, **kw)
    descs = _descs(grid.blocks, 1)
    n = sum(int(np.prod(b.n)) for b in grid.blocks)
    fd, pd = hip.DeviceArray.from_host(f), hip.DeviceArray.from_host(np.full(n, np.nan))
    rmax, iters = ctypes.c_double(0.0), ctypes.c_int32(0)
    hist = np.zeros(cfg.maxcy + 1)
    err = ctypes.create_string_buffer(256)
    rc = L.mf3d_mg_solve_device_noalloc(ctypes.byref(cfg), descs, ctypes.c_void_p(fd.ptr),
                                        ctypes.c_void_p(pd.ptr), None, ctypes.byref(rmax),
                                        ctypes.byref(iters), hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                        err, 256)
    assert rc in (0, 1), (rc, err.value)
    p = pd.to_host()
    k = iters.value
    return rc, p, hist[:k + 1].copy(), rmax.value, k


def _handle(grid, levels, f, **kw):
    n = sum(int(np.prod(b.n)) for b in grid.blocks)
    fd, pd = hip.DeviceArray.from_host(f), hip.DeviceArray(n)
    with PressureMultigrid(grid, levels, **kw) as mg:
        res = mg.solve_device(fd.ptr, pd.ptr, warn=False)
    return res.status, pd.to_host(), np.asarray(res.hist), res.rmax, res.iters


CASES = [
    # (grid, levels, solver options): a larger single block, then a smaller multi-block grid
    (lambda: T.single_block([48, 32, 40], [1 / 48, 1 / 32, 1 / 40], T.CHANNEL_BC), 4,
     dict(precision="fp64", maxcy=30, eps=1e-9)),
    (lambda: T.block_grid([8, 16], [8, 8], [16, 8], T.CHANNEL_BC, [1 / 24, 1 / 16, 1 / 24]), 3,
     dict(precision="mixed", maxcy=20, eps=1e-7)),
]


def test_noalloc_equals_handle_two_calls_second_smaller():
    outs = []
    for make_grid, levels, kw in CASES:
        grid = make_grid()
        f = np.concatenate([a.ravel() for a in W.split_blocks(grid, W.global_rhs(grid, W.BASE_SEED + 7))])
        print(f"noalloc case: {len(grid.blocks)} blocks, {f.size} cells, L={levels}, {kw}", flush=True)
        got = _noalloc(grid, levels, f, **kw)
        ref = _handle(grid, levels, f, **kw)
        assert got[0] == ref[0] and got[4] == ref[4] and got[4] > 1
        assert np.array_equal(got[1], ref[1]), "pressure differs from the handle's solve"
        assert np.array_equal(got[2], ref[2]) and got[3] == ref[3]
        outs.append(got[1].size)
    assert outs[1] < outs[0]

