"""Host-side mirror of the reference's pressure-solver interface, on top of
the C ABI.

The reference exposes three argument-less subroutines over module state:

    coeff        mgsolver.for:2-222      coefficient hierarchy
    mgkcyc       mgsolver.for:224-369    one V-cycle on su (sup) into pp
    newsolv_mg   newsolv_mg.for:2-102    outer loop: the reference's velocity
                                         form (calvel/calmas/pbound, press.for)
                                         as ``newsolv``, or the Poisson form
                                         f -> p as ``newsolv_mg``

``PressureMultigrid`` keeps those names; the module state becomes the
constructor's arguments (a block grid and the MG scalars of control.cin)
and explicit arrays.  Error behaviour follows newsolv_mg: not converging
within maxcy warns and returns (newsolv_mg.for:60-61), rmax > 100 raises
(the reference stops, newsolv_mg.for:64-83).

Block arrays are numpy float64 arrays shaped (nz, ny, nx) -- the memory
order of the reference's column-major p(i,j,k) -- either interior-only or
with a one-cell ghost layer, (nz+2, ny+2, nx+2).
"""
from __future__ import annotations

import ctypes
import warnings
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from .topology import Block, BlockGrid


class MGError(RuntimeError):
    pass


class Diverged(MGError):
    pass


class NotConverged(UserWarning):
    pass


@dataclass
class SolveResult:
    status: int          # 0 converged, 1 not converged
    rmax: float          # max |r| dV after the last cycle (reference criterion)
    iters: int
    hist: np.ndarray     # ||r_k|| / ||f||, k = 0..iters


PRECISIONS = {"fp64": _lib.MF3D_FP64, "mixed": _lib.MF3D_MIXED}
COARSE_OPS = {"reference": _lib.MF3D_COARSE_REFERENCE, "face": _lib.MF3D_COARSE_FACE}


def _descs(blocks: Sequence[Block], ghost: int):
    arr = (_lib.BlockDesc * len(blocks))()
    for i, b in enumerate(blocks):
        arr[i].n[:] = list(b.n)
        arr[i].ghost = ghost
        arr[i].ld[:] = [b.n[0] + 2 * ghost, b.n[1] + 2 * ghost, b.n[2] + 2 * ghost]
        arr[i].h[:] = list(b.h)
        arr[i].nbr[:] = list(b.nbr)
        arr[i].per[:] = list(b.per)
        arr[i].bc[:] = list(b.bc)
        arr[i].rdiv = int(getattr(b, "rdiv", 1))
    return arr


COARSE_SOLVES = {"default": _lib.MF3D_COARSEST_DEFAULT, "inner": _lib.MF3D_COARSEST_INNER,
                 "direct": _lib.MF3D_COARSEST_DIRECT}


def make_cfg(nblocks: int, levels: int, *, nu_pre=2, nu_post=2, maxcy=30, eps=1e-8,
             min_cycles=1, precision="fp64", variable_coef=False, device=0, smoother=3,
             coarse_op="reference", krylov=False, coarse_solve="default", differencing=1, pl_ex=0):
    c = _lib.Cfg()
    c.nblocks, c.levels, c.smoother = nblocks, levels, smoother
    c.nu_pre, c.nu_post, c.maxcy = nu_pre, nu_post, maxcy
    c.eps, c.min_cycles = float(eps), min_cycles
    c.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    c.variable_coef = int(bool(variable_coef))
    c.device = device
    c.coarse_op = COARSE_OPS[coarse_op] if isinstance(coarse_op, str) else int(coarse_op)
    c.krylov = int(bool(krylov))
    c.coarse_solve = COARSE_SOLVES[coarse_solve] if isinstance(coarse_solve, str) else int(coarse_solve)
    c.differencing = int(differencing)
    c.pl_ex = int(pl_ex)
    return c


def estimate_bytes(grid_or_blocks, levels: int, **kw) -> int:
    blocks = grid_or_blocks.blocks if isinstance(grid_or_blocks, BlockGrid) else list(grid_or_blocks)
    cfg = make_cfg(len(blocks), levels, **kw)
    v = _lib.lib().mf3d_mg_estimate_bytes(ctypes.byref(cfg), _descs(blocks, 1))
    if v < 0:
        raise MGError("invalid configuration")
    return int(v)


class PressureMultigrid:
    """One solver handle: device-resident level hierarchy for a block set."""

    GHOST = 1   # default ghost layers of caller arrays (the reference's pl with LMR 1, pl_ex 0)

    def __init__(self, grid_or_blocks, levels: int, *, nu_pre: int = 2, nu_post: int = 2,
                 maxcy: int = 30, eps: float = 1e-8, min_cycles: int = 1,
                 precision: str = "fp64", variable_coef: bool = False, device: int = 0,
                 smoother: int = 3, coarse_op: str = "reference", krylov: bool = False,
                 coarse_solve: str = "default", differencing: int = 1, pl_ex: int = 0,
                 ghost: Optional[int] = None):
        """ghost: layers of every caller array (the reference's pl; default 1).
        The velocity form (newsolv) works on the whole shell: exchange(1..3)
        refreshes pl_ex extra layers, differencing = 2 selects the 4th-order
        calmas.  coarse_solve: "default" (the reference's smoothing of the
        coarsest level, in FP64 in both precisions), "inner" (that smoothing
        in the inner precision) or "direct" (an exact FP64 solve)."""
        self.blocks: List[Block] = (grid_or_blocks.blocks if isinstance(grid_or_blocks, BlockGrid)
                                    else list(grid_or_blocks))
        self.levels = levels
        self.maxcy = maxcy
        self.variable_coef = variable_coef
        self.ghost = int(self.GHOST if ghost is None else ghost)
        self.cfg = make_cfg(len(self.blocks), levels, nu_pre=nu_pre, nu_post=nu_post,
                            maxcy=maxcy, eps=eps, min_cycles=min_cycles, precision=precision,
                            variable_coef=variable_coef, device=device, smoother=smoother,
                            coarse_op=coarse_op, krylov=krylov, coarse_solve=coarse_solve,
                            differencing=differencing, pl_ex=pl_ex)
        L = _lib.lib()
        h = ctypes.c_void_p()
        rc = L.mf3d_mg_create(ctypes.byref(self.cfg), _descs(self.blocks, self.ghost), ctypes.byref(h))
        self._h = h
        if rc != 0:
            msg = L.mf3d_mg_last_error(h).decode() if h.value else ""
            self.close()
            raise MGError(f"mf3d_mg_create failed ({rc}): {msg}")

    # ------------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().mf3d_mg_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _err(self) -> str:
        return _lib.lib().mf3d_mg_last_error(self._h).decode()

    @property
    def cells(self) -> int:
        return int(_lib.lib().mf3d_mg_cells(self._h))

    @property
    def device_bytes(self) -> int:
        return int(_lib.lib().mf3d_mg_device_bytes(self._h))

    # ------------------------------------------------------------------
    def _ghosted(self, arrays, copy_in=True):
        """Caller arrays with self.ghost layers: given so, or from interiors
        (nz, ny, nx) or one-layer boxes (nz+2, ny+2, nx+2), zero-padded."""
        out = []
        G = self.ghost
        for b, a in zip(self.blocks, arrays):
            nx, ny, nz = b.n
            a = np.asarray(a, dtype=np.float64)
            if a.shape == (nz + 2 * G, ny + 2 * G, nx + 2 * G):
                out.append(np.ascontiguousarray(a))
                continue
            g = np.zeros((nz + 2 * G, ny + 2 * G, nx + 2 * G))
            if a.shape == (nz, ny, nx):
                if copy_in:
                    g[G:-G, G:-G, G:-G] = a
            elif a.shape == (nz + 2, ny + 2, nx + 2):
                if copy_in:
                    g[G - 1:g.shape[0] - G + 1, G - 1:g.shape[1] - G + 1, G - 1:g.shape[2] - G + 1] = a
            else:
                raise ValueError(f"block array shape {a.shape} does not match extent {b.n}")
            out.append(g)
        return out

    @staticmethod
    def _ptrs(arrays):
        return (ctypes.c_void_p * len(arrays))(*[a.ctypes.data for a in arrays])

    def coeff(self, density=None):
        """Coefficient hierarchy; under LSM from the cell density of each block."""
        if density is None:
            if self.variable_coef:
                raise MGError("variable-coefficient handle needs a density")
            return
        dens = self._ghosted(density)
        rc = _lib.lib().mf3d_mg_set_density(self._h, self._ptrs(dens))
        if rc != 0:
            raise MGError(f"coeff failed ({rc}): {self._err()}")

    def coeff_device(self, dens_ptr: int, stream_ptr: int = 0):
        """coeff from a device-resident density: ghosted_cells doubles (all
        blocks' ghosted arrays, the layout of newsolv_device)."""
        rc = _lib.lib().mf3d_mg_set_density_device(self._h, ctypes.c_void_p(int(dens_ptr)),
                                                   ctypes.c_void_p(stream_ptr or None))
        if rc != 0:
            raise MGError(f"coeff failed ({rc}): {self._err()}")

    @property
    def shell_cells(self) -> int:
        """Length of newsolv_device's fields (blocks' arrays with their ghost layers)."""
        return int(_lib.lib().mf3d_mg_shell_cells(self._h))

    def heaviside(self, phi, densl: float, densg: float, mul: float, mug: float):
        """heaviside (lsm.for:1456-1509) + bound_lsm(18/19), then coeff from
        the new density.  phi: per-block ghosted arrays (nz+2, ny+2, nx+2)
        with their ghost layer filled.  Returns (dens, mu) as new ghosted
        arrays (interior and first ghost layer)."""
        ph = self._vel_arrays(phi, "phi", False)
        dens = [np.zeros_like(a) for a in ph]
        mu = [np.zeros_like(a) for a in ph]
        rc = _lib.lib().mf3d_mg_heaviside(self._h, self._ptrs(ph), float(densl), float(densg), float(mul),
                                          float(mug), self._ptrs(dens), self._ptrs(mu))
        if rc != 0:
            raise MGError(f"heaviside failed ({rc}): {self._err()}")
        return dens, mu

    def heaviside_device(self, phi_ptr: int, densl: float, densg: float, mul: float, mug: float,
                         dens_ptr: int = 0, mu_ptr: int = 0, stream_ptr: int = 0):
        """Device-resident heaviside + coeff: phi_ptr, dens_ptr, mu_ptr are
        ghosted_cells doubles each (dens and mu optional)."""
        rc = _lib.lib().mf3d_mg_heaviside_device(
            self._h, ctypes.c_void_p(int(phi_ptr)), float(densl), float(densg), float(mul), float(mug),
            ctypes.c_void_p(dens_ptr or None), ctypes.c_void_p(mu_ptr or None),
            ctypes.c_void_p(stream_ptr or None))
        if rc != 0:
            raise MGError(f"heaviside failed ({rc}): {self._err()}")

    def mgkcyc(self, rhs, pp):
        """One V-cycle on ``rhs``; ``pp`` (ghosted arrays) is updated in place and returned."""
        r = self._ghosted(rhs)
        G = self.ghost
        for b, a in zip(self.blocks, pp):
            if a.shape != (b.n[2] + 2 * G, b.n[1] + 2 * G, b.n[0] + 2 * G) or a.dtype != np.float64 \
                    or not a.flags.c_contiguous:
                raise ValueError(f"pp blocks must be C-contiguous float64 with {G} ghost layer(s)")
        rc = _lib.lib().mf3d_mg_vcycle(self._h, self._ptrs(r), self._ptrs(list(pp)))
        if rc != 0:
            raise MGError(f"mgkcyc failed ({rc}): {self._err()}")
        return pp

    def _finish(self, rc, rmax, iters, warn=True):
        hist = self._hist[: iters.value + 1].copy()
        res = SolveResult(status=rc, rmax=rmax.value, iters=iters.value, hist=hist)
        if rc == _lib.MF3D_DIVERGED:
            raise Diverged(f"BIG RMAX!! STOP!!!!! rmax={rmax.value!r}")
        if rc < 0:
            raise MGError(f"solve failed ({rc}): {self._err()}")
        if rc == _lib.MF3D_NOT_CONVERGED and warn:
            warnings.warn(f"not converged!! {self.maxcy} {rmax.value!r}", NotConverged)
        return res

    def newsolv_mg(self, f, p: Optional[list] = None):
        """Solve A p = f from p = 0; returns (p blocks interior, SolveResult)."""
        fg = self._ghosted(f)
        pg = [np.zeros_like(a) for a in fg]
        rmax = ctypes.c_double(0.0)
        iters = ctypes.c_int32(0)
        self._hist = np.zeros(self.maxcy + 1)
        rc = _lib.lib().mf3d_mg_solve(self._h, self._ptrs(fg), self._ptrs(pg), ctypes.byref(rmax),
                                      ctypes.byref(iters),
                                      self._hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        res = self._finish(rc, rmax, iters)
        G = self.ghost
        out = [a[G:-G, G:-G, G:-G].copy() for a in pg]
        if p is not None:
            for dst, src in zip(p, out):
                if dst.shape == src.shape:
                    dst[...] = src
                else:
                    dst[1:-1, 1:-1, 1:-1] = src
        return out, res

    @property
    def ghosted_cells(self) -> int:
        """Length of each device field of newsolv_device (all blocks' ghosted arrays)."""
        return int(_lib.lib().mf3d_mg_ghosted_cells(self._h))

    def _vel_arrays(self, arrays, name, writable):
        out = []
        G = self.ghost
        for b, a in zip(self.blocks, arrays):
            shp = (b.n[2] + 2 * G, b.n[1] + 2 * G, b.n[0] + 2 * G)
            if a.shape != shp or a.dtype != np.float64 or not a.flags.c_contiguous or \
                    (writable and not a.flags.writeable):
                raise ValueError(f"{name} blocks must be C-contiguous float64 arrays shaped {shp}")
            out.append(a)
        if len(out) != len(self.blocks):
            raise ValueError(f"{name}: one array per block")
        return out

    def newsolv(self, ustar, vstar, wstar, u, v, w, p, dt: float, alfapr: float, eps: float,
                su: Optional[list] = None, warn: bool = True) -> SolveResult:
        """newsolv_mg in the reference's velocity form (newsolv_mg.for:12-58):
        calvel/calmas, then V-cycles on su (sup under LSM), p += pp,
        exchange(4)/pbound, calvel/calmas, until rmax < eps after two cycles.
        All fields are per-block arrays with the handle's ghost layers
        (nz+2g, ny+2g, nx+2g); u, v, w and p are updated in place over the
        interior faces and the ghost shell the reference's exchanges write
        (faces on walls and untouched ghosts keep their values), su
        (optional, same shape) gets the last right-hand side in its
        interior.  Returns a SolveResult whose hist is rmax per cycle (index 0
        before the first).  Under LSM, coeff(density) must have received the
        ghosted density."""
        arrs = [self._vel_arrays(x, nm, wr) for x, nm, wr in
                ((ustar, "ustar", False), (vstar, "vstar", False), (wstar, "wstar", False),
                 (u, "u", True), (v, "v", True), (w, "w", True), (p, "p", True))]
        sup = None
        if su is not None:
            sup = self._ptrs(self._vel_arrays(su, "su", True))
        rmax = ctypes.c_double(0.0)
        iters = ctypes.c_int32(0)
        self._hist = np.zeros(self.maxcy + 1)
        rc = _lib.lib().mf3d_mg_newsolv(
            self._h, *[self._ptrs(a) for a in arrs], float(dt), float(alfapr), float(eps), sup,
            ctypes.byref(rmax), ctypes.byref(iters),
            self._hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return self._finish(rc, rmax, iters, warn)

    def newsolv_device(self, ptrs, dt: float, alfapr: float, eps: float, su_ptr: int = 0,
                       stream_ptr: int = 0, warn: bool = True) -> SolveResult:
        """Device-resident velocity form: ptrs = (ustar, vstar, wstar, u, v, w, p)
        device addresses of shell_cells doubles each (every block's array
        with the handle's ghost layers; e.g. torch tensors' data_ptr());
        su_ptr optional."""
        if len(ptrs) != 7:
            raise ValueError("ptrs: ustar, vstar, wstar, u, v, w, p")
        rmax = ctypes.c_double(0.0)
        iters = ctypes.c_int32(0)
        self._hist = np.zeros(self.maxcy + 1)
        rc = _lib.lib().mf3d_mg_newsolv_device(
            self._h, *[ctypes.c_void_p(int(x)) for x in ptrs], float(dt), float(alfapr), float(eps),
            ctypes.c_void_p(su_ptr or None), ctypes.c_void_p(stream_ptr or None), ctypes.byref(rmax),
            ctypes.byref(iters), self._hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return self._finish(rc, rmax, iters, warn)

    def time_kernel(self, kernel: int):
        """Enable the finest-level timing hook for one kernel (MF3D_K_*; 0 = off)."""
        rc = _lib.lib().mf3d_mg_time_kernel(self._h, int(kernel))
        if rc != 0:
            raise MGError(f"time_kernel failed ({rc})")

    def kernel_time(self):
        """(launches, total_ms) recorded by the timing hook."""
        n = ctypes.c_int64(0)
        ms = ctypes.c_double(0.0)
        _lib.lib().mf3d_mg_kernel_time(self._h, ctypes.byref(n), ctypes.byref(ms))
        return int(n.value), float(ms.value)

    def set_eps(self, eps: float):
        """Stop tolerance of the following solves."""
        rc = _lib.lib().mf3d_mg_set_eps(self._h, float(eps))
        if rc != 0:
            raise MGError(f"set_eps failed ({rc})")

    def graph_replay(self) -> bool:
        """Whether the last solve replayed captured cycle graphs."""
        return bool(_lib.lib().mf3d_mg_graph_replay(self._h))

    def solve_device(self, f_dev_ptr: int, p_dev_ptr: int, stream_ptr: int = 0,
                     warn: bool = True) -> SolveResult:
        """Device-resident solve on interior-packed device vectors (e.g. torch
        tensors' data_ptr()); the fields never leave HBM.  warn=False leaves
        non-convergence to the returned status (no NotConverged warning)."""
        rmax = ctypes.c_double(0.0)
        iters = ctypes.c_int32(0)
        self._hist = np.zeros(self.maxcy + 1)
        rc = _lib.lib().mf3d_mg_solve_device(
            self._h, ctypes.c_void_p(f_dev_ptr), ctypes.c_void_p(p_dev_ptr),
            ctypes.c_void_p(stream_ptr), ctypes.byref(rmax), ctypes.byref(iters),
            self._hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return self._finish(rc, rmax, iters, warn)
