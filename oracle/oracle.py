"""ctypes front end of the CPU oracle (liborc_mg.so) and of the compiled
reference driver (_ref/mg_ref).

TEST INFRASTRUCTURE ONLY.  Import it from tests/, __graft_entry__.smoke()
and bench.py's CPU-baseline leg, and nowhere else: the tests compare the
library against it; it is never the thing measured or shipped.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import tempfile
from typing import List, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc_mg.so")
LIB_OMP_PATH = os.path.join(HERE, "liborc_mg_omp.so")  # same code, OpenMP plane loops
REF_BIN = os.path.join(HERE, "_ref", "mg_ref")
REF_VEL_BIN = os.path.join(HERE, "_ref", "mg_ref_vel")


class OrcBlock(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32 * 3), ("h", ctypes.c_double * 3),
                ("nbr", ctypes.c_int32 * 6), ("per", ctypes.c_int32 * 6),
                ("bc", ctypes.c_int32 * 6), ("rdiv", ctypes.c_int32)]


_libs = {}


def lib(omp: bool = False):
    """The serial oracle, or (omp) its OpenMP build, whose results are
    bitwise the serial build's; threads follow OMP_NUM_THREADS."""
    path = LIB_OMP_PATH if omp else LIB_PATH
    if path not in _libs:
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run make -C oracle)")
        L = ctypes.CDLL(path)
        vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ip, ctypes.POINTER(OrcBlock), ip, ip, ip, ip]
        L.orc_destroy.argtypes = [vp]
        L.orc_set_density.argtypes = [vp, dp]
        L.orc_set_smoother.restype = ip
        L.orc_set_smoother.argtypes = [vp, ip]
        L.orc_set_poison.argtypes = [vp, ip]
        L.orc_coeff.argtypes = [vp]
        L.orc_vcycle.argtypes = [vp, dp, dp]
        L.orc_solve.restype = ip
        L.orc_solve.argtypes = [vp, dp, dp, ctypes.c_double, ip, ip, dp, dp]
        L.orc_solve_gcr.restype = ip
        L.orc_solve_gcr.argtypes = [vp, dp, dp, ctypes.c_double, ip, ip, dp, dp]
        L.orc_dims.argtypes = [vp, ip, ip, ctypes.POINTER(ctypes.c_int32)]
        for name in ("orc_phi", "orc_rhs", "orc_dens"):
            getattr(L, name).restype = dp
            getattr(L, name).argtypes = [vp, ip, ip]
        L.orc_cof.restype = dp
        L.orc_cof.argtypes = [vp, ip, ip, ip]
        for name in ("orc_relax",):
            getattr(L, name).argtypes = [vp, ip, ip]
        for name in ("orc_restrict", "orc_correct", "orc_exchange", "orc_bound"):
            getattr(L, name).argtypes = [vp, ip]
        L.orc_residual.argtypes = [vp, ip, dp]
        _libs[path] = L
    return _libs[path]


def _dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _concat(parts) -> np.ndarray:
    return np.ascontiguousarray(np.concatenate([np.ravel(p).astype(np.float64) for p in parts]))


class OracleMG:
    """FP64 multigrid of the reference, restated on the CPU."""

    def __init__(self, blocks, levels: int, lsm: bool = False, nu_pre: int = 2, nu_post: int = 2,
                 coarse_op: str = "reference", smoother: int = 3, omp: bool = False):
        """lsm: variable (density) coefficients; coarse_op "face" builds their
        coarse operators by face-coefficient restriction instead of the
        reference's density hierarchy; smoother: mg_itrsch (1 lexicographic
        GS, 2 ADI-TDMA, 3 red-black GS); omp: run on the OpenMP build."""
        self._L = lib(omp)
        self.blocks = list(blocks)
        self.levels = levels
        self.lsm = lsm
        arr = (OrcBlock * len(self.blocks))()
        for i, b in enumerate(self.blocks):
            arr[i].n[:] = list(b.n)
            arr[i].h[:] = list(b.h)
            arr[i].nbr[:] = list(b.nbr)
            arr[i].per[:] = list(b.per)
            arr[i].bc[:] = list(b.bc)
            arr[i].rdiv = int(getattr(b, "rdiv", 1))
        mode = (2 if coarse_op == "face" else 1) if lsm else 0
        self._h = self._L.orc_create(len(self.blocks), arr, levels, mode, nu_pre, nu_post)
        if not self._h:
            raise ValueError("oracle rejected the block set (extent not divisible by 2**(levels-1), "
                             "or refined blocks outside the covered scope)")
        if self._L.orc_set_smoother(self._h, smoother) != 0:
            raise ValueError(f"unknown smoother {smoother}")

    def close(self):
        if self._h:
            self._L.orc_destroy(self._h)
            self._h = None

    __del__ = close

    def set_poison(self, on: bool = True):
        """NaN into the 2:1 fine-side ghost rows no sender fills."""
        self._L.orc_set_poison(self._h, int(on))

    def dims(self, block: int, level: int):
        out = (ctypes.c_int32 * 3)()
        self._L.orc_dims(self._h, block, level, out)
        return tuple(out)

    def set_density(self, dens_blocks):
        d = _concat(dens_blocks)
        self._L.orc_set_density(self._h, _dptr(d))

    def coeff(self):
        self._L.orc_coeff(self._h)

    def vcycle(self, rhs_blocks, pp_blocks=None):
        """One mgkcyc; returns pp per block with its ghost layer, shaped (z+2, y+2, x+2)."""
        rhs = _concat(rhs_blocks)
        shapes = [tuple(x + 2 for x in reversed(self.dims(b, 0))) for b in range(len(self.blocks))]
        if pp_blocks is None:
            pp = np.zeros(sum(int(np.prod(s)) for s in shapes))
        else:
            pp = _concat(pp_blocks)
        self._L.orc_vcycle(self._h, _dptr(rhs), _dptr(pp))
        out, o = [], 0
        for s in shapes:
            m = int(np.prod(s))
            out.append(pp[o:o + m].reshape(s).copy())
            o += m
        return out

    def solve(self, f_blocks, tol: float, maxcy: int, min_cycles: int = 1, gcr: bool = False):
        """Poisson-form outer loop from p = 0 (GCR(1) around the V-cycle with
        gcr=True).  Returns (p_blocks, rho, rmax, cycles)."""
        f = _concat(f_blocks)
        p = np.zeros_like(f)
        rho = np.zeros(maxcy + 1)
        rmax = np.zeros(maxcy + 1)
        fn = self._L.orc_solve_gcr if gcr else self._L.orc_solve
        k = fn(self._h, _dptr(f), _dptr(p), tol, maxcy, min_cycles, _dptr(rho), _dptr(rmax))
        return self._split(p), rho[:k + 1].copy(), rmax[:k + 1].copy(), k

    def _split(self, flat):
        out, o = [], 0
        for b in range(len(self.blocks)):
            nx, ny, nz = self.dims(b, 0)
            m = nx * ny * nz
            out.append(flat[o:o + m].reshape(nz, ny, nx).copy())
            o += m
        return out

    # -- unit access (views into oracle memory; copy before the next call) --
    def phi(self, block: int, level: int) -> np.ndarray:
        nx, ny, nz = self.dims(block, level)
        p = self._L.orc_phi(self._h, block, level)
        return np.ctypeslib.as_array(p, shape=((nz + 2) * (ny + 2) * (nx + 2),)).reshape(nz + 2, ny + 2, nx + 2)

    def rhs(self, block: int, level: int) -> np.ndarray:
        nx, ny, nz = self.dims(block, level)
        p = self._L.orc_rhs(self._h, block, level)
        return np.ctypeslib.as_array(p, shape=(nz * ny * nx,)).reshape(nz, ny, nx)

    def cof(self, block: int, level: int, which: int) -> np.ndarray:
        nx, ny, nz = self.dims(block, level)
        p = self._L.orc_cof(self._h, block, level, which)
        return np.ctypeslib.as_array(p, shape=(nz * ny * nx,)).reshape(nz, ny, nx)

    def dens(self, block: int, level: int) -> np.ndarray:
        nx, ny, nz = self.dims(block, level)
        p = self._L.orc_dens(self._h, block, level)
        return np.ctypeslib.as_array(p, shape=(nz * ny * nx,)).reshape(nz, ny, nx)

    def relax(self, level: int, sweeps: int):
        self._L.orc_relax(self._h, level, sweeps)

    def restrict(self, coarse_level: int):
        self._L.orc_restrict(self._h, coarse_level)

    def correct(self, fine_level: int):
        self._L.orc_correct(self._h, fine_level)

    def exchange(self, level: int):
        self._L.orc_exchange(self._h, level)

    def bound(self, level: int):
        self._L.orc_bound(self._h, level)

    def residual(self, level: int):
        tot = sum(int(np.prod(self.dims(b, level))) for b in range(len(self.blocks)))
        r = np.zeros(tot)
        self._L.orc_residual(self._h, level, _dptr(r))
        out, o = [], 0
        for b in range(len(self.blocks)):
            nx, ny, nz = self.dims(b, level)
            m = nx * ny * nz
            out.append(r[o:o + m].reshape(nz, ny, nx).copy())
            o += m
        return out


def have_ref() -> bool:
    return os.path.exists(REF_BIN)


def run_ref(grid, levels: int, f_blocks, tol: float, maxcy: int, min_cycles: int = 1,
            lsm: bool = False, dens_blocks=None, nu_pre: int = 2, nu_post: int = 2, smoother: int = 3):
    """Run the compiled reference driver (mg_itrsch = smoother).  Returns
    (p_blocks, rho, rmax, cycles, pp1_blocks)."""
    if not have_ref():
        raise RuntimeError("compiled reference not built (make -C oracle ref)")
    with tempfile.TemporaryDirectory() as td:
        frhs = os.path.join(td, "rhs.bin")
        _concat(f_blocks).tofile(frhs)
        fden = "-"
        if lsm:
            fden = os.path.join(td, "dens.bin")
            _concat(dens_blocks).tofile(fden)
        fout = os.path.join(td, "out.bin")
        ctl = os.path.join(td, "ctl.txt")
        with open(ctl, "w") as fh:
            nbx, nby, nbz = grid.shape
            # levels is ngrd_gl; ref_driver takes ngrid_input = ngrd_gl - log2(max rdiv)
            lin = levels - (max(getattr(b, "rdiv", 1) for b in grid.blocks).bit_length() - 1)
            fh.write(f"{nbx} {nby} {nbz} {lin} {int(lsm)} {maxcy} {min_cycles} {nu_pre} {nu_post}\n")
            fh.write(f"{tol!r} {smoother}\n")
            fh.write(" ".join(str(c) for c in grid.bc) + "\n")
            fh.write(" ".join(map(str, grid.ex)) + "\n")
            fh.write(" ".join(map(str, grid.ey)) + "\n")
            fh.write(" ".join(map(str, grid.ez)) + "\n")
            fh.write(" ".join(repr(float(x)) for x in grid.h) + "\n")
            fh.write(frhs + "\n" + fden + "\n" + fout + "\n")
            if any(getattr(b, "rdiv", 1) > 1 for b in grid.blocks):
                fh.write(" ".join(str(b.rdiv) for b in grid.blocks) + "\n")
        subprocess.run([REF_BIN, ctl], check=True, capture_output=True, timeout=600)
        raw = np.fromfile(fout, dtype=np.float64)
    k = int(raw[0])
    o = 1
    rho = raw[o:o + maxcy + 1]; o += maxcy + 1
    rmax = raw[o:o + maxcy + 1]; o += maxcy + 1
    p, pp1 = [], []
    for blk in grid.blocks:
        nx, ny, nz = blk.n
        m = nx * ny * nz
        p.append(raw[o:o + m].reshape(nz, ny, nx).copy()); o += m
    for blk in grid.blocks:
        nx, ny, nz = blk.n
        m = (nx + 2) * (ny + 2) * (nz + 2)
        pp1.append(raw[o:o + m].reshape(nz + 2, ny + 2, nx + 2).copy()); o += m
    global LAST_REF_SECONDS
    LAST_REF_SECONDS = float(raw[o]) if o < len(raw) else None
    return p, rho[:k + 1].copy(), rmax[:k + 1].copy(), k, pp1


LAST_REF_SECONDS = None  # solve time of the last run_ref, measured inside ref_driver


def have_ref_vel() -> bool:
    return os.path.exists(REF_VEL_BIN)


def _ref_vel_call(grid, levels, mode, parts, maxcy, eps, dt, alfapr, props=None, nu_pre=2, nu_post=2,
                  smoother=3, cwd=None, pl_ex=0, lmr=1, differencing=1):
    """One run of mg_ref_vel in `mode` (0 constant, 1 density, 2 level set,
    3 bound_lsm only, 4 constant + checkpoint files written into cwd) on
    the concatenated input `parts` (arrays with lmr + pl_ex ghost layers);
    returns its raw f64 output."""
    if not have_ref_vel():
        raise RuntimeError("compiled reference not built (make -C oracle ref)")
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        _concat(parts).tofile(fin)
        fout = os.path.join(td, "out.bin")
        ctl = os.path.join(td, "ctl.txt")
        with open(ctl, "w") as fh:
            nbx, nby, nbz = grid.shape
            fh.write(f"{nbx} {nby} {nbz} {levels} {mode} {maxcy} {nu_pre} {nu_post}\n")
            fh.write(f"{eps!r} {smoother} {pl_ex} {lmr} {differencing}\n")
            fh.write(" ".join(str(c) for c in grid.bc) + "\n")
            fh.write(" ".join(map(str, grid.ex)) + "\n")
            fh.write(" ".join(map(str, grid.ey)) + "\n")
            fh.write(" ".join(map(str, grid.ez)) + "\n")
            fh.write(" ".join(repr(float(x)) for x in grid.h) + "\n")
            fh.write(f"{dt!r} {alfapr!r}\n")
            if mode == 2:
                fh.write(" ".join(repr(float(x)) for x in props) + "\n")
            fh.write(fin + "\n" + fout + "\n")
        subprocess.run([REF_VEL_BIN, ctl], check=True, capture_output=True, timeout=600, cwd=cwd)
        return np.fromfile(fout, dtype=np.float64)


def _boxes(grid, raw, o, keys, G=1):
    out = {key: [] for key in keys}
    for blk in grid.blocks:
        nx, ny, nz = blk.n
        m = (nx + 2 * G) * (ny + 2 * G) * (nz + 2 * G)
        for key in keys:
            out[key].append(raw[o:o + m].reshape(nz + 2 * G, ny + 2 * G, nx + 2 * G).copy())
            o += m
    return out, o


def run_ref_vel(grid, levels: int, ustar, vstar, wstar, p0, eps: float, maxcy: int, dt: float,
                alfapr: float, lsm: bool = False, dens=None, nu_pre: int = 2, nu_post: int = 2,
                phi=None, props=None, smoother: int = 3, pl_ex: int = 0, lmr: int = 1,
                differencing: int = 1):
    """Run the compiled reference's velocity-form newsolv_mg loop (calvel,
    calmas, mgkcyc, p update, exchange(4), pbound).  Every field is a list of
    full ghosted block arrays with G = lmr + pl_ex ghost layers (the
    reference's pl), shaped (z+2G, y+2G, x+2G).  Returns (iterations, rmax
    history, p, u, v, w, su) with the fields as lists of ghosted block
    arrays.  With a level set `phi` and props = (densl, densg, mul, mug)
    the reference's heaviside runs first (lsm.for:1456-1509) and dens and
    mu are returned as well."""
    mode = 2 if phi is not None else int(lsm)
    parts = []
    for b in range(len(grid.blocks)):
        parts += [ustar[b], vstar[b], wstar[b], p0[b]]
        if mode == 1:
            parts.append(dens[b])
        elif mode == 2:
            parts.append(phi[b])
    raw = _ref_vel_call(grid, levels, mode, parts, maxcy, eps, dt, alfapr, props, nu_pre, nu_post, smoother,
                        pl_ex=pl_ex, lmr=lmr, differencing=differencing)
    k = int(raw[0])
    o = 1
    rmh = raw[o:o + maxcy + 1].copy(); o += maxcy + 1
    keys = ("p", "u", "v", "w", "su") + (("dens", "mu") if mode == 2 else ())
    out, o = _boxes(grid, raw, o, keys, G=lmr + pl_ex)
    res = (k, rmh[:k + 1], out["p"], out["u"], out["v"], out["w"], out["su"])
    return res + (out["dens"], out["mu"]) if mode == 2 else res


CHECKPOINT_FIELDS = ("p", "pm", "ppm", "vis", "u", "um", "uum", "v", "vm", "vvm", "w", "wm", "wwm", "uvm", "uwm",
                     "vwm")


def run_ref_checkpoint(grid, levels: int, ustar, vstar, wstar, p0, maxcy: int, dt: float, alfapr: float,
                       workdir: str):
    """The compiled reference's velocity-form loop (eps = 0, maxcy cycles)
    followed by its own checkpoint writer tecbin (post.for:496-548) and
    final_ctime.dat, both written into workdir.  Returns, per block, the
    dict of what was checkpointed: x, y, z (1-D) and the 16 fields of
    CHECKPOINT_FIELDS as ghosted arrays shaped (z+2, y+2, x+2)."""
    parts = []
    for b in range(len(grid.blocks)):
        parts += [ustar[b], vstar[b], wstar[b], p0[b]]
    raw = _ref_vel_call(grid, levels, 4, parts, maxcy, 0.0, dt, alfapr, cwd=workdir)
    o = 1 + maxcy + 1
    _, o = _boxes(grid, raw, o, ("p", "u", "v", "w", "su"))
    out = []
    for blk in grid.blocks:
        nx, ny, nz = blk.n
        d = {}
        for key, m in (("x", nx + 2), ("y", ny + 2), ("z", nz + 2)):
            d[key] = raw[o:o + m].copy()
            o += m
        m = (nx + 2) * (ny + 2) * (nz + 2)
        for key in CHECKPOINT_FIELDS:
            d[key] = raw[o:o + m].reshape(nz + 2, ny + 2, nx + 2).copy()
            o += m
        out.append(d)
    return out


def run_ref_bound_lsm(grid, field):
    """The compiled reference's bound_lsm(19) (bounds_lsm.for:2-104, with
    exchange(19)) on a cell field given as ghosted block arrays; returns the
    field after it."""
    raw = _ref_vel_call(grid, 1, 3, list(field), 0, 0.0, 1.0, 1.0)
    return _boxes(grid, raw, 0, ("f",))[0]["f"]
