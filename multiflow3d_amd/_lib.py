"""ctypes binding of libmf3d_mg.so (the C ABI declared in include/mf3d_mg.h).

There is no fallback: if the library is missing or does not load, every
entry point raises.  Build it with ``python -m multiflow3d_amd.build``.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmf3d_mg.so")

MF3D_OK, MF3D_NOT_CONVERGED, MF3D_DIVERGED = 0, 1, 2
MF3D_EINVAL, MF3D_ENOTSUP, MF3D_ENOMEM, MF3D_EDEVICE = -22, -95, -12, -5
MF3D_FP64, MF3D_MIXED = 0, 1

# every symbol include/mf3d_mg.h declares
EXPORTS = (
    "mf3d_mg_create", "mf3d_mg_set_density", "mf3d_mg_vcycle", "mf3d_mg_solve",
    "mf3d_mg_solve_device", "mf3d_mg_cells", "mf3d_mg_device_bytes",
    "mf3d_mg_estimate_bytes", "mf3d_mg_last_error", "mf3d_mg_destroy",
    "mf3d_mg_abi_version", "mf3d_mg_time_kernel", "mf3d_mg_kernel_time",
    "mf3d_mg_graph_replay", "mf3d_mg_set_eps", "mf3d_mg_newsolv", "mf3d_mg_newsolv_device",
    "mf3d_mg_ghosted_cells", "mf3d_mg_set_density_device", "mf3d_mg_heaviside",
    "mf3d_mg_heaviside_device", "mf3d_tecbin_header", "mf3d_tecbin_read", "mf3d_final_ctime_read",
    "mf3d_io_last_error", "mf3d_mg_solve_device_noalloc", "mf3d_mg_shell_cells",
    "mf3d_mg_ghost_table",
)
MF3D_NOALLOC_SCRATCH_BYTES = 32 << 20
MF3D_ENOENT, MF3D_EFORMAT = -2, -74
MF3D_COARSE_REFERENCE, MF3D_COARSE_FACE = 0, 1
MF3D_OUTER_DEFECT, MF3D_OUTER_GCR = 0, 1
MF3D_COARSEST_DEFAULT, MF3D_COARSEST_INNER, MF3D_COARSEST_DIRECT = 0, 1, 2
MF3D_COARSEST_MAX_CELLS = 4096
MF3D_K_OFF, MF3D_K_SMOOTH, MF3D_K_OUTER, MF3D_K_RESTRICT, MF3D_K_PROLONG, MF3D_K_HEAVISIDE = 0, 1, 2, 3, 4, 5


class BlockDesc(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32 * 3), ("ghost", ctypes.c_int32),
                ("ld", ctypes.c_int32 * 3), ("h", ctypes.c_double * 3),
                ("nbr", ctypes.c_int32 * 6), ("per", ctypes.c_int32 * 6),
                ("bc", ctypes.c_int32 * 6), ("rdiv", ctypes.c_int32)]


class TecbinHdr(ctypes.Structure):
    _fields_ = [("tt", ctypes.c_int32 * 3), ("pl", ctypes.c_int32), ("ind", ctypes.c_int32 * 3)]


class FinalCtime(ctypes.Structure):
    _fields_ = [("ctime", ctypes.c_double), ("forcn", ctypes.c_double), ("qstpn", ctypes.c_double),
                ("ntav1_count", ctypes.c_double), ("ntav2_count", ctypes.c_double),
                ("ntime", ctypes.c_int32), ("count", ctypes.c_int32)]


class Cfg(ctypes.Structure):
    _fields_ = [("nblocks", ctypes.c_int32), ("levels", ctypes.c_int32),
                ("smoother", ctypes.c_int32), ("nu_pre", ctypes.c_int32),
                ("nu_post", ctypes.c_int32), ("maxcy", ctypes.c_int32),
                ("eps", ctypes.c_double), ("min_cycles", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("variable_coef", ctypes.c_int32),
                ("device", ctypes.c_int32), ("coarse_op", ctypes.c_int32),
                ("krylov", ctypes.c_int32), ("coarse_solve", ctypes.c_int32),
                ("differencing", ctypes.c_int32), ("pl_ex", ctypes.c_int32)]


_lib = None


class LibraryMissing(RuntimeError):
    pass


def lib():
    """Load libmf3d_mg.so once; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(f"{LIB_PATH} not built: run python -m multiflow3d_amd.build")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    dpp = ctypes.POINTER(ctypes.c_void_p)
    dp = ctypes.POINTER(ctypes.c_double)
    L.mf3d_mg_create.restype = ctypes.c_int
    L.mf3d_mg_create.argtypes = [ctypes.POINTER(Cfg), ctypes.POINTER(BlockDesc), ctypes.POINTER(vp)]
    L.mf3d_mg_set_density.restype = ctypes.c_int
    L.mf3d_mg_set_density.argtypes = [vp, dpp]
    L.mf3d_mg_vcycle.restype = ctypes.c_int
    L.mf3d_mg_vcycle.argtypes = [vp, dpp, dpp]
    L.mf3d_mg_solve.restype = ctypes.c_int
    L.mf3d_mg_solve.argtypes = [vp, dpp, dpp, dp, ctypes.POINTER(ctypes.c_int32), dp]
    L.mf3d_mg_solve_device.restype = ctypes.c_int
    L.mf3d_mg_solve_device.argtypes = [vp, vp, vp, vp, dp, ctypes.POINTER(ctypes.c_int32), dp]
    L.mf3d_mg_cells.restype = ctypes.c_int64
    L.mf3d_mg_cells.argtypes = [vp]
    L.mf3d_mg_device_bytes.restype = ctypes.c_int64
    L.mf3d_mg_device_bytes.argtypes = [vp]
    L.mf3d_mg_estimate_bytes.restype = ctypes.c_int64
    L.mf3d_mg_estimate_bytes.argtypes = [ctypes.POINTER(Cfg), ctypes.POINTER(BlockDesc)]
    L.mf3d_mg_last_error.restype = ctypes.c_char_p
    L.mf3d_mg_last_error.argtypes = [vp]
    L.mf3d_mg_destroy.restype = None
    L.mf3d_mg_destroy.argtypes = [vp]
    L.mf3d_mg_abi_version.restype = ctypes.c_int
    L.mf3d_mg_abi_version.argtypes = []
    L.mf3d_mg_time_kernel.restype = ctypes.c_int
    L.mf3d_mg_time_kernel.argtypes = [vp, ctypes.c_int]
    L.mf3d_mg_kernel_time.restype = ctypes.c_int
    L.mf3d_mg_kernel_time.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double)]
    L.mf3d_mg_graph_replay.restype = ctypes.c_int
    L.mf3d_mg_graph_replay.argtypes = [vp]
    L.mf3d_mg_set_eps.restype = ctypes.c_int
    L.mf3d_mg_set_eps.argtypes = [vp, ctypes.c_double]
    d = ctypes.c_double
    L.mf3d_mg_newsolv.restype = ctypes.c_int
    L.mf3d_mg_newsolv.argtypes = [vp, dpp, dpp, dpp, dpp, dpp, dpp, dpp, d, d, d, dpp, dp,
                                  ctypes.POINTER(ctypes.c_int32), dp]
    L.mf3d_mg_newsolv_device.restype = ctypes.c_int
    L.mf3d_mg_newsolv_device.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, d, d, d, vp, vp, dp,
                                         ctypes.POINTER(ctypes.c_int32), dp]
    L.mf3d_mg_ghosted_cells.restype = ctypes.c_int64
    L.mf3d_mg_ghosted_cells.argtypes = [vp]
    L.mf3d_mg_shell_cells.restype = ctypes.c_int64
    L.mf3d_mg_shell_cells.argtypes = [vp]
    L.mf3d_mg_ghost_table.restype = ctypes.c_int
    L.mf3d_mg_ghost_table.argtypes = [ctypes.POINTER(Cfg), ctypes.POINTER(BlockDesc), ctypes.c_int32,
                                      ctypes.POINTER(ctypes.c_int32), ctypes.c_int64,
                                      ctypes.POINTER(ctypes.c_int64), ctypes.c_char_p, ctypes.c_size_t]
    L.mf3d_mg_set_density_device.restype = ctypes.c_int
    L.mf3d_mg_set_density_device.argtypes = [vp, vp, vp]
    L.mf3d_mg_heaviside.restype = ctypes.c_int
    L.mf3d_mg_heaviside.argtypes = [vp, dpp, d, d, d, d, dpp, dpp]
    L.mf3d_tecbin_header.restype = ctypes.c_int
    L.mf3d_tecbin_header.argtypes = [ctypes.c_char_p, ctypes.POINTER(TecbinHdr)]
    L.mf3d_tecbin_read.restype = ctypes.c_int
    L.mf3d_tecbin_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(TecbinHdr), dpp, vp, vp, vp]
    L.mf3d_final_ctime_read.restype = ctypes.c_int
    L.mf3d_final_ctime_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(FinalCtime)]
    L.mf3d_io_last_error.restype = ctypes.c_char_p
    L.mf3d_io_last_error.argtypes = []
    L.mf3d_mg_solve_device_noalloc.restype = ctypes.c_int
    L.mf3d_mg_solve_device_noalloc.argtypes = [ctypes.POINTER(Cfg), ctypes.POINTER(BlockDesc), vp, vp, vp, dp,
                                               ctypes.POINTER(ctypes.c_int32), dp, ctypes.c_char_p,
                                               ctypes.c_size_t]
    L.mf3d_mg_heaviside_device.restype = ctypes.c_int
    L.mf3d_mg_heaviside_device.argtypes = [vp, vp, d, d, d, d, vp, vp, vp]
    _lib = L
    return L
