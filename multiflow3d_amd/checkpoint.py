"""The reference's checkpoint files, read through the library's C ABI
(mf3d_tecbin_read, mf3d_final_ctime_read), to start the solver from the
state of a real MultiFlow3D run.

- tecbin####.bin, one per block (written by tecbin, post.for:496-548; read
  at restart by initial.for:708-748): x, y, z and 16 fields with pl ghost
  layers.
- final_ctime.dat (flosol.for:296-299, initial.for:692-694): ntime, ctime,
  forcn, qstpn, count, ntav1_count, ntav2_count.

Arrays come back as numpy arrays shaped (ttc_k, ttc_j, ttc_i) -- the
reference's column-major dom(ib) arrays seen in C order, i fastest.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, List

import numpy as np

from . import _lib

FIELDS = ("p", "pm", "ppm", "vis", "u", "um", "uum", "v", "vm", "vvm", "w", "wm", "wwm", "uvm", "uwm", "vwm")


class CheckpointError(OSError):
    pass


def _raise(rc: int, what: str):
    msg = _lib.lib().mf3d_io_last_error().decode(errors="replace")
    if rc == _lib.MF3D_ENOENT:
        raise FileNotFoundError(msg)
    raise CheckpointError(f"{what} failed ({rc}): {msg}")


def tecbin_name(dom_id: int) -> str:
    """tecbin<dom_id, zero-padded to 4 digits>.bin (post.for:507-510)."""
    return f"tecbin{dom_id:04d}.bin"


def read_header(path: str) -> Dict:
    h = _lib.TecbinHdr()
    rc = _lib.lib().mf3d_tecbin_header(os.fsencode(path), ctypes.byref(h))
    if rc != 0:
        _raise(rc, "mf3d_tecbin_header")
    return {"tt": tuple(h.tt), "pl": int(h.pl), "ind": tuple(h.ind)}


def read_tecbin(path: str, fields=FIELDS) -> Dict:
    """One block's checkpoint: header keys (tt, pl, ind), x, y, z and the
    requested fields."""
    hdr = read_header(path)
    ti, tj, tk = hdr["tt"]
    out = dict(hdr)
    arrs = {f: np.empty((tk, tj, ti)) for f in fields}
    ptrs = (ctypes.c_void_p * 16)(*[arrs[f].ctypes.data if f in arrs else None for f in FIELDS])
    x, y, z = np.empty(ti), np.empty(tj), np.empty(tk)
    h = _lib.TecbinHdr()
    rc = _lib.lib().mf3d_tecbin_read(os.fsencode(path), ctypes.byref(h), ptrs, x.ctypes.data, y.ctypes.data,
                                     z.ctypes.data)
    if rc != 0:
        _raise(rc, "mf3d_tecbin_read")
    out.update(arrs)
    out.update(x=x, y=y, z=z)
    return out


def read_final_ctime(path: str) -> Dict:
    r = _lib.FinalCtime()
    rc = _lib.lib().mf3d_final_ctime_read(os.fsencode(path), ctypes.byref(r))
    if rc != 0:
        _raise(rc, "mf3d_final_ctime_read")
    return {k: getattr(r, k) for k in ("ntime", "ctime", "forcn", "qstpn", "count", "ntav1_count", "ntav2_count")}


def one_ghost_box(a: np.ndarray, pl: int) -> np.ndarray:
    """The interior and first ghost layer of a pl-ghost array: the layout of
    the solver's ghosted block arrays."""
    if pl < 1:
        raise ValueError("the array has no ghost layer")
    s = slice(pl - 1, a.shape[0] - pl + 1)
    t = slice(pl - 1, a.shape[1] - pl + 1)
    u = slice(pl - 1, a.shape[2] - pl + 1)
    return np.ascontiguousarray(a[s, t, u])


def load_checkpoint(directory: str, grid, pl: int = None, fields=("p", "u", "v", "w")) -> List[Dict]:
    """Every block's checkpoint of a run whose blocks are grid.blocks (block
    b = dom_id b): the requested fields as one-ghost-layer boxes shaped like
    the solver's block arrays, plus x, y, z and the header.  The extents
    must match the grid; pl, when given, must match the files (the
    reference stops otherwise, initial.for:717-721)."""
    out = []
    for b, blk in enumerate(grid.blocks):
        d = read_tecbin(os.path.join(directory, tecbin_name(b)), fields)
        if pl is not None and d["pl"] != pl:
            raise CheckpointError(f"&*&* different number of overlapping layers!! {pl} {d['pl']}")
        want = tuple(n + 2 * d["pl"] for n in blk.n)
        if d["tt"] != want:
            raise CheckpointError(f"block {b}: checkpoint extents {d['tt']} do not match the grid's {want}")
        for f in fields:
            d[f] = one_ghost_box(d[f], d["pl"])
        out.append(d)
    return out
