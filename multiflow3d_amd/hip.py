"""Device memory and streams through the HIP runtime the library itself links.

A process that uses this package and PyTorch holds two HIP runtimes (PyTorch
bundles its own); whichever initialises second may find no device.  Host
code of this package that needs device buffers or streams of its own (the
serving loop and its tests) therefore calls the runtime that
libmf3d_mg.so is linked against, loaded after it so the dynamic loader hands
back that same copy.  Streams are created at the default priority.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

_hip = None
H2D, D2H, D2D = 1, 2, 3


class HipError(RuntimeError):
    pass


def rt():
    global _hip
    if _hip is None:
        _lib.lib()  # the library first: its libamdhip64 is the one returned below
        h = ctypes.CDLL("libamdhip64.so.7")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        h.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        h.hipFree.argtypes = [vp]
        h.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        h.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(vp), ctypes.c_uint]
        h.hipStreamDestroy.argtypes = [vp]
        h.hipStreamSynchronize.argtypes = [vp]
        h.hipGetErrorString.restype = ctypes.c_char_p
        h.hipGetErrorString.argtypes = [ctypes.c_int]
        h.hipSetDevice.argtypes = [ctypes.c_int]
        _hip = h
    return _hip


def _check(rc, what):
    if rc != 0:
        raise HipError(f"{what}: {rt().hipGetErrorString(rc).decode()}")


class DeviceArray:
    """A 1-D FP64 device buffer."""

    def __init__(self, n: int):
        self.n = int(n)
        p = ctypes.c_void_p()
        _check(rt().hipMalloc(ctypes.byref(p), max(self.n, 1) * 8), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        d = cls(a.size)
        _check(rt().hipMemcpy(d.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        return d

    def copy_from_host(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        if a.size != self.n:
            raise ValueError(f"size {a.size} != {self.n}")
        _check(rt().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")

    def to_host(self) -> np.ndarray:
        out = np.empty(self.n)
        _check(rt().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            rt().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    """A non-blocking stream (no implicit ordering with the legacy stream, so
    a graph captured on it never conflicts with other threads' copies)."""

    def __init__(self):
        s = ctypes.c_void_p()
        _check(rt().hipStreamCreateWithFlags(ctypes.byref(s), 1), "hipStreamCreateWithFlags")
        self.handle = s.value

    def synchronize(self):
        _check(rt().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def destroy(self):
        if self.handle:
            rt().hipStreamDestroy(self.handle)
            self.handle = None
