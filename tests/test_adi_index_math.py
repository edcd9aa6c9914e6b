"""The ADI kernel's division-free split of a diagonal's cell index
(AdiDiag::split in multiflow3d_amd/csrc/mg_linesmooth.hpp): e -> (line,
cell) = divmod(e, nl) from a single-precision estimate e * (1/nl),
truncated, then corrected by at most one step either way.  The LDS kernel
runs only where a diagonal holds at most 1024 cells (so e < 1024 and
nl <= 1024).  Checked exhaustively over that range with IEEE single
precision (round to nearest, as the device computes it), and shown to need
its correction steps: the raw estimate alone is wrong for some inputs."""
import numpy as np


def _split(e, nl):
    rnl = np.float32(1.0) / nl.astype(np.float32)
    q = (e.astype(np.float32) * rnl).astype(np.int64)  # truncation toward zero, as (int) does
    q = q - (q * nl > e)
    q = q + ((q + 1) * nl <= e)
    return q, e - q * nl


def test_split_is_divmod_over_the_lds_range():
    nl = np.arange(1, 1025, dtype=np.int64)[:, None]
    e = np.arange(0, 1024, dtype=np.int64)[None, :]
    nl, e = np.broadcast_arrays(nl, e)
    q, ln = _split(e, nl)
    assert np.array_equal(q, e // nl) and np.array_equal(ln, e % nl)


def test_raw_estimate_needs_the_correction():
    nl = np.arange(1, 1025, dtype=np.int64)[:, None]
    e = np.arange(0, 1024, dtype=np.int64)[None, :]
    nl, e = np.broadcast_arrays(nl, e)
    raw = (e.astype(np.float32) * (np.float32(1.0) / nl.astype(np.float32))).astype(np.int64)
    assert np.any(raw != e // nl)
