"""The checkpoint reader (mf3d_tecbin_read, mf3d_final_ctime_read: host code
of the library, no GPU) against files the reference's own tecbin
(post.for:496-548) wrote, with the values the reference handed it
(tests/golden/make_golden_ckpt.py); malformed and missing files."""
import json
import os
import shutil

import numpy as np
import pytest

from multiflow3d_amd import checkpoint as C
from multiflow3d_amd import topology as T
from oracle.oracle import CHECKPOINT_FIELDS

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MAN = json.load(open(os.path.join(GOLDEN, "manifest_ckpt.json")))


def _split(grid, flat, ghost=1):
    out, o = [], 0
    for b in grid.blocks:
        nx, ny, nz = (x + 2 * ghost for x in b.n)
        m = nx * ny * nz
        out.append(flat[o:o + m].reshape(nz, ny, nx).copy())
        o += m
    return out


def load_ckpt(name):
    """(grid, manifest entry, directory, expected arrays split per block)."""
    m = MAN[name]
    g = T.block_grid(m["ex"], m["ey"], m["ez"], m["bc"], m["h"])
    d = os.path.join(GOLDEN, f"ckpt_{name}")
    z = np.load(os.path.join(d, "expected.npz"))
    e = {k: _split(g, z[k]) for k in CHECKPOINT_FIELDS + ("ustar", "vstar", "wstar", "r_p", "r_u", "r_v", "r_w")}
    e["r_su"] = _split(g, z["r_su"], ghost=0)
    e["iters"] = int(z["iters"][0])
    e["rmax"] = z["rmax"]
    o = {"x": 0, "y": 0, "z": 0}
    for key, ax in (("x", 0), ("y", 1), ("z", 2)):
        e[key] = []
        for blk in g.blocks:
            n = blk.n[ax] + 2
            e[key].append(z[key][o[key]:o[key] + n])
            o[key] += n
    return g, m, d, e


@pytest.mark.parametrize("name", sorted(MAN))
def test_reader_returns_what_the_reference_checkpointed(name):
    g, m, d, e = load_ckpt(name)
    for b, blk in enumerate(g.blocks):
        r = C.read_tecbin(os.path.join(d, C.tecbin_name(b)))
        assert r["tt"] == tuple(n + 2 for n in blk.n) and r["pl"] == 1
        # inind/jnind/knind: -1 at a physical non-periodic east/north/top face
        want = tuple(-1 if (blk.nbr[2 * a + 1] < 0 and blk.bc[2 * a + 1] != 5) else 0 for a in range(3))
        assert r["ind"] == want
        for key in ("x", "y", "z"):
            assert np.array_equal(r[key], e[key][b]), key
        for f in CHECKPOINT_FIELDS:
            assert np.array_equal(r[f], e[f][b]), f


@pytest.mark.parametrize("name", sorted(MAN))
def test_final_ctime(name):
    _, m, d, _ = load_ckpt(name)
    r = C.read_final_ctime(os.path.join(d, "final_ctime.dat"))
    assert r["ntime"] == m["cycles"] and r["count"] == 7
    # the values the harness wrote, at the format's six decimals
    assert r["ctime"] == round(0.0123456789 * m["cycles"], 6)
    assert (r["forcn"], r["qstpn"], r["ntav1_count"], r["ntav2_count"]) == (-1.5, 2.25, 3.0, 12345.678901)


def test_load_checkpoint_gives_solver_block_arrays():
    g, m, d, e = load_ckpt("channel21")
    blocks = C.load_checkpoint(d, g, pl=1)
    for b, blk in enumerate(g.blocks):
        for f in ("p", "u", "v", "w"):
            assert blocks[b][f].shape == (blk.n[2] + 2, blk.n[1] + 2, blk.n[0] + 2)
            assert np.array_equal(blocks[b][f], e[f][b])
    with pytest.raises(C.CheckpointError, match="different number of overlapping layers"):
        C.load_checkpoint(d, g, pl=2)
    g2 = T.block_grid([8, 8], [8], [4], T.CHANNEL_BC, [1 / 16] * 3)
    with pytest.raises(C.CheckpointError, match="extents"):
        C.load_checkpoint(d, g2)


def test_one_ghost_box_of_a_two_layer_array():
    a = np.arange(6 * 7 * 8, dtype=np.float64).reshape(6, 7, 8)  # n = (4, 3, 2), pl = 2
    box = C.one_ghost_box(a, 2)
    assert box.shape == (4, 5, 6)
    assert np.array_equal(box, a[1:5, 1:6, 1:7])
    assert np.array_equal(C.one_ghost_box(box, 1), box)


def test_missing_and_malformed_files(tmp_path):
    _, _, d, _ = load_ckpt("walls11")
    with pytest.raises(FileNotFoundError):
        C.read_tecbin(str(tmp_path / "tecbin9999.bin"))
    with pytest.raises(FileNotFoundError):
        C.read_final_ctime(str(tmp_path / "none.dat"))
    src = open(os.path.join(d, "tecbin0000.bin"), "rb").read()
    trunc = tmp_path / "trunc.bin"
    trunc.write_bytes(src[:-100])
    with pytest.raises(C.CheckpointError, match="cell records|truncated"):
        C.read_tecbin(str(trunc))
    bad = bytearray(src)
    bad[0] = 8  # the first record claims 8 bytes: not a tecbin header
    badf = tmp_path / "bad.bin"
    badf.write_bytes(bytes(bad))
    with pytest.raises(C.CheckpointError, match="expected 12"):
        C.read_header(str(badf))
    bad = bytearray(src)
    bad[len(src) // 2] ^= 0xFF  # somewhere in the cell records: payload or a marker
    off = 12 + 8 + 4 + 8 + 12 + 8 + 160 * 3  # the start marker of cell record 3
    bad[off] = 77
    badf.write_bytes(bytes(bad))
    with pytest.raises(C.CheckpointError, match="cell record 3"):
        C.read_tecbin(str(badf))
    short = tmp_path / "final_ctime.dat"
    short.write_text("       3       0.100000      -1.500000       2.250000\n")
    with pytest.raises(C.CheckpointError, match="two lines"):
        C.read_final_ctime(str(short))
    short.write_text("       3       0.100000      -1.500000       2.250000\n       7    ***********\n")
    with pytest.raises(C.CheckpointError, match="not a number"):
        C.read_final_ctime(str(short))


def test_read_selected_fields_only(tmp_path):
    _, _, d, e = load_ckpt("walls11")
    r = C.read_tecbin(os.path.join(d, "tecbin0000.bin"), fields=("p", "vwm"))
    assert set(C.FIELDS) & set(r) == {"p", "vwm"}
    assert np.array_equal(r["vwm"], e["vwm"][0])
    shutil.copy(os.path.join(d, "tecbin0000.bin"), tmp_path / "tecbin0000.bin")
    assert C.read_header(str(tmp_path / "tecbin0000.bin"))["tt"] == (10, 8, 6)
