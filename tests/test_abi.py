"""CPU-side checks of the drop-in boundary and the host logic (no GPU calls):
the C-ABI library loads and exports every symbol include/mf3d_mg.h declares,
argument validation, topology tables and the seeded workload recipes."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from multiflow3d_amd import MGError, PressureMultigrid, estimate_bytes
from multiflow3d_amd import _lib
from multiflow3d_amd import topology as T
from multiflow3d_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mf3d_mg.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(mf3d_\w+)\s*\(", text))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _declared()
    assert declared == set(_lib.EXPORTS)
    for name in declared:
        assert hasattr(L, name), name
    nm = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                        text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(mf3d_\w+)", nm))
    assert declared <= exported
    assert L.mf3d_mg_abi_version() == 7


def test_product_library_does_not_contain_the_oracle():
    nm = subprocess.run(["nm", "-D", _lib.LIB_PATH], capture_output=True, text=True, check=True).stdout
    assert "orc_" not in nm
    ldd = subprocess.run(["ldd", _lib.LIB_PATH], capture_output=True, text=True).stdout
    assert "orc_mg" not in ldd


def test_struct_layout_matches_header(tmp_path):
    """ctypes mirrors of the header structs agree with the C compiler's layout."""
    fields = {"mf3d_block_desc": [f for f, _ in _lib.BlockDesc._fields_],
              "mf3d_mg_cfg": [f for f, _ in _lib.Cfg._fields_]}
    src = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for st, fl in fields.items():
        src.append(f'printf("%zu\\n", sizeof({st}));')
        src += [f'printf("%zu\\n", offsetof({st}, {f}));' for f in fl]
    src.append("return 0;}")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", str(c), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    want = []
    for cls in (_lib.BlockDesc, _lib.Cfg):
        want.append(ctypes.sizeof(cls))
        want += [getattr(cls, f).offset for f, _ in cls._fields_]
    assert got == want


def test_estimate_bytes_and_validation():
    g = T.single_block(256, [1 / 256] * 3, T.CHANNEL_BC)
    b64 = estimate_bytes(g, 7)
    b32 = estimate_bytes(g, 7, precision="mixed")
    assert 16.7e6 * 8 * 5 < b64 < 16.7e6 * 8 * 9
    assert b32 < b64
    with pytest.raises(MGError):
        estimate_bytes(T.single_block(24, [1 / 24] * 3, T.CHANNEL_BC), 5)   # 24 % 16 != 0


def test_create_rejects_bad_input_before_touching_a_device():
    bad = T.block_grid([8, 8], [8], [8], T.CHANNEL_BC, [0.1] * 3)
    bad.blocks[1].n = [8, 16, 8]          # no longer conforms across the x face
    with pytest.raises(MGError, match="conform"):
        PressureMultigrid(bad, 2)
    g = T.single_block(12, [0.1] * 3, T.CHANNEL_BC)
    with pytest.raises(MGError, match="not correct for mgsolver"):
        PressureMultigrid(g, 4)
    with pytest.raises(MGError):
        PressureMultigrid(T.single_block(8, [0.1] * 3, T.CHANNEL_BC), 2, smoother=2)
    asym = T.block_grid([8, 8], [8], [8], T.CHANNEL_BC, [0.1] * 3)
    asym.blocks[1].nbr[0] = -1
    with pytest.raises(MGError, match="symmetric"):
        PressureMultigrid(asym, 2)


def test_block_grid_tables():
    g = T.block_grid([8, 12, 4], [16, 8], [4], T.CHANNEL_BC, [0.1] * 3)
    assert len(g.blocks) == 6 and g.global_n == (24, 24, 4)
    for i, b in enumerate(g.blocks):
        for f in range(6):
            if b.nbr[f] >= 0:
                assert g.blocks[b.nbr[f]].nbr[f ^ 1] == i
            if b.per[f] >= 0:
                assert g.blocks[b.per[f]].per[f ^ 1] == i
    # x periodic across 3 columns; z periodic with one layer wraps onto itself
    assert g.blocks[0].per[0] == 2 and g.blocks[2].per[1] == 0
    assert all(b.per[4] == -1 and b.per[5] == -1 for b in g.blocks)
    assert T.max_levels(g) == 3


def test_splitmix64_known_answer():
    # published first outputs of SplitMix64 seeded with 0
    x = W.splitmix64(0, 0, 2)
    assert int(x[0]) == 0xE220A8397B1DCDAF and int(x[1]) == 0x6E789E6AA1B965F4
    u = W.uniform_pm1(W.BASE_SEED, 0, 100000)
    assert -1.0 <= u.min() and u.max() < 1.0 and abs(u.mean()) < 0.01


def test_rhs_recipe():
    g = T.single_block([8, 6, 4], [1 / 8, 1 / 6, 1 / 4], T.CHANNEL_BC)
    f = W.global_rhs(g, 1)
    assert f.shape == (4, 6, 8) and abs(f.mean()) < 1e-14
    # the divergence of a constant velocity field vanishes, of u = x is 1
    u = np.ones((4, 6, 8)); v = np.ones((4, 7, 8)); w = np.ones((4, 6, 8))
    assert np.allclose(W.divergence(u, v, w, g.h, [True, False, True]), 0.0)
    v = np.broadcast_to(np.arange(7.0)[None, :, None] / 6, (4, 7, 8))
    assert np.allclose(W.divergence(u * 0, v, w * 0, g.h, [True, False, True]), 1.0)


def test_sphere_density_recipe():
    g = T.single_block(32, [1 / 32] * 3, T.CHANNEL_BC)
    rho = W.sphere_density(g)
    assert rho.min() == 1.0 and rho.max() == 1000.0
    assert rho[16, 16, 16] == 1.0 and rho[0, 0, 0] == 1000.0
    assert W.heaviside(np.array([0.0]), 0.1)[0] == 0.5


def test_noalloc_solve_rejects_before_device_work():
    """The allocation-free solve refuses a hierarchy larger than its fixed
    scratch array, and a variable-coefficient problem, with a message and
    before any device call (the pointers are never dereferenced)."""
    from multiflow3d_amd.solver import _descs, make_cfg
    L = _lib.lib()
    cases = ((T.single_block([128] * 3, [1 / 128] * 3, T.CHANNEL_BC), {}, _lib.MF3D_ENOMEM, "scratch"),
             (T.single_block([16] * 3, [1 / 16] * 3, T.CHANNEL_BC), {"variable_coef": True},
              _lib.MF3D_ENOTSUP, "constant-coefficient"))
    for grid, kw, code, text in cases:
        cfg = make_cfg(1, 4, **kw)
        err = ctypes.create_string_buffer(256)
        rc = L.mf3d_mg_solve_device_noalloc(ctypes.byref(cfg), _descs(grid.blocks, 1), ctypes.c_void_p(8),
                                            ctypes.c_void_p(8), None, None, None, None, err, 256)
        assert rc == code, (rc, err.value)
        assert text in err.value.decode()
    assert _lib.MF3D_NOALLOC_SCRATCH_BYTES == 32 << 20
