"""Build the HIP library in-tree for gfx950.

    python -m multiflow3d_amd.build          # libmf3d_mg.so next to this file

hipcc flags: -ffp-contract=off keeps every FP64 multiply and add separately
rounded, as in the reference (no FMA contraction); packed FP32 instructions
are disabled (-packed-fp32-ops) for all device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libmf3d_mg.so")
SRC = [os.path.join(HERE, "csrc", "mf3d_mg.hip")]
DEPS = SRC + [os.path.join(HERE, "csrc", h) for h in ("mg_common.hpp", "mg_outer.hpp", "mg_setup.hpp", "mg_sweep.hpp", "mg_tail.hpp", "mg_levelset.hpp", "mg_linesmooth.hpp", "mf3d_tecbin.hpp", "mg_transfer.hpp", "mg_velocity.hpp", "mg_coarse.hpp", "mg_ghostmap.hpp", "mg_refine.hpp")] + [os.path.join(ROOT, "include", "mf3d_mg.h")]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MF3D_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    f"-I{os.path.join(ROOT, 'include')}",
]


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, *SRC, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libmf3d_mg.so")
    os.replace(tmp, LIB)
    return LIB


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FORTRAN_DIR = os.path.join(HERE, "fortran")
# executables in fortran/bin (they travel with the tree); objects and
# module files in fortran/_build
FORTRAN_HOST = os.path.join(FORTRAN_DIR, "bin", "mf3d_host")
FORTRAN_HOST_VEL = os.path.join(FORTRAN_DIR, "bin", "mf3d_host_vel")


def build_fortran(verbose: bool = False) -> str:
    """The ISO_C_BINDING module and the small Fortran hosts that drive the
    library through it (fortran/bin/mf3d_host for the Poisson form,
    mf3d_host_vel for the velocity form; linked to libmf3d_mg.so)."""
    out = os.path.join(FORTRAN_DIR, "_build")
    os.makedirs(out, exist_ok=True)
    os.makedirs(os.path.dirname(FORTRAN_HOST), exist_ok=True)
    steps = [
        [FLANG, "-c", "-O2", "-J", out, os.path.join(FORTRAN_DIR, "mf3d_mg_f.f90"),
         "-o", os.path.join(out, "mf3d_mg_f.o")],
        [FLANG, "-c", "-O2", "-J", out, "-I", out, os.path.join(FORTRAN_DIR, "mf3d_host.f90"),
         "-o", os.path.join(out, "mf3d_host.o")],
        [FLANG, "-o", FORTRAN_HOST, os.path.join(out, "mf3d_host.o"), os.path.join(out, "mf3d_mg_f.o"),
         f"-L{HERE}", "-lmf3d_mg", "-Wl,-rpath,$ORIGIN/../.."],
        [FLANG, "-c", "-O2", "-J", out, "-I", out, os.path.join(FORTRAN_DIR, "mf3d_host_vel.f90"),
         "-o", os.path.join(out, "mf3d_host_vel.o")],
        [FLANG, "-o", FORTRAN_HOST_VEL, os.path.join(out, "mf3d_host_vel.o"), os.path.join(out, "mf3d_mg_f.o"),
         f"-L{HERE}", "-lmf3d_mg", "-Wl,-rpath,$ORIGIN/../.."],
    ]
    for cmd in steps:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    return FORTRAN_HOST


def build_oracle(with_ref: bool = True) -> None:
    """Build the CPU oracle and, where the reference sources are present, the
    compiled reference driver (test infrastructure, never shipped)."""
    odir = os.path.join(ROOT, "oracle")
    subprocess.run(["make", "-s", "-C", odir], check=True)
    if with_ref and os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", odir, "ref"], check=True)


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran(verbose=True))
