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
DEPS = SRC + [os.path.join(HERE, "csrc", "mg_kernels.hpp"), os.path.join(ROOT, "include", "mf3d_mg.h")]

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


def build_oracle(with_ref: bool = True) -> None:
    """Build the CPU oracle and, where the reference sources are present, the
    compiled reference driver (test infrastructure, never shipped)."""
    odir = os.path.join(ROOT, "oracle")
    subprocess.run(["make", "-s", "-C", odir], check=True)
    if with_ref and os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", odir, "ref"], check=True)


if __name__ == "__main__":
    print(build(verbose=True))
