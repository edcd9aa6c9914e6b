"""The OpenMP build of the CPU oracle (liborc_mg_omp.so, the multi-core CPU
baseline) gives bitwise the serial build's results: its parallel plane
loops write disjoint cells and its reductions stay serial.  Checked on a
multi-block grid with periodic partners, constant and variable
coefficients, with the threads the machine offers."""
import os
import subprocess
import sys

import numpy as np
import pytest

from multiflow3d_amd import topology as T
from multiflow3d_amd import workloads as W
from oracle.oracle import LIB_OMP_PATH, OracleMG

HERE = os.path.dirname(os.path.abspath(__file__))


def _solve(g, L, f, lsm, omp, dens=None):
    o = OracleMG(g.blocks, L, lsm=lsm, omp=omp)
    if lsm:
        o.set_density(dens)
    o.coeff()
    p, rho, rmax, k = o.solve(f, 0.0, 3)
    pp = o.vcycle(f)
    o.close()
    return p, rho, rmax, k, pp


@pytest.mark.skipif(not os.path.exists(LIB_OMP_PATH), reason="make -C oracle builds liborc_mg_omp.so")
@pytest.mark.parametrize("lsm", [False, True])
def test_openmp_oracle_is_bitwise_the_serial_oracle(lsm):
    g = T.block_grid([16, 16], [8, 16], [16, 16], T.CHANNEL_BC, [1 / 32, 1 / 24, 1 / 32])
    f = W.split_blocks(g, W.global_rhs(g, W.BASE_SEED + 31, fv_volume=lsm))
    rng = np.random.default_rng(31)
    dens = [rng.uniform(1.0, 100.0, x.shape) for x in f] if lsm else None
    a = _solve(g, 3, f, lsm, False, dens)
    b = _solve(g, 3, f, lsm, True, dens)
    assert a[3] == b[3]
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    for x, y in zip(a[0], b[0]):
        assert np.array_equal(x, y)
    for x, y in zip(a[4], b[4]):
        assert np.array_equal(x, y)


@pytest.mark.skipif(not os.path.exists(LIB_OMP_PATH), reason="make -C oracle builds liborc_mg_omp.so")
@pytest.mark.parametrize("omp", [False, True])
def test_only_the_openmp_build_brings_the_openmp_runtime(omp):
    """Loading the OpenMP build maps libgomp into the process (its loops run
    on a thread team); the serial build does not."""
    code = ("import sys; sys.path.insert(0, %r); from oracle.oracle import lib; lib(%r); "
            "print(any('libgomp' in l for l in open('/proc/self/maps')))" % (os.path.dirname(HERE), omp))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == str(omp)
