"""Host side of the serving loop: request stream and level choice (no GPU)."""
import numpy as np
import pytest

from multiflow3d_amd import serving as S


def test_levels_for_sizes():
    # coarsest 4^3 for powers of two, 3^3 for 3 * 2^k
    assert [S.levels_for(n) for n in S.SIZES] == [4, 5, 6, 6, 7, 7]
    for n in S.SIZES:
        L = S.levels_for(n)
        assert n % (1 << (L - 1)) == 0 and (n >> (L - 1)) in (3, 4)


def test_poisson_requests_deterministic_and_rate():
    a = S.poisson_requests(200.0, 4000, 11)
    b = S.poisson_requests(200.0, 4000, 11)
    assert a == b
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert abs(gaps.mean() * 200.0 - 1.0) < 0.05          # mean inter-arrival 1/rate
    assert abs(np.std(gaps) / np.mean(gaps) - 1.0) < 0.06  # exponential: cv = 1
    sizes = np.array([r.n for r in a])
    for n in S.SIZES:                                     # sizes drawn uniformly
        assert abs(np.mean(sizes == n) - 1 / 6) < 0.03
    assert abs(np.mean([r.two_phase for r in a]) - 0.5) < 0.03
    assert set(r.tol for r in a) == set(S.TOLERANCES)


def test_request_larger_than_budget_is_refused():
    """A handle that cannot fit even an empty pool is refused with an error
    (waiting for memory to free up would never end), before any device work."""
    pool = S.SolverPool(S.Inputs(), budget_bytes=1 << 20, precision="mixed", maxcy=20)
    need = pool._bytes(64, False)
    assert need > (1 << 20)
    with pytest.raises(S.AdmissionError, match="more than the pool's whole budget"):
        pool.acquire(64, False)
    assert pool.used == 0 and pool.created == 0
