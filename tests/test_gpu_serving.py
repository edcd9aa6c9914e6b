"""Serving loop on the GPU: continuous batching over 4 streams.

- Every request's pressure is bit for bit the same in a mix of overlapping
  solves as when solved alone on a fresh handle (deterministic kernels; the
  stream, slot and neighbours on the device do not enter the arithmetic):
  a random small mix, and the full size mix (every size of the serving
  workload, 32^3 .. 256^3, with both coefficient kinds and all three
  tolerances).
- HBM-budgeted admission: with room for one 64^3 hierarchy, requests still
  all complete and the pool never holds more than the budget.
"""
import numpy as np
import pytest

from multiflow3d_amd import PressureMultigrid, hip
from multiflow3d_amd import serving as S

pytestmark = pytest.mark.gpu


def _alone(req, precision, inputs):
    g = S.Inputs.grid(req.n)
    with PressureMultigrid(g, S.levels_for(req.n), maxcy=60, eps=req.tol, precision=precision,
                           variable_coef=req.two_phase, coarse_op="face" if req.two_phase else "reference",
                           krylov=req.two_phase, coarse_solve="direct") as mg:
        if req.two_phase:
            mg.coeff(inputs.density(req.n))
        f = inputs.rhs(req.n, req.seed, req.two_phase)
        p = hip.DeviceArray(f.n)
        res = mg.solve_device(f.ptr, p.ptr, warn=False)
    return p.to_host(), res


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_mix_equals_alone(precision):
    reqs = S.poisson_requests(1e6, 14, 99, sizes=(32, 64, 96), two_phase_share=0.4)
    srv = S.Server(slots=4, precision=precision)
    try:
        out = srv.run(reqs, realtime=False)
        assert [r.rid for r in out] == [q.rid for q in reqs]
        for q, r in zip(reqs, out):
            p, res = _alone(q, precision, srv.inputs)
            assert r.status == 0 and res.status == 0 and r.rho <= q.tol
            assert r.iters == res.iters, (q, r.iters, res.iters)
            assert np.array_equal(r.p.to_host(), p), f"request {q.rid} differs between the mix and alone"
    finally:
        srv.close()


def test_full_size_mix_equals_alone():
    """Every serving size (32 .. 256 per axis) in both coefficient kinds,
    tolerances cycling 1e-6, 1e-8, 1e-10, all submitted at once to 4 slots
    (mixed precision): each converges to its tolerance and is bitwise its
    solve alone."""
    reqs = []
    for i, (n, tp) in enumerate([(n, tp) for n in S.SIZES for tp in (False, True)]):
        reqs.append(S.Request(rid=i, arrival=0.0, n=n, tol=S.TOLERANCES[i % 3], two_phase=tp,
                              seed=7000 + i))
    assert {q.n for q in reqs} == set(S.SIZES) and sum(q.two_phase for q in reqs) == len(reqs) // 2
    srv = S.Server(slots=4, precision="mixed")
    try:
        out = srv.run(reqs, realtime=False)
        assert [r.rid for r in out] == [q.rid for q in reqs]
        for q, r in zip(reqs, out):
            print(f"[serving full mix] n={q.n} two_phase={q.two_phase} tol={q.tol}", flush=True)
            p, res = _alone(q, "mixed", srv.inputs)
            assert r.status == 0 and res.status == 0 and r.rho <= q.tol
            assert r.iters == res.iters, (q, r.iters, res.iters)
            assert np.array_equal(r.p.to_host(), p), f"request {q.rid} differs between the mix and alone"
    finally:
        srv.close()


def test_budget_admission():
    reqs = S.poisson_requests(1e6, 10, 5, sizes=(32, 64), two_phase_share=0.0)
    need64 = S.SolverPool(None, 0, "mixed", 60)._bytes(64, False)
    srv = S.Server(slots=4, budget_bytes=need64 + (1 << 20), precision="mixed")
    try:
        out = srv.run(reqs, realtime=False)
        assert len(out) == len(reqs) and all(r.status == 0 for r in out)
        assert srv.pool.peak <= need64 + (1 << 20)
        summ = S.summarize(out, srv.elapsed)
        assert summ["requests"] == len(reqs) and summ["p99_ms"] >= summ["p50_ms"] > 0
    finally:
        srv.close()
