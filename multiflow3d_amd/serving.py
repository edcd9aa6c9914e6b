"""Serving loop for independent pressure-Poisson solves on one GPU.

Requests -- a cubic grid of n^3 cells with the channel boundary conditions,
a stop tolerance, a constant or two-phase coefficient and an input seed --
arrive over time.  The server

- keeps solver handles (device-resident level hierarchies, coefficients
  included) per (n, coefficient, precision) and reuses them, so a request
  finds its hierarchy already in HBM;
- admits a request when one of its slots is free and the handle it needs
  is idle or fits the HBM budget (idle handles of other shapes are evicted,
  least recently used first, to make room); requests that do not fit wait
  while later, smaller ones proceed;
- runs every admitted solve on its slot's own HIP stream from its own host
  thread, so solves of different sizes overlap on the device (continuous
  batching: a request arriving while others run is admitted as soon as a
  slot and memory free up).

Inputs are device-resident when a request arrives (the caller's fields live
in HBM); a request's latency runs from its arrival to the end of its solve.
Each solve is deterministic: alone or in a mix, the same request gives the
same pressure bit for bit (the stream, the slot and the neighbours on the
device do not enter the arithmetic).

Nothing of this exists in the reference, which solves one pressure field
per time step inside its MPI time loop (newsolv_mg.for:2-102).
"""
from __future__ import annotations

import collections
import queue
import threading
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import hip
from . import topology as T
from . import workloads as W
from .solver import PressureMultigrid, estimate_bytes

SIZES = (32, 64, 96, 128, 192, 256)
TOLERANCES = (1e-6, 1e-8, 1e-10)


def levels_for(n: int) -> int:
    """Levels of an n^3 request: coarsest 4^3 for powers of two (L =
    log2(n/4) + 1), 3^3 for 3 * 2^k (96 -> 6, 192 -> 7)."""
    L = 1
    while n % (1 << L) == 0 and (n >> L) >= 3:
        L += 1
    return L


@dataclass
class Request:
    rid: int
    arrival: float      # seconds after the start of the run
    n: int
    tol: float
    two_phase: bool
    seed: int


@dataclass
class Result:
    rid: int
    latency_ms: float
    queue_ms: float
    solve_ms: float
    iters: int
    status: int
    rho: float
    p: object = field(repr=False)   # hip.DeviceArray (interior-packed pressure)


def poisson_requests(rate: float, count: int, seed: int, sizes=SIZES, tolerances=TOLERANCES,
                     two_phase_share: float = 0.5, seeds_per_size: int = 2) -> List[Request]:
    """`count` requests with exponential inter-arrival times of mean 1/rate;
    size, tolerance, coefficient and input seed drawn per request from a
    SplitMix64 stream of `seed`."""
    u = W.uniform_pm1(seed, 0, 5 * count) * 0.5 + 0.5      # uniform [0, 1)
    out, t = [], 0.0
    for i in range(count):
        a, b, c, d, e = u[5 * i:5 * i + 5]
        t += -np.log(1.0 - a) / rate
        out.append(Request(i, t, sizes[int(b * len(sizes)) % len(sizes)],
                           tolerances[int(c * len(tolerances)) % len(tolerances)],
                           bool(d < two_phase_share),
                           W.BASE_SEED + 5 + int(e * seeds_per_size) % seeds_per_size))
    return out


class Inputs:
    """Device-resident right-hand sides and densities, one per (n, seed) and n."""

    def __init__(self):
        self._f: Dict[Tuple[int, int, bool], hip.DeviceArray] = {}
        self._dens: Dict[int, List[np.ndarray]] = {}

    @staticmethod
    def grid(n: int):
        return T.single_block(n, [1.0 / n] * 3, T.CHANNEL_BC)

    def rhs(self, n: int, seed: int, two_phase: bool) -> hip.DeviceArray:
        key = (n, seed, two_phase)
        if key not in self._f:
            g = self.grid(n)
            self._f[key] = hip.DeviceArray.from_host(W.global_rhs(g, seed, fv_volume=two_phase))
        return self._f[key]

    def density(self, n: int):
        if n not in self._dens:
            g = self.grid(n)
            self._dens[n] = W.split_blocks(g, W.sphere_density(g))
        return self._dens[n]

    def prepare(self, reqs: List[Request]):
        for r in reqs:
            self.rhs(r.n, r.seed, r.two_phase)
            if r.two_phase:
                self.density(r.n)


class AdmissionError(ValueError):
    """A request whose solver handle can never fit the HBM budget."""


class SolverPool:
    """Resident handles per (n, two_phase); HBM-budgeted, LRU eviction of idle ones."""

    def __init__(self, inputs: Inputs, budget_bytes: int, precision: str, maxcy: int):
        self.inputs = inputs
        self.budget = budget_bytes
        self.precision = precision
        self.maxcy = maxcy
        self.idle: Dict[Tuple[int, bool], List[PressureMultigrid]] = collections.defaultdict(list)
        self.used = 0
        self.peak = 0
        self.created = 0
        self._lru: "collections.OrderedDict[int, Tuple[Tuple[int, bool], PressureMultigrid]]" = \
            collections.OrderedDict()

    def _bytes(self, n: int, two_phase: bool) -> int:
        g = Inputs.grid(n)
        return estimate_bytes(g, levels_for(n), precision=self.precision, variable_coef=two_phase,
                              coarse_op="face" if two_phase else "reference", krylov=two_phase,
                              coarse_solve="direct")

    def acquire(self, n: int, two_phase: bool) -> Optional[PressureMultigrid]:
        key = (n, two_phase)
        if self.idle[key]:
            mg = self.idle[key].pop()
            self._lru.pop(id(mg), None)
            return mg
        need = self._bytes(n, two_phase)
        if need > self.budget:  # would wait forever: nothing to evict frees enough
            raise AdmissionError(f"a {n}^3 {'two-phase' if two_phase else 'constant-density'} solve "
                                 f"needs {need} bytes of HBM, more than the pool's whole budget "
                                 f"of {self.budget} bytes")
        while self.used + need > self.budget and self._lru:
            _, (k, victim) = self._lru.popitem(last=False)
            self.idle[k].remove(victim)
            self.used -= victim.pool_bytes
            victim.close()
        if self.used + need > self.budget:
            return None
        # two-phase requests: face-restricted coarse operators and GCR
        # (the reference's density hierarchy stalls on a 1000:1 sphere);
        # every request: the FP64 direct solve of the coarsest level, which
        # cuts the cycle count of the small sizes (64^3: 11 -> 7) -- each
        # request only has to meet its own tolerance
        mg = PressureMultigrid(Inputs.grid(n), levels_for(n), maxcy=self.maxcy, eps=1e-8,
                               precision=self.precision, variable_coef=two_phase,
                               coarse_op="face" if two_phase else "reference", krylov=two_phase,
                               coarse_solve="direct")
        if two_phase:
            mg.coeff(self.inputs.density(n))
        mg.pool_bytes = max(mg.device_bytes, need)
        self.used += mg.pool_bytes
        self.peak = max(self.peak, self.used)
        self.created += 1
        return mg

    def reserve(self, shapes, per_shape: int):
        """Create `per_shape` idle handles of each (n, two_phase) shape up front
        (within the budget), so a burst of same-shape requests finds them."""
        for n, tp in shapes:
            made = [self.acquire(n, tp) for _ in range(per_shape - len(self.idle[(n, tp)]))]
            for mg in made:
                if mg is not None:
                    self.release(n, tp, mg)

    def release(self, n: int, two_phase: bool, mg: PressureMultigrid):
        self.idle[(n, two_phase)].append(mg)
        self._lru[id(mg)] = ((n, two_phase), mg)

    def close(self):
        for lst in self.idle.values():
            for mg in lst:
                mg.close()
        self.idle.clear()
        self._lru.clear()
        self.used = 0


class Server:
    """`slots` worker threads, each with its own HIP stream."""

    def __init__(self, slots: int = 4, budget_bytes: int = 64 << 30, precision: str = "mixed",
                 maxcy: int = 60):
        self.inputs = Inputs()
        self.pool = SolverPool(self.inputs, budget_bytes, precision, maxcy)
        self.slots = slots
        self.streams = [hip.Stream() for _ in range(slots)]

    def _run_one(self, slot: int, req: Request, mg: PressureMultigrid):
        f = self.inputs.rhs(req.n, req.seed, req.two_phase)
        p = hip.DeviceArray(f.n)
        mg.set_eps(req.tol)
        res = mg.solve_device(f.ptr, p.ptr, self.streams[slot].handle, warn=False)
        return p, res

    def run(self, reqs: List[Request], realtime: bool = True) -> List[Result]:
        """Serve `reqs` (sorted by arrival).  With realtime=False every
        request counts as arrived at time 0 (a closed batch)."""
        self.inputs.prepare(reqs)
        # one handle per shape before the clock starts: its sweep tiling is
        # timed on an idle device, and handles created later in the run find
        # that choice in the library's per-process cache instead of timing
        # while other slots' solves run
        self.pool.reserve(sorted({(r.n, r.two_phase) for r in reqs}), 1)
        work: List["queue.Queue"] = [queue.Queue() for _ in range(self.slots)]
        done_q: "queue.Queue" = queue.Queue()
        t0 = time.perf_counter()

        def worker(slot: int):
            while True:
                item = work[slot].get()
                if item is None:
                    return
                req, mg, t_adm = item
                try:
                    p, res = self._run_one(slot, req, mg)
                    done_q.put((slot, req, mg, t_adm, time.perf_counter(), p, res, None))
                except Exception as e:  # reported to the caller below
                    done_q.put((slot, req, mg, t_adm, time.perf_counter(), None, None, e))

        threads = [threading.Thread(target=worker, args=(s,), daemon=True) for s in range(self.slots)]
        for th in threads:
            th.start()
        pending: "collections.deque[Request]" = collections.deque()
        free = list(range(self.slots))
        results: List[Result] = []
        nxt, total = 0, len(reqs)
        error = None
        try:
            while len(results) < total:
                now = time.perf_counter() - t0
                while nxt < total and (not realtime or reqs[nxt].arrival <= now):
                    pending.append(reqs[nxt])
                    nxt += 1
                # admit in arrival order; a request whose handle does not fit waits
                i = 0
                while free and i < len(pending):
                    req = pending[i]
                    mg = self.pool.acquire(req.n, req.two_phase)
                    if mg is None:
                        i += 1
                        continue
                    del pending[i]
                    work[free.pop()].put((req, mg, time.perf_counter() - t0))
                timeout = 0.0005
                if nxt < total and realtime and not pending:
                    timeout = max(0.0, min(0.01, reqs[nxt].arrival - (time.perf_counter() - t0)))
                try:
                    item = done_q.get(timeout=timeout) if timeout > 0 else done_q.get_nowait()
                except queue.Empty:
                    continue
                while True:
                    slot, req, mg, t_adm, t_end, p, res, err = item
                    free.append(slot)
                    self.pool.release(req.n, req.two_phase, mg)
                    if err is not None:
                        error = err
                        raise err
                    arr = req.arrival if realtime else 0.0
                    results.append(Result(req.rid, (t_end - t0 - arr) * 1e3, (t_adm - arr) * 1e3,
                                          (t_end - t0 - t_adm) * 1e3, res.iters, res.status,
                                          float(res.hist[-1]), p))
                    try:
                        item = done_q.get_nowait()
                    except queue.Empty:
                        break
        finally:
            for s in range(self.slots):
                work[s].put(None)
            for th in threads:
                th.join(timeout=60)
        if error is not None:
            raise error
        self.elapsed = time.perf_counter() - t0
        results.sort(key=lambda r: r.rid)
        return results

    def close(self):
        self.pool.close()
        for s in self.streams:
            s.destroy()


def summarize(results: List[Result], elapsed: float) -> dict:
    lat = np.array([r.latency_ms for r in results])
    return {
        "requests": len(results),
        "solves_per_s": round(len(results) / elapsed, 2),
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "mean_solve_ms": round(float(np.mean([r.solve_ms for r in results])), 3),
        "not_converged": int(sum(1 for r in results if r.status != 0)),
    }
