"""Benchmark: pressure-Poisson multigrid solves/s to 1e-8 relative residual on
one MI355X, FP64 (bitwise the reference's V-cycle) and mixed precision.

    python bench.py [--steps K] [--warmup W] [--n 256] [--levels 7] [--no-cpu-baseline]

A step is one complete solve -- from p = 0 to ||f - A p|| / ||f|| <= 1e-8 --
of the 256^3 single-block channel problem (periodic x, z; Neumann walls in
y; RHS = divergence of seeded face velocities, mean removed), with the RHS
already resident in HBM.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
W_SEED = 20261202  # the workloads' base seed
METRIC = "MG solves/s to 1e-8 resid (mixed vs FP64), HBM GB/s % peak, p50/p99 ms"


def algorithmic_bytes_per_cycle(n_cells_levels, w, c=0, gcr=False):
    """Streaming model per V(2,2) cycle (w: bytes of the inner precision, c:
    coefficient arrays per cell, 3 under variable coefficients).  On every
    level 4 sweeps of (3+c) w N_l less w N_l for the first, which starts
    from zero and reads no potential; restriction and prolongation
    (4.25+c) w N_l below the coarsest.  Finest-level FP64 outer step:
    (24+2w+8c) N_0 for the defect correction (p, f, e read; p, r written;
    FP64 coefficients), (80+3w+24c) N_0 for GCR(1)'s three passes
    (A z . w; d = z - beta d, w = A d; p += alpha d, r = f - A p)."""
    nu = 4
    L = len(n_cells_levels)
    b_inner = sum((nu * (3 + c) - 1) * n * w for n in n_cells_levels)
    b_inner += sum((4.25 + c) * n_cells_levels[l] * w for l in range(L - 1))
    per = (80 + 3 * w + 24 * c) if gcr else (24 + 2 * w + 8 * c)
    return b_inner + per * n_cells_levels[0]


def sweep_bytes_per_launch(w, c=0):
    """Finest-level sweep launches of a V(2,2) cycle: three full sweeps
    (read phi, the RHS and c coefficient arrays, write phi: (3+c) w per
    cell) and one from zero ((2+c) w), (11+4c)/4 w per cell on average."""
    return (3 * (3 + c) + (2 + c)) / 4.0 * w


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_rbs_256cubed_L7_traffic.md")
VEL_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_velmas_256cubed_L7.md")


def cpu_model():
    """The host CPU's model name (first /proc/cpuinfo entry), or None."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def measured_traffic(form, path=TRAFFIC_FILE, key="traffic_bytes_per_launch:"):
    """HBM bytes per launch of the 256^3, 7-level workload from a committed
    rocprofv3 PMC summary: FETCH_SIZE x 2 + WRITE_SIZE, the factors fixed by
    a 1 GiB copy in the same runs.  For k_rbs (default) the finest-level
    launch mix (3 full sweeps + 1 zero-start sweep); for k_velmas the key
    velocity_traffic_bytes_per_launch of its summary.  None when absent."""
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith(key):
                    kv = dict(item.split("=") for item in line.split(":", 1)[1].split())
                    return float(kv[form])
    except (OSError, ValueError, KeyError):
        return None
    return None


def run_gpu(grid, levels, f_host, precision, steps, warmup, tol, maxcy, time_kernel, dens=None,
            warn=True, **kw):
    """Device-resident solves of one problem: dens (per-block cell density)
    selects the variable-coefficient operator; kw goes to the handle."""
    import torch

    from multiflow3d_amd import PressureMultigrid
    from multiflow3d_amd import _lib

    dev = torch.device("cuda:0")
    f_dev = torch.from_numpy(np.ascontiguousarray(np.concatenate([a.ravel() for a in f_host]))).to(dev)
    p_dev = torch.zeros_like(f_dev)
    torch.cuda.synchronize()  # the solver runs on its own non-blocking stream
    mg = PressureMultigrid(grid, levels, maxcy=maxcy, eps=tol, precision=precision,
                           variable_coef=dens is not None, **kw)
    if dens is not None:
        mg.coeff(dens)
    if time_kernel:
        mg.time_kernel(_lib.MF3D_K_SMOOTH)
    for _ in range(warmup):
        mg.solve_device(f_dev.data_ptr(), p_dev.data_ptr(), warn=warn)
    torch.cuda.synchronize()
    if time_kernel:
        mg.time_kernel(_lib.MF3D_K_SMOOTH)  # resets the counters
    lat = []
    res = None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        res = mg.solve_device(f_dev.data_ptr(), p_dev.data_ptr(), warn=warn)
        lat.append((time.perf_counter() - ts) * 1e3)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    launches, kms = mg.kernel_time() if time_kernel else (0, 0.0)
    replay = mg.graph_replay()
    mg.time_kernel(0)
    p = p_dev.cpu().numpy()
    out = dict(ms=(t1 - t0) * 1e3 / steps, lat=lat, iters=res.iters, rho=float(res.hist[-1]),
               status=res.status, p=p, launches=launches, kernel_ms=kms, graph_replay=replay,
               hook_replay=replay if time_kernel else None)
    mg.close()
    return out


def run_pair(grid, levels, f_host, steps, warmup, tol, maxcy, label, dens=None, **kw):
    """FP64 and mixed solves of one problem, each with the finest sweep's
    launches timed inside the timed region (graph replay) where the runtime
    captures the timing events, else timed eagerly and the solve rate taken
    from a second, graph-replayed pass."""
    out = {}
    for prec in ("fp64", "mixed"):
        r = run_gpu(grid, levels, f_host, prec, steps, warmup, tol, maxcy, True, dens=dens, **kw)
        if not r["graph_replay"]:
            t = run_gpu(grid, levels, f_host, prec, steps, warmup, tol, maxcy, False, dens=dens, **kw)
            r.update(ms=t["ms"], lat=t["lat"], graph_replay=t["graph_replay"])
        print(f"# {label} {prec}: {r['ms']:.3f} ms/solve, {r['iters']} cycles, rho={r['rho']:.3e}",
              file=sys.stderr, flush=True)
        out[prec] = r
    return out["fp64"], out["mixed"]


def rel_l2_mean_removed(a, b):
    a = a - a.mean()
    b = b - b.mean()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def sweep_roofline(r, w, n0, c=0, traffic=None, traffic_source=None, kernel="k_rbs"):
    """Roofline of the finest-level sweep from the hook's launch times:
    sweep_bytes_per_launch(w, c) x n0 interior cells per launch."""
    if not r["launches"]:
        return None
    avg_s = r["kernel_ms"] / r["launches"] * 1e-3
    per_launch = sweep_bytes_per_launch(w, c) * n0
    achieved = per_launch / avg_s / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": traffic, "traffic_source": traffic_source if traffic else None,
            "kernel": f"{kernel} (fused red-black sweep, finest level)", "avg_us": round(avg_s * 1e6, 2),
            "launches_timed": r["launches"], "bytes_per_launch": int(per_launch),
            "timed_with": "HIP events on the solver's stream around each launch, inside the timed "
                          "region (graph replay)" if r["hook_replay"] else
                          "HIP events on the solver's stream around each launch, over the same "
                          "solves launched eagerly, the step time taken from a second pass with "
                          "graph replay (events recorded by a graph cannot be timed here)"}


def solve_leg(r, w, levels_cells, c=0, gcr=False, roof=None):
    """The per-precision record of one timed solve."""
    b = algorithmic_bytes_per_cycle(levels_cells, w, c, gcr) * r["iters"]
    gbps = b / (r["ms"] * 1e-3) / 1e9
    return {"solves_per_s": round(1e3 / r["ms"], 4), "ms_per_solve": round(r["ms"], 4),
            "cycles": r["iters"], "final_rho": r["rho"], "status": r["status"],
            "alg_GBps": round(gbps, 1), "alg_frac_of_peak": round(gbps / HBM_PEAK_GBPS, 4),
            "p50_ms": round(float(np.percentile(r["lat"], 50)), 3),
            "p99_ms": round(float(np.percentile(r["lat"], 99)), 3),
            "graph_replay": r["graph_replay"], "roofline": roof}


def direct_variants(grid, levels, f_host, steps, warmup, tol, maxcy, p64, cells, c=0, gcr=False, dens=None,
                    **kw):
    """The same problem with the coarsest level solved exactly in FP64
    (coarse_solve = "direct") in both precisions: solves/s, cycles and the
    pressure's distance from the FP64 reference-form solve p64.  Returns
    the records and the pressures."""
    out, ps = {}, {}
    for prec, w in (("fp64", 8), ("mixed", 4)):
        r = run_gpu(grid, levels, f_host, prec, steps, warmup, tol, maxcy, False, dens=dens,
                    coarse_solve="direct", **kw)
        print(f"# direct coarse {prec}: {r['ms']:.3f} ms/solve, {r['iters']} cycles, rho={r['rho']:.3e}",
              file=sys.stderr, flush=True)
        out[prec] = dict(solve_leg(r, w, cells, c, gcr), rel_l2_vs_fp64=rel_l2_mean_removed(r["p"], p64))
        ps[prec] = r["p"]
    return out, ps


def accuracy(grid, levels, f_host, pressures, dens=None, tol=1e-13, maxcy=200, **kw):
    """Distance of each solve's pressure (mean removed) from the FP64 solve
    taken to tol (coarsest level solved exactly): at a 1e-8 relative
    residual, what is left in the pressure is mostly the smoothest error
    component, which a residual norm weights by the operator's smallest
    eigenvalue, so two solves that both stop at 1e-8 can differ by far
    more than 1e-8."""
    r = run_gpu(grid, levels, f_host, "fp64", 1, 0, tol, maxcy, False, dens=dens, warn=False,
                coarse_solve="direct", **kw)
    print(f"# converged fp64: {r['iters']} cycles, rho={r['rho']:.3e}", file=sys.stderr, flush=True)
    return {"reference": f"FP64, coarsest level solved exactly, to {tol:g} relative residual "
                         f"({r['iters']} cycles, rho {r['rho']:.2e})",
            "rel_l2_vs_converged": {k: rel_l2_mean_removed(p, r["p"]) for k, p in pressures.items()}}


def run_uneven(steps, warmup, tol, maxcy):
    """The uneven 2 x 2 x 2 block grid: per-axis extents drawn from
    {96, 128, 160, 192} by SplitMix64 of seed BASE_SEED + 3, channel
    boundaries, 6 levels, all blocks of a level in one launch."""
    from multiflow3d_amd import topology as T
    from multiflow3d_amd import workloads as W

    ex, ey, ez = W.uneven_extents(W.BASE_SEED + 3)
    gn = [sum(ex), sum(ey), sum(ez)]
    grid = T.block_grid(ex, ey, ez, T.CHANNEL_BC, [1.0 / g for g in gn])
    L = 6
    f = W.split_blocks(grid, W.global_rhs(grid, W.BASE_SEED + 3))
    cells = [sum(int(np.prod([x >> l for x in b.n])) for b in grid.blocks) for l in range(L)]
    r64, rmx = run_pair(grid, L, f, steps, warmup, tol, maxcy, "uneven 2x2x2")
    n0 = cells[0]
    direct, pd = direct_variants(grid, L, f, steps, warmup, tol, maxcy, r64["p"], cells)
    acc = accuracy(grid, L, f, {"fp64": r64["p"], "mixed": rmx["p"], "fp64_direct_coarse": pd["fp64"],
                                "mixed_direct_coarse": pd["mixed"]})
    return {"workload": f"2 x 2 x 2 blocks, extents x {ex}, y {ey}, z {ez} (seed BASE_SEED+3), "
                        f"periodic x/z, Neumann y, V(2,2) RBGS, {L} levels, to {tol:g} relative residual, "
                        f"ghosts by the on-device region-table exchange",
            "cells": n0, "levels": L, "blocks": [list(b.n) for b in grid.blocks],
            "fp64": solve_leg(r64, 8, cells, roof=sweep_roofline(r64, 8, n0)),
            "mixed": solve_leg(rmx, 4, cells, roof=sweep_roofline(rmx, 4, n0)),
            "mixed_rel_l2_vs_fp64": rel_l2_mean_removed(rmx["p"], r64["p"]),
            "speedup_mixed_vs_fp64": round(r64["ms"] / rmx["ms"], 3),
            "coarsest_cells": cells[-1], "direct_coarse": direct, "accuracy": acc}


def run_two_phase(steps, warmup, tol, maxcy, n=384):
    """The n^3 two-phase problem: sphere r = 0.2 at the centre, density
    1000:1 smoothed over 1.5 dx, RHS scaled by the cell volume; face-
    restricted coarse operators, GCR(1) outer iteration, 7 levels.  Also
    the reference's own form (its coarse operators, defect correction) in
    FP64 for maxcy cycles, to show where it stops."""
    from multiflow3d_amd import topology as T
    from multiflow3d_amd import workloads as W

    grid = T.single_block(n, [1.0 / n] * 3, T.CHANNEL_BC)
    L = 7
    f = W.split_blocks(grid, W.global_rhs(grid, W.BASE_SEED + 4, fv_volume=True))
    dens = W.split_blocks(grid, W.sphere_density(grid))
    cells = [(n >> l) ** 3 for l in range(L)]
    kw = dict(coarse_op="face", krylov=True)
    r64, rmx = run_pair(grid, L, f, steps, warmup, tol, maxcy, f"{n}^3 two-phase", dens=dens, **kw)
    ref = run_gpu(grid, L, f, "fp64", 1, 0, tol, maxcy, False, dens=dens, warn=False)
    print(f"# {n}^3 two-phase reference form fp64: {ref['iters']} cycles, rho={ref['rho']:.3e}",
          file=sys.stderr, flush=True)
    n0 = cells[0]
    direct, pd = direct_variants(grid, L, f, steps, warmup, tol, maxcy, r64["p"], cells, 3, True, dens=dens, **kw)
    acc = accuracy(grid, L, f, {"fp64": r64["p"], "mixed": rmx["p"], "fp64_direct_coarse": pd["fp64"],
                                "mixed_direct_coarse": pd["mixed"], "reference_form_fp64": ref["p"]},
                   dens=dens, **kw)
    return {"workload": f"single {n}^3 block, periodic x/z, Neumann y, div((1/rho) grad p), sphere r 0.2, "
                        f"rho 1000:1 over 1.5 dx, face-restricted coarse operators, GCR(1) outer, "
                        f"V(2,2) RBGS, {L} levels, to {tol:g} relative residual",
            "cells": n0, "levels": L,
            "fp64": solve_leg(r64, 8, cells, 3, True, sweep_roofline(r64, 8, n0, 3)),
            "mixed": solve_leg(rmx, 4, cells, 3, True, sweep_roofline(rmx, 4, n0, 3)),
            "mixed_rel_l2_vs_fp64": rel_l2_mean_removed(rmx["p"], r64["p"]),
            "speedup_mixed_vs_fp64": round(r64["ms"] / rmx["ms"], 3),
            "coarsest_cells": cells[-1], "direct_coarse": direct, "accuracy": acc,
            "reference_form_fp64": {
                "cycles": ref["iters"], "final_rho": ref["rho"], "status": ref["status"],
                "ms_per_solve": round(ref["ms"], 4),
                "note": "the reference's variable-coefficient V-cycle (coarse operators from the "
                        "finest spacing, defect correction) stalls on this problem: it stops "
                        f"near rho = {ref['rho']:.2e}, not at {tol:g}"}}


def run_velocity(grid, levels, precision, steps, warmup, rel_eps, maxcy, time_kernel):
    """The velocity form of newsolv_mg (calvel/calmas around each V-cycle,
    the reference's own outer loop) on resident u*, v*, w*: each step
    starts from u = u*, p = 0 (restored outside the timed call) and runs
    until rmax < rel_eps * rmax_0.  Steps are timed one by one on the host
    around the synchronous call."""
    import torch

    from multiflow3d_amd import PressureMultigrid
    from multiflow3d_amd import _lib
    from multiflow3d_amd import workloads as W

    dev = torch.device("cuda:0")
    us, vs, ws = W.velocity_inputs(grid, W.BASE_SEED + 40)
    init = [torch.from_numpy(np.ascontiguousarray(np.concatenate([a.ravel() for a in x]))).to(dev)
            for x in (us, vs, ws)]
    work = [t.clone() for t in init] + [torch.zeros_like(init[0])]
    torch.cuda.synchronize()
    mg = PressureMultigrid(grid, levels, maxcy=maxcy, precision=precision)

    def reset():
        for dst, src in zip(work[:3], init):
            dst.copy_(src)
        work[3].zero_()
        torch.cuda.synchronize()

    ptrs = [t.data_ptr() for t in init] + [t.data_ptr() for t in work]
    reset()
    r0 = mg.newsolv_device(ptrs, 0.01, 1.0, 0.0, warn=False)  # rmax_0 from step 0's calmas
    eps = rel_eps * float(r0.hist[0])
    if time_kernel:
        mg.time_kernel(_lib.MF3D_K_OUTER)
    for _ in range(warmup):
        reset()
        mg.newsolv_device(ptrs, 0.01, 1.0, eps, warn=False)
    if time_kernel:
        mg.time_kernel(_lib.MF3D_K_OUTER)
    lat = []
    res = None
    for _ in range(steps):
        reset()
        ts = time.perf_counter()
        res = mg.newsolv_device(ptrs, 0.01, 1.0, eps, warn=False)
        lat.append((time.perf_counter() - ts) * 1e3)
    launches, kms = mg.kernel_time() if time_kernel else (0, 0.0)
    replay = mg.graph_replay()
    mg.time_kernel(0)
    u = work[0].cpu().numpy()
    mg.close()
    return dict(ms=float(np.mean(lat)), lat=lat, iters=res.iters, rmax=res.rmax, rmax0=float(r0.hist[0]),
                eps=eps, status=res.status, u=u, launches=launches, kernel_ms=kms, graph_replay=replay)


def levelset_phi(grid):
    """Signed distance to a sphere of radius 0.3 at the domain centre, over
    the ghosted box of each block (the transition band |phi| <= 2 dx is
    about 1.7 % of the cells at 256^3)."""
    out = []
    for blk in grid.blocks:
        nx, ny, nz = blk.n
        ox, oy, oz = blk.origin
        ax = [(o + np.arange(m + 2) - 0.5) * h - 0.5 for o, m, h in zip((ox, oy, oz), (nx, ny, nz), grid.h)]
        r = np.sqrt(ax[2][:, None, None] ** 2 + ax[1][None, :, None] ** 2 + ax[0][None, None, :] ** 2)
        out.append(0.3 - r)
    return out


def run_levelset(grid, levels, precision, steps, warmup):
    """The LSM density update of every time step on a resident phi:
    heaviside + bound_lsm(18/19) + coeff (mf3d_mg_heaviside_device, dens and
    mu written to caller buffers too).  Steps are timed one by one on the
    host around the synchronous call; k_heaviside by the kernel hook."""
    import torch

    from multiflow3d_amd import PressureMultigrid
    from multiflow3d_amd import _lib

    dev = torch.device("cuda:0")
    phi = torch.from_numpy(np.concatenate([a.ravel() for a in levelset_phi(grid)])).to(dev)
    dens = torch.empty_like(phi)
    mu = torch.empty_like(phi)
    torch.cuda.synchronize()
    mg = PressureMultigrid(grid, levels, maxcy=4, precision=precision, variable_coef=True)
    props = (1000.0, 1.2, 1.0e-3, 1.8e-5)
    for _ in range(warmup):
        mg.heaviside_device(phi.data_ptr(), *props, dens_ptr=dens.data_ptr(), mu_ptr=mu.data_ptr())
    lat = []
    for _ in range(steps):
        ts = time.perf_counter()
        mg.heaviside_device(phi.data_ptr(), *props, dens_ptr=dens.data_ptr(), mu_ptr=mu.data_ptr())
        lat.append((time.perf_counter() - ts) * 1e3)
    mg.time_kernel(_lib.MF3D_K_HEAVISIDE)
    for _ in range(steps):
        mg.heaviside_device(phi.data_ptr(), *props, dens_ptr=dens.data_ptr(), mu_ptr=mu.data_ptr())
    launches, kms = mg.kernel_time()
    mg.time_kernel(0)
    mg.close()
    band = int((phi.abs() <= 2 * max(grid.h)).sum().item())
    return dict(ms=float(np.mean(lat)), lat=lat, launches=launches, kernel_ms=kms, cells=phi.numel(), band=band)


def shipped_grid():
    """The layout of the reference's shipped deck (infodom.cin, control.cin):
    80 blocks of 40 x 10 x 25 cells lined up in x, dx = 0.005, periodic in
    x, walls south/north/bottom, slip top."""
    from multiflow3d_amd import topology as T
    return T.block_grid([40] * 80, [10], [25], [T.PERIODIC, T.PERIODIC, T.NEUMANN, T.NEUMANN, T.NEUMANN, 3],
                        [0.005] * 3)


def run_shipped(precision, steps, warmup, time_kernel):
    """newsolv_mg in the velocity form with the shipped deck's multigrid
    settings (control.cin: ngrid 1, mg_itrsch 2 = ADI-TDMA, maxcy 30,
    irestr 2, iproln 1) on resident seeded u*, v*, w*, p from 0, eps = 0
    so every call runs the deck's 30 cycles.  Steps timed one by one on the
    host around the synchronous call."""
    import torch

    from multiflow3d_amd import PressureMultigrid
    from multiflow3d_amd import _lib
    from multiflow3d_amd import workloads as W

    grid = shipped_grid()
    dev = torch.device("cuda:0")
    us, vs, ws = W.velocity_inputs(grid, W.BASE_SEED + 95)
    init = [torch.from_numpy(np.ascontiguousarray(np.concatenate([a.ravel() for a in x]))).to(dev)
            for x in (us, vs, ws)]
    work = [t.clone() for t in init] + [torch.zeros_like(init[0])]
    torch.cuda.synchronize()
    mg = PressureMultigrid(grid, 1, nu_pre=2, nu_post=1, smoother=2, maxcy=30, precision=precision)

    def reset():
        for dst, src in zip(work[:3], init):
            dst.copy_(src)
        work[3].zero_()
        torch.cuda.synchronize()

    ptrs = [t.data_ptr() for t in init] + [t.data_ptr() for t in work]
    if time_kernel:
        mg.time_kernel(_lib.MF3D_K_SMOOTH)
    for _ in range(warmup):
        reset()
        mg.newsolv_device(ptrs, 0.0025, 1.0, 0.0, warn=False)
    if time_kernel:
        mg.time_kernel(_lib.MF3D_K_SMOOTH)
    lat, res = [], None
    for _ in range(steps):
        reset()
        ts = time.perf_counter()
        res = mg.newsolv_device(ptrs, 0.0025, 1.0, 0.0, warn=False)
        lat.append((time.perf_counter() - ts) * 1e3)
    launches, kms = mg.kernel_time() if time_kernel else (0, 0.0)
    mg.time_kernel(0)
    mg.close()
    return dict(ms=float(np.mean(lat)), lat=lat, iters=res.iters, rmax=res.rmax, launches=launches, kernel_ms=kms,
                cells=sum(b.n[0] * b.n[1] * b.n[2] for b in grid.blocks))


def run_serving(count):
    """Serving sweep: Poisson arrivals of independent solves, n^3 with n
    uniform in {32,64,96,128,192,256}, tolerance uniform in {1e-6,1e-8,1e-10},
    constant or two-phase coefficient 50/50, mixed precision, 4 slots
    (streams), hierarchies resident after one warm-up request per shape."""
    from multiflow3d_amd import serving as S
    srv = S.Server(slots=4, precision="mixed")
    out = {"slots": 4, "precision": "mixed", "requests_per_rate": count, "rates": []}
    try:
        # warm pool: one resident hierarchy per slot and shape
        shapes = [(n, tp) for n in S.SIZES for tp in (False, True)]
        warm = [S.Request(i, 0.0, n, 1e-6, tp, W_SEED + 5) for i, (n, tp) in
                enumerate(sh for sh in shapes for _ in range(4))]
        srv.pool.reserve(shapes, 4)
        srv.run(warm, realtime=False)
        for rate in (100.0, 400.0, 1600.0):
            reqs = S.poisson_requests(rate, count, W_SEED + 50)
            res = srv.run(reqs)
            summ = S.summarize(res, srv.elapsed)
            summ["rate_per_s"] = rate
            out["rates"].append(summ)
            print(f"# serving rate {rate:g}/s: {summ}", file=sys.stderr, flush=True)
        reqs = S.poisson_requests(1.0, count, W_SEED + 50)
        res = srv.run(reqs, realtime=False)
        out["closed_batch"] = S.summarize(res, srv.elapsed)
        out["pool_peak_bytes"] = srv.pool.peak
        out["handles_created"] = srv.pool.created
    finally:
        srv.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--levels", type=int, default=7)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--maxcy", type=int, default=60)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-velocity", action="store_true",
                    help="skip the velocity-form (calvel/calmas) leg")
    ap.add_argument("--no-shipped", action="store_true",
                    help="skip the shipped-deck leg (80 blocks, one level, ADI smoother)")
    ap.add_argument("--no-levelset", action="store_true",
                    help="skip the level-set density update (heaviside + coeff) leg")
    ap.add_argument("--no-uneven", action="store_true",
                    help="skip the uneven 2x2x2 block-grid leg")
    ap.add_argument("--no-two-phase", action="store_true",
                    help="skip the 384^3 two-phase (1000:1) leg")
    ap.add_argument("--only", choices=("uneven", "two_phase"), default=None,
                    help="run one workload leg alone and print its record (for profiling it apart "
                         "from the others)")
    ap.add_argument("--serving-requests", type=int, default=64,
                    help="requests per arrival rate in the serving sweep (0: skip it)")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="after the solves, copy MIB MiB device-to-device once (a copy of known size "
                         "to scale HBM byte counters read by a profiler)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(independent solves shard across GPUs by request)", file=sys.stderr)
        sys.exit(2)

    if args.only:
        leg = {"uneven": run_uneven, "two_phase": run_two_phase}[args.only]
        print(json.dumps({"leg": args.only, **leg(args.steps, args.warmup, args.tol, args.maxcy)}), flush=True)
        return

    from multiflow3d_amd import topology as T
    from multiflow3d_amd import workloads as W

    n, L = args.n, args.levels
    grid = T.single_block(n, [1.0 / n] * 3, T.CHANNEL_BC)
    seed = W.BASE_SEED + 2
    f_host = W.split_blocks(grid, W.global_rhs(grid, seed))
    levels_cells = [(n >> l) ** 3 for l in range(L)]

    r64, rmx = run_pair(grid, L, f_host, args.steps, args.warmup, args.tol, args.maxcy, f"{n}^3")
    mixed_err = rel_l2_mean_removed(rmx["p"], r64["p"])
    acc = accuracy(grid, L, f_host, {"fp64": r64["p"], "mixed": rmx["p"]})

    if args.calib_copy > 0:
        import torch
        nbytes = args.calib_copy << 20
        a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(f"# calibration copy: {nbytes} bytes read, {nbytes} bytes written", file=sys.stderr, flush=True)
        del a, b

    # Dominant kernel: k_rbs, one fused red-black sweep (both colours) of the
    # finest level; sweep_bytes_per_launch gives its algorithmic bytes.
    N0 = levels_cells[0]
    src = ("profiles/k_rbs_256cubed_L7_traffic.md (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE per launch, "
           "calibrated on a 1 GiB copy)")

    def roofline(r, w):
        traffic = measured_traffic("fp64" if w == 8 else "mixed") if (n, L) == (256, 7) else None
        return sweep_roofline(r, w, N0, 0, traffic, src)

    cpu = None
    if not args.no_cpu_baseline:
        from oracle.oracle import OracleMG
        o = OracleMG(grid.blocks, L)
        o.coeff()
        t0 = time.perf_counter()
        p_o, rho_o, _, k_o = o.solve(f_host, args.tol, args.maxcy)
        t_cpu = time.perf_counter() - t0
        o.close()
        fp64_bitwise = bool(np.array_equal(p_o[0].ravel(), r64["p"]))
        cpu = {"value": round(1.0 / t_cpu, 5), "unit": "solves/s", "cores": 1, "kind": "port",
               "sample": f"one full FP64 solve of the same {n}^3 problem ({k_o} cycles, "
                         f"{t_cpu:.1f} s) by the C restatement of the reference (oracle/mg_oracle.c), "
                         f"single thread",
               "fp64_gpu_bitwise_equal": fp64_bitwise,
               "cpu_model": cpu_model(), "host_cpus": os.cpu_count()}
        # the same solve on every core the process may use (the OpenMP build
        # of the same source; its results are bitwise the serial build's)
        threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or len(os.sched_getaffinity(0))
        os.environ["OMP_NUM_THREADS"] = str(threads)  # read when libgomp loads
        o = OracleMG(grid.blocks, L, omp=True)
        o.coeff()
        t0 = time.perf_counter()
        p_m, _, _, k_m = o.solve(f_host, args.tol, args.maxcy)
        t_mc = time.perf_counter() - t0
        o.close()
        cpu["multicore"] = {"value": round(1.0 / t_mc, 5), "unit": "solves/s", "cores": threads,
                            "sample": f"the same solve ({k_m} cycles, {t_mc:.2f} s) by the OpenMP build "
                                      f"of oracle/mg_oracle.c (plane loops parallel, reductions serial)",
                            "bitwise_equal_to_single_thread": bool(np.array_equal(p_m[0], p_o[0]))}

    velocity = None
    if not args.no_velocity:
        vel = {}
        for prec in ("fp64", "mixed"):
            v = run_velocity(grid, L, prec, args.steps, args.warmup, args.tol, args.maxcy, False)
            t = run_velocity(grid, L, prec, 2, 1, args.tol, args.maxcy, True)  # eager, kernel hook on
            w = 8 if prec == "fp64" else 4
            per_launch = (56 + w) * N0
            avg_s = t["kernel_ms"] / t["launches"] * 1e-3 if t["launches"] else None
            vel[prec] = {"solves_per_s": round(1e3 / v["ms"], 4), "ms_per_solve": round(v["ms"], 4),
                         "cycles": v["iters"], "rmax0": v["rmax0"], "eps": v["eps"], "final_rmax": v["rmax"],
                         "status": v["status"], "graph_replay": v["graph_replay"],
                         "p50_ms": round(float(np.percentile(v["lat"], 50)), 3),
                         "p99_ms": round(float(np.percentile(v["lat"], 99)), 3),
                         "k_velmas": None if avg_s is None else {
                             "bound": "hbm", "avg_us": round(avg_s * 1e6, 2), "launches_timed": t["launches"],
                             "bytes_per_launch": int(per_launch),
                             "achieved": round(per_launch / avg_s / 1e9, 1), "peak": HBM_PEAK_GBPS,
                             "unit": "GB/s", "frac": round(per_launch / avg_s / 1e9 / HBM_PEAK_GBPS, 4),
                             "traffic": measured_traffic(prec, VEL_TRAFFIC_FILE,
                                                         "velocity_traffic_bytes_per_launch:")
                             if (n, L) == (256, 7) else None,
                             "model": "reads u*, v*, w*, p (32 B), writes u, v, w (24 B) and the "
                                      "V-cycle RHS (w B) per cell"}}
            vel[prec + "_u"] = v["u"]
            print(f"# velocity {prec}: {v['ms']:.3f} ms/solve, {v['iters']} cycles, rmax {v['rmax']:.3e} "
                  f"(eps {v['eps']:.3e})", file=sys.stderr, flush=True)
        ua, ub = vel.pop("mixed_u"), vel.pop("fp64_u")
        vel["mixed_u_rel_l2_vs_fp64"] = float(np.linalg.norm(ua - ub) / np.linalg.norm(ub))
        vel["speedup_mixed_vs_fp64"] = round(vel["fp64"]["ms_per_solve"] / vel["mixed"]["ms_per_solve"], 3)
        vel["workload"] = (f"newsolv_mg velocity form (calvel/calmas/exchange around each V-cycle) on the "
                           f"same {n}^3 channel, u* seeded, p from 0, dt 0.01, alfapr 1, until "
                           f"rmax < {args.tol:g} x rmax_0")
        velocity = vel

    levelset = None
    if not args.no_levelset:
        levelset = {}
        for prec in ("fp64", "mixed"):
            r = run_levelset(grid, L, prec, args.steps, args.warmup)
            cg = r["cells"]
            per_launch = 32 * cg  # reads phi, writes the level-0 density, dens and mu (8 B each)
            avg_s = r["kernel_ms"] / r["launches"] * 1e-3 if r["launches"] else None
            levelset[prec] = {
                "ms_per_update": round(r["ms"], 4),
                "p50_ms": round(float(np.percentile(r["lat"], 50)), 3),
                "k_heaviside": None if avg_s is None else {
                    "bound": "hbm", "avg_us": round(avg_s * 1e6, 2), "launches_timed": r["launches"],
                    "bytes_per_launch": int(per_launch),
                    "achieved": round(per_launch / avg_s / 1e9, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                    "frac": round(per_launch / avg_s / 1e9 / HBM_PEAK_GBPS, 4),
                    "model": "reads phi, writes the handle's level-0 density and the caller's dens and "
                             "mu, 8 B each per ghosted cell"}}
            print(f"# levelset {prec}: {r['ms']:.3f} ms/update (heaviside + coeff), band {r['band']} of "
                  f"{cg} cells", file=sys.stderr, flush=True)
        levelset["workload"] = (f"heaviside + bound_lsm(18/19) + coeff per time step on the same {n}^3 "
                                f"grid, {L} levels: phi = signed distance to a sphere r = 0.3, "
                                f"densl/densg 1000/1.2, mul/mug 1e-3/1.8e-5, dens and mu returned")
    shipped = None
    if not args.no_shipped:
        shipped = {}
        for prec in ("fp64", "mixed"):
            r = run_shipped(prec, args.steps, args.warmup, False)
            t = run_shipped(prec, 2, 1, True)
            sweep_us = t["kernel_ms"] / t["launches"] * 1e3 if t["launches"] else None
            shipped[prec] = {"ms_per_call": round(r["ms"], 4), "cycles": r["iters"],
                             "ms_per_cycle": round(r["ms"] / max(r["iters"], 1), 4),
                             "p50_ms": round(float(np.percentile(r["lat"], 50)), 3),
                             "adi_sweep_us": None if sweep_us is None else round(sweep_us, 2),
                             "adi_sweeps_timed": t["launches"]}
            print(f"# shipped deck {prec}: {r['ms']:.3f} ms per newsolv_mg call ({r['iters']} cycles), "
                  f"ADI sweep {sweep_us} us", file=sys.stderr, flush=True)
        cells = r["cells"]
        if not args.no_cpu_baseline:
            from oracle.oracle import OracleMG
            from multiflow3d_amd import workloads as W
            g = shipped_grid()
            o = OracleMG(g.blocks, 1, nu_pre=2, nu_post=1, smoother=2)
            o.coeff()
            f = W.split_blocks(g, W.global_rhs(g, W.BASE_SEED + 96))
            t0 = time.perf_counter()
            o.solve(f, 0.0, 30)
            t_cpu = time.perf_counter() - t0
            o.close()
            shipped["cpu_oracle_ms_per_call"] = round(t_cpu * 1e3, 1)
            shipped["cpu_sample"] = ("the C restatement (oracle/mg_oracle.c), 1 thread, 30 ADI cycles of the "
                                     "Poisson form on the same layout")
        shipped["workload"] = (f"the reference's shipped deck layout (80 blocks of 40x10x25 = {cells} cells, "
                               f"dx 0.005, periodic x, walls S/N/B, slip top) with its multigrid settings: "
                               f"1 level, ADI-TDMA (mg_itrsch 2), irestr 2, iproln 1, 30 cycles per call, "
                               f"velocity form, u* seeded, p from 0")
    uneven = None if args.no_uneven else run_uneven(args.steps, args.warmup, args.tol, args.maxcy)
    two_phase = None if args.no_two_phase else run_two_phase(args.steps, args.warmup, args.tol, args.maxcy)
    serving = run_serving(args.serving_requests) if args.serving_requests > 0 else None

    line = {
        "metric": METRIC,
        "value": round(1e3 / rmx["ms"], 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(rmx["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner cycle, f64 outer residual/update",
        "data": "synthetic (SplitMix64 face velocities -> divergence, mean removed)",
        "config": {"workload": f"single {n}^3 block, periodic x/z, Neumann y, V(2,2) RBGS, "
                               f"{L} levels, to {args.tol:g} relative residual",
                   "cells": N0, "levels": L, "tol": args.tol},
        "roofline": roofline(rmx, 4),
        "cpu_baseline": cpu,
        "fp64": solve_leg(r64, 8, levels_cells, roof=roofline(r64, 8)),
        "mixed": dict(solve_leg(rmx, 4, levels_cells), rel_l2_vs_fp64=mixed_err),
        "speedup_mixed_vs_fp64": round(r64["ms"] / rmx["ms"], 3),
        "accuracy": acc,
        "velocity_form": velocity,
        "levelset_update": levelset,
        "shipped_deck": shipped,
        "uneven_blocks": uneven,
        "two_phase": two_phase,
        "serving": serving,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
