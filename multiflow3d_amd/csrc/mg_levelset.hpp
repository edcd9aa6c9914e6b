// mg_levelset.hpp -- the two-phase density from a level set on the device
// (heaviside, lsm.for:1456-1509), so that the LSM coefficients can be
// rebuilt every time step from a resident phi without a host round trip.
//
// Per block epsl = n_epsl*max(dx, dy, dz) with n_epsl = 2; per ghosted cell
//   h = 0                                         phi < -epsl
//   h = 1                                         phi >  epsl
//   h = 0.5*((1 + phi/epsl) + (1/pi)*sin((pi*phi)/epsl))   |phi| <= epsl
// with pi = 3.14159265359 (the reference's constant, not M_PI), then
//   dens = densg + (densl - densg)*h,   mu = mug + (mul - mug)*h,
// each multiply, add and divide rounded on its own (no contraction), as
// the reference evaluates them.  sin is the device library's FP64 sin; the
// reference takes the host libm's, and the two may round differently in
// the last place.
//
// bound_lsm(18/19) and exchange(18/19) only copy values, so the ghost and
// boundary-row entries they rewrite are a fixed map {destination, source}
// over the level-0 arena (composed on the host).  Since the density is a
// pointwise function of phi, a rewritten entry is simply the function of
// phi at its source: k_heaviside fills every cell, k_heaviside_map then
// overwrites the mapped ones, reading only phi, so the map needs no
// staging.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mf3d {

// a block of the level-0 arena: first entry and the block's epsl
struct LsBlock {
    long long off;
    long long cells;
    double epsl;
};

struct LsProps {
    double densl, densg, mul, mug;
};

constexpr double kLsPi = 3.14159265359;  // lsm.for:1466

__device__ __forceinline__ double ls_h(double phi, double epsl) {
    if (phi < -1.0 * epsl) return 0.0;
    if (phi > epsl) return 1.0;
    if (fabs(phi) <= epsl) return 0.5 * ((1.0 + phi / epsl) + (1.0 / kLsPi) * sin((kLsPi * phi) / epsl));
    return phi;  // NaN: the reference leaves h unset; propagate it
}

// dens (and the optional copies / viscosity) of one entry
__device__ __forceinline__ void ls_store(long long c, double h, const LsProps &pr, double *dens, double *dens_out,
                                         double *mu) {
    const double d = pr.densg + (pr.densl - pr.densg) * h;
    dens[c] = d;
    if (dens_out) dens_out[c] = d;
    if (mu) mu[c] = pr.mug + (pr.mul - pr.mug) * h;
}

// Every ghosted cell of block blockIdx.y; a workgroup takes 4 x 256
// consecutive entries.
__global__ __launch_bounds__(256) void k_heaviside(const double *__restrict__ phi, double *__restrict__ dens,
                                                   double *__restrict__ dens_out, double *__restrict__ mu,
                                                   LsProps pr, const LsBlock *__restrict__ lb) {
    const LsBlock B = lb[blockIdx.y];
    constexpr int kPer = 4;
    const long long base = (long long)blockIdx.x * (256 * kPer) + threadIdx.x;
    if (base >= B.cells) return;
    double v[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const long long c = base + u * 256;
        v[u] = c < B.cells ? phi[B.off + c] : 0.0;
    }
    // every h before the first store (a load consumed after a store makes
    // the wait for it wait for the store too)
#pragma unroll
    for (int u = 0; u < kPer; ++u) v[u] = ls_h(v[u], B.epsl);
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const long long c = base + u * 256;
        if (c < B.cells) ls_store(B.off + c, v[u], pr, dens, dens_out, mu);
    }
}

// The entries bound_lsm + exchange rewrite: dst takes the function of phi
// at src, with the epsl of the block that holds src.
__global__ __launch_bounds__(256) void k_heaviside_map(const double *__restrict__ phi, const uint2 *__restrict__ ent,
                                                       long long n, double *__restrict__ dens,
                                                       double *__restrict__ dens_out, double *__restrict__ mu,
                                                       LsProps pr, const LsBlock *__restrict__ lb, int nb) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const uint2 m = ent[e];
        int lo = 0, hi = nb - 1;  // last block whose offset is <= src
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (lb[mid].off <= (long long)m.y) lo = mid;
            else hi = mid - 1;
        }
        ls_store(m.x, ls_h(phi[m.y], lb[lo].epsl), pr, dens, dens_out, mu);
    }
}

}  // namespace mf3d
