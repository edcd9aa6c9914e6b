// mg_transfer.hpp -- residual + restriction and prolongation + correction.
#pragma once
#include "mg_common.hpp"

namespace mf3d {

template <typename T, bool VAR>
__device__ __forceinline__ T residual_at(const T *__restrict__ phi, const T *__restrict__ rhs,
                                         const T *__restrict__ cx, const T *__restrict__ cy,
                                         const T *__restrict__ cz, const BlockLevel &g, int i,
                                         int j, int k) {
    const long long q = gidx(g, i, j, k);
    T aw, ae, as, an, ab, at, ap;
    coefs<T, VAR>(g, cx, cy, cz, q, i, j, k, aw, ae, as, an, ab, at, ap);
    T s = aw * phi[q - 1];
    s = s + ae * phi[q + 1];
    s = s + as * phi[q - g.sy];
    s = s + an * phi[q + g.sy];
    s = s + ab * phi[q - g.sz];
    s = s + at * phi[q + g.sz];
    s = s + ap * phi[q];
    return rhs[q] - s;
}

// Fine residual + 8-cell restriction (mgsolver.for:688-751); one thread per
// coarse interior cell, children summed in the reference's order.  The
// coarse potential is not touched: the first coarse sweep treats it as zero.
template <typename T, bool VAR>
__global__ __launch_bounds__(256) void k_restrict(
    const T *__restrict__ phif, const T *__restrict__ rhsf, const T *__restrict__ cx,
    const T *__restrict__ cy, const T *__restrict__ cz, const BlockLevel *__restrict__ blf,
    T *__restrict__ rhsc, const BlockLevel *__restrict__ blc, const Tile *__restrict__ tiles,
    const int *__restrict__ done) {
    if (done && *done) return;
    const Tile t = tiles[blockIdx.x];
    const BlockLevel gc = blc[t.b];
    const int i = t.x0 + threadIdx.x, j = t.y0 + threadIdx.y, k = t.z0;
    if (i > gc.n0 || j > gc.n1) return;
    const BlockLevel gf = blf[t.b];
    const int x0 = 2 * i - 1, y0 = 2 * j - 1, z0 = 2 * k - 1;
    T s = residual_at<T, VAR>(phif, rhsf, cx, cy, cz, gf, x0, y0, z0);
    s = s + residual_at<T, VAR>(phif, rhsf, cx, cy, cz, gf, x0 + 1, y0, z0);
    s = s + residual_at<T, VAR>(phif, rhsf, cx, cy, cz, gf, x0, y0 + 1, z0);
    s = s + residual_at<T, VAR>(phif, rhsf, cx, cy, cz, gf, x0 + 1, y0 + 1, z0);
    s = s + residual_at<T, VAR>(phif, rhsf, cx, cy, cz, gf, x0, y0, z0 + 1);
    s = s + residual_at<T, VAR>(phif, rhsf, cx, cy, cz, gf, x0 + 1, y0, z0 + 1);
    s = s + residual_at<T, VAR>(phif, rhsf, cx, cy, cz, gf, x0, y0 + 1, z0 + 1);
    s = s + residual_at<T, VAR>(phif, rhsf, cx, cy, cz, gf, x0 + 1, y0 + 1, z0 + 1);
    rhsc[gidx(gc, i, j, k)] = T(0.125) * s;
}

// Restriction into a level where a block repeats its grid (a refined grid's
// block past its own level count): the coarse RHS is the fine residual,
// cell for cell (mgsolver.for:719-732).  Tiles: 64 x 4 interior cells of
// one plane of such blocks.
template <typename T, bool VAR>
__global__ __launch_bounds__(256) void k_restrict_rep(
    const T *__restrict__ phif, const T *__restrict__ rhsf, const T *__restrict__ cx,
    const T *__restrict__ cy, const T *__restrict__ cz, const BlockLevel *__restrict__ blf,
    T *__restrict__ rhsc, const BlockLevel *__restrict__ blc, const Tile *__restrict__ tiles,
    const int *__restrict__ done) {
    if (done && *done) return;
    const Tile t = tiles[blockIdx.x];
    const BlockLevel gc = blc[t.b];
    const int i = t.x0 + threadIdx.x, j = t.y0 + threadIdx.y, k = t.z0;
    if (i > gc.n0 || j > gc.n1) return;
    rhsc[gidx(gc, i, j, k)] = residual_at<T, VAR>(phif, rhsf, cx, cy, cz, blf[t.b], i, j, k);
}

// Trilinear prolongation added to every fine point including the ghost
// layer (mgsolver.for:823-891), fused with the Neumann copy that follows it
// (mgbound, mgsolver.for:902): a ghost on a physical Neumann face takes the
// corrected value of the cell it reflects.  Reads fin, writes fout, so the
// reflected source is never one this launch has already updated.  The eight
// coarse terms are summed in the reference's order c, c+j, c+k, c+jk, c+i,
// c+ij, c+ik, c+ijk, weight 3 per axis on the child's side and 1 otherwise.
//
// A thread owns the fine octet (2m .. 2m+1) on every axis (ghosted fine
// indices 0 .. n+1, m = 0 .. n/2): all eight read the same eight coarse
// values (coarse index m on each axis; a reflected ghost maps to the same
// m), loaded once.  Tiles: block b, coarse rows m_y = y0 .. y0+3 of the
// coarse plane m_z = z0; threads stride along m_x.  In FP64 the octet's
// two x columns are read and written as one 16-byte access where they are
// contiguous and aligned (256^3: 81 -> 68 us); FP32 keeps per-cell
// accesses, which measured faster there (27 against 35 us).
template <typename T>
__global__ __launch_bounds__(256) void k_prolong(const T *__restrict__ fin, T *__restrict__ fout,
                                                 const BlockLevel *__restrict__ blf,
                                                 const T *__restrict__ phic,
                                                 const BlockLevel *__restrict__ blc,
                                                 const Tile *__restrict__ tiles,
                                                 const int *__restrict__ done) {
    if (done && *done) return;
    const Tile t = tiles[blockIdx.x];
    const BlockLevel gf = blf[t.b];
    const int my = t.y0 + threadIdx.y, mz = t.z0;
    if (2 * my > gf.n1 + 1) return;
    const BlockLevel gc = blc[t.b];
    // fine rows / planes of the octet, their reflected sources and weights
    int sj[2], sk[2];
    T wj0[2], wj1[2], wk0[2], wk1[2];
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        const int j = 2 * my + d, k = 2 * mz + d;
        int a = j, b = k;
        if (j == 0 && (gf.neu & kNeuS)) a = 1;
        if (j == gf.n1 + 1 && (gf.neu & kNeuN)) a = gf.n1;
        if (k == 0 && (gf.neu & kNeuB)) b = 1;
        if (k == gf.n2 + 1 && (gf.neu & kNeuT)) b = gf.n2;
        sj[d] = a;
        sk[d] = b;
        wj0[d] = (a & 1) ? T(1) : T(3);
        wj1[d] = (a & 1) ? T(3) : T(1);
        wk0[d] = (b & 1) ? T(1) : T(3);
        wk1[d] = (b & 1) ? T(3) : T(1);
    }
    if constexpr (sizeof(T) == 4) {  // FP32: per cell (the paired form measured slower)
        for (int mx = threadIdx.x; 2 * mx <= gf.n0 + 1; mx += blockDim.x) {
            const long long c = gidx(gc, mx, my, mz);
            const T c000 = phic[c], c010 = phic[c + gc.sy], c001 = phic[c + gc.sz],
                    c011 = phic[c + gc.sy + gc.sz];
            const T c100 = phic[c + 1], c110 = phic[c + 1 + gc.sy], c101 = phic[c + 1 + gc.sz],
                    c111 = phic[c + 1 + gc.sy + gc.sz];
#pragma unroll
            for (int dz = 0; dz < 2; ++dz)
#pragma unroll
                for (int dy = 0; dy < 2; ++dy)
#pragma unroll
                    for (int dx = 0; dx < 2; ++dx) {
                        const int i = 2 * mx + dx;
                        int si = i;
                        if (i == 0 && (gf.neu & kNeuW)) si = 1;
                        if (i == gf.n0 + 1 && (gf.neu & kNeuE)) si = gf.n0;
                        const T wi0 = (si & 1) ? T(1) : T(3), wi1 = (si & 1) ? T(3) : T(1);
                        T s = (wi0 * wj0[dy] * wk0[dz]) * c000;
                        s = s + (wi0 * wj1[dy] * wk0[dz]) * c010;
                        s = s + (wi0 * wj0[dy] * wk1[dz]) * c001;
                        s = s + (wi0 * wj1[dy] * wk1[dz]) * c011;
                        s = s + (wi1 * wj0[dy] * wk0[dz]) * c100;
                        s = s + (wi1 * wj1[dy] * wk0[dz]) * c110;
                        s = s + (wi1 * wj0[dy] * wk1[dz]) * c101;
                        s = s + (wi1 * wj1[dy] * wk1[dz]) * c111;
                        fout[gidx(gf, i, 2 * my + dy, 2 * mz + dz)] = fin[gidx(gf, si, sj[dy], sk[dz])] + s / T(64);
                    }
        }
        return;
    }
    using T2 = T __attribute__((ext_vector_type(2)));
    for (int mx = threadIdx.x; 2 * mx <= gf.n0 + 1; mx += blockDim.x) {
        const long long c = gidx(gc, mx, my, mz);
        const T c000 = phic[c], c010 = phic[c + gc.sy], c001 = phic[c + gc.sz], c011 = phic[c + gc.sy + gc.sz];
        const T c100 = phic[c + 1], c110 = phic[c + 1 + gc.sy], c101 = phic[c + 1 + gc.sz],
                c111 = phic[c + 1 + gc.sy + gc.sz];
        // the octet's two columns i0 = 2 mx and i0 + 1 (a fine level's extent
        // is even, so both exist) and their reflected sources
        const int i0 = 2 * mx;
        const int si0 = (i0 == 0 && (gf.neu & kNeuW)) ? 1 : i0;
        const int si1 = (i0 + 1 == gf.n0 + 1 && (gf.neu & kNeuE)) ? gf.n0 : i0 + 1;
        const T wa0 = (si0 & 1) ? T(1) : T(3), wa1 = (si0 & 1) ? T(3) : T(1);
        const T wb0 = (si1 & 1) ? T(1) : T(3), wb1 = (si1 & 1) ? T(3) : T(1);
#pragma unroll
        for (int dz = 0; dz < 2; ++dz)
#pragma unroll
            for (int dy = 0; dy < 2; ++dy) {
                auto term = [&](T wi0, T wi1) {
                    T s = (wi0 * wj0[dy] * wk0[dz]) * c000;
                    s = s + (wi0 * wj1[dy] * wk0[dz]) * c010;
                    s = s + (wi0 * wj0[dy] * wk1[dz]) * c001;
                    s = s + (wi0 * wj1[dy] * wk1[dz]) * c011;
                    s = s + (wi1 * wj0[dy] * wk0[dz]) * c100;
                    s = s + (wi1 * wj1[dy] * wk0[dz]) * c110;
                    s = s + (wi1 * wj0[dy] * wk1[dz]) * c101;
                    s = s + (wi1 * wj1[dy] * wk1[dz]) * c111;
                    return s;
                };
                const T s0 = term(wa0, wa1), s1 = term(wb0, wb1);
                const long long q = gidx(gf, i0, 2 * my + dy, 2 * mz + dz);
                const long long r0 = gidx(gf, si0, sj[dy], sk[dz]), r1 = gidx(gf, si1, sj[dy], sk[dz]);
                // both columns in one 2-element access where they are
                // contiguous and aligned (no reflection in x, even offsets)
                if (r1 == r0 + 1 && !((q | r0) & 1)) {
                    const T2 a = *reinterpret_cast<const T2 *>(fin + r0);
                    T2 v;
                    v.x = a.x + s0 / T(64);
                    v.y = a.y + s1 / T(64);
                    *reinterpret_cast<T2 *>(fout + q) = v;
                } else {
                    fout[q] = fin[r0] + s0 / T(64);
                    fout[q + 1] = fin[r1] + s1 / T(64);
                }
            }
    }
}

// Correction of a block whose next level repeats its grid (mgsolver.for:
// 807-819): every cell, ghosts included, adds the coarse value of the same
// cell, fused with the Neumann copy that follows as in k_prolong (a ghost on
// a physical Neumann face takes the corrected value of the cell it
// reflects).  Reads fin, writes fout.  Tiles: 4 full ghosted rows of one
// ghosted plane of such blocks.
template <typename T>
__global__ __launch_bounds__(256) void k_prolong_rep(const T *__restrict__ fin, T *__restrict__ fout,
                                                     const BlockLevel *__restrict__ blf,
                                                     const T *__restrict__ phic,
                                                     const BlockLevel *__restrict__ blc,
                                                     const Tile *__restrict__ tiles,
                                                     const int *__restrict__ done) {
    if (done && *done) return;
    const Tile t = tiles[blockIdx.x];
    const BlockLevel gf = blf[t.b], gc = blc[t.b];
    const int j = t.y0 + threadIdx.y, k = t.z0;
    if (j > gf.n1 + 1) return;
    auto refl = [](int v, int n, bool lo, bool hi) { return (v == 0 && lo) ? 1 : ((v == n + 1 && hi) ? n : v); };
    const int sj = refl(j, gf.n1, gf.neu & kNeuS, gf.neu & kNeuN);
    const int sk = refl(k, gf.n2, gf.neu & kNeuB, gf.neu & kNeuT);
    for (int i = threadIdx.x; i <= gf.n0 + 1; i += blockDim.x) {
        const int si = refl(i, gf.n0, gf.neu & kNeuW, gf.neu & kNeuE);
        fout[gidx(gf, i, j, k)] = fin[gidx(gf, si, sj, sk)] + phic[gidx(gc, si, sj, sk)];
    }
}

}  // namespace mf3d
