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

// Trilinear prolongation added to every fine point including the ghost
// layer (mgsolver.for:823-891), fused with the Neumann copy that follows it
// (mgbound, mgsolver.for:902): a ghost on a physical Neumann face takes the
// corrected value of the cell it reflects.  Reads fin, writes fout, so the
// reflected source is never one this launch has already updated.  The eight
// coarse terms are summed in the reference's order c, c+j, c+k, c+jk, c+i,
// c+ij, c+ik, c+ijk, weight 3 per axis on the child's side and 1 otherwise.
// Row tiles: block b, 4 rows from y0, plane z0; threads stride along x.
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
    const int j = t.y0 + threadIdx.y, k = t.z0;
    if (j > gf.n1 + 1) return;
    const BlockLevel gc = blc[t.b];
    int sj = j, sk = k;
    if (j == 0 && (gf.neu & kNeuS)) sj = 1;
    if (j == gf.n1 + 1 && (gf.neu & kNeuN)) sj = gf.n1;
    if (k == 0 && (gf.neu & kNeuB)) sk = 1;
    if (k == gf.n2 + 1 && (gf.neu & kNeuT)) sk = gf.n2;
    const int jc = sj >> 1, kc = sk >> 1, fj = sj & 1, fk = sk & 1;
    const T wj0 = fj ? T(1) : T(3), wj1 = fj ? T(3) : T(1);
    const T wk0 = fk ? T(1) : T(3), wk1 = fk ? T(3) : T(1);
    for (int i = threadIdx.x; i <= gf.n0 + 1; i += blockDim.x) {
        int si = i;
        if (i == 0 && (gf.neu & kNeuW)) si = 1;
        if (i == gf.n0 + 1 && (gf.neu & kNeuE)) si = gf.n0;
        const int ic = si >> 1, fi = si & 1;
        const T wi0 = fi ? T(1) : T(3), wi1 = fi ? T(3) : T(1);
        const long long c = gidx(gc, ic, jc, kc);
        T s = (wi0 * wj0 * wk0) * phic[c];
        s = s + (wi0 * wj1 * wk0) * phic[c + gc.sy];
        s = s + (wi0 * wj0 * wk1) * phic[c + gc.sz];
        s = s + (wi0 * wj1 * wk1) * phic[c + gc.sy + gc.sz];
        s = s + (wi1 * wj0 * wk0) * phic[c + 1];
        s = s + (wi1 * wj1 * wk0) * phic[c + 1 + gc.sy];
        s = s + (wi1 * wj0 * wk1) * phic[c + 1 + gc.sz];
        s = s + (wi1 * wj1 * wk1) * phic[c + 1 + gc.sy + gc.sz];
        fout[gidx(gf, i, j, k)] = fin[gidx(gf, si, sj, sk)] + s / T(64);
    }
}

}  // namespace mf3d
