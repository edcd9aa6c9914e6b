// mg_linesmooth.hpp -- the two other smoothers of mgrelax
// (mgsolver.for:403-545): lexicographic Gauss-Seidel (mg_itrsch = 1) and
// ADI line relaxation with the Thomas algorithm (mg_itrsch = 2).
//
// Both update the level's potential in place, in an order whose every
// read sees what the reference's sequential loop would see, so the FP64
// results are the reference's bitwise:
//
// * Lexicographic GS runs its loops k, j, i (i innermost), so cell (i,j,k)
//   reads this sweep's W, S and B neighbours and the previous E, N and T.
//   All cells with i + j + k = s depend only on cells of s - 1 and s + 1:
//   one launch per hyperplane s, in increasing s, every cell of the
//   hyperplane (of every block) one thread.
// * ADI solves j lines in the order (k, i), then i lines in the order
//   (k, j), then k lines in the order (i, j).  A line reads the cross
//   neighbours of the lines before it in its pass (new) and after it
//   (old): with lines indexed (u, v) in loop order, line (u, v) needs
//   (u-1, v) and (u, v-1) finished and (u+1, v), (u, v+1) untouched --
//   the diagonals u + v in increasing order, the lines of a diagonal in
//   parallel.  Blocks are independent within a sweep (a block reads only
//   its own ghosts), so one workgroup per block walks the diagonals of a
//   pass with a barrier between them.  Where the level's blocks fit in
//   LDS, one launch runs the whole sweep there, the RHS staged in LDS too
//   when it fits.  The forward recurrence keeps c(m) in the line's own
//   potential entries (no other line reads them before the diagonal ends)
//   and a(m) in a scratch; the backward pass overwrites c with phi.
//
// The per-cell arithmetic is the reference's: GS as the red-black sweep
// (accumulated aw*W + ae*E + as*S + an*N + ab*B + at*T, then
// (rhs - s)/ap, or times 1/ap in the FP32 inner cycle); TDMA as
// apr = 1/(ap - a_lo*a(m-1)), a(m) = a_hi*apr,
// c(m) = (rhs - x1 - x2 - x3 - x4 - a_lo*c(m-1))*apr with the cross terms
// in the reference's order per pass, phi(m) = c(m) - a(m)*phi(m+1).
#pragma once

#include <hip/hip_runtime.h>

#include "mg_common.hpp"

namespace mf3d {

// One hyperplane i + j + k = s of every block (grid: x over j, y over k in
// 16 x 16 tiles, z = block).
template <typename T, bool VAR, bool RCP>
__global__ __launch_bounds__(256) void k_lex(T *__restrict__ p, const T *__restrict__ rhs,
                                             const T *__restrict__ cx, const T *__restrict__ cy,
                                             const T *__restrict__ cz, const BlockLevel *__restrict__ bls, int s,
                                             const int *__restrict__ done) {
    if (done && *done) return;
    const BlockLevel g = bls[blockIdx.z];
    const int j = blockIdx.x * 16 + (int)(threadIdx.x & 15) + 1;
    const int k = blockIdx.y * 16 + (int)(threadIdx.x >> 4) + 1;
    const int i = s - j - k;
    if (j > g.n1 || k > g.n2 || i < 1 || i > g.n0) return;
    const long long q = gidx(g, i, j, k);
    T aw, ae, as, an, ab, at, ap;
    coefs<T, VAR>(g, cx, cy, cz, q, i, j, k, aw, ae, as, an, ab, at, ap);
    T t = aw * p[q - 1];
    t = t + ae * p[q + 1];
    t = t + as * p[q - g.sy];
    t = t + an * p[q + g.sy];
    t = t + ab * p[q - g.sz];
    t = t + at * p[q + g.sz];
    const T r = rhs[q];
    p[q] = RCP ? (r - t) * (T(1) / ap) : (r - t) / ap;
}

// The cross terms and coefficients of cell (i,j,k) of a line along axis AX:
// t = rhs - the four cross terms in the reference's order for that pass
// (mgsolver.for:450-458, 484-492, 520-528), lo / hi the coefficients
// towards the previous / next cell of the line, ap the diagonal.  The
// potential is read through P at offset lp and the RHS through R at
// offset lr (global memory or LDS); the operator at arena offset q.
template <typename T, bool VAR, int AX>
__device__ __forceinline__ void adi_terms(const BlockLevel &g, const T *P, long long lp, long long sy, long long sz,
                                          const T *R, long long lr, const T *__restrict__ cx,
                                          const T *__restrict__ cy, const T *__restrict__ cz, long long q, int i,
                                          int j, int k, T &t, T &lo, T &hi, T &ap) {
    T aw, ae, as, an, ab, at;
    coefs<T, VAR>(g, cx, cy, cz, q, i, j, k, aw, ae, as, an, ab, at, ap);
    if (AX == 1) {
        lo = as;
        hi = an;
        t = R[lr] - ae * P[lp + 1];
        t = t - aw * P[lp - 1];
        t = t - at * P[lp + sz];
        t = t - ab * P[lp - sz];
    } else if (AX == 0) {
        lo = aw;
        hi = ae;
        t = R[lr] - an * P[lp + sy];
        t = t - as * P[lp - sy];
        t = t - at * P[lp + sz];
        t = t - ab * P[lp - sz];
    } else {
        lo = ab;
        hi = at;
        t = R[lr] - an * P[lp + sy];
        t = t - as * P[lp - sy];
        t = t - ae * P[lp + 1];
        t = t - aw * P[lp - 1];
    }
}

// The lines of one diagonal u + v = d of a pass (lines along AX) over
// block g.  LDS: P is the block's ghosted potential in LDS and A a
// per-thread LDS scratch for a(m) (entry m*blockDim + thread); otherwise P
// is the level arena and A a level-sized scratch at the cells' offsets.
// RLDS: R is the block's RHS in LDS, else the arena.  Each line runs in
// chunks of U cells whose loads are all issued before that part of the
// recurrence.
template <typename T, bool VAR, int AX, bool LDS, bool RLDS>
__device__ __forceinline__ void adi_diagonal(const BlockLevel &g, T *P, const T *R, T *A, const T *__restrict__ cx,
                                             const T *__restrict__ cy, const T *__restrict__ cz, int d) {
    constexpr int U = VAR ? 4 : 8;
    const int nm = AX == 0 ? g.n0 : (AX == 1 ? g.n1 : g.n2);
    const int nu = AX == 1 ? g.n0 : g.n1;
    const int nv = AX == 2 ? g.n0 : g.n2;
    const long long sy = g.sy, sz = g.sz;
    const long long st = AX == 0 ? 1 : (AX == 1 ? sy : sz);  // stride along the line
    const long long pb = LDS ? 0 : g.off, rb = RLDS ? 0 : g.off;
    const int nt = (int)blockDim.x, tid = (int)threadIdx.x;
    const int ulo = max(1, d - nv), uhi = min(nu, d - 1);
    for (int u = ulo + tid; u <= uhi; u += nt) {
        const int v = d - u;
        const int i0 = AX == 1 ? u : (AX == 0 ? 0 : v);
        const int j0 = AX == 1 ? 0 : u;
        const int k0 = AX == 2 ? 0 : v;
        const long long l0 = i0 + sy * j0 + sz * k0;  // the ghost before the line, block-local
        auto aix = [&](int m) -> long long { return LDS ? (long long)m * nt + tid : g.off + l0 + st * m; };
        T a_prev = T(0), c_prev = P[pb + l0];
        for (int m0 = 1; m0 <= nm; m0 += U) {  // forward: a(m) into A, c(m) into the line's own cells
            T tt[U], lo[U], hi[U], ap[U];
#pragma unroll
            for (int e = 0; e < U; ++e) {
                const int m = min(m0 + e, nm);  // past the end: a repeat of the last cell, not used
                const int i = AX == 0 ? m : i0, j = AX == 1 ? m : j0, k = AX == 2 ? m : k0;
                const long long l = l0 + st * m;
                adi_terms<T, VAR, AX>(g, P, pb + l, sy, sz, R, rb + l, cx, cy, cz, g.off + l, i, j, k, tt[e], lo[e],
                                      hi[e], ap[e]);
            }
#pragma unroll
            for (int e = 0; e < U; ++e) {
                if (m0 + e > nm) break;
                const T apr = T(1) / (ap[e] - lo[e] * a_prev);
                a_prev = hi[e] * apr;
                c_prev = (tt[e] - lo[e] * c_prev) * apr;
                A[aix(m0 + e)] = a_prev;
                P[pb + l0 + st * (m0 + e)] = c_prev;
            }
        }
        T nxt = P[pb + l0 + st * (nm + 1)];  // the ghost after the line
        for (int m1 = nm; m1 >= 1; m1 -= U) {  // backward: phi(m) = c(m) - a(m)*phi(m+1)
            T cc[U], aa[U];
#pragma unroll
            for (int e = 0; e < U; ++e) {
                const int m = max(m1 - e, 1);
                cc[e] = P[pb + l0 + st * m];
                aa[e] = A[aix(m)];
            }
#pragma unroll
            for (int e = 0; e < U; ++e) {
                if (m1 - e < 1) break;
                nxt = cc[e] - aa[e] * nxt;
                P[pb + l0 + st * (m1 - e)] = nxt;
            }
        }
    }
}

// One pass of ADI (lines along AX) on a level in global memory, one
// workgroup per block walking the diagonals with a barrier between them;
// sa is a level-sized scratch for a(m).
template <typename T, bool VAR, int AX>
__global__ __launch_bounds__(256) void k_adi(T *__restrict__ p, T *__restrict__ sa, const T *__restrict__ rhs,
                                             const T *__restrict__ cx, const T *__restrict__ cy,
                                             const T *__restrict__ cz, const BlockLevel *__restrict__ bls,
                                             const int *__restrict__ done) {
    if (done && *done) return;
    const BlockLevel g = bls[blockIdx.x];
    const int nu = AX == 1 ? g.n0 : g.n1;
    const int nv = AX == 2 ? g.n0 : g.n2;
    for (int d = 2; d <= nu + nv; ++d) {
        adi_diagonal<T, VAR, AX, false, false>(g, p, rhs, sa, cx, cy, cz, d);
        __syncthreads();
    }
}

// One diagonal of a pass in LDS, in three phases:
//   1. every thread: for the cells of all lines of the diagonal, the cross
//      terms t (into the cell itself in P -- no other line reads it before
//      the diagonal ends) and the coefficients lo, hi, ap (into the
//      scratch S, entry (m, line) at m*nl + line);
//   2. one lane per line: the forward recurrence from LDS, c(m) into P and
//      a(m) into S's fourth array;
//   3. the same lane: phi(m) = c(m) - a(m)*phi(m+1), backwards.
// Phase 1 is parallel over all cells, so the sequential part of a line
// runs only the recurrence's own operations.
template <typename T, bool VAR, int AX, bool RLDS>
__device__ __forceinline__ void adi_diagonal_lds(const BlockLevel &g, T *P, const T *R, T *S, int sstride,
                                                 const T *__restrict__ cx, const T *__restrict__ cy,
                                                 const T *__restrict__ cz, int d) {
    const int nm = AX == 0 ? g.n0 : (AX == 1 ? g.n1 : g.n2);
    const int nu = AX == 1 ? g.n0 : g.n1;
    const int nv = AX == 2 ? g.n0 : g.n2;
    const long long sy = g.sy, sz = g.sz;
    const long long st = AX == 0 ? 1 : (AX == 1 ? sy : sz);
    const long long rb = RLDS ? 0 : g.off;
    const int ulo = max(1, d - nv), uhi = min(nu, d - 1);
    const int nl = uhi - ulo + 1;
    const int nt = (int)blockDim.x, tid = (int)threadIdx.x;
    T *LO = S, *HI = S + sstride, *AP = S + 2 * sstride, *A = S + 3 * sstride;
    auto line0 = [&](int u) -> long long {  // the ghost before line u, block-local
        const int v = d - u;
        const int i0 = AX == 1 ? u : (AX == 0 ? 0 : v), j0 = AX == 1 ? 0 : u, k0 = AX == 2 ? 0 : v;
        return i0 + sy * j0 + sz * k0;
    };
    for (int e = tid; e < nl * nm; e += nt) {  // phase 1
        const int ln = e % nl, m = e / nl + 1;
        const int u = ulo + ln, v = d - u;
        const int i0 = AX == 1 ? u : (AX == 0 ? 0 : v), j0 = AX == 1 ? 0 : u, k0 = AX == 2 ? 0 : v;
        const int i = AX == 0 ? m : i0, j = AX == 1 ? m : j0, k = AX == 2 ? m : k0;
        const long long l = i + sy * j + sz * k;
        T t, lo, hi, ap;
        adi_terms<T, VAR, AX>(g, P, l, sy, sz, R, rb + l, cx, cy, cz, g.off + l, i, j, k, t, lo, hi, ap);
        P[l] = t;
        LO[e] = lo;
        HI[e] = hi;
        AP[e] = ap;
    }
    __syncthreads();
    constexpr int U = 8;  // LDS reads of a chunk issued before its recurrence steps
    for (int ln = tid; ln < nl; ln += nt) {  // phases 2 and 3
        const long long l0 = line0(ulo + ln);
        T a_prev = T(0), c_prev = P[l0];
        for (int m0 = 1; m0 <= nm; m0 += U) {
            T lo[U], hi[U], ap[U], tt[U];
#pragma unroll
            for (int q = 0; q < U; ++q) {
                const int m = min(m0 + q, nm);
                const int e = (m - 1) * nl + ln;
                lo[q] = LO[e];
                hi[q] = HI[e];
                ap[q] = AP[e];
                tt[q] = P[l0 + st * m];
            }
#pragma unroll
            for (int q = 0; q < U; ++q) {
                if (m0 + q > nm) break;
                const T apr = T(1) / (ap[q] - lo[q] * a_prev);
                a_prev = hi[q] * apr;
                c_prev = (tt[q] - lo[q] * c_prev) * apr;
                P[l0 + st * (m0 + q)] = c_prev;
                A[(m0 + q - 1) * nl + ln] = a_prev;
            }
        }
        T nxt = P[l0 + st * (nm + 1)];
        for (int m1 = nm; m1 >= 1; m1 -= U) {
            T cc[U], aa[U];
#pragma unroll
            for (int q = 0; q < U; ++q) {
                const int m = max(m1 - q, 1);
                cc[q] = P[l0 + st * m];
                aa[q] = A[(m - 1) * nl + ln];
            }
#pragma unroll
            for (int q = 0; q < U; ++q) {
                if (m1 - q < 1) break;
                nxt = cc[q] - aa[q] * nxt;
                P[l0 + st * (m1 - q)] = nxt;
            }
        }
    }
}

// A whole ADI sweep (j, i, then k lines) of one block per workgroup with
// the block in LDS: its ghosted potential, with RLDS its RHS, and the
// per-diagonal scratch S of 4 arrays of sstride entries (the most cells a
// diagonal holds); loaded once, written back once.  Dynamic LDS:
// (cells + (RLDS ? cells : 0) + 4*sstride) entries, cells = the largest
// block's ghosted box.
template <typename T, bool VAR, bool RLDS>
__global__ __launch_bounds__(256) void k_adi_lds(T *__restrict__ p, const T *__restrict__ rhs,
                                                 const T *__restrict__ cx, const T *__restrict__ cy,
                                                 const T *__restrict__ cz, const BlockLevel *__restrict__ bls,
                                                 int cells_max, int sstride, const int *__restrict__ done) {
    if (done && *done) return;
    extern __shared__ __align__(16) unsigned char adi_smem[];
    T *P = reinterpret_cast<T *>(adi_smem);
    T *R = P + cells_max;
    T *S = RLDS ? R + cells_max : R;
    const BlockLevel g = bls[blockIdx.x];
    const int cells = (int)(g.sz * (g.n2 + 2));
    for (int c = (int)threadIdx.x; c < cells; c += (int)blockDim.x) {
        P[c] = p[g.off + c];
        if (RLDS) R[c] = rhs[g.off + c];
    }
    __syncthreads();
    const T *Rs = RLDS ? (const T *)R : rhs;
    for (int d = 2; d <= g.n0 + g.n2; ++d) {
        adi_diagonal_lds<T, VAR, 1, RLDS>(g, P, Rs, S, sstride, cx, cy, cz, d);
        __syncthreads();
    }
    for (int d = 2; d <= g.n1 + g.n2; ++d) {
        adi_diagonal_lds<T, VAR, 0, RLDS>(g, P, Rs, S, sstride, cx, cy, cz, d);
        __syncthreads();
    }
    for (int d = 2; d <= g.n1 + g.n0; ++d) {
        adi_diagonal_lds<T, VAR, 2, RLDS>(g, P, Rs, S, sstride, cx, cy, cz, d);
        __syncthreads();
    }
    for (int c = (int)threadIdx.x; c < cells; c += (int)blockDim.x) p[g.off + c] = P[c];
}

}  // namespace mf3d
