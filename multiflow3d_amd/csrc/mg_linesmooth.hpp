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

// The Thomas coefficients of every line along AX: apr(m) =
// 1/(ap - lo*a(m-1)), a(m) = hi*apr(m), a(0) = 0 -- the reference's
// operations (mgsolver.for:447-449, 481-483, 517-519).  They depend on the
// operator only, so they are built once per coefficient build and the
// sweeps' recurrence keeps no division.  Grid: x over the block's lines in
// 256s, y = block.
template <typename T, bool VAR, int AX>
__global__ __launch_bounds__(256) void k_adi_coef(T *__restrict__ apr_out, T *__restrict__ a_out,
                                                  const T *__restrict__ cx, const T *__restrict__ cy,
                                                  const T *__restrict__ cz, const BlockLevel *__restrict__ bls) {
    const BlockLevel g = bls[blockIdx.y];
    const int nm = AX == 0 ? g.n0 : (AX == 1 ? g.n1 : g.n2);
    const int nu = AX == 1 ? g.n0 : g.n1;
    const int nv = AX == 2 ? g.n0 : g.n2;
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= nu * nv) return;
    const int u = t % nu + 1, v = t / nu + 1;
    const int i0 = AX == 1 ? u : (AX == 0 ? 0 : v), j0 = AX == 1 ? 0 : u, k0 = AX == 2 ? 0 : v;
    const long long st = AX == 0 ? 1 : (AX == 1 ? g.sy : g.sz);
    const long long q0 = gidx(g, i0, j0, k0);
    T a_prev = T(0);
    for (int m = 1; m <= nm; ++m) {
        const long long q = q0 + st * m;
        const int i = AX == 0 ? m : i0, j = AX == 1 ? m : j0, k = AX == 2 ? m : k0;
        T aw, ae, as, an, ab, at, ap;
        coefs<T, VAR>(g, cx, cy, cz, q, i, j, k, aw, ae, as, an, ab, at, ap);
        const T lo = AX == 0 ? aw : (AX == 1 ? as : ab), hi = AX == 0 ? ae : (AX == 1 ? an : at);
        const T apr = T(1) / (ap - lo * a_prev);
        a_prev = hi * apr;
        apr_out[q] = apr;
        a_out[q] = a_prev;
    }
}

// Pointers to a level's ADI line coefficients, apr and a per line axis.
template <typename T>
struct AdiCoef {
    const T *apr[3];
    const T *a[3];
};

// Geometry of one diagonal u + v = d of a pass over block g.
template <int AX>
struct AdiDiag {
    int nm, ulo, nl;
    long long st;
    __device__ AdiDiag(const BlockLevel &g, int d) {
        nm = AX == 0 ? g.n0 : (AX == 1 ? g.n1 : g.n2);
        const int nu = AX == 1 ? g.n0 : g.n1, nv = AX == 2 ? g.n0 : g.n2;
        ulo = max(1, d - nv);
        nl = max(0, min(nu, d - 1) - ulo + 1);
        st = AX == 0 ? 1 : (AX == 1 ? g.sy : g.sz);
    }
    // block-local offset of cell m of the diagonal's line ln (m = 0: the
    // ghost before the line), and its (i, j, k)
    __device__ long long cell(const BlockLevel &g, int d, int ln, int m, int &i, int &j, int &k) const {
        const int u = ulo + ln, v = d - u;
        const int i0 = AX == 1 ? u : (AX == 0 ? 0 : v), j0 = AX == 1 ? 0 : u, k0 = AX == 2 ? 0 : v;
        i = AX == 0 ? m : i0;
        j = AX == 1 ? m : j0;
        k = AX == 2 ? m : k0;
        return i + g.sy * j + g.sz * k;
    }
    // entry e = (m-1)*nl + ln of the diagonal's cell list -> (ln, m), without
    // an integer division: a float estimate (e < 2^20, nl <= 1024) corrected
    // by one step either way
    __device__ void split(int e, float rnl, int &ln, int &m) const {
        int q = (int)((float)e * rnl);
        q -= q * nl > e ? 1 : 0;
        q += (q + 1) * nl <= e ? 1 : 0;
        ln = e - q * nl;
        m = q + 1;
    }
};

// The apr and a of diagonal d's cells (entry e = (m-1)*nl + line), up to
// two per thread of a 256-thread workgroup: loaded into registers now,
// stored into LDS later, so the loads fly while the previous diagonal runs.
template <typename T, int AX>
struct AdiPrefetch {
    static constexpr int K = 4;  // entries per thread (the scratch holds at most 256*K)
    T ap[K], aa[K];
    __device__ void load(const BlockLevel &g, const AdiCoef<T> &ac, int d, int dmax) {
        if (d > dmax) return;
        const AdiDiag<AX> D(g, d);
        const float rnl = 1.0f / (float)max(D.nl, 1);
        const T *apr = ac.apr[AX] + g.off, *a = ac.a[AX] + g.off;
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const int e = (int)threadIdx.x + q * 256;
            if (e < D.nl * D.nm) {
                int i, j, k, ln, m;
                D.split(e, rnl, ln, m);
                const int l = (int)D.cell(g, d, ln, m, i, j, k);
                ap[q] = apr[l];
                aa[q] = a[l];
            }
        }
    }
    __device__ void store(const BlockLevel &g, int d, int dmax, T *APR, T *A) const {
        if (d > dmax) return;
        const AdiDiag<AX> D(g, d);
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const int e = (int)threadIdx.x + q * 256;
            if (e < D.nl * D.nm) {
                APR[e] = ap[q];
                A[e] = aa[q];
            }
        }
    }
};

// One pass (lines along AX) of a block in LDS.  Per diagonal:
//   1. every thread: the cross terms t of the diagonal's cells (into the
//      cells themselves in P -- no other line reads them before the
//      diagonal ends) and lo (into LO); and the loads of the next
//      diagonal's apr and a;
//   2. one lane per line: c(m) = (t - lo*c(m-1))*apr(m) into P;
//   3. the same lane: phi(m) = c(m) - a(m)*phi(m+1), backwards;
//   then the next diagonal's apr and a go to the other LDS buffer.
template <typename T, bool VAR, int AX, bool RLDS>
__device__ __forceinline__ void adi_pass_lds(const BlockLevel &g, T *P, const T *R, T *LO, T *APR0, T *A0,
                                             T *APR1, T *A1, const AdiCoef<T> &ac, const T *__restrict__ cx,
                                             const T *__restrict__ cy, const T *__restrict__ cz) {
    const int nu = AX == 1 ? g.n0 : g.n1, nv = AX == 2 ? g.n0 : g.n2;
    const int dmax = nu + nv;
    const long long rb = RLDS ? 0 : g.off;
    const int tid = (int)threadIdx.x;
    AdiPrefetch<T, AX> pf;
    pf.load(g, ac, 2, dmax);
    pf.store(g, 2, dmax, APR0, A0);
    __syncthreads();
    for (int d = 2; d <= dmax; ++d) {
        T *APR = (d & 1) ? APR1 : APR0, *A = (d & 1) ? A1 : A0;
        const AdiDiag<AX> D(g, d);
        const float rnl = 1.0f / (float)max(D.nl, 1);
        for (int e = tid; e < D.nl * D.nm; e += 256) {  // phase 1
            int i, j, k, ln, m;
            D.split(e, rnl, ln, m);
            const int l = (int)D.cell(g, d, ln, m, i, j, k);
            T t, lo, hi, ap;
            adi_terms<T, VAR, AX>(g, P, l, g.sy, g.sz, R, rb + l, cx, cy, cz, g.off + l, i, j, k, t, lo, hi, ap);
            P[l] = t;
            LO[e] = lo;
        }
        pf.load(g, ac, d + 1, dmax);
        __syncthreads();
        constexpr int U = 8;  // full chunks run without branches, the remainder one cell at a time
        for (int ln = tid; ln < D.nl; ln += 256) {  // phases 2 and 3
            int i, j, k;
            const int l0 = (int)D.cell(g, d, ln, 0, i, j, k);  // LDS offsets fit in 32 bits
            const int st = (int)D.st, nm = D.nm, nl = D.nl;
            T c_prev = P[l0];
            int m = 1;
            for (; m + U - 1 <= nm; m += U) {
                const int l = l0 + st * m, e = (m - 1) * nl + ln;
                T lo[U], ap[U], tt[U];
#pragma unroll
                for (int q = 0; q < U; ++q) {
                    lo[q] = LO[e + q * nl];
                    ap[q] = APR[e + q * nl];
                    tt[q] = P[l + q * st];
                }
#pragma unroll
                for (int q = 0; q < U; ++q) {
                    c_prev = (tt[q] - lo[q] * c_prev) * ap[q];
                    P[l + q * st] = c_prev;
                }
            }
            for (; m <= nm; ++m) {
                const int e = (m - 1) * nl + ln;
                c_prev = (P[l0 + st * m] - LO[e] * c_prev) * APR[e];
                P[l0 + st * m] = c_prev;
            }
            T nxt = P[l0 + st * (nm + 1)];
            m = nm;
            for (; m - U + 1 >= 1; m -= U) {
                const int l = l0 + st * m, e = (m - 1) * nl + ln;
                T cc[U], aa[U];
#pragma unroll
                for (int q = 0; q < U; ++q) {
                    cc[q] = P[l - q * st];
                    aa[q] = A[e - q * nl];
                }
#pragma unroll
                for (int q = 0; q < U; ++q) {
                    nxt = cc[q] - aa[q] * nxt;
                    P[l - q * st] = nxt;
                }
            }
            for (; m >= 1; --m) {
                nxt = P[l0 + st * m] - A[(m - 1) * nl + ln] * nxt;
                P[l0 + st * m] = nxt;
            }
        }
        pf.store(g, d + 1, dmax, (d & 1) ? APR0 : APR1, (d & 1) ? A0 : A1);
        __syncthreads();
    }
}

// A whole ADI sweep (j, i, then k lines) of one block per 256-thread
// workgroup with the block in LDS: its ghosted potential, with RLDS its
// RHS, and per-diagonal scratch of 5 arrays of sstride entries (lo, and
// two buffers of apr and a); loaded once, written back once.  Dynamic
// LDS: (cells + (RLDS ? cells : 0) + 5*sstride) entries, cells = the
// largest block's ghosted box; sstride <= 1024 (the prefetch holds 4
// entries per thread).
template <typename T, bool VAR, bool RLDS>
__global__ __launch_bounds__(256) void k_adi_lds(T *__restrict__ p, const T *__restrict__ rhs,
                                                 const T *__restrict__ cx, const T *__restrict__ cy,
                                                 const T *__restrict__ cz, AdiCoef<T> ac,
                                                 const BlockLevel *__restrict__ bls, int cells_max, int sstride,
                                                 const int *__restrict__ done) {
    if (done && *done) return;
    extern __shared__ __align__(16) unsigned char adi_smem[];
    T *P = reinterpret_cast<T *>(adi_smem);
    T *R = P + cells_max;
    T *LO = RLDS ? R + cells_max : R;
    T *APR0 = LO + sstride, *A0 = APR0 + sstride, *APR1 = A0 + sstride, *A1 = APR1 + sstride;
    const BlockLevel g = bls[blockIdx.x];
    const int cells = (int)(g.sz * (g.n2 + 2));
    for (int c = (int)threadIdx.x; c < cells; c += (int)blockDim.x) {
        P[c] = p[g.off + c];
        if (RLDS) R[c] = rhs[g.off + c];
    }
    __syncthreads();
    const T *Rs = RLDS ? (const T *)R : rhs;
    adi_pass_lds<T, VAR, 1, RLDS>(g, P, Rs, LO, APR0, A0, APR1, A1, ac, cx, cy, cz);
    adi_pass_lds<T, VAR, 0, RLDS>(g, P, Rs, LO, APR0, A0, APR1, A1, ac, cx, cy, cz);
    adi_pass_lds<T, VAR, 2, RLDS>(g, P, Rs, LO, APR0, A0, APR1, A1, ac, cx, cy, cz);
    for (int c = (int)threadIdx.x; c < cells; c += (int)blockDim.x) p[g.off + c] = P[c];
}

}  // namespace mf3d
