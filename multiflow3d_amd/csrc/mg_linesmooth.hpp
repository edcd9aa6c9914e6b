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
//   one launch per diagonal u + v, every line of it one thread.  The
//   forward recurrence keeps c(m) in the line's own potential entries (no
//   other line reads them before the launch ends) and a(m) in a scratch
//   arena at the same offsets; the backward pass then overwrites c with
//   phi.
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

// One diagonal u + v = d of the lines along axis AX (0: i, 1: j, 2: k) of
// every block; grid: x over the lines of the diagonal in 64s, y = block.
// Line coordinates in loop order: AX = 1: (u, v) = (i, k); AX = 0:
// (u, v) = (j, k); AX = 2: (u, v) = (j, i).
template <typename T, bool VAR, int AX>
__global__ __launch_bounds__(64) void k_adi(T *__restrict__ p, T *__restrict__ sa, const T *__restrict__ rhs,
                                            const T *__restrict__ cx, const T *__restrict__ cy,
                                            const T *__restrict__ cz, const BlockLevel *__restrict__ bls, int d,
                                            const int *__restrict__ done) {
    if (done && *done) return;
    const BlockLevel g = bls[blockIdx.y];
    // extents of the line axis (m) and of the two line coordinates (u, v)
    const int nm = AX == 0 ? g.n0 : (AX == 1 ? g.n1 : g.n2);
    const int nu = AX == 1 ? g.n0 : g.n1;
    const int nv = AX == 2 ? g.n0 : g.n2;
    const int u = max(1, d - nv) + (int)(blockIdx.x * 64 + threadIdx.x);
    const int v = d - u;
    if (u > nu || u > d - 1 || v < 1) return;
    const int i0 = AX == 1 ? u : (AX == 0 ? 0 : v);
    const int j0 = AX == 1 ? 0 : u;
    const int k0 = AX == 2 ? 0 : v;
    const long long st = AX == 0 ? 1 : (AX == 1 ? g.sy : g.sz);  // stride along the line
    long long q = gidx(g, i0, j0, k0);                            // the ghost before the line
    T a_prev = T(0), c_prev = p[q];
    for (int m = 1; m <= nm; ++m) {
        q += st;
        const int i = AX == 0 ? m : i0, j = AX == 1 ? m : j0, k = AX == 2 ? m : k0;
        T aw, ae, as, an, ab, at, ap;
        coefs<T, VAR>(g, cx, cy, cz, q, i, j, k, aw, ae, as, an, ab, at, ap);
        T lo, hi, t;
        if (AX == 1) {  // mgsolver.for:450-458
            lo = as;
            hi = an;
            t = rhs[q] - ae * p[q + 1];
            t = t - aw * p[q - 1];
            t = t - at * p[q + g.sz];
            t = t - ab * p[q - g.sz];
        } else if (AX == 0) {  // :484-492
            lo = aw;
            hi = ae;
            t = rhs[q] - an * p[q + g.sy];
            t = t - as * p[q - g.sy];
            t = t - at * p[q + g.sz];
            t = t - ab * p[q - g.sz];
        } else {  // :520-528
            lo = ab;
            hi = at;
            t = rhs[q] - an * p[q + g.sy];
            t = t - as * p[q - g.sy];
            t = t - ae * p[q + 1];
            t = t - aw * p[q - 1];
        }
        const T apr = T(1) / (ap - lo * a_prev);
        a_prev = hi * apr;
        c_prev = (t - lo * c_prev) * apr;
        sa[q] = a_prev;
        p[q] = c_prev;
    }
    // q is the last cell; phi(m) = c(m) - a(m)*phi(m+1) from the ghost after the line
    T nxt = p[q + st];
    for (int m = nm; m >= 1; --m) {
        nxt = p[q] - sa[q] * nxt;
        p[q] = nxt;
        q -= st;
    }
}

}  // namespace mf3d
