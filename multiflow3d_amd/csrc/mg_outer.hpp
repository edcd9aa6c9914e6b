// mg_outer.hpp -- FP64 outer step of the defect-correction loop and the
// residual norms.
#pragma once
#include "mg_common.hpp"
#include "mg_sweep.hpp"

namespace mf3d {

// FP64 outer step (newsolv_mg.for:38-54 in Poisson form):
// p_new = p_old + e over every ghosted cell (both operands carry mapped
// ghosts, so p_new does too), and on interior cells r = f - A p_new.  r
// becomes the next cycle's level-0 RHS (rounded to the inner precision TE);
// per-workgroup partial sums of r^2 and max |r| dV go to `partials`.
//
// Same tiling as the sweep: 64 x TY columns marching over planes, with
// pn = p_old + e formed once per cell into a 4-plane LDS ring (1-cell halo
// in x and y).  Each workgroup writes p_new for its own cells plus the ghost
// cells next to them on the block boundary.  Plane c+2 of p_old and e and
// the next plane's f are loaded while plane c is computed.
//
// ADD = false (refined block grids, whose 2:1 ghosts are not the sum of the
// ghosts of p_old and e): pold already holds p_new with its ghosts; only r
// and the partials are formed (e and pnew unused).
template <typename TE, bool VAR, int TY, bool ADD = true>
__global__ __launch_bounds__(SW_TX *TY) void k_outer_lds(
    const double *__restrict__ pold, double *__restrict__ pnew, const TE *__restrict__ e,
    const double *__restrict__ f, TE *__restrict__ rhs0, const double *__restrict__ cx,
    const double *__restrict__ cy, const double *__restrict__ cz,
    const BlockLevel *__restrict__ bls, const Tile *__restrict__ tiles,
    double *__restrict__ partials, const int *__restrict__ done) {
    if (done && *done) return;
    constexpr int LX = SW_TX + 2, LY = TY + 2, NS = 4, NT = SW_TX * TY, LXY = LX * LY;
    constexpr int KL = (LXY + NT - 1) / NT;
    __shared__ double sp[NS][LXY];
    const Tile t = tiles[blockIdx.x];
    const BlockLevel g = bls[t.b];
    const int tid = threadIdx.x + SW_TX * threadIdx.y;
    const int xb = t.x0 - 1, yb = t.y0 - 1;
    const int xhi = t.x0 + SW_TX - 1 < g.n0 ? t.x0 + SW_TX - 1 : g.n0;
    const int yhi = t.y0 + TY - 1 < g.n1 ? t.y0 + TY - 1 : g.n1;
    const int wx0 = t.x0 == 1 ? 0 : t.x0, wx1 = xhi == g.n0 ? g.n0 + 1 : xhi;
    const int wy0 = t.y0 == 1 ? 0 : t.y0, wy1 = yhi == g.n1 ? g.n1 + 1 : yhi;
    const int wz0 = t.z0 == 1 ? 0 : t.z0, wz1 = t.z1 == g.n2 ? g.n2 + 1 : t.z1;

    long long lofs[KL];
    bool lok[KL], lown[KL];
#pragma unroll
    for (int k = 0; k < KL; ++k) {
        const int p = tid + k * NT;
        const int ly = p / LX, lx = p - ly * LX;
        const int x = xb + lx, y = yb + ly;
        lok[k] = p < LXY && x >= 0 && x <= g.n0 + 1 && y >= 0 && y <= g.n1 + 1;
        lown[k] = lok[k] && x >= wx0 && x <= wx1 && y >= wy0 && y <= wy1;
        lofs[k] = g.off + (long long)x + g.sy * y;
    }
    // unconditional loads from clamped addresses, masked afterwards, in a
    // 3-deep pipeline of register sets (see k_rbs)
    long long lsrc[KL];
#pragma unroll
    for (int k = 0; k < KL; ++k) lsrc[k] = lok[k] ? lofs[k] : g.off;
    const int bx = t.x0 + threadIdx.x, by = t.y0 + threadIdx.y;
    const bool has_col = bx <= g.n0 && by <= g.n1;
    struct Regs {
        double pp[KL];
        TE pe[KL];
        double ff;
    };
    auto load_regs = [&](Regs &R, int c) {  // plane c+1 of p_old and e, f on plane c
        const long long zo = g.sz * (long long)(c + 1 < t.z1 + 1 ? c + 1 : t.z1 + 1);
#pragma unroll
        for (int k = 0; k < KL; ++k) {
            const double a = pold[lsrc[k] + zo];
            R.pp[k] = lok[k] ? a : 0.0;
            if constexpr (ADD) {
                const TE b = e[lsrc[k] + zo];
                R.pe[k] = lok[k] ? b : TE(0);
            }
        }
        const bool ok = has_col && c >= t.z0 && c <= t.z1;
        const double v = f[ok ? gidx(g, bx, by, c) : g.off];
        R.ff = ok ? v : 0.0;
    };
    auto store_plane = [&](int c, const double *pp, const TE *pe) {
        double *P = sp[c % NS];
        const bool zown = c >= wz0 && c <= wz1;
#pragma unroll
        for (int k = 0; k < KL; ++k) {
            if (tid + k * NT >= LXY) continue;
            if constexpr (ADD) {
                const double v = pp[k] + (double)pe[k];
                if (zown && lown[k]) pnew[lofs[k] + g.sz * c] = v;
                P[tid + k * NT] = v;
            } else {
                P[tid + k * NT] = pp[k];
            }
        }
    };
    double r2 = 0.0, ra = 0.0;
    {
        Regs R;
        load_regs(R, t.z0 - 2);
        store_plane(t.z0 - 1, R.pp, R.pe);
        load_regs(R, t.z0 - 1);
        store_plane(t.z0, R.pp, R.pe);
    }
    auto step = [&](int c, Regs &R) {
        store_plane(c + 1, R.pp, R.pe);
        const double f_cur = R.ff;
        load_regs(R, c + 3);
        __syncthreads();
        if (has_col) {
            const double *P = sp[c % NS];
            const double *Pm = sp[(c + NS - 1) % NS];
            const double *Pp = sp[(c + 1) % NS];
            const int l = (bx - xb) + LX * (by - yb);
            const long long q = gidx(g, bx, by, c);
            double aw, ae, as, an, ab, at, ap;
            coefs<double, VAR>(g, cx, cy, cz, q, bx, by, c, aw, ae, as, an, ab, at, ap);
            double s = aw * P[l - 1];
            s = s + ae * P[l + 1];
            s = s + as * P[l - LX];
            s = s + an * P[l + LX];
            s = s + ab * Pm[l];
            s = s + at * Pp[l];
            s = s + ap * P[l];
            const double r = f_cur - s;
            rhs0[q] = (TE)r;
            r2 = r2 + r * r;
            ra = fmax(ra, g.lsm ? fabs(r) : fabs(r * g.h0 * g.h1 * g.h2));
        }
    };
    Regs R0, R1, R2;
    load_regs(R0, t.z0);
    load_regs(R1, t.z0 + 1);
    load_regs(R2, t.z0 + 2);
    int c = t.z0;
    for (; c + 2 <= t.z1; c += 3) {
        step(c, R0);
        step(c + 1, R1);
        step(c + 2, R2);
    }
    if (c <= t.z1) step(c, R0);
    if (c + 1 <= t.z1) step(c + 1, R1);
    block_sum_max<SW_TX * TY>(r2, ra, partials + 2 * (long long)blockIdx.x);
}

// p_new = p_old + e on every cell of the level-0 arena (a refined grid's
// outer step; its ghosts are then refilled from p_new by the exchange).
template <typename TE>
__global__ __launch_bounds__(256) void k_padd(double *__restrict__ pnew, const double *__restrict__ pold,
                                              const TE *__restrict__ e, long long n, const int *__restrict__ done) {
    if (done && *done) return;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        pnew[i] = pold[i] + (double)e[i];
}

// ---------------------------------------------------------------------------
// GCR(1) outer iteration (krylov = MF3D_OUTER_GCR; no counterpart in the
// reference, whose outer loop is the defect correction above).  With
// z_k = V(r_k), the V-cycle from zero on the residual:
//   beta_k = (A z_k, w_{k-1}) / (w_{k-1}, w_{k-1})      (beta_1 = 0)
//   d_k = z_k - beta_k d_{k-1},   w_k = A d_k
//   alpha_k = (r_k, w_k) / (w_k, w_k)
//   p_{k+1} = p_k + alpha_k d_k,  r_{k+1} = f - A p_{k+1}  (true residual)
// The residual-minimising direction update needs no symmetry (the
// reference's variable-coefficient operator is not symmetric across block
// faces, where each block takes its own cell's density) and tolerates a
// V-cycle that changes from one step to the next.  All of it in FP64; d and
// p carry mapped ghosts like the V-cycle output.
//
// k_krylov_lds is the z-marching LDS pass of the outer step with a plane
// value v formed per cell:
//   MODE 0:  v = p + alpha d -> p_new (ghosts included), r = f - A v, written
//            in FP64 and in the inner precision; partials (r^2, max|r| dV)
//   MODE 1:  v = z - beta d_prev -> d (ghosts included), w = A v on interior
//            cells; partials (w . w, r . w)
//   MODE 2:  v = z; partials (A v . w_prev, 0)
//   MODE 3:  v = p_new as given (a refined grid's, its 2:1 ghosts from the
//            exchange after k_paxpy); r = f - A v and partials as MODE 0
template <int MODE, typename TA, typename TR, bool VAR, int TY>
__global__ __launch_bounds__(SW_TX *TY) void k_krylov_lds(
    const TA *__restrict__ fa, const double *__restrict__ fb, double *__restrict__ vout,
    const double *__restrict__ f, const double *__restrict__ r64in, double *__restrict__ out64,
    TR *__restrict__ rhs0, const double *__restrict__ cx, const double *__restrict__ cy,
    const double *__restrict__ cz, const SolveState *__restrict__ st,
    const BlockLevel *__restrict__ bls, const Tile *__restrict__ tiles,
    double *__restrict__ partials, const int *__restrict__ done) {
    if (done && *done) return;
    constexpr int LX = SW_TX + 2, LY = TY + 2, NS = 4, NT = SW_TX * TY, LXY = LX * LY;
    constexpr int KL = (LXY + NT - 1) / NT;
    __shared__ double sp[NS][LXY];
    const double sc = MODE == 1 ? -st->beta : (MODE == 0 ? st->alpha : 0.0);
    const Tile t = tiles[blockIdx.x];
    const BlockLevel g = bls[t.b];
    const int tid = threadIdx.x + SW_TX * threadIdx.y;
    const int xb = t.x0 - 1, yb = t.y0 - 1;
    const int xhi = t.x0 + SW_TX - 1 < g.n0 ? t.x0 + SW_TX - 1 : g.n0;
    const int yhi = t.y0 + TY - 1 < g.n1 ? t.y0 + TY - 1 : g.n1;
    const int wx0 = t.x0 == 1 ? 0 : t.x0, wx1 = xhi == g.n0 ? g.n0 + 1 : xhi;
    const int wy0 = t.y0 == 1 ? 0 : t.y0, wy1 = yhi == g.n1 ? g.n1 + 1 : yhi;
    const int wz0 = t.z0 == 1 ? 0 : t.z0, wz1 = t.z1 == g.n2 ? g.n2 + 1 : t.z1;
    long long lofs[KL], lsrc[KL];
    bool lok[KL], lown[KL];
#pragma unroll
    for (int k = 0; k < KL; ++k) {
        const int p = tid + k * NT;
        const int ly = p / LX, lx = p - ly * LX;
        const int x = xb + lx, y = yb + ly;
        lok[k] = p < LXY && x >= 0 && x <= g.n0 + 1 && y >= 0 && y <= g.n1 + 1;
        lown[k] = lok[k] && x >= wx0 && x <= wx1 && y >= wy0 && y <= wy1;
        lofs[k] = g.off + (long long)x + g.sy * y;
        lsrc[k] = lok[k] ? lofs[k] : g.off;
    }
    const int bx = t.x0 + threadIdx.x, by = t.y0 + threadIdx.y;
    const bool has_col = bx <= g.n0 && by <= g.n1;
    auto plane = [&](int c) {  // form plane c of v in LDS, write owned cells
        double *P = sp[(c + NS) % NS];
        const bool zown = c >= wz0 && c <= wz1;
        const long long zo = g.sz * (long long)c;
#pragma unroll
        for (int k = 0; k < KL; ++k) {
            if (tid + k * NT >= LXY) continue;
            double v;
            if (MODE == 2 || MODE == 3) {
                const double a = (double)fa[lsrc[k] + zo];
                v = lok[k] ? a : 0.0;
            } else {
                const double a = (double)fa[lsrc[k] + zo], b = fb[lsrc[k] + zo];
                v = lok[k] ? a + sc * b : 0.0;
                if (zown && lown[k]) vout[lofs[k] + zo] = v;
            }
            P[tid + k * NT] = v;
        }
    };
    double s1 = 0.0, s2 = 0.0;
    plane(t.z0 - 1);
    plane(t.z0);
    for (int c = t.z0; c <= t.z1; ++c) {
        plane(c + 1);
        __syncthreads();
        if (has_col) {
            const double *P = sp[c % NS];
            const double *Pm = sp[(c + NS - 1) % NS];
            const double *Pp = sp[(c + 1) % NS];
            const int l = (bx - xb) + LX * (by - yb);
            const long long q = gidx(g, bx, by, c);
            double aw, ae, as, an, ab, at, ap;
            coefs<double, VAR>(g, cx, cy, cz, q, bx, by, c, aw, ae, as, an, ab, at, ap);
            double s = aw * P[l - 1];
            s = s + ae * P[l + 1];
            s = s + as * P[l - LX];
            s = s + an * P[l + LX];
            s = s + ab * Pm[l];
            s = s + at * Pp[l];
            s = s + ap * P[l];
            if (MODE == 1) {
                out64[q] = s;  // w = A d
                s1 = s1 + s * s;
                s2 = s2 + r64in[q] * s;
            } else if (MODE == 2) {
                s1 = s1 + s * r64in[q];
            } else {
                const double r = f[q] - s;
                out64[q] = r;
                rhs0[q] = (TR)r;
                s1 = s1 + r * r;
                s2 = fmax(s2, g.lsm ? fabs(r) : fabs(r * g.h0 * g.h1 * g.h2));
            }
        }
        __syncthreads();
    }
    if (MODE == 0 || MODE == 3)
        block_sum_max<SW_TX * TY>(s1, s2, partials + 2 * (long long)blockIdx.x);
    else
        block_sum2<SW_TX * TY>(s1, s2, partials + 2 * (long long)blockIdx.x);
}

// p_new = p_old + alpha d on every cell of the level-0 arena, alpha from
// the solve state (a refined grid's GCR step; its ghosts are then refilled
// from p_new by the exchange).
__global__ __launch_bounds__(256) void k_paxpy(double *__restrict__ pnew, const double *__restrict__ pold,
                                               const double *__restrict__ d, const SolveState *__restrict__ st,
                                               long long n, const int *__restrict__ done) {
    if (done && *done) return;
    const double a = st->alpha;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        pnew[i] = pold[i] + a * d[i];
}

// Fixed-order reduction of the partials into the GCR scalars:
// BETA: beta = (A z, w_prev) / gamma_prev with gamma_prev = (w_prev, w_prev)
// (0 on the first step); !BETA: gamma = (w, w), alpha = (r, w) / gamma.
template <bool BETA>
__global__ __launch_bounds__(1024) void k_gcr_scalar(const double *__restrict__ partials, int nparts,
                                                     SolveState *st) {
    if (st->done) return;
    const int t = threadIdx.x;
    double a = 0.0, b = 0.0;
    for (int p = t; p < nparts; p += 1024) {
        a = a + partials[2 * (long long)p];
        b = b + partials[2 * (long long)p + 1];
    }
    wg_reduce2<1024, false>(a, b);
    if (t == 0) {
        if (BETA) {
            st->beta = (st->fk > 0 && st->gamma != 0.0) ? a / st->gamma : 0.0;
        } else {
            st->gamma = a;
            st->alpha = a != 0.0 ? b / a : 0.0;
            st->fk = st->fk + 1;
        }
    }
}

// Reduce the partials in a fixed order; record rho_k = ||r_k|| / ||f|| and
// rmax_k for k = st->k; raise `done` when converged (k >= 1 and
// rho_k <= tol and k >= min_cycles).
__global__ __launch_bounds__(1024) void k_finalize(const double *__restrict__ partials, int nparts,
                                                   SolveState *st, double *hist, double *rmaxh,
                                                   int min_cycles) {
    if (st->done) return;
    const int t = threadIdx.x;
    double s = 0.0, m = 0.0;
    for (int p = t; p < nparts; p += 1024) {
        s = s + partials[2 * (long long)p];
        m = fmax(m, partials[2 * (long long)p + 1]);
    }
    wg_reduce2<1024, true>(s, m);
    if (t == 0) {
        const int k = st->k;
        const double nr = sqrt(s);
        if (k == 0) st->fnorm = nr > 0.0 ? nr : 1.0;
        const double rho = nr / st->fnorm;
        hist[k] = rho;
        rmaxh[k] = m;
        st->last = k;
        st->k = k + 1;
        if (k >= 1 && rho <= st->tol && k >= min_cycles) st->done = 1;
    }
}

__global__ void k_reset_state(SolveState *st, double tol) {
    st->tol = tol;
    st->alpha = 0.0;
    st->beta = 0.0;
    st->gamma = 1.0;
    st->fk = 0;
    st->fnorm = 1.0;
    st->done = 0;
    st->k = 0;
    st->last = -1;
}


}  // namespace mf3d
