// mg_sweep.hpp -- red-black Gauss-Seidel sweep and ghost exchange.
#pragma once
#include "mg_common.hpp"

namespace mf3d {

// ---------------------------------------------------------------------------
// One red-black Gauss-Seidel sweep, both colours, in one pass
// (mgsolver.for:549-601).  Reads pin, whose ghost layer holds the values of
// the last exchange, and writes every interior cell of pout; as in the
// reference no ghost is refreshed between the colours.  Colour: the
// reference's first colour is mod(i+j+k,2)==0 on its 1-based ghosted
// indices, i.e. (i+j+k) odd on the interior indices used here.
//
// A workgroup owns SW_TX x TY columns of one block and marches over the
// planes z0..z1.  LDS keeps a ring of 5 planes of pin, each with a 2-cell
// halo in x and y.  At step c the first colour is computed on plane c over
// the tile plus a 1-cell ring (in place in LDS), then the second colour on
// plane c-1 over the tile, which is written out.  Red cells in the ring are
// recomputed by both neighbouring workgroups from the same inputs, so they
// are bitwise the values the reference computes once.
//
// Loads run one step ahead: while step c computes, the registers already
// hold plane c+2 of pin and the RHS values of the next red and black cells,
// so global-memory latency overlaps the LDS phases instead of stalling them.
// ZERO: pin is identically zero (the first sweep after restriction, where
// the reference has just zeroed phi, mgsolver.for:650-657).
constexpr int SW_TX = 64;

template <typename T, bool VAR, bool ZERO, int TY>
__global__ __launch_bounds__(SW_TX *TY) void k_rbs(const T *__restrict__ pin, T *__restrict__ pout,
                                                  const T *__restrict__ rhs,
                                                  const T *__restrict__ cx,
                                                  const T *__restrict__ cy,
                                                  const T *__restrict__ cz,
                                                  const BlockLevel *__restrict__ bls,
                                                  const Tile *__restrict__ tiles,
                                                  const int *__restrict__ done) {
    if (done && *done) return;
    constexpr int LX = SW_TX + 4, LY = TY + 4, NS = 5, NT = SW_TX * TY, LXY = LX * LY;
    constexpr int KL = (LXY + NT - 1) / NT;
    constexpr int HALF = (SW_TX + 2) / 2;
    constexpr int PD = 3;  // load pipeline depth in plane steps
    __shared__ T sp[NS][LXY];
    const Tile t = tiles[blockIdx.x];
    const BlockLevel g = bls[t.b];
    const int tid = threadIdx.x + SW_TX * threadIdx.y;
    const int xb = t.x0 - 2, yb = t.y0 - 2;
    const int zlo = t.z0 - 2 < 0 ? 0 : t.z0 - 2;
    const int zhi = t.z1 + 2 > g.n2 + 1 ? g.n2 + 1 : t.z1 + 2;

    // this thread's share of a plane load
    long long lofs[KL];
    bool lok[KL];
#pragma unroll
    for (int k = 0; k < KL; ++k) {
        const int p = tid + k * NT;
        const int ly = p / LX, lx = p - ly * LX;
        const int x = xb + lx, y = yb + ly;
        lok[k] = p < LXY && x >= 0 && x <= g.n0 + 1 && y >= 0 && y <= g.n1 + 1;
        lofs[k] = g.off + (long long)x + g.sy * y;
    }
    // Loads are issued unconditionally from a clamped address and masked
    // afterwards: loads under divergent branches make the compiler wait for
    // all outstanding loads instead of just the one it needs.
#pragma unroll
    for (int k = 0; k < KL; ++k)
        if (!lok[k]) lofs[k] = g.off + g.sy + 1;

    // this thread's red cell (the ring position depends on the plane's parity)
    const bool has_red = tid < HALF * (TY + 2);
    const int r_row = tid / HALF, r_m = tid - r_row * HALF;
    const int ry = t.y0 - 1 + r_row, rxs = t.x0 - 1;
    auto red_x = [&](int c) { return rxs + ((1 - ry - c - rxs) & 1) + 2 * r_m; };
    auto red_ok = [&](int c) {
        const int x = red_x(c);
        return has_red && c >= 1 && c <= g.n2 && x >= 1 && x <= g.n0 && ry >= 1 && ry <= g.n1;
    };
    // this thread's black-phase column
    const int bx = t.x0 + threadIdx.x, by = t.y0 + threadIdx.y;
    const bool has_col = bx <= g.n0 && by <= g.n1;

    // Register sets for a PD-deep load pipeline: set s serves the steps
    // c == c0 + s (mod PD).  At step c it supplies plane c+1 of pin (stored
    // to LDS), the RHS of this thread's red cell on plane c and of its black
    // cell on plane c-1, and is then reloaded with the values step c+PD
    // needs -- consumed, then overwritten, never copied.
    struct Regs {
        T pre[KL];
        T rr, rb;
    };
    auto load_regs = [&](Regs &R, int c) {  // the values step c needs
        const int zp = c + 1 < zhi ? c + 1 : zhi;
#pragma unroll
        for (int k = 0; k < KL; ++k) {
            if (ZERO) {
                R.pre[k] = T(0);
            } else {
                const T v = pin[lofs[k] + g.sz * zp];
                R.pre[k] = lok[k] ? v : T(0);
            }
        }
        const bool ok = red_ok(c);
        const T v = rhs[ok ? gidx(g, red_x(c), ry, c) : g.off];
        R.rr = ok ? v : T(0);
        const bool okb = has_col && c - 1 >= t.z0 && c - 1 <= t.z1;
        const T w = rhs[okb ? gidx(g, bx, by, c - 1) : g.off];
        R.rb = okb ? w : T(0);
    };

    // prologue: planes z0-2 and z0-1 straight into LDS
    for (int c = t.z0 - 2; c < t.z0; ++c)
        if (c >= zlo) {
            T *P = sp[c % NS];
#pragma unroll
            for (int k = 0; k < KL; ++k) {
                if (tid + k * NT >= LXY) continue;
                P[tid + k * NT] = (!ZERO && lok[k]) ? pin[lofs[k] + g.sz * c] : T(0);
            }
        }

    auto step = [&](int c, Regs &R) {
        {
            T *P = sp[(c + 1) % NS];  // beyond zhi the slot holds a plane nobody reads
#pragma unroll
            for (int k = 0; k < KL; ++k)
                if (tid + k * NT < LXY) P[tid + k * NT] = R.pre[k];
        }
        const T rr_cur = R.rr, rb_cur = R.rb;
        load_regs(R, c + PD);
        __syncthreads();
        if (red_ok(c)) {
            T *P = sp[c % NS];
            const T *Pm = sp[(c + NS - 1) % NS];
            const T *Pp = sp[(c + 1) % NS];
            const int x = red_x(c);
            const int l = (x - xb) + LX * (ry - yb);
            T aw, ae, as, an, ab, at, ap;
            coefs<T, VAR>(g, cx, cy, cz, gidx(g, x, ry, c), x, ry, c, aw, ae, as, an, ab, at, ap);
            T s = aw * P[l - 1];
            s = s + ae * P[l + 1];
            s = s + as * P[l - LX];
            s = s + an * P[l + LX];
            s = s + ab * Pm[l];
            s = s + at * Pp[l];
            P[l] = (rr_cur - s) / ap;
        }
        __syncthreads();
        const int z = c - 1;
        if (has_col && z >= t.z0 && z <= t.z1) {
            const T *P = sp[z % NS];
            const int l = (bx - xb) + LX * (by - yb);
            const long long q = gidx(g, bx, by, z);
            T v;
            if (((bx + by + z) & 1) == 0) {
                const T *Pm = sp[(z + NS - 1) % NS];
                const T *Pp = sp[(z + 1) % NS];
                T aw, ae, as, an, ab, at, ap;
                coefs<T, VAR>(g, cx, cy, cz, q, bx, by, z, aw, ae, as, an, ab, at, ap);
                T s = aw * P[l - 1];
                s = s + ae * P[l + 1];
                s = s + as * P[l - LX];
                s = s + an * P[l + LX];
                s = s + ab * Pm[l];
                s = s + at * Pp[l];
                v = (rb_cur - s) / ap;
            } else {
                v = P[l];
            }
            pout[q] = v;
        }
    };

    const int c0 = t.z0 - 1, c1 = t.z1 + 1;
    Regs R0, R1, R2;
    load_regs(R0, c0);
    load_regs(R1, c0 + 1);
    load_regs(R2, c0 + 2);
    int c = c0;
    for (; c + 2 <= c1; c += 3) {
        step(c, R0);
        step(c + 1, R1);
        step(c + 2, R2);
    }
    if (c <= c1) step(c, R0);
    if (c + 1 <= c1) step(c + 1, R1);
}

// Ghost fill from a region table: exchangepp + mgbound + exchbc_mgpp of the
// reference in one gather (every ghost takes the value of the interior cell
// it stands for).  Sources are interior cells, which this kernel never
// writes.
template <typename T>
__global__ __launch_bounds__(256) void k_gather(T *__restrict__ phi,
                                                const BlockLevel *__restrict__ bls,
                                                const Region *__restrict__ regs,
                                                const int *__restrict__ done) {
    if (done && *done) return;
    const Region r = regs[blockIdx.y];
    const BlockLevel gd = bls[r.dst_b], gs = bls[r.src_b];
    const int c0 = r.cnt[0], c1 = r.cnt[1];
    const int tot = c0 * c1 * r.cnt[2];
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += gridDim.x * blockDim.x) {
        const int x = t % c0, y = (t / c0) % c1, z = t / (c0 * c1);
        const int di = r.lo[0] + x, dj = r.lo[1] + y, dk = r.lo[2] + z;
        const int si = r.map[0] < 0 ? di : r.map[0];
        const int sj = r.map[1] < 0 ? dj : r.map[1];
        const int sk = r.map[2] < 0 ? dk : r.map[2];
        phi[gidx(gd, di, dj, dk)] = phi[gidx(gs, si, sj, sk)];
    }
}

}  // namespace mf3d
