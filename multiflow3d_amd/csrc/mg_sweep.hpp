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
// plane c-1: half the threads compute its cells, the other half copy the
// first-colour cells of that plane out of LDS.  Red cells in the ring are
// recomputed by both neighbouring workgroups from the same inputs, so they
// are bitwise the values the reference computes once.
//
// Loads run ahead of use: two register sets alternate between even and odd
// steps, so while step c computes, plane c+2 of pin and the RHS of the next
// cells are in flight; loads are unconditional (clamped address, masked
// result) so the compiler waits only for the one it needs.  Addresses and
// LDS indices are precomputed per thread for both plane parities.  Where no
// Neumann-zeroed coefficient is in reach of the tile (the common case) the
// constant coefficients and ap are hoisted.
// ZERO: pin is identically zero (the first sweep after restriction, where
// the reference has just zeroed phi, mgsolver.for:650-657).
// RCP: multiply by 1/ap instead of dividing (inner cycle of the mixed
// precision form only; the FP64 form divides, as the reference does).
constexpr int SW_TX = 64;

template <typename T, bool VAR, bool ZERO, bool RCP, int TY>
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
    constexpr int HALF = (SW_TX + 2) / 2;  // first-colour cells per ring row
    constexpr int NB = NT / 2;             // second-colour cells per tile plane
    __shared__ T sp[NS][LXY];
    const Tile t = tiles[blockIdx.x];
    const BlockLevel g = bls[t.b];
    const int tid = threadIdx.x + SW_TX * threadIdx.y;
    const int xb = t.x0 - 2, yb = t.y0 - 2;
    const int zlo = t.z0 - 2 < 0 ? 0 : t.z0 - 2;
    const int zhi = t.z1 + 2 > g.n2 + 1 ? g.n2 + 1 : t.z1 + 2;
    const long long sz = g.sz;

    // this thread's share of a plane load (clamped address when outside)
    long long lofs[KL];
    bool lok[KL];
#pragma unroll
    for (int k = 0; k < KL; ++k) {
        const int p = tid + k * NT;
        const int ly = p / LX, lx = p - ly * LX;
        const int x = xb + lx, y = yb + ly;
        lok[k] = p < LXY && x >= 0 && x <= g.n0 + 1 && y >= 0 && y <= g.n1 + 1;
        lofs[k] = lok[k] ? g.off + (long long)x + g.sy * y : g.off + g.sy + 1;
    }

    // first-colour (red) ring cell of this thread, per plane parity
    const int r_row = tid / HALF, r_m = tid - r_row * HALF;
    const int ry = t.y0 - 1 + r_row, rxs = t.x0 - 1;
    bool rok[2];
    int rl[2];
    long long rg[2];
    int rbnd[2];  // x/y boundary position bits for the coefficients
#pragma unroll
    for (int par = 0; par < 2; ++par) {
        const int x = rxs + ((1 - ry - par - rxs) & 1) + 2 * r_m;
        rok[par] = tid < HALF * (TY + 2) && x >= 1 && x <= g.n0 && ry >= 1 && ry <= g.n1;
        rl[par] = (x - xb) + LX * (ry - yb);
        rg[par] = rok[par] ? g.off + (long long)x + g.sy * ry : g.off;
        rbnd[par] = x | (ry << 16);
    }
    // second phase: threads < NB compute second-colour cells, the others
    // copy first-colour cells of the same plane out of LDS
    const bool blk = tid < NB;
    const int bt = blk ? tid : tid - NB;
    const int b_row = bt / (SW_TX / 2), b_m = bt - b_row * (SW_TX / 2);
    const int by = t.y0 + b_row;
    bool bok[2];
    int bl[2];
    long long bg[2];
    int bbnd[2];
#pragma unroll
    for (int par = 0; par < 2; ++par) {
        const int e = (1 + by + par) & 1;  // offset of the second-colour cell in its pair
        const int x = t.x0 + 2 * b_m + (blk ? e : 1 - e);
        bok[par] = x <= g.n0 && by <= g.n1;
        bl[par] = (x - xb) + LX * (by - yb);
        bg[par] = bok[par] ? g.off + (long long)x + g.sy * by : g.off;
        bbnd[par] = x | (by << 16);
    }

    // hoisted constant coefficients where no Neumann zero is in reach
    const bool xy_clean = !VAR && (!(g.neu & kNeuW) || t.x0 > 2) && (!(g.neu & kNeuE) || t.x0 + SW_TX < g.n0) &&
                          (!(g.neu & kNeuS) || t.y0 > 2) && (!(g.neu & kNeuN) || t.y0 + TY < g.n1);
    const T ca = cconst<T>(g, 0), cb = cconst<T>(g, 1), cc = cconst<T>(g, 2);
    T ap0;
    {
        T s0 = ca + ca;
        s0 = s0 + cb;
        s0 = s0 + cb;
        s0 = s0 + cc;
        s0 = s0 + cc;
        ap0 = -s0;
    }
    const T rap0 = T(1) / ap0;
    auto z_clean = [&](int z) {
        return (!(g.neu & kNeuB) || z > 1) && (!(g.neu & kNeuT) || z < g.n2);
    };
    // one Gauss-Seidel update from the LDS planes P (this), Pm, Pp
    auto update = [&](const T *P, const T *Pm, const T *Pp, int l, int bnd, int z, long long q, T r,
                      bool fast) -> T {
        if (fast) {
            T s = ca * P[l - 1];
            s = s + ca * P[l + 1];
            s = s + cb * P[l - LX];
            s = s + cb * P[l + LX];
            s = s + cc * Pm[l];
            s = s + cc * Pp[l];
            return RCP ? (r - s) * rap0 : (r - s) / ap0;
        }
        T aw, ae, as, an, ab, at, ap;
        coefs<T, VAR>(g, cx, cy, cz, q, bnd & 0xffff, bnd >> 16, z, aw, ae, as, an, ab, at, ap);
        T s = aw * P[l - 1];
        s = s + ae * P[l + 1];
        s = s + as * P[l - LX];
        s = s + an * P[l + LX];
        s = s + ab * Pm[l];
        s = s + at * Pp[l];
        return RCP ? (r - s) * (T(1) / ap) : (r - s) / ap;
    };

    struct Regs {
        T pre[KL];
        T rr, rb;
    };
    // the values step c needs: plane c+1 of pin, the RHS of this thread's red
    // cell on plane c and of its second-phase cell on plane c-1
    auto load_regs = [&](Regs &R, int c) {
        const long long zo = sz * (long long)(c + 1 < zhi ? c + 1 : zhi);
#pragma unroll
        for (int k = 0; k < KL; ++k) {
            if (ZERO) {
                R.pre[k] = T(0);
            } else {
                const T v = pin[lofs[k] + zo];
                R.pre[k] = lok[k] ? v : T(0);
            }
        }
        const int pr = c & 1;
        const bool okr = rok[pr] && c >= 1 && c <= g.n2;
        const T v = rhs[(okr ? rg[pr] : g.off) + (okr ? sz * c : 0)];
        R.rr = okr ? v : T(0);
        const int pb = (c - 1) & 1;
        const bool okb = blk && bok[pb] && c - 1 >= t.z0 && c - 1 <= t.z1;
        const T w = rhs[(okb ? bg[pb] : g.off) + (okb ? sz * (c - 1) : 0)];
        R.rb = okb ? w : T(0);
    };

    // prologue: planes z0-2 and z0-1 straight into LDS
    for (int c = t.z0 - 2; c < t.z0; ++c)
        if (c >= zlo) {
            T *P = sp[c % NS];
#pragma unroll
            for (int k = 0; k < KL; ++k) {
                if (tid + k * NT >= LXY) continue;
                P[tid + k * NT] = (!ZERO && lok[k]) ? pin[lofs[k] + sz * c] : T(0);
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
        load_regs(R, c + 2);
        __syncthreads();
        {
            const int pr = c & 1;
            if (rok[pr] && c >= 1 && c <= g.n2) {
                T *P = sp[c % NS];
                const int l = rl[pr];
                P[l] = update(P, sp[(c + NS - 1) % NS], sp[(c + 1) % NS], l, rbnd[pr], c,
                              rg[pr] + sz * c, rr_cur, xy_clean && z_clean(c));
            }
        }
        __syncthreads();
        const int z = c - 1, pz = z & 1;
        if (bok[pz] && z >= t.z0 && z <= t.z1) {
            const T *P = sp[z % NS];
            const int l = bl[pz];
            const long long q = bg[pz] + sz * z;
            T v;
            if (blk)
                v = update(P, sp[(z + NS - 1) % NS], sp[(z + 1) % NS], l, bbnd[pz], z, q, rb_cur,
                           xy_clean && z_clean(z));
            else
                v = P[l];
            pout[q] = v;
        }
    };

    const int c0 = t.z0 - 1, c1 = t.z1 + 1;
    Regs R0, R1;
    load_regs(R0, c0);
    load_regs(R1, c0 + 1);
    int c = c0;
    for (; c + 1 <= c1; c += 2) {
        step(c, R0);
        step(c + 1, R1);
    }
    if (c <= c1) step(c, R0);
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
