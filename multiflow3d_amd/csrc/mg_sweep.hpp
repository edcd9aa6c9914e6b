// mg_sweep.hpp -- red-black Gauss-Seidel sweep and ghost exchange.
#pragma once
#include <type_traits>

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
// planes z0..z1 (z0 odd: chunks start at 1 and have even depth).  LDS keeps
// a ring of planes of pin, each with a 2-cell halo in x and y.  At step c
// the first colour is computed on plane c over the tile plus a 1-cell ring
// (in place in LDS), then the second colour on plane c-LAG: half the
// threads compute its cells, the other half copy the first-colour cells of
// that plane out of LDS.  Red cells in the ring are recomputed by both
// neighbouring workgroups from the same inputs, so they are bitwise the
// values the reference computes once.  LAG 1 (5 slots) needs two barriers
// a step (plane c+1 stored before the first colour reads it; the first
// colour of plane c done before the second colour of plane c-1 reads it);
// LAG 2 (6 slots) needs one, at the end of the step, since everything a
// step reads was written in earlier steps -- at the price of one more step
// a chunk.
//
// The kernel is bound by instruction issue before HBM, so the march keeps
// scalar and address work out of the loop.  The loop is unrolled by whole
// turns of the ring (10 steps at 5 slots, 6 at 6: colours alternate, so
// the turn is even), which makes every LDS slot, colour parity and
// register set compile-time.  A chunk whose step count is not a multiple
// of the turn starts up to a turn early: those steps load clamped planes
// and compute first-colour cells that no stored cell reads (their slots
// are overwritten by whole planes before use), and store nothing.  Global
// traffic goes through buffer resources: a per-thread 32-bit offset and a
// per-plane scalar offset, no address arithmetic per load; loads are
// unconditional (clamped in-bounds offsets) and never masked, since values
// outside the block feed no update.  NR register sets rotate, so the loads
// for step c+NR are in flight while step c computes: 2 for a full sweep
// (four loads a step keep the queue deep), one per slot for the
// zero-start sweep, which loads only the two RHS values and would
// otherwise wait on the stores of the steps before.  Where no
// Neumann-zeroed coefficient is in reach of the tile the constant
// coefficients and ap are hoisted.
// ZERO: pin is identically zero (the first sweep after restriction, where
// the reference has just zeroed phi, mgsolver.for:650-657).
// RCP: multiply by 1/ap instead of dividing (inner cycle of the mixed
// precision form only; the FP64 form divides, as the reference does).
// `bytes` is the arena length in bytes (below 2^32, checked by the host).
constexpr int SW_TX = 64;

template <int J>
using Step = std::integral_constant<int, J>;

// Ghost images of a lone block (own_ghosts: no neighbour, no periodic
// partner, the same kind of face at both ends of an axis, every extent at
// least 2).  On each axis the ghost at 0 holds the interior cell m0 (1 on a
// Neumann face, n on a periodic one) and the ghost at n+1 holds mN (n or
// 1): the values mgbound + exchangepp + exchbc_mgpp leave there after a
// sweep (mgsolver.for:610-611), edges and corners taking the image on every
// axis.  So a cell has at most one image coordinate per axis; the thread
// that stores it also stores the copies in the (up to 7) ghosts that map
// to it, and no gather runs.
__device__ __forceinline__ int ghost_image(int v, int n, bool neu_lo, bool neu_hi) {
    const int m0 = neu_lo ? 1 : n, mn = neu_hi ? n : 1;
    return v == m0 ? 0 : (v == mn ? n + 1 : -1);
}

// FP32 instances are held to 64 VGPRs (8 waves a SIMD: four 512-thread
// workgroups a CU, LDS permitting), which they fit without spilling; FP64
// ones would spill there.
template <typename T, bool VAR, bool ZERO>
constexpr int rbs_waves_per_eu() { return sizeof(T) == 4 && !(VAR && ZERO) ? 8 : 1; }

template <typename T, bool VAR, bool ZERO, bool RCP, int TY, int LAG>
__global__ __launch_bounds__(SW_TX *TY, (rbs_waves_per_eu<T, VAR, ZERO>())) void k_rbs(const T *__restrict__ pin, T *__restrict__ pout,
                                                  const T *__restrict__ rhs,
                                                  const T *__restrict__ cx,
                                                  const T *__restrict__ cy,
                                                  const T *__restrict__ cz, uint32_t bytes, int own_ghosts,
                                                  const BlockLevel *__restrict__ bls,
                                                  const Tile *__restrict__ tiles,
                                                  const int *__restrict__ done) {
    if (done && *done) return;
    static_assert(LAG == 1 || LAG == 2, "the second colour trails the first by 1 or 2 planes");
    constexpr int LX = SW_TX + 4, LY = TY + 4, NT = SW_TX * TY, LXY = LX * LY;
    constexpr int NS = LAG == 2 ? 6 : 5;        // ring slots
    constexpr int UNROLL = LAG == 2 ? 6 : 10;   // whole turns of the ring, even
    constexpr int KL = (LXY + NT - 1) / NT;
    constexpr int PS = KL * NT;            // slot stride: every thread stores KL values
    constexpr int HALF = (SW_TX + 2) / 2;  // first-colour cells per ring row
    constexpr int NB = NT / 2;             // second-colour cells per tile plane
    constexpr uint32_t W = sizeof(T);
    __shared__ T sp[NS * PS];
    const Tile t = tiles[blockIdx.x];
    const BlockLevel g = bls[t.b];
    const auto rs_in = make_rsrc(pin, bytes), rs_rhs = make_rsrc(rhs, bytes), rs_out = make_rsrc(pout, bytes);
    const int tid = threadIdx.x + SW_TX * threadIdx.y;
    const int xb = t.x0 - 2, yb = t.y0 - 2;
    const int zhi = t.z1 + 2 > g.n2 + 1 ? g.n2 + 1 : t.z1 + 2;
    const int sy = (int)g.sy;
    const uint32_t pb = (uint32_t)g.sz * W;   // plane bytes
    const uint32_t ob = (uint32_t)g.off * W;  // block offset bytes

    // this thread's share of a plane (in-bounds offset when outside)
    uint32_t lofs[KL];
#pragma unroll
    for (int k = 0; k < KL; ++k) {
        const int p = tid + k * NT;
        const int ly = p / LX, lx = p - ly * LX;
        const int x = xb + lx, y = yb + ly;
        const bool ok = p < LXY && x >= 0 && x <= g.n0 + 1 && y >= 0 && y <= g.n1 + 1;
        lofs[k] = ob + (uint32_t)(ok ? x + sy * y : sy + 1) * W;
    }

    // first-colour (red) ring cell of this thread, per plane parity
    const int r_row = tid / HALF, r_m = tid - r_row * HALF;
    const int ry = t.y0 - 1 + r_row, rxs = t.x0 - 1;
    bool rok[2];
    int rl[2], rx[2];
    uint32_t ro[2];
#pragma unroll
    for (int par = 0; par < 2; ++par) {
        const int x = rxs + ((1 - ry - par - rxs) & 1) + 2 * r_m;
        rok[par] = tid < HALF * (TY + 2) && x >= 1 && x <= g.n0 && ry >= 1 && ry <= g.n1;
        rl[par] = (x - xb) + LX * (ry - yb);
        ro[par] = ob + (uint32_t)(rok[par] ? x + sy * ry : 0) * W;
        rx[par] = x;
    }
    // second phase: threads < NB (whole waves) compute second-colour
    // cells, the others copy first-colour cells of the same plane
    const bool blk = tid < NB;
    const int bt = blk ? tid : tid - NB;
    const int b_row = bt / (SW_TX / 2), b_m = bt - b_row * (SW_TX / 2);
    const int by = t.y0 + b_row;
    bool bok[2];
    int bl[2], bx[2];
    uint32_t bo[2];
#pragma unroll
    for (int par = 0; par < 2; ++par) {
        const int e = (1 + by + par) & 1;  // offset of the second-colour cell in its pair
        const int x = t.x0 + 2 * b_m + (blk ? e : 1 - e);
        bok[par] = x <= g.n0 && by <= g.n1;
        bl[par] = (x - xb) + LX * (by - yb);
        bo[par] = ob + (uint32_t)(bok[par] ? x + sy * by : 0) * W;
        bx[par] = x;
    }
    // ghost images of this thread's stored cells (own_ghosts), per parity
    int bix[2];
#pragma unroll
    for (int par = 0; par < 2; ++par)
        bix[par] = own_ghosts && bok[par] ? ghost_image(bx[par], g.n0, g.neu & kNeuW, g.neu & kNeuE) : -1;
    const int biy = own_ghosts && by <= g.n1 ? ghost_image(by, g.n1, g.neu & kNeuS, g.neu & kNeuN) : -1;

    // hoisted constant coefficients where no Neumann zero is in reach:
    // planes zf0..zf1 of this tile
    const bool xy_clean = !VAR && (!(g.neu & kNeuW) || t.x0 > 2) && (!(g.neu & kNeuE) || t.x0 + SW_TX < g.n0) &&
                          (!(g.neu & kNeuS) || t.y0 > 2) && (!(g.neu & kNeuN) || t.y0 + TY < g.n1);
    const int zf0 = !xy_clean ? 1 << 30 : ((g.neu & kNeuB) ? 2 : -(1 << 30));
    const int zf1 = (g.neu & kNeuT) ? g.n2 - 1 : 1 << 30;
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
    // one Gauss-Seidel update; P, Pm, Pp: this plane and the planes below/above
    auto update = [&](const T *P, const T *Pm, const T *Pp, int l, int x, int y, int z, uint32_t o, T r) -> T {
        if (z >= zf0 && z <= zf1) {
            T s = ca * P[l - 1];
            s = madd(s, ca, P[l + 1]);
            s = madd(s, cb, P[l - LX]);
            s = madd(s, cb, P[l + LX]);
            s = madd(s, cc, Pm[l]);
            s = madd(s, cc, Pp[l]);
            return RCP ? (r - s) * rap0 : (r - s) / ap0;
        }
        T aw, ae, as, an, ab, at, ap;
        const long long q = (long long)(o / W) + g.sz * z;
        coefs<T, VAR>(g, cx, cy, cz, q, x, y, z, aw, ae, as, an, ab, at, ap);
        T s = aw * P[l - 1];
        s = madd(s, ae, P[l + 1]);
        s = madd(s, as, P[l - LX]);
        s = madd(s, an, P[l + LX]);
        s = madd(s, ab, Pm[l]);
        s = madd(s, at, Pp[l]);
        return RCP ? (r - s) * (T(1) / ap) : (r - s) / ap;
    };

    struct Regs {
        T pre[KL];
        T rr, rb;
    };
    // the values step c needs: plane c+LAG of pin, the RHS of this thread's
    // red cell on plane c (parity PR) and of its second-phase cell on plane
    // c-LAG (parity index BP)
    auto load_regs = [&](Regs &R, int c, auto parc) {
        constexpr int PR = decltype(parc)::value, BP = LAG == 2 ? PR : 1 - PR;
        if (!ZERO) {
            const int zc = c + LAG < zhi ? c + LAG : zhi;
            const uint32_t so = (uint32_t)(zc < 0 ? 0 : zc) * pb;
#pragma unroll
            for (int k = 0; k < KL; ++k) R.pre[k] = bload<T>(rs_in, lofs[k], so);
        }
        const int cr = c < 0 ? 0 : (c > g.n2 + 1 ? g.n2 + 1 : c);
        R.rr = bload<T>(rs_rhs, ro[PR], (uint32_t)cr * pb);
        const int cb = c - LAG < 0 ? 0 : (c - LAG > g.n2 + 1 ? g.n2 + 1 : c - LAG);
        R.rb = bload<T>(rs_rhs, bo[BP], (uint32_t)cb * pb);
    };

    const int c0 = t.z0 - 1;  // even: chunks start at odd z0
    int nsteps = t.z1 - t.z0 + 2 + LAG;  // steps c0 .. z1+LAG
    nsteps += nsteps & 1;
    const int pad = (UNROLL - nsteps % UNROLL) % UNROLL;  // even
    const int cm = c0 - pad;  // first step
    // plane z lives in slot (z - cm + 1 + LAG) mod NS
    auto slot = [&](int z) { return ((z - cm + 1 + LAG) % NS + NS) % NS; };

    // prologue: planes z0-2 .. z0-2+LAG
#pragma unroll
    for (int d = 0; d <= LAG; ++d) {
        const int c = t.z0 - 2 + d;
        if (c >= 0) {
            T *P = sp + slot(c) * PS;
#pragma unroll
            for (int k = 0; k < KL; ++k) P[tid + k * NT] = ZERO ? T(0) : bload<T>(rs_in, lofs[k], (uint32_t)c * pb);
        }
    }
    __syncthreads();

    constexpr int NR = ZERO ? NS : 2;  // divides the unroll
    Regs RS[NR];
    if (ZERO) {
#pragma unroll
        for (int i = 0; i < NR; ++i)
#pragma unroll
            for (int k = 0; k < KL; ++k) RS[i].pre[k] = T(0);
    }
    load_regs(RS[0], cm, Step<0>{});
    load_regs(RS[1], cm + 1, Step<1>{});
    if constexpr (NR >= 5) {
        load_regs(RS[2 % NR], cm + 2, Step<0>{});
        load_regs(RS[3 % NR], cm + 3, Step<1>{});
        load_regs(RS[4 % NR], cm + 4, Step<0>{});
    }
    if constexpr (NR == 6) load_regs(RS[5 % NR], cm + 5, Step<1>{});

    // Step J of a turn: the first colour on plane c = cm + UNROLL m + J
    // (parity J & 1), the second on plane c - LAG.  Plane c+d sits in slot
    // (J + 1 + LAG + d) mod NS.
    // LAG 1: two barriers, after plane c+1 is stored and after the first
    // colour of plane c, which the second colour of plane c-1 reads.
    // LAG 2: one barrier, at the end: the step reads planes c-3 .. c+1, whose
    // stores and first-colour updates all come from earlier steps, and
    // writes plane c+2 (over plane c-4) and the first colour of plane c,
    // neither of which it reads.
    auto step = [&](int c, auto jc) {
        constexpr int J = decltype(jc)::value, PR = J & 1, BP = LAG == 2 ? PR : 1 - PR;
        constexpr auto sl = [](int d) { return ((J + 1 + LAG + d) % NS + NS) % NS; };
        Regs &R = RS[J % NR];
        {
            T *P = sp + sl(LAG) * PS;  // beyond zhi the slot holds a plane nobody reads
#pragma unroll
            for (int k = 0; k < KL; ++k) P[tid + k * NT] = R.pre[k];
        }
        const T rr_cur = R.rr, rb_cur = R.rb;
        load_regs(R, c + NR, Step<(PR + NR) & 1>{});
        if constexpr (LAG == 1) __syncthreads();
        if (c >= 1 && c <= g.n2 && rok[PR]) {
            T *P = sp + sl(0) * PS;
            const int l = rl[PR];
            P[l] = update(P, sp + sl(-1) * PS, sp + sl(1) * PS, l, rx[PR], ry, c, ro[PR], rr_cur);
        }
        if constexpr (LAG == 1) __syncthreads();
        const int z = c - LAG;
        if (z >= t.z0 && z <= t.z1 && bok[BP]) {
            const T *P = sp + sl(-LAG) * PS;
            const int l = bl[BP];
            T v;
            if (blk)
                v = update(P, sp + sl(-LAG - 1) * PS, sp + sl(1 - LAG) * PS, l, bx[BP], by, z, bo[BP], rb_cur);
            else
                v = P[l];
            bstore<T>(rs_out, v, bo[BP], (uint32_t)z * pb);
            const int iz = own_ghosts ? ghost_image(z, g.n2, g.neu & kNeuB, g.neu & kNeuT) : -1;
            const int ix = bix[BP], X = bx[BP];
            if (ix >= 0 || biy >= 0 || iz >= 0) {
                auto img = [&](int xx, int yy, int zz) {
                    bstore<T>(rs_out, v, ob + (uint32_t)(xx + sy * yy) * W, (uint32_t)zz * pb);
                };
                if (ix >= 0) img(ix, by, z);
                if (biy >= 0) img(X, biy, z);
                if (ix >= 0 && biy >= 0) img(ix, biy, z);
                if (iz >= 0) {
                    img(X, by, iz);
                    if (ix >= 0) img(ix, by, iz);
                    if (biy >= 0) img(X, biy, iz);
                    if (ix >= 0 && biy >= 0) img(ix, biy, iz);
                }
            }
        }
        if constexpr (LAG == 2) __syncthreads();
    };
    for (int c = cm; c < c0 + nsteps; c += UNROLL) {
        step(c + 0, Step<0>{});
        step(c + 1, Step<1>{});
        step(c + 2, Step<2>{});
        step(c + 3, Step<3>{});
        step(c + 4, Step<4>{});
        step(c + 5, Step<5>{});
        if constexpr (UNROLL == 10) {
            step(c + 6, Step<6>{});
            step(c + 7, Step<7>{});
            step(c + 8, Step<8>{});
            step(c + 9, Step<9>{});
        }
    }
}

// The zero-start sweep (the first sweep after restriction, or of a cycle
// from pp = 0): the potential is zero, ghosts included, so the first colour
// is rhs/ap on every cell and the second colour reads only those values.
// No potential is loaded: a z-march over the RHS alone that writes every
// interior cell of pout (with its ghost images on a lone block), bitwise
// what k_rbs<ZERO> and k_rb3d<ZERO> store (same expressions: (r - 0)/ap,
// then the second colour's update in the usual operand order).
//
// A workgroup owns SW_TX x TY columns of one block and a chunk z0..z1 of
// planes.  At step c it computes the first-colour values of plane c over
// the tile plus a 1-cell ring into an LDS ring of 4 planes (zero at
// second-colour and out-of-block positions), then, after one barrier,
// stores plane c-1: first-colour cells from LDS, second-colour cells from
// their six LDS neighbours (planes c-2, c-1, c).  Four slots make one
// barrier per step enough: the slot written at step c+1 holds plane c-3,
// which step c does not read, and no thread writes at step c+2 before all
// have passed the barrier of step c+1.  The RHS loads run 4 steps ahead
// in 4 register sets (the loop is unrolled by 4, so slots and sets are
// compile-time).
template <typename T, bool VAR, bool RCP, int TY>
__global__ __launch_bounds__(SW_TX *TY) void k_rbz(T *__restrict__ pout, const T *__restrict__ rhs,
                                                  const T *__restrict__ cx, const T *__restrict__ cy,
                                                  const T *__restrict__ cz, uint32_t bytes, int own_ghosts,
                                                  const BlockLevel *__restrict__ bls,
                                                  const Tile *__restrict__ tiles, const int *__restrict__ done) {
    if (done && *done) return;
    constexpr int LX = SW_TX + 2, LY = TY + 2, LXY = LX * LY, NT = SW_TX * TY, NS = 4;
    constexpr int NRING = 2 * LX + 2 * (LY - 2);
    static_assert(NRING <= NT, "one ring cell per thread at most");
    constexpr uint32_t W = sizeof(T);
    __shared__ T red[NS][LXY];
    const Tile t = tiles[blockIdx.x];
    const BlockLevel g = bls[t.b];
    const auto rs_rhs = make_rsrc(rhs, bytes), rs_out = make_rsrc(pout, bytes);
    const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + SW_TX * ty;
    const int sy = (int)g.sy;
    const uint32_t pb = (uint32_t)g.sz * W, ob = (uint32_t)g.off * W;
    // the own tile cell, and the ring cell of threads below NRING
    const int x = t.x0 + tx, y = t.y0 + ty;
    const bool in = x <= g.n0 && y <= g.n1;
    const int lo = (tx + 1) + LX * (ty + 1);
    int rlx, rly;
    {
        const int q = tid < NRING ? tid : 0;
        if (q < LX) { rlx = q; rly = 0; }
        else if (q < 2 * LX) { rlx = q - LX; rly = LY - 1; }
        else { const int u = q - 2 * LX; rlx = (u & 1) ? LX - 1 : 0; rly = 1 + (u >> 1); }
    }
    const bool has_ring = tid < NRING;
    const int xr = t.x0 - 1 + rlx, yr = t.y0 - 1 + rly, lr = rlx + LX * rly;
    auto inside = [&](int xx, int yy) { return xx >= 1 && xx <= g.n0 && yy >= 1 && yy <= g.n1; };
    const bool in_r = has_ring && inside(xr, yr);
    const uint32_t o_own = ob + (uint32_t)(in ? x + sy * y : sy + 1) * W;
    const uint32_t o_ring = ob + (uint32_t)(in_r ? xr + sy * yr : sy + 1) * W;
    const int bix = own_ghosts && in ? ghost_image(x, g.n0, g.neu & kNeuW, g.neu & kNeuE) : -1;
    const int biy = own_ghosts && in ? ghost_image(y, g.n1, g.neu & kNeuS, g.neu & kNeuN) : -1;
    // Constant coefficients with reciprocals (the mixed form): 1/ap of the
    // own and the ring cell in the three plane classes (first, inner, last
    // plane), formed once; 1/ap is the same correctly rounded quotient
    // wherever it is formed, so the results are unchanged.
    constexpr bool HOIST = RCP && !VAR;
    T iao[3] = {T(0), T(0), T(0)}, iar[3] = {T(0), T(0), T(0)};
    if constexpr (HOIST) {
        const int kc[3] = {1, g.n2 > 2 ? 2 : 1, g.n2};
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            T aw, ae, as, an, ab, at, ap;
            coefs<T, false>(g, cx, cy, cz, 0, in ? x : 1, in ? y : 1, kc[m], aw, ae, as, an, ab, at, ap);
            iao[m] = T(1) / ap;
            coefs<T, false>(g, cx, cy, cz, 0, in_r ? xr : 1, in_r ? yr : 1, kc[m], aw, ae, as, an, ab, at, ap);
            iar[m] = T(1) / ap;
        }
    }
    auto pcls = [&](int cc) { return cc == 1 ? 0 : (cc == g.n2 ? 2 : 1); };
    // first-colour value of (xx, yy, cc) from its RHS r, 0 elsewhere
    auto redval = [&](bool ok, int xx, int yy, int cc, T r, const T *ia) -> T {
        if (!ok || cc < 1 || cc > g.n2 || !((xx + yy + cc) & 1)) return T(0);
        if constexpr (HOIST) {
            const int m = pcls(cc);
            return (r - T(0)) * (m == 0 ? ia[0] : (m == 2 ? ia[2] : ia[1]));
        }
        T aw, ae, as, an, ab, at, ap;
        coefs<T, VAR>(g, cx, cy, cz, gidx(g, xx, yy, cc), xx, yy, cc, aw, ae, as, an, ab, at, ap);
        return RCP ? (r - T(0)) * (T(1) / ap) : (r - T(0)) / ap;
    };
    auto ld = [&](uint32_t o, int cc) {
        const int cl = cc < 0 ? 0 : (cc > g.n2 + 1 ? g.n2 + 1 : cc);
        return bload<T>(rs_rhs, o, (uint32_t)cl * pb);
    };
    const int cs = t.z0 - 1;
    int nst = t.z1 + 2 - cs;  // steps c = z0-1 .. z1+1
    nst = (nst + NS - 1) / NS * NS;
    T R[NS], RR[NS];
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        R[j] = ld(o_own, cs + j);
        RR[j] = ld(o_ring, cs + j);
    }
    T rprev = T(0);
    for (int c0 = cs; c0 < cs + nst; c0 += NS) {
#pragma unroll
        for (int J = 0; J < NS; ++J) {
            const int c = c0 + J;
            const T rcur = R[J], rring = RR[J];
            R[J] = ld(o_own, c + NS);
            RR[J] = ld(o_ring, c + NS);
            red[J][lo] = redval(in, x, y, c, rcur, iao);
            if (has_ring) red[J][lr] = redval(in_r, xr, yr, c, rring, iar);
            __syncthreads();
            const int z = c - 1;
            if (in && z >= t.z0 && z <= t.z1) {
                const T *P = red[(J + NS - 1) % NS];
                T v;
                if ((x + y + z) & 1) {
                    v = P[lo];
                } else {
                    T aw, ae, as, an, ab, at, ap;
                    coefs<T, VAR>(g, cx, cy, cz, gidx(g, x, y, z), x, y, z, aw, ae, as, an, ab, at, ap);
                    T s = aw * P[lo - 1];
                    s = madd(s, ae, P[lo + 1]);
                    s = madd(s, as, P[lo - LX]);
                    s = madd(s, an, P[lo + LX]);
                    s = madd(s, ab, red[(J + NS - 2) % NS][lo]);
                    s = madd(s, at, red[J][lo]);
                    if constexpr (HOIST) {
                        const int m = pcls(z);
                        v = (rprev - s) * (m == 0 ? iao[0] : (m == 2 ? iao[2] : iao[1]));
                    } else {
                        v = RCP ? (rprev - s) * (T(1) / ap) : (rprev - s) / ap;
                    }
                }
                bstore<T>(rs_out, v, o_own, (uint32_t)z * pb);
                const int iz = own_ghosts ? ghost_image(z, g.n2, g.neu & kNeuB, g.neu & kNeuT) : -1;
                if (bix >= 0 || biy >= 0 || iz >= 0) {
                    auto img = [&](int xx, int yy, int zz) {
                        bstore<T>(rs_out, v, ob + (uint32_t)(xx + sy * yy) * W, (uint32_t)zz * pb);
                    };
                    if (bix >= 0) img(bix, y, z);
                    if (biy >= 0) img(x, biy, z);
                    if (bix >= 0 && biy >= 0) img(bix, biy, z);
                    if (iz >= 0) {
                        img(x, y, iz);
                        if (bix >= 0) img(bix, y, iz);
                        if (biy >= 0) img(x, biy, iz);
                        if (bix >= 0 && biy >= 0) img(bix, biy, iz);
                    }
                }
            }
            rprev = rcur;
        }
    }
}

// The same red-black sweep (mgsolver.for:549-601) over 3-D tiles, without
// a march, for levels where the march's per-step latency dominates (few
// columns, few planes).  A workgroup loads its R3_TX x R3_TY x R3_TZ tile of
// pin with a 2-cell halo into LDS, updates the first colour on the tile plus
// a 1-cell ring in place, then computes the second colour on the tile, and
// writes every interior cell of the tile (with its ghost images on a lone
// block, as k_rbs does).  First-colour cells in the ring are recomputed by
// the neighbouring tiles from the same inputs, so every stored value is the
// one k_rbs stores: same update expression, same operand order.
constexpr int R3_TX = 32, R3_TY = 8, R3_TZ = 8, R3_NT = 1024;

//
// PRO: the sweep after a prolongation, fused with it: pin is the level's
// potential before the correction, and every tile and halo value is formed
// as k_prolong forms it -- pin at the (Neumann-reflected) source plus the
// trilinear sum of the eight coarse values of phic (level blc), in the same
// operand order -- instead of being read from the corrected field, which is
// never stored.  Saves the prolongation's launch and its pass over the level
// (the levels with 3-D tiles are small: launch-bound).
template <typename T, bool VAR, bool ZERO, bool RCP, bool PRO = false>
__global__ __launch_bounds__(R3_NT) void k_rb3d(const T *__restrict__ pin, T *__restrict__ pout,
                                                const T *__restrict__ rhs, const T *__restrict__ cx,
                                                const T *__restrict__ cy, const T *__restrict__ cz, uint32_t bytes,
                                                int own_ghosts, const BlockLevel *__restrict__ bls,
                                                const Tile *__restrict__ tiles, const int *__restrict__ done,
                                                const T *__restrict__ phic = nullptr,
                                                const BlockLevel *__restrict__ blc = nullptr) {
    static_assert(!(PRO && ZERO), "a prolongation is followed by a sweep from its result");
    if (done && *done) return;
    constexpr int LX = R3_TX + 4, LY = R3_TY + 4, LZ = R3_TZ + 4, LXY = LX * LY, LN = LXY * LZ;
    constexpr int KL = (LN + R3_NT - 1) / R3_NT;
    constexpr int RX = (R3_TX + 2) / 2;                       // first-colour cells per ring row
    constexpr int NRED = RX * (R3_TY + 2) * (R3_TZ + 2);      // first-colour cells of tile + ring
    constexpr int KR = (NRED + R3_NT - 1) / R3_NT;
    constexpr int NPAIR = (R3_TX / 2) * R3_TY * R3_TZ;        // (second, first) colour pairs of the tile
    constexpr int KP = NPAIR / R3_NT;
    static_assert(NPAIR % R3_NT == 0, "whole pairs per thread");
    constexpr uint32_t W = sizeof(T);
    __shared__ T sp[LN];
    const Tile t = tiles[blockIdx.x];
    const BlockLevel g = bls[t.b];
    const auto rs_in = make_rsrc(pin, bytes), rs_rhs = make_rsrc(rhs, bytes), rs_out = make_rsrc(pout, bytes);
    const int tid = threadIdx.x;
    const int xb = t.x0 - 2, yb = t.y0 - 2, zb = t.z0 - 2;
    const int sy = (int)g.sy;
    const uint32_t pb = (uint32_t)g.sz * W, ob = (uint32_t)g.off * W;
    // cell (x, y, z) of the ghosted block: in-bounds byte offset of its row
    auto roff = [&](int x, int y) { return ob + (uint32_t)(x + sy * y) * W; };

    // the tile with its halo (positions outside the ghosted block feed no
    // update and stay zero)
    {
        T v[KL];
#pragma unroll
        for (int k = 0; k < KL; ++k) {
            const int p = tid + k * R3_NT;
            const int lz = p / LXY, r = p - lz * LXY, ly = r / LX, lx = r - ly * LX;
            const int x = xb + lx, y = yb + ly, z = zb + lz;
            const bool ok = p < LN && x >= 0 && x <= g.n0 + 1 && y >= 0 && y <= g.n1 + 1 && z >= 0 && z <= g.n2 + 1;
            v[k] = T(0);
            if constexpr (PRO) {
                if (ok) {
                    const BlockLevel gc = blc[t.b];
                    auto refl = [](int a, int n, bool lo, bool hi) { return (a == 0 && lo) ? 1 : ((a == n + 1 && hi) ? n : a); };
                    const int si = refl(x, g.n0, g.neu & kNeuW, g.neu & kNeuE);
                    const int sj = refl(y, g.n1, g.neu & kNeuS, g.neu & kNeuN);
                    const int sk = refl(z, g.n2, g.neu & kNeuB, g.neu & kNeuT);
                    const T wi0 = (si & 1) ? T(1) : T(3), wi1 = (si & 1) ? T(3) : T(1);
                    const T wj0 = (sj & 1) ? T(1) : T(3), wj1 = (sj & 1) ? T(3) : T(1);
                    const T wk0 = (sk & 1) ? T(1) : T(3), wk1 = (sk & 1) ? T(3) : T(1);
                    const long long c = gidx(gc, x >> 1, y >> 1, z >> 1);
                    T s = (wi0 * wj0 * wk0) * phic[c];
                    s = s + (wi0 * wj1 * wk0) * phic[c + gc.sy];
                    s = s + (wi0 * wj0 * wk1) * phic[c + gc.sz];
                    s = s + (wi0 * wj1 * wk1) * phic[c + gc.sy + gc.sz];
                    s = s + (wi1 * wj0 * wk0) * phic[c + 1];
                    s = s + (wi1 * wj1 * wk0) * phic[c + 1 + gc.sy];
                    s = s + (wi1 * wj0 * wk1) * phic[c + 1 + gc.sz];
                    s = s + (wi1 * wj1 * wk1) * phic[c + 1 + gc.sy + gc.sz];
                    v[k] = bload<T>(rs_in, roff(si, sj), (uint32_t)sk * pb) + s / T(64);
                }
            } else if (!ZERO && ok) {
                v[k] = bload<T>(rs_in, roff(x, y), (uint32_t)z * pb);
            }
        }
#pragma unroll
        for (int k = 0; k < KL; ++k)
            if (tid + k * R3_NT < LN) sp[tid + k * R3_NT] = v[k];
    }
    // one update from LDS, as k_rbs's (coefficients from coefs)
    auto update = [&](int l, int x, int y, int z, T r) -> T {
        T aw, ae, as, an, ab, at, ap;
        const long long q = g.off + x + (long long)sy * y + g.sz * z;
        coefs<T, VAR>(g, cx, cy, cz, q, x, y, z, aw, ae, as, an, ab, at, ap);
        T s = aw * sp[l - 1];
        s = madd(s, ae, sp[l + 1]);
        s = madd(s, as, sp[l - LX]);
        s = madd(s, an, sp[l + LX]);
        s = madd(s, ab, sp[l - LXY]);
        s = madd(s, at, sp[l + LXY]);
        return RCP ? (r - s) * (T(1) / ap) : (r - s) / ap;
    };
    // first colour ((x+y+z) odd) on the tile and its ring, interior cells only
    int rx[KR], ry[KR], rz[KR];
    bool rv[KR];
    T rr[KR];
#pragma unroll
    for (int k = 0; k < KR; ++k) {
        const int m = tid + k * R3_NT;
        const int row = m / RX, i = m - row * RX;
        const int z = t.z0 - 1 + row / (R3_TY + 2), y = t.y0 - 1 + row % (R3_TY + 2);
        const int xs = t.x0 - 1;
        const int x = xs + ((1 - y - z - xs) & 1) + 2 * i;
        rx[k] = x; ry[k] = y; rz[k] = z;
        rv[k] = m < NRED && x >= 1 && x <= g.n0 && y >= 1 && y <= g.n1 && z >= 1 && z <= g.n2;
        rr[k] = rv[k] ? bload<T>(rs_rhs, roff(x, y), (uint32_t)z * pb) : T(0);
    }
    // second-colour cells of this thread's pairs and their RHS
    int px[KP], py[KP], pz[KP];
    T pr[KP];
#pragma unroll
    for (int k = 0; k < KP; ++k) {
        const int m = tid + k * R3_NT;
        const int row = m / (R3_TX / 2), i = m - row * (R3_TX / 2);
        const int z = t.z0 + row / R3_TY, y = t.y0 + row % R3_TY;
        const int x = t.x0 + 2 * i + ((y + z + t.x0) & 1);  // (x+y+z) even
        px[k] = x; py[k] = y; pz[k] = z;
        const bool v = x <= g.n0 && y <= g.n1 && z <= g.n2;
        pr[k] = v ? bload<T>(rs_rhs, roff(x, y), (uint32_t)z * pb) : T(0);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KR; ++k)
        if (rv[k]) {
            const int l = (rx[k] - xb) + LX * ((ry[k] - yb) + LY * (rz[k] - zb));
            sp[l] = update(l, rx[k], ry[k], rz[k], rr[k]);
        }
    __syncthreads();
    auto store = [&](T v, int x, int y, int z) {
        bstore<T>(rs_out, v, roff(x, y), (uint32_t)z * pb);
        if (!own_ghosts) return;
        const int ix = ghost_image(x, g.n0, g.neu & kNeuW, g.neu & kNeuE);
        const int iy = ghost_image(y, g.n1, g.neu & kNeuS, g.neu & kNeuN);
        const int iz = ghost_image(z, g.n2, g.neu & kNeuB, g.neu & kNeuT);
        if (ix < 0 && iy < 0 && iz < 0) return;
        auto img = [&](int xx, int yy, int zz) { bstore<T>(rs_out, v, roff(xx, yy), (uint32_t)zz * pb); };
        if (ix >= 0) img(ix, y, z);
        if (iy >= 0) img(x, iy, z);
        if (ix >= 0 && iy >= 0) img(ix, iy, z);
        if (iz >= 0) {
            img(x, y, iz);
            if (ix >= 0) img(ix, y, iz);
            if (iy >= 0) img(x, iy, iz);
            if (ix >= 0 && iy >= 0) img(ix, iy, iz);
        }
    };
#pragma unroll
    for (int k = 0; k < KP; ++k) {
        const int x = px[k], y = py[k], z = pz[k];
        if (y > g.n1 || z > g.n2) continue;
        // the first-colour partner: the other cell of the pair (x0+2i, x0+2i+1)
        const int xr = ((x - t.x0) & 1) ? x - 1 : x + 1;
        const int l = (x - xb) + LX * ((y - yb) + LY * (z - zb));
        if (x <= g.n0) store(update(l, x, y, z, pr[k]), x, y, z);
        if (xr <= g.n0) store(sp[l + (xr - x)], xr, y, z);
    }
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

// Ghost fill of a refined block grid's level from its expression table
// (every ghost is a copy, an 8-cell average or a 2:1 extrapolation of
// interior cells, which this kernel never writes).
template <typename T>
__global__ __launch_bounds__(256) void k_gather_expr(T *__restrict__ phi, const GhostExpr *__restrict__ ex, int n,
                                                     const int *__restrict__ done) {
    if (done && *done) return;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const GhostExpr e = ex[t];
        T v;
        if (e.kind == 0) {
            v = phi[e.src[0]];
        } else if (e.kind == 1) {
            T s = phi[e.src[0]];
#pragma unroll
            for (int q = 1; q < 8; ++q) s = s + phi[e.src[q]];
            v = T(0.125) * s;
        } else {
            v = ((T(2) * phi[e.src[0]] + T(3) * phi[e.src[1]]) - phi[e.src[2]]) / T(4);
        }
        phi[e.dst] = v;
    }
}

}  // namespace mf3d
