// mg_velocity.hpp -- the velocity form of newsolv_mg on gfx950: pressure
// update, the projection of the predicted velocity (calvel) fused with its
// divergence (calmas), and the ghost maps of exchange(1..4) + pbound.
//
// Fields u*, v*, w*, u, v, w, p and the cell density live in level-0
// arenas (one ghost layer, i fastest, blocks one after another).  A face
// field at index i holds the face east of (north of, above) cell i.
//
// Arithmetic follows press.for operand by operand (the library is built
// with -ffp-contract=off):
//   calvel (press.for:150-188): u = u* - dt*alfapr*(p(i+1)-p(i))/dx
//           [/(0.5*(rho(i)+rho(i+1))) under LSM], dt*alfapr rounded first;
//   calmas (press.for:42-62):   fact = ((du/dx + dv/dy) + dw/dz),
//           su = fact/dt, LSM sup = fact*dx*dy*dz/dt, rmax = max|fact*dx*dy*dz|.
#pragma once
#include "mg_common.hpp"
#include "mg_sweep.hpp"

namespace mf3d {

// Per block and axis: the last face the block computes (n, or n-1 before a
// physical non-periodic wall, localparameters.for:185-190) and the block
// whose last face its ghost face below the first cell receives in
// exchange(1..3) (face neighbour, periodic partner, or the block itself
// when it is alone along a periodic axis; -1 before a wall, where the
// caller's face stays).
struct VelBlock {
    int ie[3];
    int lo_src[3];
};

// p += pp over every ghosted entry (newsolv_mg.for:38-51; ppref is 0).
template <typename TE>
__global__ __launch_bounds__(256) void k_vel_pupdate(double *__restrict__ p, const TE *__restrict__ pp,
                                                     long long n, const int *__restrict__ done) {
    if (done && *done) return;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = p[i] + (double)pp[i];
}

// A composed ghost map: entry t copies the value its source held before
// the map into its destination (both element offsets into one arena).
// k_map_stage reads every source, k_map_scatter then writes every
// destination, so a source that is itself a destination is read first.
// Up to three arenas share one launch: entries [0, c1) go to a0,
// [c1, c2) to a1, [c2, c3) to a2.
struct MapArgs {
    double *a[3];
    long long c[3];  // cumulative entry counts
};

__global__ __launch_bounds__(256) void k_map_stage(MapArgs m, const uint2 *__restrict__ ent,
                                                   double *__restrict__ stage, const int *__restrict__ done) {
    if (done && *done) return;
    const long long n = m.c[2];
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const double *a = t < m.c[0] ? m.a[0] : (t < m.c[1] ? m.a[1] : m.a[2]);
        stage[t] = a[ent[t].y];
    }
}

__global__ __launch_bounds__(256) void k_map_scatter(MapArgs m, const uint2 *__restrict__ ent,
                                                     const double *__restrict__ stage, const int *__restrict__ done) {
    if (done && *done) return;
    const long long n = m.c[2];
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        double *a = t < m.c[0] ? m.a[0] : (t < m.c[1] ? m.a[1] : m.a[2]);
        a[ent[t].x] = stage[t];
    }
}

struct VelFields {
    const double *us, *vs, *ws;  // u*, v*, w*
    double *u, *v, *w;
    const double *p;
    const double *dens;  // cell density (LSM), ghosted
};

// One face of component A of block g at the cell with arena offset c and
// index i along A: computed from u* and p when i <= ie, else the stored
// face (a wall face the caller owns).
template <int A, bool LSM>
__device__ __forceinline__ double vel_face(const VelFields &F, const BlockLevel &g, int ie, long long c, int i,
                                           double dta) {
    const double *st = A == 0 ? F.us : (A == 1 ? F.vs : F.ws);
    const double *cur = A == 0 ? F.u : (A == 1 ? F.v : F.w);
    if (i > ie) return cur[c];
    const long long s = A == 0 ? 1 : (A == 1 ? g.sy : g.sz);
    const double h = A == 0 ? g.h0 : (A == 1 ? g.h1 : g.h2);
    double gr = (dta * (F.p[c + s] - F.p[c])) / h;
    if (LSM) gr = gr / (0.5 * (F.dens[c] + F.dens[c + s]));
    return st[c] - gr;
}

// The ghost face below the first cell of component A at transverse cell
// (i1, i2) (the two other indices, in x, y, z order): the source block's
// last face, recomputed exactly as that block computes it, or the stored
// face before a wall.
template <int A, bool LSM>
__device__ __forceinline__ double vel_low_face(const VelFields &F, const BlockLevel *__restrict__ bls,
                                               const VelBlock *__restrict__ vbs, const BlockLevel &g, int b,
                                               int i1, int i2, double dta) {
    const int s = vbs[b].lo_src[A];
    if (s < 0) {
        const double *cur = A == 0 ? F.u : (A == 1 ? F.v : F.w);
        const long long c0 = A == 0 ? gidx(g, 0, i1, i2) : (A == 1 ? gidx(g, i1, 0, i2) : gidx(g, i1, i2, 0));
        return cur[c0];
    }
    const BlockLevel gs = bls[s];
    const int ns = A == 0 ? gs.n0 : (A == 1 ? gs.n1 : gs.n2);
    const long long cs = A == 0 ? gidx(gs, ns, i1, i2) : (A == 1 ? gidx(gs, i1, ns, i2) : gidx(gs, i1, i2, ns));
    return vel_face<A, LSM>(F, gs, vbs[s].ie[A], cs, ns, dta);
}

// calvel fused with calmas over the finest level: each thread owns one
// (x, y) column of a 64 x TY tile and marches its z chunk.  Writes the
// faces the block computes (its own east, north and top faces), the
// V-cycle right-hand side su (sup under LSM) in the inner precision TR,
// optionally su in FP64 (su64), and per workgroup max|fact*dx*dy*dz|.
// The west face of a column comes from the lane to its left, the south
// face is recomputed, the bottom face is carried from the plane below.
template <typename TR, bool LSM, int TY>
__global__ __launch_bounds__(SW_TX *TY) void k_velmas(VelFields F, TR *__restrict__ rhs, double *__restrict__ su64,
                                                      const SolveState *__restrict__ st,
                                                      const BlockLevel *__restrict__ bls,
                                                      const VelBlock *__restrict__ vbs,
                                                      const Tile *__restrict__ tiles, double *__restrict__ partials,
                                                      const int *__restrict__ done) {
    if (done && *done) return;
    const double dta = st->dta, dt = st->dt;
    const Tile t = tiles[blockIdx.x];
    const BlockLevel g = bls[t.b];
    const VelBlock vb = vbs[t.b];
    const int x = t.x0 + threadIdx.x, y = t.y0 + threadIdx.y;
    const bool act = x <= g.n0 && y <= g.n1;
    // inactive lanes work on a clamped column so the lane shuffle sees
    // every lane; nothing they compute is stored
    const int xc = act ? x : 1, yc = act ? y : 1;
    const double dx = g.h0, dy = g.h1, dz = g.h2;
    double rm = 0.0;
    double wB;
    {
        const int z = t.z0;
        if (z == 1) wB = vel_low_face<2, LSM>(F, bls, vbs, g, t.b, xc, yc, dta);
        else wB = vel_face<2, LSM>(F, g, vb.ie[2], gidx(g, xc, yc, z - 1), z - 1, dta);
    }
    const int lane = threadIdx.x;  // SW_TX == wavefront size: one row per wave
    for (int z = t.z0; z <= t.z1; ++z) {
        const long long c = gidx(g, xc, yc, z);
        const double uE = vel_face<0, LSM>(F, g, vb.ie[0], c, xc, dta);
        const double vN = vel_face<1, LSM>(F, g, vb.ie[1], c, yc, dta);
        const double wT = vel_face<2, LSM>(F, g, vb.ie[2], c, z, dta);
        double uW = __shfl_up(uE, 1);
        if (lane == 0) {
            uW = xc == 1 ? vel_low_face<0, LSM>(F, bls, vbs, g, t.b, yc, z, dta)
                         : vel_face<0, LSM>(F, g, vb.ie[0], c - 1, xc - 1, dta);
        }
        const double vS = yc == 1 ? vel_low_face<1, LSM>(F, bls, vbs, g, t.b, xc, z, dta)
                                  : vel_face<1, LSM>(F, g, vb.ie[1], c - g.sy, yc - 1, dta);
        const double fact = ((uE - uW) / dx + (vN - vS) / dy) + (wT - wB) / dz;
        const double vol = ((fact * dx) * dy) * dz;
        const double su = LSM ? vol / dt : fact / dt;
        if (act) {
            if (xc <= vb.ie[0]) F.u[c] = uE;
            if (yc <= vb.ie[1]) F.v[c] = vN;
            if (z <= vb.ie[2]) F.w[c] = wT;
            if (rhs) rhs[c] = (TR)su;
            if (su64) su64[c] = su;
            rm = fmax(rm, fabs(vol));
        }
        wB = wT;
    }
    if (!partials) return;
    __shared__ double sm[SW_TX * TY];
    const int tid = threadIdx.x + SW_TX * threadIdx.y;
    sm[tid] = rm;
    __syncthreads();
    for (int w = SW_TX * TY / 2; w > 0; w >>= 1) {
        if (tid < w) sm[tid] = fmax(sm[tid], sm[tid + w]);
        __syncthreads();
    }
    if (tid == 0) partials[blockIdx.x] = sm[0];
}

__global__ void k_vel_reset(SolveState *st, double eps, double dta, double dt) {
    st->tol = eps;
    st->dta = dta;
    st->dt = dt;
    st->done = 0;
    st->k = 0;
    st->last = -1;
}

// rmax of outer step st->k from the workgroup maxima; the loop ends once
// rmax < eps (st->tol) after at least two cycles (newsolv_mg.for:55).
__global__ __launch_bounds__(1024) void k_vel_finalize(const double *__restrict__ partials, int nparts,
                                                       SolveState *st, double *rmaxh) {
    if (st->done) return;
    __shared__ double sm[1024];
    const int t = threadIdx.x;
    double m = 0.0;
    for (int q = t; q < nparts; q += 1024) m = fmax(m, partials[q]);
    sm[t] = m;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (t < w) sm[t] = fmax(sm[t], sm[t + w]);
        __syncthreads();
    }
    if (t == 0) {
        const int k = st->k;
        rmaxh[k] = sm[0];
        st->last = k;
        st->k = k + 1;
        if (sm[0] < st->tol && k > 1) st->done = 1;
    }
}

}  // namespace mf3d
