// mg_velocity.hpp -- the velocity form of newsolv_mg on gfx950: pressure
// update, the projection of the predicted velocity (calvel) fused with its
// divergence (calmas), and the ghost maps of exchange(1..4) + pbound.
//
// Fields u*, v*, w*, u, v, w and p live in shell arenas: every block's
// array with the caller's G = pl ghost layers, i fastest, blocks one after
// another (a BlockLevel of the shell geometry addresses coordinates
// 1-G .. n+G).  The cell density and the V-cycle's RHS live in level-0
// arenas (one ghost layer).  A face field at index i holds the face east
// of (north of, above) cell i.
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

// p += pp over every ghosted entry (newsolv_mg.for:38-51; ppref is 0),
// shell and level-0 layouts equal (one ghost layer).
template <typename TE>
__global__ __launch_bounds__(256) void k_vel_pupdate(double *__restrict__ p, const TE *__restrict__ pp,
                                                     long long n, const int *__restrict__ done) {
    if (done && *done) return;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = p[i] + (double)pp[i];
}

// The same over the one-cell box of a wider shell: row tiles of the
// level-0 geometry (4 whole ghosted rows of one plane), p addressed in the
// shell geometry.
template <typename TE>
__global__ __launch_bounds__(256) void k_vel_pupdate_shell(double *__restrict__ p, const TE *__restrict__ pp,
                                                           const BlockLevel *__restrict__ bls,
                                                           const BlockLevel *__restrict__ wbls,
                                                           const Tile *__restrict__ tiles,
                                                           const int *__restrict__ done) {
    if (done && *done) return;
    const Tile t = tiles[blockIdx.x];
    const BlockLevel g = bls[t.b], gw = wbls[t.b];
    const int y = t.y0 + threadIdx.y, z = t.z0;
    if (y > g.n1 + 1) return;
    for (int x = threadIdx.x; x <= g.n0 + 1; x += blockDim.x) {
        const long long cw = gidx(gw, x, y, z);
        p[cw] = p[cw] + (double)pp[gidx(g, x, y, z)];
    }
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

// Operands of one face: value = s - (dt*alfapr*(p1 - p0))/h, divided by
// 0.5*(d0 + d1) under LSM (press.for:150-188).  A face the block does not
// compute (a wall face the caller owns) is loaded as s with `stored` set.
struct FaceOp {
    double s, p0, p1, d0, d1, h;
    bool stored;
};

template <bool LSM>
__device__ __forceinline__ double face_val(double s, double p0, double p1, double d0, double d1, double h,
                                           double dta) {
    double gr = (dta * (p1 - p0)) / h;
    if (LSM) gr = gr / (0.5 * (d0 + d1));
    return s - gr;
}

template <bool LSM>
__device__ __forceinline__ double face_val(const FaceOp &f, double dta) {
    return f.stored ? f.s : face_val<LSM>(f.s, f.p0, f.p1, f.d0, f.d1, f.h, dta);
}

// Operands of face A of a block at index i along A: shell offset c of the
// fields (geometry gw), level-0 offset cn of the density (geometry g).
template <int A, bool LSM>
__device__ __forceinline__ FaceOp own_face(const VelFields &F, const BlockLevel &g, const BlockLevel &gw, int ie,
                                           long long c, long long cn, int i) {
    FaceOp f;
    const long long st = A == 0 ? 1 : (A == 1 ? gw.sy : gw.sz);
    const long long sn = A == 0 ? 1 : (A == 1 ? g.sy : g.sz);
    f.h = A == 0 ? g.h0 : (A == 1 ? g.h1 : g.h2);
    f.stored = i > ie;
    const double *sv = f.stored ? (A == 0 ? F.u : (A == 1 ? F.v : F.w)) : (A == 0 ? F.us : (A == 1 ? F.vs : F.ws));
    f.s = sv[c];
    f.p0 = F.p[c];
    f.p1 = F.p[c + st];
    f.d0 = LSM ? F.dens[cn] : 1.0;
    f.d1 = LSM ? F.dens[cn + sn] : 1.0;
    return f;
}

template <int A, bool LSM>
__device__ __forceinline__ FaceOp own_face_at(const VelFields &F, const BlockLevel &g, const BlockLevel &gw, int ie,
                                              int x, int y, int z) {
    const int i = A == 0 ? x : (A == 1 ? y : z);
    return own_face<A, LSM>(F, g, gw, ie, gidx(gw, x, y, z), gidx(g, x, y, z), i);
}

// The ghost face below the first cell of component A at transverse cell
// (i1, i2) (the two other indices in x, y, z order): the source block's
// last face, computed exactly as that block computes it, or the stored
// face before a wall.
template <int A, bool LSM>
__device__ __forceinline__ FaceOp low_face(const VelFields &F, const BlockLevel *__restrict__ bls,
                                           const BlockLevel *__restrict__ wbls, const VelBlock *__restrict__ vbs,
                                           const BlockLevel &gw, int b, int i1, int i2) {
    const int s = vbs[b].lo_src[A];
    if (s < 0) {
        FaceOp f;
        const long long c0 = A == 0 ? gidx(gw, 0, i1, i2) : (A == 1 ? gidx(gw, i1, 0, i2) : gidx(gw, i1, i2, 0));
        f.s = (A == 0 ? F.u : (A == 1 ? F.v : F.w))[c0];
        f.p0 = f.p1 = 0.0;
        f.d0 = f.d1 = 1.0;
        f.h = 1.0;
        f.stored = true;
        return f;
    }
    const BlockLevel gs = bls[s], gsw = wbls[s];
    const int ns = A == 0 ? gs.n0 : (A == 1 ? gs.n1 : gs.n2);
    if (A == 0) return own_face_at<A, LSM>(F, gs, gsw, vbs[s].ie[A], ns, i1, i2);
    if (A == 1) return own_face_at<A, LSM>(F, gs, gsw, vbs[s].ie[A], i1, ns, i2);
    return own_face_at<A, LSM>(F, gs, gsw, vbs[s].ie[A], i1, i2, ns);
}

// calvel fused with calmas over the finest level.  A persistent grid walks
// the tile table (64 x TY columns, a chunk of planes); each thread owns one
// (x, y) column and marches the chunk.  Per tile the faces west of the
// tile's first column and south of its first row are computed up front
// into LDS (the block's own faces, a source block's last face, or the
// caller's wall face).  Per plane each thread computes its own east, north
// and top faces; the west face comes from the lane to its left, the south
// face from the row below through LDS, the bottom face from the plane
// below.  Writes the faces the block computes, the right-hand side su (sup
// under LSM) -- into the V-cycle's RHS in the inner precision TR, or in
// FP64 into su64 when SU64 -- and per workgroup max|fact*dx*dy*dz|.
//
// Every plane issues the same memory operations: the next plane's loads
// (the last plane of a chunk loads its own plane again) go out before this
// plane's stores, in two register sets used in turn, and a store a lane
// must not make (a wall face, a column outside the block) goes to the
// lane's own slot of `sink` instead.  With no path-dependent loads or
// stores the wait for a set covers exactly that set.
constexpr int VEL_TZ_MAX = 32;

template <typename TR, bool LSM, bool SU64, int TY>
__global__ __launch_bounds__(SW_TX *TY, LSM ? 4 : 6) void k_velmas(
    VelFields F, TR *__restrict__ rhs, double *__restrict__ su64, double *__restrict__ sink,
    const SolveState *__restrict__ st, const BlockLevel *__restrict__ bls, const BlockLevel *__restrict__ wbls,
    const VelBlock *__restrict__ vbs, const Tile *__restrict__ tiles, int ntiles, double *__restrict__ partials,
    const int *__restrict__ done) {
    if (done && *done) return;
    constexpr int NT = SW_TX * TY;
    __shared__ double sW[VEL_TZ_MAX][TY];      // west faces of the tile's first column
    __shared__ double sS[VEL_TZ_MAX][SW_TX];   // south faces of the tile's first row
    __shared__ double sN[2][TY][SW_TX];        // north faces of each row, per plane parity
    const double dta = st->dta, dt = st->dt;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + SW_TX * ty;
    double *const my_sink = sink + (long long)blockIdx.x * NT + tid;
    double rm = 0.0;
    for (int ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        const Tile t = tiles[ti];
        const BlockLevel g = bls[t.b], gw = wbls[t.b];
        const VelBlock vb = vbs[t.b];
        const int nz = t.z1 - t.z0 + 1;
        __syncthreads();  // the previous tile's halo reads are done
        // west halo: TY rows x nz planes
        for (int q = tid; q < TY * nz; q += NT) {
            const int yy = t.y0 + q % TY, zz = t.z0 + q / TY;
            if (yy <= g.n1) {
                const FaceOp f = t.x0 == 1 ? low_face<0, LSM>(F, bls, wbls, vbs, gw, t.b, yy, zz)
                                           : own_face_at<0, LSM>(F, g, gw, vb.ie[0], t.x0 - 1, yy, zz);
                sW[q / TY][q % TY] = face_val<LSM>(f, dta);
            }
        }
        // south halo: 64 columns x nz planes
        for (int q = tid; q < SW_TX * nz; q += NT) {
            const int xx = t.x0 + q % SW_TX, zz = t.z0 + q / SW_TX;
            if (xx <= g.n0) {
                const FaceOp f = t.y0 == 1 ? low_face<1, LSM>(F, bls, wbls, vbs, gw, t.b, xx, zz)
                                           : own_face_at<1, LSM>(F, g, gw, vb.ie[1], xx, t.y0 - 1, zz);
                sS[q / SW_TX][q % SW_TX] = face_val<LSM>(f, dta);
            }
        }
        const int x = t.x0 + tx, y = t.y0 + ty;
        const bool act = x <= g.n0 && y <= g.n1;
        // inactive lanes march a clamped column; what they compute goes to the sink
        const int xc = act ? x : g.n0, yc = act ? y : g.n1;
        const bool eown = xc == g.n0 || tx == SW_TX - 1;  // p(x+1) not in the lane to the right
        const bool estored = xc > vb.ie[0], nstored = yc > vb.ie[1];
        const double *esrc = estored ? F.u : F.us;
        const double *nsrc = nstored ? F.v : F.vs;
        double wB;  // bottom face of the first plane
        {
            const FaceOp f = t.z0 == 1 ? low_face<2, LSM>(F, bls, wbls, vbs, gw, t.b, xc, yc)
                                       : own_face_at<2, LSM>(F, g, gw, vb.ie[2], xc, yc, t.z0 - 1);
            wB = face_val<LSM>(f, dta);
        }
        // operands of one plane; p and rho at the cell come from the plane below
        struct Ops {
            double us, vs, ws, pe, pn, pz, de, dn, dz;
        };
        auto load = [&](Ops &o, int z) {
            const long long c = gidx(gw, xc, yc, z);
            const double *tsrc = z > vb.ie[2] ? F.w : F.ws;
            o.us = esrc[c];
            o.vs = nsrc[c];
            o.ws = tsrc[c];
            o.pe = F.p[c + 1];
            o.pn = F.p[c + gw.sy];
            o.pz = F.p[c + gw.sz];
            if (LSM) {
                const long long cn = gidx(g, xc, yc, z);
                o.de = F.dens[cn + 1];
                o.dn = F.dens[cn + g.sy];
                o.dz = F.dens[cn + g.sz];
            }
        };
        double pc, dc = 1.0;
        {
            pc = F.p[gidx(gw, xc, yc, t.z0)];
            if (LSM) dc = F.dens[gidx(g, xc, yc, t.z0)];
        }
        auto step = [&](const Ops &o, int z) {
            const long long c = gidx(gw, xc, yc, z), cn = gidx(g, xc, yc, z);
            double pe = __shfl_down(pc, 1);
            double de = LSM ? __shfl_down(dc, 1) : 1.0;
            pe = eown ? o.pe : pe;
            if (LSM) de = eown ? o.de : de;
            const bool tstored = z > vb.ie[2];
            const double uE = estored ? o.us : face_val<LSM>(o.us, pc, pe, dc, de, g.h0, dta);
            const double vN = nstored ? o.vs : face_val<LSM>(o.vs, pc, o.pn, dc, o.dn, g.h1, dta);
            const double wT = tstored ? o.ws : face_val<LSM>(o.ws, pc, o.pz, dc, o.dz, g.h2, dta);
            double uW = __shfl_up(uE, 1);
            uW = tx == 0 ? sW[z - t.z0][ty] : uW;
            sN[z & 1][ty][tx] = vN;
            __syncthreads();
            const double vS = ty == 0 ? sS[z - t.z0][tx] : sN[z & 1][ty - 1][tx];
            const double fact = ((uE - uW) / g.h0 + (vN - vS) / g.h1) + (wT - wB) / g.h2;
            const double vol = ((fact * g.h0) * g.h1) * g.h2;
            const double su = LSM ? vol / dt : fact / dt;
            *((act && !estored) ? F.u + c : my_sink) = uE;
            *((act && !nstored) ? F.v + c : my_sink) = vN;
            *((act && !tstored) ? F.w + c : my_sink) = wT;
            if (SU64) *(act ? su64 + c : my_sink) = su;
            else *(act ? rhs + cn : (TR *)my_sink) = (TR)su;
            rm = act ? fmax(rm, fabs(vol)) : rm;
            wB = wT;
            pc = o.pz;
            if (LSM) dc = o.dz;
        };
        Ops A, B;
        load(A, t.z0);
        __syncthreads();  // halos written
        for (int z = t.z0;; z += 2) {
            load(B, z + 1 <= t.z1 ? z + 1 : t.z1);
            step(A, z);
            if (z + 1 > t.z1) break;
            load(A, z + 2 <= t.z1 ? z + 2 : t.z1);
            step(B, z + 1);
            if (z + 2 > t.z1) break;
        }
    }
    if (!partials) return;
    double unused = 0.0;
    wg_reduce2<NT, true>(unused, rm);
    if (tid == 0) partials[blockIdx.x] = rm;
}

// calmas, 4th-order branch (differencing = 2, press.for:13-36), after
// calvel and exchange(1..3) have refreshed the faces it reads (u(i+1) of
// the last cell is a ghost face, u(i-2) of the first the second ghost
// layer): fact = sum over the axes of (-u(i+1) + 27 u(i) - 27 u(i-1) +
// u(i-2)) / (24 dx), in the reference's operand order; su = fact/dt into
// the V-cycle's RHS (inner precision TR, level-0 geometry) or, SU64, into
// su64 (shell geometry); per workgroup max|fact*dx*dy*dz| into partials.
// Workgroups stride over the level-0 cell tiles (64 x 4 cells of a plane).
template <typename TR, bool SU64>
__global__ __launch_bounds__(256) void k_calmas4(VelFields F, TR *__restrict__ rhs, double *__restrict__ su64,
                                                 const SolveState *__restrict__ st,
                                                 const BlockLevel *__restrict__ bls,
                                                 const BlockLevel *__restrict__ wbls, const Tile *__restrict__ tiles,
                                                 int ntiles, double *__restrict__ partials,
                                                 const int *__restrict__ done) {
    if (done && *done) return;
    __shared__ double sm[256];
    const double dt = st->dt;
    const int tid = threadIdx.x + 64 * threadIdx.y;
    double rm = 0.0;
    for (int ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        const Tile t = tiles[ti];
        const BlockLevel g = bls[t.b], gw = wbls[t.b];
        const int x = t.x0 + threadIdx.x, y = t.y0 + threadIdx.y, z = t.z0;
        if (x > g.n0 || y > g.n1) continue;
        const long long c = gidx(gw, x, y, z);
        const long long sy = gw.sy, sz = gw.sz;
        const double dx = g.h0, dy = g.h1, dz = g.h2;
        const double fx = (((-F.u[c + 1] + 27.0 * F.u[c]) - 27.0 * F.u[c - 1]) + F.u[c - 2]) / (24.0 * dx);
        const double fy =
            (((-F.v[c + sy] + 27.0 * F.v[c]) - 27.0 * F.v[c - sy]) + F.v[c - 2 * sy]) / (24.0 * dy);
        const double fz =
            (((-F.w[c + sz] + 27.0 * F.w[c]) - 27.0 * F.w[c - sz]) + F.w[c - 2 * sz]) / (24.0 * dz);
        const double fact = (fx + fy) + fz;
        const double vol = ((fact * dx) * dy) * dz;
        if (SU64) su64[c] = fact / dt;
        else rhs[gidx(g, x, y, z)] = (TR)(fact / dt);
        rm = fmax(rm, fabs(vol));
    }
    if (!partials) return;
    sm[tid] = rm;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
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
    const int t = threadIdx.x;
    double m = 0.0, unused = 0.0;
    for (int q = t; q < nparts; q += 1024) m = fmax(m, partials[q]);
    wg_reduce2<1024, true>(unused, m);
    if (t == 0) {
        const int k = st->k;
        rmaxh[k] = m;
        st->last = k;
        st->k = k + 1;
        if (m < st->tol && k > 1) st->done = 1;
    }
}

}  // namespace mf3d
