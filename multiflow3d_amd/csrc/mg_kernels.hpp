// mg_kernels.hpp -- gfx950 kernels of the pressure-Poisson multigrid.
//
// Field layout in HBM: per level, one arena per field holding every block's
// ghosted array ((n0+2)*(n1+2)*(n2+2) values, i fastest), blocks one after
// another.  A BlockLevel record gives a block's extent, strides and arena
// offset at one level.  Kernels walk a per-level tile table, so one launch
// covers every block of a batch whatever their extents.
//
// Arithmetic follows the reference operand by operand (mgsolver.for:563-569
// smoother, :702-709 residual, :743-747 restriction, :833-887 prolongation);
// this file is compiled with -ffp-contract=off so the FP64 path rounds
// exactly as the reference does.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mf3d {

enum : int { kNeuW = 1, kNeuE = 2, kNeuS = 4, kNeuN = 8, kNeuB = 16, kNeuT = 32 };

struct BlockLevel {
    int n0, n1, n2;  // interior cells
    int neu;         // faces that are physical Neumann walls (kNeu* bits)
    long long sy;    // ghosted stride of j
    long long sz;    // ghosted stride of k
    long long off;   // element offset of this block in the level arena
    long long poff;  // offset of this block in interior-packed vectors
    double c0, c1, c2;  // 1/h_l^2 per axis (constant coefficients)
    float f0, f1, f2;   // the same in FP32
    int lsm;            // rmax measured without the cell volume
    double h0, h1, h2;  // finest dx, dy, dz
    double gx, gy, gz;  // LSM face factors dy*dz/dx, dx*dz/dy, dx*dy/dz (finest h)
};

struct Tile {
    int b, x0, y0, z;
};

// One ghost region of one block: every destination cell in the box takes
// the value of a source cell; per axis the source coordinate is either the
// destination coordinate (map < 0) or the constant map.
struct Region {
    int dst_b, src_b;
    int lo[3];
    int cnt[3];
    int map[3];
};

struct SolveState {
    double fnorm;
    int done;
    int last;
};

__device__ __forceinline__ long long gidx(const BlockLevel &g, int i, int j, int k) {
    return g.off + (long long)i + g.sy * j + g.sz * k;
}

template <typename T> __device__ __forceinline__ T cconst(const BlockLevel &g, int d);
template <> __device__ __forceinline__ double cconst<double>(const BlockLevel &g, int d) {
    return d == 0 ? g.c0 : (d == 1 ? g.c1 : g.c2);
}
template <> __device__ __forceinline__ float cconst<float>(const BlockLevel &g, int d) {
    return d == 0 ? g.f0 : (d == 1 ? g.f1 : g.f2);
}

// The six neighbour coefficients of cell (i,j,k) and ap (mgsolver.for:184-208).
// Constant form: 1/h^2 with a zero on a physical Neumann face.  Variable
// form: face arrays cx, cy, cz where cx at (i,j,k) is aw of cell i and cx at
// (n0+1,j,k) is ae of the last cell (aw(i+1) == ae(i) bitwise inside a block).
template <typename T, bool VAR>
__device__ __forceinline__ void coefs(const BlockLevel &g, const T *__restrict__ cx,
                                      const T *__restrict__ cy, const T *__restrict__ cz,
                                      long long q, int i, int j, int k, T &aw, T &ae, T &as,
                                      T &an, T &ab, T &at, T &ap) {
    if (VAR) {
        aw = cx[q];
        ae = cx[q + 1];
        as = cy[q];
        an = cy[q + g.sy];
        ab = cz[q];
        at = cz[q + g.sz];
    } else {
        const T a = cconst<T>(g, 0), b = cconst<T>(g, 1), c = cconst<T>(g, 2);
        aw = (i == 1 && (g.neu & kNeuW)) ? T(0) : a;
        ae = (i == g.n0 && (g.neu & kNeuE)) ? T(0) : a;
        as = (j == 1 && (g.neu & kNeuS)) ? T(0) : b;
        an = (j == g.n1 && (g.neu & kNeuN)) ? T(0) : b;
        ab = (k == 1 && (g.neu & kNeuB)) ? T(0) : c;
        at = (k == g.n2 && (g.neu & kNeuT)) ? T(0) : c;
    }
    T s = aw + ae;
    s = s + as;
    s = s + an;
    s = s + ab;
    s = s + at;
    ap = -s;
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill(T *__restrict__ a, long long n, T v) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; t < n; t += stride) a[t] = v;
}

// One colour of a red-black sweep (mgsolver.for:549-601).  Tiles are 128
// cells along i by 4 rows; each thread updates the one cell of its pair that
// has the colour.  parity 1 is the reference's first colour
// (mod(i+j+k,2)==0 on its 1-based ghosted indices).
template <typename T, bool VAR>
__global__ __launch_bounds__(256) void k_rb(T *__restrict__ phi, const T *__restrict__ rhs,
                                            const T *__restrict__ cx, const T *__restrict__ cy,
                                            const T *__restrict__ cz,
                                            const BlockLevel *__restrict__ bls,
                                            const Tile *__restrict__ tiles, int parity) {
    const Tile t = tiles[blockIdx.x];
    const BlockLevel g = bls[t.b];
    const int j = t.y0 + threadIdx.y, k = t.z;
    if (j > g.n1) return;
    const int i = t.x0 + 2 * threadIdx.x + ((parity - j - k - t.x0) & 1);
    if (i > g.n0) return;
    const long long q = gidx(g, i, j, k);
    T aw, ae, as, an, ab, at, ap;
    coefs<T, VAR>(g, cx, cy, cz, q, i, j, k, aw, ae, as, an, ab, at, ap);
    T s = aw * phi[q - 1];
    s = s + ae * phi[q + 1];
    s = s + as * phi[q - g.sy];
    s = s + an * phi[q + g.sy];
    s = s + ab * phi[q - g.sz];
    s = s + at * phi[q + g.sz];
    phi[q] = (rhs[q] - s) / ap;
}

// Ghost fill from a region table: exchangepp + mgbound + exchbc_mgpp of the
// reference in one gather (every ghost takes the value of the interior cell
// it stands for), or the mgbound-only Neumann copy after prolongation.
template <typename T>
__global__ __launch_bounds__(256) void k_gather(T *__restrict__ phi,
                                                const BlockLevel *__restrict__ bls,
                                                const Region *__restrict__ regs) {
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
// coarse interior cell, children summed in the reference's order.
template <typename T, bool VAR>
__global__ __launch_bounds__(256) void k_restrict(
    const T *__restrict__ phif, const T *__restrict__ rhsf, const T *__restrict__ cx,
    const T *__restrict__ cy, const T *__restrict__ cz, const BlockLevel *__restrict__ blf,
    T *__restrict__ rhsc, const BlockLevel *__restrict__ blc, const Tile *__restrict__ tiles) {
    const Tile t = tiles[blockIdx.x];
    const BlockLevel gc = blc[t.b];
    const int i = t.x0 + threadIdx.x, j = t.y0 + threadIdx.y, k = t.z;
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
// layer (mgsolver.for:823-891); one thread per fine ghosted cell.  The eight
// coarse terms are summed in the reference's order c, c+j, c+k, c+jk, c+i,
// c+ij, c+ik, c+ijk with weight 3 per axis where the term sits on the
// child's side and 1 otherwise.
template <typename T>
__global__ __launch_bounds__(256) void k_prolong(T *__restrict__ phif,
                                                 const BlockLevel *__restrict__ blf,
                                                 const T *__restrict__ phic,
                                                 const BlockLevel *__restrict__ blc,
                                                 const Tile *__restrict__ tiles) {
    const Tile t = tiles[blockIdx.x];
    const BlockLevel gf = blf[t.b];
    const int i = t.x0 + threadIdx.x, j = t.y0 + threadIdx.y, k = t.z;
    if (i > gf.n0 + 1 || j > gf.n1 + 1) return;
    const BlockLevel gc = blc[t.b];
    const int ic = i >> 1, jc = j >> 1, kc = k >> 1;
    const int fi = i & 1, fj = j & 1, fk = k & 1;
    const long long c = gidx(gc, ic, jc, kc);
    const T wi0 = fi ? T(1) : T(3), wi1 = fi ? T(3) : T(1);
    const T wj0 = fj ? T(1) : T(3), wj1 = fj ? T(3) : T(1);
    const T wk0 = fk ? T(1) : T(3), wk1 = fk ? T(3) : T(1);
    T s = (wi0 * wj0 * wk0) * phic[c];
    s = s + (wi0 * wj1 * wk0) * phic[c + gc.sy];
    s = s + (wi0 * wj0 * wk1) * phic[c + gc.sz];
    s = s + (wi0 * wj1 * wk1) * phic[c + gc.sy + gc.sz];
    s = s + (wi1 * wj0 * wk0) * phic[c + 1];
    s = s + (wi1 * wj1 * wk0) * phic[c + 1 + gc.sy];
    s = s + (wi1 * wj0 * wk1) * phic[c + 1 + gc.sz];
    s = s + (wi1 * wj1 * wk1) * phic[c + 1 + gc.sy + gc.sz];
    const long long q = gidx(gf, i, j, k);
    phif[q] = phif[q] + s / T(64);
}

}  // namespace mf3d

namespace mf3d {

// Fixed-shape tree reduction of (sum, max) over a 256-thread workgroup;
// the order never changes, so results are bitwise reproducible run to run.
__device__ __forceinline__ void block_sum_max(double s, double m, double *out) {
    __shared__ double ss[256], sm[256];
    const int t = threadIdx.x + threadIdx.y * blockDim.x;
    ss[t] = s;
    sm[t] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            ss[t] = ss[t] + ss[t + w];
            sm[t] = fmax(sm[t], sm[t + w]);
        }
        __syncthreads();
    }
    if (t == 0) {
        out[0] = ss[0];
        out[1] = sm[0];
    }
}

// FP64 outer step of the defect-correction loop (newsolv_mg.for:38-54 in
// Poisson form): p_new = p_old + e over every ghosted cell (both operands
// carry mapped ghosts, so p_new does too), and on interior cells
// r = f - A p_new with each neighbour's p_new formed on the fly.  r becomes
// the next cycle's level-0 RHS (rounded to the inner precision TE);
// per-workgroup partial sums of r^2 and max |r| dV go to `partials`.
template <typename TE, bool VAR>
__global__ __launch_bounds__(256) void k_outer(
    const double *__restrict__ pold, double *__restrict__ pnew, const TE *__restrict__ e,
    const double *__restrict__ f, TE *__restrict__ rhs0, const double *__restrict__ cx,
    const double *__restrict__ cy, const double *__restrict__ cz,
    const BlockLevel *__restrict__ bls, const Tile *__restrict__ tiles,
    double *__restrict__ partials) {
    const Tile t = tiles[blockIdx.x];
    const BlockLevel g = bls[t.b];
    const int i = t.x0 + threadIdx.x, j = t.y0 + threadIdx.y, k = t.z;
    double r2 = 0.0, ra = 0.0;
    if (i <= g.n0 + 1 && j <= g.n1 + 1) {
        const long long q = gidx(g, i, j, k);
        const double pn = pold[q] + (double)e[q];
        pnew[q] = pn;
        if (i >= 1 && i <= g.n0 && j >= 1 && j <= g.n1 && k >= 1 && k <= g.n2) {
            double aw, ae, as, an, ab, at, ap;
            coefs<double, VAR>(g, cx, cy, cz, q, i, j, k, aw, ae, as, an, ab, at, ap);
            double s = aw * (pold[q - 1] + (double)e[q - 1]);
            s = s + ae * (pold[q + 1] + (double)e[q + 1]);
            s = s + as * (pold[q - g.sy] + (double)e[q - g.sy]);
            s = s + an * (pold[q + g.sy] + (double)e[q + g.sy]);
            s = s + ab * (pold[q - g.sz] + (double)e[q - g.sz]);
            s = s + at * (pold[q + g.sz] + (double)e[q + g.sz]);
            s = s + ap * pn;
            const double r = f[q] - s;
            rhs0[q] = (TE)r;
            r2 = r * r;
            ra = g.lsm ? fabs(r) : fabs(r * g.h0 * g.h1 * g.h2);
        }
    }
    block_sum_max(r2, ra, partials + 2 * (long long)blockIdx.x);
}

// Reduce the partials in a fixed order; record rho_k = ||r_k|| / ||f|| and
// rmax_k; raise `done` when converged.
__global__ __launch_bounds__(1024) void k_finalize(const double *__restrict__ partials, int nparts,
                                                   SolveState *st, double *hist, double *rmaxh,
                                                   int k, double tol, int min_cycles) {
    __shared__ double ss[1024], sm[1024];
    const int t = threadIdx.x;
    double s = 0.0, m = 0.0;
    for (int p = t; p < nparts; p += 1024) {
        s = s + partials[2 * (long long)p];
        m = fmax(m, partials[2 * (long long)p + 1]);
    }
    ss[t] = s;
    sm[t] = m;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (t < w) {
            ss[t] = ss[t] + ss[t + w];
            sm[t] = fmax(sm[t], sm[t + w]);
        }
        __syncthreads();
    }
    if (t == 0) {
        const double nr = sqrt(ss[0]);
        if (k == 0) {
            st->fnorm = nr > 0.0 ? nr : 1.0;
            st->done = 0;
        }
        const double rho = nr / st->fnorm;
        hist[k] = rho;
        rmaxh[k] = sm[0];
        st->last = k;
        if (k >= 1 && rho <= tol && k >= min_cycles) st->done = 1;
    }
}

// LSM density hierarchy (mgsolver.for:81-102): 8-cell average in the
// reference's child order (0,0,0) (0,1,0) (1,0,0) (1,1,0) (0,0,1) (0,1,1)
// (1,0,1) (1,1,1) as (di,dj,dk).
__global__ __launch_bounds__(256) void k_dens_restrict(const double *__restrict__ df,
                                                       const BlockLevel *__restrict__ blf,
                                                       double *__restrict__ dc,
                                                       const BlockLevel *__restrict__ blc,
                                                       const Tile *__restrict__ tiles) {
    const Tile t = tiles[blockIdx.x];
    const BlockLevel gc = blc[t.b];
    const int i = t.x0 + threadIdx.x, j = t.y0 + threadIdx.y, k = t.z;
    if (i > gc.n0 || j > gc.n1) return;
    const BlockLevel gf = blf[t.b];
    const long long q = gidx(gf, 2 * i - 1, 2 * j - 1, 2 * k - 1);
    double s = df[q];
    s = s + df[q + gf.sy];
    s = s + df[q + 1];
    s = s + df[q + 1 + gf.sy];
    s = s + df[q + gf.sz];
    s = s + df[q + gf.sy + gf.sz];
    s = s + df[q + 1 + gf.sz];
    s = s + df[q + 1 + gf.sy + gf.sz];
    dc[gidx(gc, i, j, k)] = 0.125 * s;
}

// LSM face coefficients (mgsolver.for:120-203) on one level: cx at
// (i,j,k), i = 1..n0, is aw of cell i = (dy*dz/dx)/rho_face with rho_face
// the mean of the two cells (the cell itself at the block edge); cx at
// i = n0+1 is ae of the last cell (one-sided).  A physical Neumann face
// gets 0.  The finest dx, dy, dz are used on every level, as the
// reference does.
__global__ __launch_bounds__(256) void k_lsm_coef(const double *__restrict__ d,
                                                  double *__restrict__ cx, double *__restrict__ cy,
                                                  double *__restrict__ cz,
                                                  const BlockLevel *__restrict__ bls,
                                                  const Tile *__restrict__ tiles) {
    const Tile t = tiles[blockIdx.x];
    const BlockLevel g = bls[t.b];
    const int i = t.x0 + threadIdx.x, j = t.y0 + threadIdx.y, k = t.z;
    if (i > g.n0 + 1 || j > g.n1 + 1) return;
    const long long q = gidx(g, i, j, k);
    const bool jin = j >= 1 && j <= g.n1, kin = k >= 1 && k <= g.n2, iin = i >= 1 && i <= g.n0;
    if (jin && kin && i >= 1) {
        double v;
        if (i <= g.n0) {
            const double d0 = d[q];
            const double dm = (i == 1) ? d0 : 0.5 * (d0 + d[q - 1]);
            v = (i == 1 && (g.neu & kNeuW)) ? 0.0 : g.gx / dm;
        } else {
            v = (g.neu & kNeuE) ? 0.0 : g.gx / d[q - 1];
        }
        cx[q] = v;
    }
    if (iin && kin && j >= 1) {
        double v;
        if (j <= g.n1) {
            const double d0 = d[q];
            const double dm = (j == 1) ? d0 : 0.5 * (d0 + d[q - g.sy]);
            v = (j == 1 && (g.neu & kNeuS)) ? 0.0 : g.gy / dm;
        } else {
            v = (g.neu & kNeuN) ? 0.0 : g.gy / d[q - g.sy];
        }
        cy[q] = v;
    }
    if (iin && jin && k >= 1) {
        double v;
        if (k <= g.n2) {
            const double d0 = d[q];
            const double dm = (k == 1) ? d0 : 0.5 * (d0 + d[q - g.sz]);
            v = (k == 1 && (g.neu & kNeuB)) ? 0.0 : g.gz / dm;
        } else {
            v = (g.neu & kNeuT) ? 0.0 : g.gz / d[q - g.sz];
        }
        cz[q] = v;
    }
}

template <typename TD, typename TS>
__global__ __launch_bounds__(256) void k_convert(TD *__restrict__ dst, const TS *__restrict__ src,
                                                 long long n) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; t < n; t += stride) dst[t] = (TD)src[t];
}

// Interior-packed vector <-> ghosted arena.
template <typename T>
__global__ __launch_bounds__(256) void k_unpack(const double *__restrict__ packed,
                                                T *__restrict__ arena,
                                                const BlockLevel *__restrict__ bls,
                                                const Tile *__restrict__ tiles) {
    const Tile t = tiles[blockIdx.x];
    const BlockLevel g = bls[t.b];
    const int i = t.x0 + threadIdx.x, j = t.y0 + threadIdx.y, k = t.z;
    if (i > g.n0 || j > g.n1) return;
    arena[gidx(g, i, j, k)] =
        (T)packed[g.poff + (i - 1) + (long long)g.n0 * ((j - 1) + (long long)g.n1 * (k - 1))];
}

template <typename T>
__global__ __launch_bounds__(256) void k_pack(const T *__restrict__ arena,
                                              double *__restrict__ packed,
                                              const BlockLevel *__restrict__ bls,
                                              const Tile *__restrict__ tiles) {
    const Tile t = tiles[blockIdx.x];
    const BlockLevel g = bls[t.b];
    const int i = t.x0 + threadIdx.x, j = t.y0 + threadIdx.y, k = t.z;
    if (i > g.n0 || j > g.n1) return;
    packed[g.poff + (i - 1) + (long long)g.n0 * ((j - 1) + (long long)g.n1 * (k - 1))] =
        (double)arena[gidx(g, i, j, k)];
}

}  // namespace mf3d
