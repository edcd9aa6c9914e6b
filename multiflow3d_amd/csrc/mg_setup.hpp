// mg_setup.hpp -- coefficient hierarchy, fills and layout conversions.
#pragma once
#include "mg_common.hpp"

namespace mf3d {

template <typename T>
__global__ __launch_bounds__(256) void k_fill(T *__restrict__ a, long long n, T v,
                                              const int *__restrict__ done) {
    if (done && *done) return;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; t < n; t += stride) a[t] = v;
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
    const int i = t.x0 + threadIdx.x, j = t.y0 + threadIdx.y, k = t.z0;
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

// Face-coefficient restriction (coarse_op = MF3D_COARSE_FACE; no
// counterpart in the reference): the coarse x face west of coarse cell
// (i,j,k) covers the fine x faces west of fine cells (2i-1, 2j-1..2j,
// 2k-1..2k); its coefficient is their sum / 16, summed with the lower
// tangential axis fastest.  i runs to n0+1 (the east face of the last
// cell), likewise j and k for the y and z faces, so the one-sided block-edge
// faces and the zero Neumann faces carry over.  Row tiles over the coarse
// ghosted range.
__global__ __launch_bounds__(256) void k_face_restrict(const double *__restrict__ fcx,
                                                       const double *__restrict__ fcy,
                                                       const double *__restrict__ fcz,
                                                       const BlockLevel *__restrict__ blf,
                                                       double *__restrict__ ccx, double *__restrict__ ccy,
                                                       double *__restrict__ ccz,
                                                       const BlockLevel *__restrict__ blc,
                                                       const Tile *__restrict__ tiles) {
    const Tile t = tiles[blockIdx.x];
    const BlockLevel gc = blc[t.b];
    const int j = t.y0 + threadIdx.y, k = t.z0;
    if (j > gc.n1 + 1) return;
    const BlockLevel gf = blf[t.b];
    for (int i = threadIdx.x; i <= gc.n0 + 1; i += blockDim.x) {
        const long long q = gidx(gc, i, j, k);
        const bool jin = j >= 1 && j <= gc.n1, kin = k >= 1 && k <= gc.n2, iin = i >= 1 && i <= gc.n0;
        const long long p = gidx(gf, 2 * i - 1, 2 * j - 1, 2 * k - 1);
        if (jin && kin && i >= 1) {
            double s = fcx[p] + fcx[p + gf.sy];
            s = s + fcx[p + gf.sz];
            s = s + fcx[p + gf.sy + gf.sz];
            ccx[q] = s * 0.0625;
        }
        if (iin && kin && j >= 1) {
            double s = fcy[p] + fcy[p + 1];
            s = s + fcy[p + gf.sz];
            s = s + fcy[p + 1 + gf.sz];
            ccy[q] = s * 0.0625;
        }
        if (iin && jin && k >= 1) {
            double s = fcz[p] + fcz[p + 1];
            s = s + fcz[p + gf.sy];
            s = s + fcz[p + 1 + gf.sy];
            ccz[q] = s * 0.0625;
        }
    }
}

// LSM face coefficients (mgsolver.for:120-203) on one level: cx at
// (i,j,k), i = 1..n0, is aw of cell i = (dy*dz/dx)/rho_face with rho_face
// the mean of the two cells (the cell itself at the block edge); cx at
// i = n0+1 is ae of the last cell (one-sided).  A physical Neumann face
// gets 0.  The finest dx, dy, dz are used on every level, as the
// reference does.  Entries outside these ranges are not written (they stay
// zero).
//
// A workgroup streams kCoefChunk consecutive entries of one block's
// ghosted arena (whole rows and planes, i fastest), each thread four
// entries 256 apart with all loads issued first; the chunks are ordered so
// that each XCD walks a contiguous range of planes, and the plane below
// (read for cz) is still in that XCD's L2.  With F32 the single-precision
// copies of the inner cycle are written in the same pass, every entry (0
// where the FP64 coefficient is not written), as a separate conversion
// of the whole arrays would.
constexpr int kCoefChunk = 1024;

template <bool F32>
__global__ __launch_bounds__(256) void k_lsm_coef(const double *__restrict__ d, double *__restrict__ cx,
                                                  double *__restrict__ cy, double *__restrict__ cz,
                                                  float *__restrict__ fx, float *__restrict__ fy,
                                                  float *__restrict__ fz, const BlockLevel *__restrict__ bls,
                                                  const Tile *__restrict__ chunks) {
    constexpr int U = kCoefChunk / 256;
    const Tile t = chunks[blockIdx.x];
    const BlockLevel g = bls[t.b];
    const int sy = (int)g.sy, ny2 = g.n1 + 2;
    const long long sz = g.sz;
    const int cells = (int)(g.sz * (g.n2 + 2));
    const int c0 = t.x0 + (int)threadIdx.x;
    int i = c0 % sy, r = c0 / sy;
    int j = r % ny2, k = r / ny2;
    int ci[U], cj[U], ck[U];
    double v0[U], vx[U], vy[U], vz[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        ci[u] = i;
        cj[u] = j;
        ck[u] = k;
        const int c = c0 + 256 * u;
        const bool ok = c < cells;
        const long long q = g.off + c;
        v0[u] = ok ? d[q] : 1.0;
        vx[u] = ok && i >= 1 ? d[q - 1] : 1.0;
        vy[u] = ok && j >= 1 ? d[q - sy] : 1.0;
        vz[u] = ok && k >= 1 ? d[q - sz] : 1.0;
        i += 256;
        while (i >= sy) {
            i -= sy;
            if (++j == ny2) {
                j = 0;
                ++k;
            }
        }
    }
    // all arithmetic before the first store: a load consumed after a store
    // would make the wait for it wait for that store as well
    double ox[U], oy[U], oz[U];
    bool wx[U], wy[U], wz[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int ii = ci[u], jj = cj[u], kk = ck[u];
        const bool ok = c0 + 256 * u < cells;
        const bool jin = jj >= 1 && jj <= g.n1, kin = kk >= 1 && kk <= g.n2, iin = ii >= 1 && ii <= g.n0;
        wx[u] = ok && jin && kin && ii >= 1;
        wy[u] = ok && iin && kin && jj >= 1;
        wz[u] = ok && iin && jin && kk >= 1;
        // i <= n0: the face mean (the cell itself at i = 1); i = n0+1: the
        // last cell (one-sided); 0 on a physical Neumann face
        const double mx = ii <= g.n0 ? (ii == 1 ? v0[u] : 0.5 * (v0[u] + vx[u])) : vx[u];
        const double my = jj <= g.n1 ? (jj == 1 ? v0[u] : 0.5 * (v0[u] + vy[u])) : vy[u];
        const double mz = kk <= g.n2 ? (kk == 1 ? v0[u] : 0.5 * (v0[u] + vz[u])) : vz[u];
        const bool zx = ii <= g.n0 ? (ii == 1 && (g.neu & kNeuW)) : (g.neu & kNeuE) != 0;
        const bool zy = jj <= g.n1 ? (jj == 1 && (g.neu & kNeuS)) : (g.neu & kNeuN) != 0;
        const bool zz = kk <= g.n2 ? (kk == 1 && (g.neu & kNeuB)) : (g.neu & kNeuT) != 0;
        ox[u] = zx ? 0.0 : g.gx / mx;
        oy[u] = zy ? 0.0 : g.gy / my;
        oz[u] = zz ? 0.0 : g.gz / mz;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const long long q = g.off + c0 + 256 * u;
        if (wx[u]) cx[q] = ox[u];
        if (wy[u]) cy[q] = oy[u];
        if (wz[u]) cz[q] = oz[u];
        if (F32 && c0 + 256 * u < cells) {
            fx[q] = wx[u] ? (float)ox[u] : 0.0f;
            fy[q] = wy[u] ? (float)oy[u] : 0.0f;
            fz[q] = wz[u] ? (float)oz[u] : 0.0f;
        }
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
    const int i = t.x0 + threadIdx.x, j = t.y0 + threadIdx.y, k = t.z0;
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
    const int i = t.x0 + threadIdx.x, j = t.y0 + threadIdx.y, k = t.z0;
    if (i > g.n0 || j > g.n1) return;
    packed[g.poff + (i - 1) + (long long)g.n0 * ((j - 1) + (long long)g.n1 * (k - 1))] =
        (double)arena[gidx(g, i, j, k)];
}

}  // namespace mf3d
