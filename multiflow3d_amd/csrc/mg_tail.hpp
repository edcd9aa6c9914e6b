// mg_tail.hpp -- the coarse-level tail of the V-cycle in LDS.
//
// One workgroup per block runs everything from level lt down to the
// coarsest and back up to the post-smoothing of level lt -- sweeps, ghost
// exchange, restriction, prolongation + Neumann copy -- on level arrays held
// in LDS, then writes level lt's potential (with ghosts) to global memory.
// Only for blocks without neighbours or periodic partners (a lone block
// wraps onto itself), where the exchange is local to the workgroup.  Same
// operations, same order and same rounding as the per-level kernels.
#pragma once
#include "mg_coarse.hpp"
#include "mg_common.hpp"

namespace mf3d {

constexpr int TAIL_MAXL = 16;
constexpr int TAIL_NT = 1024;

template <typename T>
struct TailArgs {
    const T *cx[TAIL_MAXL];
    const T *cy[TAIL_MAXL];
    const T *cz[TAIL_MAXL];
};

// One level of one block inside the tail: extents, strides and the LDS
// offsets of its potential and RHS (both in the ghosted shape).
struct TLv {
    int n0, n1, n2, neu, sy, sz, po, ro;
};

__device__ __forceinline__ TLv tlv_of(const BlockLevel &g, int po, int ro) {
    TLv v;
    v.n0 = g.n0; v.n1 = g.n1; v.n2 = g.n2; v.neu = g.neu;
    v.sy = (int)g.sy; v.sz = (int)g.sz; v.po = po; v.ro = ro;
    return v;
}

// Visits cells c = start, start+step, ... of an a x b x d box in x-fastest
// order, keeping (i, j, k) by carries instead of divisions.
struct Walk {
    int i, j, k, si, sj, sk;
    __device__ Walk(int start, int step, int a, int b) {
        k = start / (a * b);
        const int r = start - k * a * b;
        j = r / a;
        i = r - j * a;
        sk = step / (a * b);
        const int rs = step - sk * a * b;
        sj = rs / a;
        si = rs - sj * a;
    }
    __device__ void next(int a, int b) {
        i += si;
        if (i >= a) { i -= a; ++j; }
        j += sj;
        if (j >= b) { j -= b; ++k; }
        k += sk;
    }
};

// Cells of one colour only: row (j,k) holds the cells i = i0, i0+2, ...
// with i0 = 1 + ((1 + j + k + parity) & 1); the items of all rows are
// walked as (row, m) with m < half = (n0 + 1) / 2.
template <typename T, bool VAR>
__device__ __forceinline__ void tail_colour(T *sm, const TLv v, const BlockLevel &g, const T *cx,
                                            const T *cy, const T *cz, int parity) {
    T *P = sm + v.po;
    const T *R = sm + v.ro;
    const int half = (v.n0 + 1) >> 1;
    const int N = half * v.n1 * v.n2;
    Walk w(threadIdx.x, TAIL_NT, half, v.n1);
    for (int c = threadIdx.x; c < N; c += TAIL_NT, w.next(half, v.n1)) {
        const int j = w.j + 1, k = w.k + 1;
        const int i = 1 + ((1 + j + k + parity) & 1) + 2 * w.i;
        if (i > v.n0) continue;
        const int q = i + v.sy * j + v.sz * k;
        T aw, ae, as, an, ab, at, ap;
        coefs<T, VAR>(g, cx, cy, cz, g.off + q, i, j, k, aw, ae, as, an, ab, at, ap);
        T s = aw * P[q - 1];
        s = madd(s, ae, P[q + 1]);
        s = madd(s, as, P[q - v.sy]);
        s = madd(s, an, P[q + v.sy]);
        s = madd(s, ab, P[q - v.sz]);
        s = madd(s, at, P[q + v.sz]);
        // FP64 divides as the reference does; the FP32 inner cycle multiplies by 1/ap
        P[q] = sizeof(T) == 8 ? (R[q] - s) / ap : (R[q] - s) * (T(1) / ap);
    }
}

// Every ghost takes the interior cell it stands for: wrap on a periodic
// axis, reflect on a Neumann one.
// Only ghost cells are visited: the two z planes (whole ghosted planes),
// then the two y rows of the interior planes, then the two x columns of the
// interior rows.  Sources are interior cells, so the order is free.
template <typename T>
__device__ __forceinline__ void tail_exchange(T *sm, const TLv v) {
    T *P = sm + v.po;
    const int a = v.n0 + 2, b = v.n1 + 2;
    const int G1 = 2 * a * b, G2 = G1 + 2 * a * v.n2, G3 = G2 + 2 * v.n1 * v.n2;
    for (int t = threadIdx.x; t < G3; t += TAIL_NT) {
        int i, j, k;
        if (t < G1) {
            const int u = t < a * b ? t : t - a * b;
            k = t < a * b ? 0 : v.n2 + 1;
            j = u / a;
            i = u - j * a;
        } else if (t < G2) {
            const int u0 = t - G1, side = u0 >= a * v.n2, u = side ? u0 - a * v.n2 : u0;
            j = side ? v.n1 + 1 : 0;
            k = 1 + u / a;
            i = u - (k - 1) * a;
        } else {
            const int u0 = t - G2, side = u0 >= v.n1 * v.n2, u = side ? u0 - v.n1 * v.n2 : u0;
            i = side ? v.n0 + 1 : 0;
            k = 1 + u / v.n1;
            j = 1 + u - (k - 1) * v.n1;
        }
        const int c = i + v.sy * j + v.sz * k;
        int si = i, sj = j, sk = k;
        if (i == 0) si = (v.neu & kNeuW) ? 1 : v.n0;
        if (i == v.n0 + 1) si = (v.neu & kNeuE) ? v.n0 : 1;
        if (j == 0) sj = (v.neu & kNeuS) ? 1 : v.n1;
        if (j == v.n1 + 1) sj = (v.neu & kNeuN) ? v.n1 : 1;
        if (k == 0) sk = (v.neu & kNeuB) ? 1 : v.n2;
        if (k == v.n2 + 1) sk = (v.neu & kNeuT) ? v.n2 : 1;
        P[c] = P[si + v.sy * sj + v.sz * sk];
    }
}

// Residual of the fine level restricted into the coarse RHS (:688-751); the
// coarse potential is zeroed including its ghosts (:650-657).
template <typename T, bool VAR>
__device__ __forceinline__ void tail_restrict(T *sm, const TLv vf, const BlockLevel &gf, const TLv vc,
                                              const T *cx, const T *cy, const T *cz) {
    const T *P = sm + vf.po;
    const T *R = sm + vf.ro;
    T *RC = sm + vc.ro;
    T *PC = sm + vc.po;
    const int MC = vc.sz * (vc.n2 + 2);
    for (int c = threadIdx.x; c < MC; c += TAIL_NT) PC[c] = T(0);
    const int N = vc.n0 * vc.n1 * vc.n2;
    Walk w(threadIdx.x, TAIL_NT, vc.n0, vc.n1);
    for (int c = threadIdx.x; c < N; c += TAIL_NT, w.next(vc.n0, vc.n1)) {
        const int i = w.i + 1, j = w.j + 1, k = w.k + 1;
        T acc = T(0);
        // children in the order (0,0,0) (1,0,0) (0,1,0) (1,1,0) (0,0,1) (1,0,1) (0,1,1) (1,1,1)
#pragma unroll
        for (int ch = 0; ch < 8; ++ch) {
            const int x = 2 * i - 1 + (ch & 1), y = 2 * j - 1 + ((ch >> 1) & 1), z = 2 * k - 1 + (ch >> 2);
            const int q = x + vf.sy * y + vf.sz * z;
            T aw, ae, as, an, ab, at, ap;
            coefs<T, VAR>(gf, cx, cy, cz, gf.off + q, x, y, z, aw, ae, as, an, ab, at, ap);
            T s = aw * P[q - 1];
            s = s + ae * P[q + 1];
            s = s + as * P[q - vf.sy];
            s = s + an * P[q + vf.sy];
            s = s + ab * P[q - vf.sz];
            s = s + at * P[q + vf.sz];
            s = s + ap * P[q];
            const T r = R[q] - s;
            acc = ch == 0 ? r : acc + r;
        }
        RC[i + vc.sy * j + vc.sz * k] = T(0.125) * acc;
    }
}

// Prolongation + correction of every fine cell including ghosts, then the
// Neumann copy (mgsolver.for:823-902).
template <typename T>
__device__ __forceinline__ void tail_prolong(T *sm, const TLv vf, const TLv vc) {
    T *PF = sm + vf.po;
    const T *PC = sm + vc.po;
    const int a = vf.n0 + 2, b = vf.n1 + 2, M = a * b * (vf.n2 + 2);
    {
        Walk w(threadIdx.x, TAIL_NT, a, b);
        for (int c = threadIdx.x; c < M; c += TAIL_NT, w.next(a, b)) {
            const int i = w.i, j = w.j, k = w.k;
            const int fi = i & 1, fj = j & 1, fk = k & 1;
            const T wi0 = fi ? T(1) : T(3), wi1 = fi ? T(3) : T(1);
            const T wj0 = fj ? T(1) : T(3), wj1 = fj ? T(3) : T(1);
            const T wk0 = fk ? T(1) : T(3), wk1 = fk ? T(3) : T(1);
            const int q = (i >> 1) + vc.sy * (j >> 1) + vc.sz * (k >> 1);
            T s = (wi0 * wj0 * wk0) * PC[q];
            s = s + (wi0 * wj1 * wk0) * PC[q + vc.sy];
            s = s + (wi0 * wj0 * wk1) * PC[q + vc.sz];
            s = s + (wi0 * wj1 * wk1) * PC[q + vc.sy + vc.sz];
            s = s + (wi1 * wj0 * wk0) * PC[q + 1];
            s = s + (wi1 * wj1 * wk0) * PC[q + 1 + vc.sy];
            s = s + (wi1 * wj0 * wk1) * PC[q + 1 + vc.sz];
            s = s + (wi1 * wj1 * wk1) * PC[q + 1 + vc.sy + vc.sz];
            PF[c] = PF[c] + s / T(64);
        }
    }
    __syncthreads();
    Walk w(threadIdx.x, TAIL_NT, a, b);
    for (int c = threadIdx.x; c < M; c += TAIL_NT, w.next(a, b)) {
        const int i = w.i, j = w.j, k = w.k;
        int si = i, sj = j, sk = k;
        if (i == 0 && (vf.neu & kNeuW)) si = 1;
        if (i == vf.n0 + 1 && (vf.neu & kNeuE)) si = vf.n0;
        if (j == 0 && (vf.neu & kNeuS)) sj = 1;
        if (j == vf.n1 + 1 && (vf.neu & kNeuN)) sj = vf.n1;
        if (k == 0 && (vf.neu & kNeuB)) sk = 1;
        if (k == vf.n2 + 1 && (vf.neu & kNeuT)) sk = vf.n2;
        if (si != i || sj != j || sk != k) PF[c] = PF[si + vf.sy * sj + vf.sz * sk];
    }
}

template <typename T, bool VAR>
__device__ __forceinline__ void tail_sweeps(T *sm, const TLv v, const BlockLevel &g, const T *cx,
                                            const T *cy, const T *cz, int n) {
    for (int s = 0; s < n; ++s) {
        tail_colour<T, VAR>(sm, v, g, cx, cy, cz, 1);
        __syncthreads();
        tail_colour<T, VAR>(sm, v, g, cx, cy, cz, 0);
        __syncthreads();
        tail_exchange<T>(sm, v);
        __syncthreads();
    }
}

// FP64 direct solve of the coarsest level inside the tail (one block): the
// level's RHS in cell order (i fastest) widened into sb, phi = Minv * rhs
// by one wave per row (coarse_row, as k_coarse_solve computes it), then the
// ghost fill.
template <typename T>
__device__ __forceinline__ void tail_direct(T *sm, const TLv v, const double *__restrict__ cinv, int nc,
                                            double *sb) {
    Walk w(threadIdx.x, TAIL_NT, v.n0, v.n1);
    for (int c = threadIdx.x; c < nc; c += TAIL_NT, w.next(v.n0, v.n1))
        sb[c] = (double)sm[v.ro + (w.i + 1) + v.sy * (w.j + 1) + v.sz * (w.k + 1)];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int r = wave; r < nc; r += TAIL_NT / 64) {
        const double s = coarse_row(cinv + (long long)r * nc, sb, nc, lane);
        if (lane == 0) {
            const int k = r / (v.n0 * v.n1), rem = r - k * v.n0 * v.n1, j = rem / v.n0, i = rem - j * v.n0;
            sm[v.po + (i + 1) + v.sy * (j + 1) + v.sz * (k + 1)] = (T)s;
        }
    }
    __syncthreads();
    tail_exchange<T>(sm, v);
    __syncthreads();
}

// The reference's sweeps of the coarsest level computed in FP64 inside an
// FP32 tail: the level's RHS widened into sd (ghosted layout, after the
// potential), the potential from zero, n sweeps with FP64 coefficients
// (cx..cz: the level's FP64 face coefficients under VAR), then every
// entry, ghosts included, narrowed back.
template <typename T, bool VAR>
__device__ __forceinline__ void tail_sweeps_fp64(T *sm, const TLv v, const BlockLevel &g, const double *cx,
                                                 const double *cy, const double *cz, int n, double *sd) {
    const int M = v.sz * (v.n2 + 2);
    TLv vd = v;
    vd.po = 0;
    vd.ro = M;
    for (int c = threadIdx.x; c < M; c += TAIL_NT) {
        sd[c] = 0.0;
        sd[M + c] = (double)sm[v.ro + c];
    }
    __syncthreads();
    tail_sweeps<double, VAR>(sd, vd, g, cx, cy, cz, n);
    for (int c = threadIdx.x; c < M; c += TAIL_NT) sm[v.po + c] = (T)sd[c];
    __syncthreads();
}

// cinv (nullptr: sweeps): the coarsest level's inverse operator, nc cells;
// c64: the coarsest level's sweeps in FP64, with its FP64 coefficients
// c64c[0..2] under VAR.  Their FP64 staging goes after the level arrays in
// LDS.
template <typename T>
struct TailC64 {
    int on;
    const double *cx, *cy, *cz;
};

template <typename T, bool VAR>
__global__ __launch_bounds__(TAIL_NT) void k_tail(const BlockLevel *__restrict__ tl, int nb, int nlev,
                                                  const T *__restrict__ rhs_in,
                                                  const T *__restrict__ phi_in, T *__restrict__ phi_out,
                                                  TailArgs<T> co, int nu_pre, int nu_post,
                                                  const double *__restrict__ cinv, int nc, TailC64<T> c64,
                                                  const int *__restrict__ done) {
    if (done && *done) return;
    extern __shared__ unsigned char smem_raw[];
    T *sm = reinterpret_cast<T *>(smem_raw);
    __shared__ TLv lv[TAIL_MAXL];
    const int b = blockIdx.x;
    if (threadIdx.x == 0) {
        int o = 0;
        for (int li = 0; li < nlev; ++li) {
            const BlockLevel &g = tl[li * nb + b];
            lv[li] = tlv_of(g, o, 0);
            o += (int)(g.sz * (g.n2 + 2));
        }
        for (int li = 0; li < nlev; ++li) {
            lv[li].ro = o;
            o += lv[li].sz * (lv[li].n2 + 2);
        }
    }
    __syncthreads();
    {
        const TLv v = lv[0];
        const long long goff = tl[b].off;
        const int M = v.sz * (v.n2 + 2);
        for (int c = threadIdx.x; c < M; c += TAIL_NT) {
            sm[v.ro + c] = rhs_in[goff + c];
            sm[v.po + c] = phi_in ? phi_in[goff + c] : T(0);
        }
    }
    __syncthreads();
    for (int li = 0; li < nlev; ++li) {
        if (li > 0) {
            tail_restrict<T, VAR>(sm, lv[li - 1], tl[(li - 1) * nb + b], lv[li], co.cx[li - 1], co.cy[li - 1],
                                  co.cz[li - 1]);
            __syncthreads();
        }
        if ((cinv || c64.on) && li == nlev - 1) {
            const TLv &v = lv[nlev - 1];
            const int end = v.ro + v.sz * (v.n2 + 2);  // the last level array ends the T arrays
            double *sb = reinterpret_cast<double *>(smem_raw + ((size_t)end * sizeof(T) + 7) / 8 * 8);
            if (cinv) tail_direct<T>(sm, v, cinv, nc, sb);
            else tail_sweeps_fp64<T, VAR>(sm, v, tl[li * nb + b], c64.cx, c64.cy, c64.cz, nu_pre + nu_post, sb);
        } else {
            tail_sweeps<T, VAR>(sm, lv[li], tl[li * nb + b], co.cx[li], co.cy[li], co.cz[li],
                                nu_pre + (li == nlev - 1 ? nu_post : 0));
        }
    }
    for (int li = nlev - 2; li >= 0; --li) {
        tail_prolong<T>(sm, lv[li], lv[li + 1]);
        __syncthreads();
        tail_sweeps<T, VAR>(sm, lv[li], tl[li * nb + b], co.cx[li], co.cy[li], co.cz[li], nu_post);
    }
    {
        const TLv v = lv[0];
        const long long goff = tl[b].off;
        const int M = v.sz * (v.n2 + 2);
        for (int c = threadIdx.x; c < M; c += TAIL_NT) phi_out[goff + c] = sm[v.po + c];
    }
}

}  // namespace mf3d
