// mg_common.hpp -- data layout and shared device helpers of the gfx950
// pressure-Poisson multigrid kernels.
//
// Field layout in HBM: per level, one arena per field holding every block's
// ghosted array ((n0+2)*(n1+2)*(n2+2) values, i fastest), blocks one after
// another.  A BlockLevel record gives a block's extent, strides and arena
// offset at one level.  Kernels walk a per-level tile table, so one launch
// covers every block of a batch whatever their extents.  The potential has
// two arenas per level (ping-pong): a sweep reads one and writes the other.
//
// Arithmetic follows the reference operand by operand (mgsolver.for:563-569
// smoother, :702-709 residual, :743-747 restriction, :833-887 prolongation);
// the library is compiled with -ffp-contract=off so the FP64 path rounds
// exactly as the reference does.
//
// Every kernel that runs inside a captured cycle takes `done`, the device
// flag raised once the solve has converged: later cycles of a speculative
// launch then return at once.
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

// A unit of work: block b, x from x0, rows from y0, planes z0..z1.
struct Tile {
    int b, x0, y0, z0, z1;
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

// One ghost of a refined block grid's level: kind 0 copies src[0]; 1 takes
// 0.125 x (src[0] + ... + src[7]) summed in that order; 2 takes
// ((2 src[0] + 3 src[1]) - src[2]) / 4.  Offsets index the level arena;
// every src is an interior cell.
struct GhostExpr {
    int dst, kind;
    int src[8];
};

struct SolveState {
    double fnorm;
    int done;   // converged: later cycle kernels return at once
    int k;      // index of the next outer step
    int last;   // last outer step recorded
    int fk;     // GCR steps taken
    double tol; // stop tolerance of this solve (read by captured cycles)
    double alpha, beta, gamma;  // GCR: step, direction weight, (w, w)
    double dta, dt;             // velocity form: dt*alfapr and dt of this solve
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

// s + a*b in the inner cycle's arithmetic.  FP64 rounds the product and
// the sum separately, as the reference's expression does (bitwise parity);
// FP32, the mixed form's inner cycle, which only has to converge, contracts
// them into one fused multiply-add.  Every FP32 sweep kernel accumulates
// its stencil through this, in the same operand order, so the sweep
// variants stay bitwise equal to each other.
template <typename T> __device__ __forceinline__ T madd(T s, T a, T b);
template <> __device__ __forceinline__ double madd<double>(double s, double a, double b) { return s + a * b; }
template <> __device__ __forceinline__ float madd<float>(float s, float a, float b) {
    return __builtin_fmaf(a, b, s);
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
    ap = -s;   // -1.0*(...) of the reference: negation is exact
}

// Raw buffer resources (gfx9 layout, word 3 = 0x00020000): loads and
// stores take a per-thread 32-bit byte offset and a scalar byte offset, and
// the hardware returns zero for (and drops) accesses at or beyond `bytes`.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
template <typename T> __device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff);
template <> __device__ __forceinline__ double bload<double>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
template <> __device__ __forceinline__ float bload<float>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <typename T> __device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, T v, uint32_t voff, uint32_t soff);
template <> __device__ __forceinline__ void bstore<double>(__amdgpu_buffer_rsrc_t r, double v, uint32_t voff, uint32_t soff) {
    using V = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, v), r, voff, soff, 0);
}
template <> __device__ __forceinline__ void bstore<float>(__amdgpu_buffer_rsrc_t r, float v, uint32_t voff, uint32_t soff) {
    using V = decltype(__builtin_amdgcn_raw_buffer_load_b32(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(V, v), r, voff, soff, 0);
}

// Wavefront reductions over the 64 lanes by a butterfly: each step combines
// lanes l and l^o, so every lane ends with the same bits (the pairwise
// operations commute) and the order never changes -- results are bitwise
// reproducible run to run.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Workgroup reduction of two values over NT threads (a multiple of 64, at
// most 4096): a wavefront reduction, one LDS slot per wave, one barrier,
// and a wavefront reduction of those slots.  a is summed; b is summed, or
// (MAX) maximised -- of values >= 0.  The results are valid in the first
// wavefront (threads 0..63).  Every thread of the workgroup must call it.
template <int NT, bool MAX>
__device__ __forceinline__ void wg_reduce2(double &a, double &b) {
    constexpr int NW = NT / 64;
    static_assert(NW >= 1 && NW <= 64 && NW * 64 == NT, "workgroup of whole wavefronts");
    __shared__ double wa[NW], wb[NW];
    const int t = threadIdx.x + threadIdx.y * blockDim.x;
    a = wave_sum(a);
    b = MAX ? wave_max(b) : wave_sum(b);
    if ((t & 63) == 0) {
        wa[t >> 6] = a;
        wb[t >> 6] = b;
    }
    __syncthreads();
    if (t < 64) {
        a = t < NW ? wa[t] : 0.0;
        b = t < NW ? wb[t] : 0.0;
        a = wave_sum(a);
        b = MAX ? wave_max(b) : wave_sum(b);
    }
}

// Per-workgroup partials: (sum, sum) and (sum, max) of NT threads into out.
template <int NT>
__device__ __forceinline__ void block_sum2(double a, double b, double *out) {
    wg_reduce2<NT, false>(a, b);
    if (threadIdx.x + threadIdx.y * blockDim.x == 0) {
        out[0] = a;
        out[1] = b;
    }
}

template <int NT>
__device__ __forceinline__ void block_sum_max(double s, double m, double *out) {
    wg_reduce2<NT, true>(s, m);
    if (threadIdx.x + threadIdx.y * blockDim.x == 0) {
        out[0] = s;
        out[1] = m;
    }
}

}  // namespace mf3d
