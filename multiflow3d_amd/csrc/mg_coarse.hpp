// mg_coarse.hpp -- FP64 direct solve of the coarsest level.
//
// The reference only smooths its coarsest grid (mgsolver.for:224-369 runs
// irestr + iproln sweeps there), which leaves most of the smoothest error
// in place and sets the V-cycle's contraction.  The mixed-precision form
// replaces those sweeps by an exact FP64 solve (the FP64 form keeps the
// reference's sweeps unless asked otherwise):
//   - once per coefficient build, the coarsest operator -- the same
//     7-point rows the smoother relaxes, with every ghost replaced by the
//     interior cell the exchange copies into it -- is assembled as a dense
//     nc x nc FP64 matrix and inverted in place by Gauss-Jordan elimination;
//   - every cycle, phi = Minv * rhs (FP64 products, the coarse RHS widened
//     from the inner precision), then the usual ghost fill.
// Pressure has only periodic and Neumann faces, so the operator is singular
// (constants): one cell per connected block group is pinned to zero (its
// row becomes the identity row for the elimination and is zeroed in the
// inverse, its column is dropped).  What remains is an
// irreducibly diagonally dominant M-matrix, for which elimination without
// pivoting is stable.  A right-hand side slightly outside the range (from
// rounding) only leaves its defect in the pinned rows.
#pragma once
#include "mg_common.hpp"

namespace mf3d {

// One coarsest-level cell: its block and interior index, the matrix columns
// of its W, E, S, N, B, T neighbours (a ghost neighbour is the cell the
// exchange copies into it), and whether it is pinned.
struct CCell {
    int b, i, j, k;
    int nb[6];
    int pin;
};

// Row r of the operator into a zeroed matrix.
template <bool VAR>
__global__ __launch_bounds__(256) void k_coarse_assemble(double *__restrict__ M, int nc,
                                                         const CCell *__restrict__ cells,
                                                         const BlockLevel *__restrict__ bls,
                                                         const double *__restrict__ cx,
                                                         const double *__restrict__ cy,
                                                         const double *__restrict__ cz) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nc) return;
    double *row = M + (long long)r * nc;
    const CCell e = cells[r];
    if (e.pin) {
        row[r] = 1.0;
        return;
    }
    const BlockLevel g = bls[e.b];
    double a[6], ap;
    coefs<double, VAR>(g, cx, cy, cz, gidx(g, e.i, e.j, e.k), e.i, e.j, e.k, a[0], a[1], a[2], a[3], a[4], a[5],
                       ap);
    row[r] = ap;
    for (int d = 0; d < 6; ++d) {
        const int c = e.nb[d];
        if (!cells[c].pin) row[c] = row[c] + a[d];
    }
}

// After the inversion: a pinned cell's row of the inverse is the identity
// row of the assembled matrix; zero it, so the pinned unknown is zero
// whatever the right-hand side holds there.
__global__ __launch_bounds__(256) void k_coarse_unpin(double *__restrict__ M, int nc,
                                                      const CCell *__restrict__ cells) {
    const int r = blockIdx.y;
    if (!cells[r].pin) return;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x)
        M[(long long)r * nc + c] = 0.0;
}

// Gauss-Jordan inversion in place, small matrices: one workgroup, a
// barrier per pivot.  col holds the pivot column of the current step.
__global__ __launch_bounds__(1024) void k_gj_small(double *__restrict__ M, int nc, double *__restrict__ col) {
    const int t = threadIdx.x;
    const long long nn = (long long)nc * nc;
    for (int k = 0; k < nc; ++k) {
        for (int i = t; i < nc; i += 1024) col[i] = M[(long long)i * nc + k];
        __syncthreads();
        const double piv = 1.0 / col[k];
        for (int j = t; j < nc; j += 1024) M[(long long)k * nc + j] = j == k ? piv : M[(long long)k * nc + j] * piv;
        __syncthreads();
        for (long long q = t; q < nn; q += 1024) {
            const int i = (int)(q / nc), j = (int)(q - (long long)i * nc);
            if (i == k) continue;
            const double f = col[i];
            M[q] = j == k ? -f * piv : M[q] - f * M[(long long)k * nc + j];
        }
        __syncthreads();
    }
}

// The same, one pivot step in three launches, for larger matrices.
__global__ __launch_bounds__(256) void k_gj_col(const double *__restrict__ M, int nc, int k,
                                                double *__restrict__ col) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nc) col[i] = M[(long long)i * nc + k];
}

__global__ __launch_bounds__(256) void k_gj_row(double *__restrict__ M, int nc, int k,
                                                const double *__restrict__ col) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nc) return;
    const double piv = 1.0 / col[k];
    M[(long long)k * nc + j] = j == k ? piv : M[(long long)k * nc + j] * piv;
}

__global__ __launch_bounds__(256) void k_gj_update(double *__restrict__ M, int nc, int k,
                                                   const double *__restrict__ col) {
    const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == k || j >= nc) return;
    const double piv = 1.0 / col[k], f = col[i];
    const long long q = (long long)i * nc + j;
    M[q] = j == k ? -f * piv : M[q] - f * M[(long long)k * nc + j];
}

// Row r of Minv times b, by one wavefront: lane l sums the columns
// l, l+64, ... in order, then a butterfly over the lanes (every lane ends
// with the same bits, since each pairwise sum is commutative).  The tail
// kernel and k_coarse_solve both use it, so their results agree bitwise.
__device__ __forceinline__ double coarse_row(const double *__restrict__ mrow, const double *b, int nc, int lane) {
    double s = 0.0;
    for (int c = lane; c < nc; c += 64) s = s + mrow[c] * b[c];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s = s + __shfl_xor(s, o, 64);
    return s;
}

// phi = Minv * rhs on the coarsest level of every block (interior cells;
// the ghost fill follows).  The RHS is staged in LDS as FP64; each wave
// takes rows blockIdx.x*4 + wave, stepping by the grid.
template <typename T>
__global__ __launch_bounds__(256) void k_coarse_solve(const double *__restrict__ Minv, int nc,
                                                      const CCell *__restrict__ cells,
                                                      const BlockLevel *__restrict__ bls, const T *__restrict__ rhs,
                                                      T *__restrict__ phi, const int *__restrict__ done) {
    if (done && *done) return;
    extern __shared__ double sb[];
    for (int c = threadIdx.x; c < nc; c += 256) {
        const CCell e = cells[c];
        sb[c] = (double)rhs[gidx(bls[e.b], e.i, e.j, e.k)];
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int r = blockIdx.x * 4 + wave; r < nc; r += gridDim.x * 4) {
        const double s = coarse_row(Minv + (long long)r * nc, sb, nc, lane);
        if (lane == 0) {
            const CCell e = cells[r];
            phi[gidx(bls[e.b], e.i, e.j, e.k)] = (T)s;
        }
    }
}

}  // namespace mf3d
