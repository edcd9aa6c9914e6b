// mf3d_mg.hip -- host side of the MI355X pressure-Poisson multigrid: the
// C ABI of include/mf3d_mg.h, device memory, the per-level block and tile
// tables, and the launch sequence of one V-cycle and of the outer
// defect-correction loop.
//
// V-cycle (mgkcyc, mgsolver.for:224-369 with kcycle = 1):
//   level 0: nu_pre sweeps; then down to the coarsest: restrict, nu_pre
//   sweeps (nu_pre + nu_post on the coarsest, which is only smoothed);
//   then up: prolong + Neumann copy, nu_post sweeps.
// One sweep = red colour, black colour, then the ghost gather (mgbound +
// exchangepp + exchbc_mgpp, mgsolver.for:610-611).  No ghost refresh
// between the two colours, and none of the periodic or inter-block ghosts
// after prolongation: the reference refreshes neither.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mf3d_mg.h"
#include "mg_kernels.hpp"

using namespace mf3d;

namespace {

constexpr int kTileX = 64, kTileY = 4, kRbTileX = 128;

struct Level {
    std::vector<BlockLevel> hb;
    BlockLevel *db = nullptr;
    Tile *t_rb = nullptr, *t_in = nullptr, *t_gh = nullptr;
    int n_rb = 0, n_in = 0, n_gh = 0;
    Region *r_full = nullptr, *r_neu = nullptr;
    int n_full = 0, n_neu = 0, max_full = 0, max_neu = 0;
    long long cells_gh = 0;  // arena length (all blocks, ghosted)
    void *phi = nullptr, *rhs = nullptr;
    void *cx = nullptr, *cy = nullptr, *cz = nullptr;  // inner precision (variable coef)
    double *dcx = nullptr, *dcy = nullptr, *dcz = nullptr;  // FP64 (variable coef)
    double *dens = nullptr;
};

}  // namespace

struct mf3d_mg {
    mf3d_mg_cfg cfg{};
    std::vector<mf3d_block_desc> blk;
    hipStream_t stream = nullptr;
    std::vector<Level> lv;
    double *p_a = nullptr, *p_b = nullptr, *f_g = nullptr;
    double *partials = nullptr;
    int nparts = 0;
    SolveState *st = nullptr;
    double *d_hist = nullptr, *d_rmax = nullptr;
    double *d_io = nullptr;   // device staging for host transfers (packed)
    double *h_io = nullptr;   // pinned host staging
    long long io_len = 0;
    long long cells = 0;
    long long bytes = 0;
    bool density_set = false;
    std::vector<void *> allocs;
    std::string err;
    // kernel timing hook
    int tk = 0;
    std::vector<hipEvent_t> ev;
    size_t ev_used = 0;
    long long tk_launches = 0;
    double tk_ms = 0.0;
};

namespace {

#define HIPCHK(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                        \
            mg->err = std::string(#expr) + ": " + hipGetErrorString(e_);               \
            return MF3D_EDEVICE;                                                       \
        }                                                                              \
    } while (0)

std::string g_create_err;

int dev_alloc(mf3d_mg *mg, void **p, size_t bytes) {
    bytes = (bytes + 255) & ~size_t(255);
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        mg->err = std::string("hipMalloc: ") + hipGetErrorString(e);
        *p = nullptr;
        return MF3D_ENOMEM;
    }
    mg->allocs.push_back(*p);
    mg->bytes += (long long)bytes;
    return 0;
}

template <typename X>
int upload(mf3d_mg *mg, X **dst, const std::vector<X> &src) {
    int rc = dev_alloc(mg, (void **)dst, sizeof(X) * std::max<size_t>(src.size(), 1));
    if (rc) return rc;
    if (!src.empty()) HIPCHK(hipMemcpy(*dst, src.data(), sizeof(X) * src.size(), hipMemcpyHostToDevice));
    return 0;
}

inline int dim_at(const mf3d_block_desc &b, int d, int l) { return b.n[d] >> l; }

bool is_neumann(const mf3d_block_desc &b, int f) {
    return b.nbr[f] < 0 && b.bc[f] != 5;
}

// Validate the block set: extents divisible by 2^(L-1), neighbours in range,
// symmetric and conforming (equal extents across a shared face).
std::string validate(const mf3d_mg_cfg &c, const mf3d_block_desc *b) {
    if (c.nblocks <= 0) return "nblocks must be positive";
    if (c.levels <= 0 || c.levels > 16) return "levels out of range";
    if (c.smoother != 3) return "only the red-black smoother (mg_itrsch = 3) is supported";
    if (c.nu_pre < 0 || c.nu_post < 0 || c.maxcy < 0) return "negative sweep or cycle count";
    if (c.precision != MF3D_FP64 && c.precision != MF3D_MIXED) return "unknown precision";
    const int div = 1 << (c.levels - 1);
    for (int i = 0; i < c.nblocks; ++i) {
        for (int d = 0; d < 3; ++d) {
            if (b[i].n[d] <= 0) return "block " + std::to_string(i) + ": empty extent";
            if (b[i].n[d] % div) return "block " + std::to_string(i) +
                                       ": the given number not correct for mgsolver (extent not "
                                       "divisible by 2^(levels-1))";
            if (!(b[i].h[d] > 0.0)) return "block " + std::to_string(i) + ": non-positive spacing";
            const int ld = b[i].ld[d] ? b[i].ld[d] : b[i].n[d] + 2 * b[i].ghost;
            if (ld < b[i].n[d] + 2 * b[i].ghost) return "block " + std::to_string(i) + ": ld too small";
        }
        if (b[i].ghost < 0) return "negative ghost width";
        for (int f = 0; f < 6; ++f) {
            const int d = f / 2, opp = f ^ 1;
            for (int kind = 0; kind < 2; ++kind) {
                const int nb = kind == 0 ? b[i].nbr[f] : b[i].per[f];
                if (nb < -1 || nb >= c.nblocks) return "neighbour index out of range";
                if (nb < 0) continue;
                if (kind == 1 && b[i].nbr[f] >= 0) continue;  // partner unused behind a neighbour
                if (kind == 1 && b[i].bc[f] != 5) return "periodic partner on a non-periodic face";
                const int back = kind == 0 ? b[nb].nbr[opp] : b[nb].per[opp];
                if (back != i) return "neighbour table not symmetric at block " + std::to_string(i);
                for (int e = 0; e < 3; ++e)
                    if (e != d && b[nb].n[e] != b[i].n[e])
                        return "blocks " + std::to_string(i) + " and " + std::to_string(nb) +
                               " do not conform on their shared face";
            }
        }
    }
    return "";
}

// Ghost regions of one level.  Full exchange: each ghost maps, axis by axis
// (z, then y, then x, the order in which the reference's face passes
// propagate edges and corners), to the interior cell it stands for --
// across a face neighbour, across a periodic partner, onto the block
// itself for a lone periodic block, or reflected at a Neumann wall.
// Neumann-only (mgbound after prolongation): reflect the Neumann axes,
// keep every other coordinate as it is.
std::vector<Region> build_regions(const std::vector<mf3d_block_desc> &blk, int l,
                                  bool neumann_only, int *max_cnt) {
    std::vector<Region> out;
    *max_cnt = 0;
    for (int b = 0; b < (int)blk.size(); ++b) {
        int n[3] = {dim_at(blk[b], 0, l), dim_at(blk[b], 1, l), dim_at(blk[b], 2, l)};
        for (int oz = -1; oz <= 1; ++oz)
            for (int oy = -1; oy <= 1; ++oy)
                for (int ox = -1; ox <= 1; ++ox) {
                    if (!ox && !oy && !oz) continue;
                    const int off[3] = {ox, oy, oz};
                    Region r{};
                    r.dst_b = b;
                    for (int d = 0; d < 3; ++d) {
                        if (off[d] < 0) { r.lo[d] = 0; r.cnt[d] = 1; }
                        else if (off[d] > 0) { r.lo[d] = n[d] + 1; r.cnt[d] = 1; }
                        else { r.lo[d] = 1; r.cnt[d] = n[d]; }
                        r.map[d] = -1;
                    }
                    if (neumann_only) {
                        bool any = false;
                        for (int d = 0; d < 3; ++d) {
                            if (!off[d]) continue;
                            const int f = 2 * d + (off[d] > 0);
                            if (is_neumann(blk[b], f)) {
                                r.map[d] = off[d] < 0 ? 1 : n[d];
                                any = true;
                            }
                        }
                        if (!any) continue;
                        r.src_b = b;
                    } else {
                        int cur = b;
                        for (int d = 2; d >= 0; --d) {
                            if (!off[d]) continue;
                            const int f = 2 * d + (off[d] > 0);
                            const mf3d_block_desc &B = blk[cur];
                            if (B.nbr[f] >= 0 || (B.bc[f] == 5 && B.per[f] >= 0)) {
                                cur = B.nbr[f] >= 0 ? B.nbr[f] : B.per[f];
                                r.map[d] = off[d] < 0 ? dim_at(blk[cur], d, l) : 1;
                            } else if (B.bc[f] == 5) {
                                r.map[d] = off[d] < 0 ? dim_at(B, d, l) : 1;
                            } else {
                                r.map[d] = off[d] < 0 ? 1 : dim_at(B, d, l);
                            }
                        }
                        r.src_b = cur;
                    }
                    *max_cnt = std::max(*max_cnt, r.cnt[0] * r.cnt[1] * r.cnt[2]);
                    out.push_back(r);
                }
    }
    return out;
}

std::vector<Tile> build_tiles(const std::vector<BlockLevel> &hb, int tx, bool ghosted) {
    std::vector<Tile> t;
    for (int b = 0; b < (int)hb.size(); ++b) {
        const BlockLevel &g = hb[b];
        const int lo = ghosted ? 0 : 1;
        const int h0 = ghosted ? g.n0 + 1 : g.n0, h1 = ghosted ? g.n1 + 1 : g.n1,
                  h2 = ghosted ? g.n2 + 1 : g.n2;
        for (int z = lo; z <= h2; ++z)
            for (int y = lo; y <= h1; y += kTileY)
                for (int x = lo; x <= h0; x += tx) t.push_back(Tile{b, x, y, z});
    }
    return t;
}

inline int inner_bytes(const mf3d_mg_cfg &c) { return c.precision == MF3D_MIXED ? 4 : 8; }

long long estimate(const mf3d_mg_cfg &c, const mf3d_block_desc *b) {
    long long tot = 0;
    const long long w = inner_bytes(c);
    for (int l = 0; l < c.levels; ++l) {
        long long gh = 0;
        for (int i = 0; i < c.nblocks; ++i)
            gh += (long long)(dim_at(b[i], 0, l) + 2) * (dim_at(b[i], 1, l) + 2) * (dim_at(b[i], 2, l) + 2);
        tot += gh * w * 2;                       // phi, rhs
        if (c.variable_coef) tot += gh * (w * 3 + 8 * 4);  // coefficients + FP64 copies + density
    }
    long long gh0 = 0, in0 = 0;
    for (int i = 0; i < c.nblocks; ++i) {
        gh0 += (long long)(b[i].n[0] + 2) * (b[i].n[1] + 2) * (b[i].n[2] + 2);
        in0 += (long long)b[i].n[0] * b[i].n[1] * b[i].n[2];
    }
    tot += gh0 * 8 * 3;   // p (two buffers) and f
    tot += std::max(gh0, in0) * 8;  // I/O staging
    return tot;
}

// ------------------------------------------------------------ timing hook
inline bool tk_on(const mf3d_mg *mg, int kernel, int level) { return mg->tk == kernel && level == 0; }

void tk_mark(mf3d_mg *mg) {
    if (mg->ev_used == mg->ev.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        mg->ev.push_back(e);
    }
    (void)hipEventRecord(mg->ev[mg->ev_used++], mg->stream);
}

// Call after the stream is synchronised: fold the recorded pairs into the totals.
void tk_flush(mf3d_mg *mg) {
    for (size_t i = 0; i + 1 < mg->ev_used; i += 2) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, mg->ev[i], mg->ev[i + 1]) == hipSuccess) {
            mg->tk_ms += ms;
            mg->tk_launches += 1;
        }
    }
    mg->ev_used = 0;
}

// ---------------------------------------------------------------- launches
template <typename T>
void fill(mf3d_mg *mg, T *a, long long n) {
    if (n <= 0) return;
    long long blocks = std::min<long long>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)blocks), dim3(256), 0, mg->stream, a, n, T(0));
}

template <typename T, bool VAR>
void sweep(mf3d_mg *mg, int l) {
    Level &L = mg->lv[l];
    if (!L.n_rb) return;
    const bool tk = tk_on(mg, MF3D_K_SMOOTH, l);
    for (int parity = 1; parity >= 0; --parity) {
        if (tk) tk_mark(mg);
        hipLaunchKernelGGL((k_rb<T, VAR>), dim3(L.n_rb), dim3(64, kTileY), 0, mg->stream, (T *)L.phi,
                           (const T *)L.rhs, (const T *)L.cx, (const T *)L.cy, (const T *)L.cz, L.db,
                           L.t_rb, parity);
        if (tk) tk_mark(mg);
    }
}

template <typename T>
void gather(mf3d_mg *mg, int l, bool neumann_only) {
    Level &L = mg->lv[l];
    const int nr = neumann_only ? L.n_neu : L.n_full;
    const int mc = neumann_only ? L.max_neu : L.max_full;
    if (!nr) return;
    const int gx = std::max(1, std::min(64, (mc + 255) / 256));
    hipLaunchKernelGGL(k_gather<T>, dim3(gx, nr), dim3(256), 0, mg->stream, (T *)L.phi, L.db,
                       neumann_only ? L.r_neu : L.r_full);
}

template <typename T, bool VAR>
void relax(mf3d_mg *mg, int l, int sweeps) {
    for (int s = 0; s < sweeps; ++s) {
        sweep<T, VAR>(mg, l);
        gather<T>(mg, l, false);
    }
}

template <typename T, bool VAR>
void restrict_to(mf3d_mg *mg, int lc) {
    Level &C = mg->lv[lc], &F = mg->lv[lc - 1];
    fill<T>(mg, (T *)C.phi, C.cells_gh);
    const bool tk = tk_on(mg, MF3D_K_RESTRICT, lc - 1);
    if (tk) tk_mark(mg);
    hipLaunchKernelGGL((k_restrict<T, VAR>), dim3(C.n_in), dim3(kTileX, kTileY), 0, mg->stream,
                       (const T *)F.phi, (const T *)F.rhs, (const T *)F.cx, (const T *)F.cy,
                       (const T *)F.cz, F.db, (T *)C.rhs, C.db, C.t_in);
    if (tk) tk_mark(mg);
}

template <typename T>
void prolong_to(mf3d_mg *mg, int lf) {
    Level &C = mg->lv[lf + 1], &F = mg->lv[lf];
    const bool tk = tk_on(mg, MF3D_K_PROLONG, lf);
    if (tk) tk_mark(mg);
    hipLaunchKernelGGL(k_prolong<T>, dim3(F.n_gh), dim3(kTileX, kTileY), 0, mg->stream, (T *)F.phi,
                       F.db, (const T *)C.phi, C.db, F.t_gh);
    if (tk) tk_mark(mg);
    gather<T>(mg, lf, true);
}

template <typename T, bool VAR>
void vcycle_dev(mf3d_mg *mg, bool zero_start) {
    const int L = mg->cfg.levels;
    if (zero_start) fill<T>(mg, (T *)mg->lv[0].phi, mg->lv[0].cells_gh);
    relax<T, VAR>(mg, 0, mg->cfg.nu_pre);
    if (L > 1) {
        for (int l = 1; l < L; ++l) {
            restrict_to<T, VAR>(mg, l);
            relax<T, VAR>(mg, l, mg->cfg.nu_pre + (l == L - 1 ? mg->cfg.nu_post : 0));
        }
        for (int l = L - 2; l >= 1; --l) {
            prolong_to<T>(mg, l);
            relax<T, VAR>(mg, l, mg->cfg.nu_post);
        }
        prolong_to<T>(mg, 0);
    }
    relax<T, VAR>(mg, 0, mg->cfg.nu_post);
}

template <typename T, bool VAR>
void outer_step(mf3d_mg *mg, int k) {
    Level &L0 = mg->lv[0];
    const double *cx = VAR ? L0.dcx : nullptr, *cy = VAR ? L0.dcy : nullptr, *cz = VAR ? L0.dcz : nullptr;
    const bool tk = tk_on(mg, MF3D_K_OUTER, 0);
    if (tk) tk_mark(mg);
    hipLaunchKernelGGL((k_outer<T, VAR>), dim3(L0.n_gh), dim3(kTileX, kTileY), 0, mg->stream, mg->p_a,
                       mg->p_b, (const T *)L0.phi, mg->f_g, (T *)L0.rhs, cx, cy, cz, L0.db, L0.t_gh,
                       mg->partials);
    if (tk) tk_mark(mg);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, mg->stream, mg->partials, L0.n_gh, mg->st,
                       mg->d_hist, mg->d_rmax, k, mg->cfg.eps, mg->cfg.min_cycles);
    std::swap(mg->p_a, mg->p_b);
}

template <typename T, bool VAR>
int solve_dev(mf3d_mg *mg, const double *f_packed, double *p_packed, double *rmax_out,
              int32_t *iters_out, double *hist) {
    Level &L0 = mg->lv[0];
    fill<double>(mg, mg->p_a, L0.cells_gh);
    fill<double>(mg, mg->f_g, L0.cells_gh);
    fill<T>(mg, (T *)L0.phi, L0.cells_gh);
    hipLaunchKernelGGL(k_unpack<double>, dim3(L0.n_in), dim3(kTileX, kTileY), 0, mg->stream, f_packed,
                       mg->f_g, L0.db, L0.t_in);
    outer_step<T, VAR>(mg, 0);
    int k = 0, done = 0;
    SolveState hs{};
    for (k = 1; k <= mg->cfg.maxcy; ++k) {
        vcycle_dev<T, VAR>(mg, true);
        outer_step<T, VAR>(mg, k);
        HIPCHK(hipMemcpyAsync(&hs, mg->st, sizeof(SolveState), hipMemcpyDeviceToHost, mg->stream));
        HIPCHK(hipStreamSynchronize(mg->stream));
        tk_flush(mg);
        if (hs.done) { done = 1; break; }
    }
    if (k > mg->cfg.maxcy) k = mg->cfg.maxcy;
    hipLaunchKernelGGL(k_pack<double>, dim3(L0.n_in), dim3(kTileX, kTileY), 0, mg->stream, mg->p_a,
                       p_packed, L0.db, L0.t_in);
    std::vector<double> hh(k + 1), rm(k + 1);
    HIPCHK(hipMemcpyAsync(hh.data(), mg->d_hist, sizeof(double) * (k + 1), hipMemcpyDeviceToHost, mg->stream));
    HIPCHK(hipMemcpyAsync(rm.data(), mg->d_rmax, sizeof(double) * (k + 1), hipMemcpyDeviceToHost, mg->stream));
    HIPCHK(hipStreamSynchronize(mg->stream));
    HIPCHK(hipGetLastError());
    if (hist) std::memcpy(hist, hh.data(), sizeof(double) * (k + 1));
    if (iters_out) *iters_out = k;
    const double rm_last = rm[k];
    if (rmax_out) *rmax_out = rm_last;
    if (!(rm_last <= 100.0)) {  // newsolv_mg.for:64 (a NaN counts as diverged here)
        mg->err = "BIG RMAX (rmax > 100): diverged";
        return MF3D_DIVERGED;
    }
    return done ? MF3D_OK : MF3D_NOT_CONVERGED;
}

template <typename T, bool VAR>
int vcycle_host_t(mf3d_mg *mg) {
    vcycle_dev<T, VAR>(mg, false);
    return 0;
}

// Host arrays in the caller's layout <-> packed staging.
void host_to_packed(const mf3d_mg *mg, const double *const *src, double *dst, int ring) {
    long long o = 0;
    for (int b = 0; b < mg->cfg.nblocks; ++b) {
        const mf3d_block_desc &B = mg->blk[b];
        const int ld0 = B.ld[0] ? B.ld[0] : B.n[0] + 2 * B.ghost;
        const int ld1 = B.ld[1] ? B.ld[1] : B.n[1] + 2 * B.ghost;
        const int e0 = B.n[0] + 2 * ring, e1 = B.n[1] + 2 * ring, e2 = B.n[2] + 2 * ring;
        const int s = B.ghost - ring;
        for (int k = 0; k < e2; ++k)
            for (int j = 0; j < e1; ++j) {
                const double *row = src[b] + (size_t)(s) + (size_t)ld0 * ((size_t)(j + s) + (size_t)ld1 * (k + s));
                std::memcpy(dst + o, row, sizeof(double) * e0);
                o += e0;
            }
    }
}

void packed_to_host(const mf3d_mg *mg, const double *src, double *const *dst, int ring) {
    long long o = 0;
    for (int b = 0; b < mg->cfg.nblocks; ++b) {
        const mf3d_block_desc &B = mg->blk[b];
        const int ld0 = B.ld[0] ? B.ld[0] : B.n[0] + 2 * B.ghost;
        const int ld1 = B.ld[1] ? B.ld[1] : B.n[1] + 2 * B.ghost;
        const int e0 = B.n[0] + 2 * ring, e1 = B.n[1] + 2 * ring, e2 = B.n[2] + 2 * ring;
        const int s = B.ghost - ring;
        for (int k = 0; k < e2; ++k)
            for (int j = 0; j < e1; ++j) {
                double *row = dst[b] + (size_t)(s) + (size_t)ld0 * ((size_t)(j + s) + (size_t)ld1 * (k + s));
                std::memcpy(row, src + o, sizeof(double) * e0);
                o += e0;
            }
    }
}

// Ghosted-packed (n+2)^3 per block <-> a level-0 arena (same shape, so a
// block's packed ghosted array is exactly its arena slice).
template <typename T>
void ghosted_in(mf3d_mg *mg, const double *packed_dev, T *arena) {
    Level &L0 = mg->lv[0];
    hipLaunchKernelGGL((k_convert<T, double>), dim3((unsigned)std::min<long long>((L0.cells_gh + 255) / 256, 16384)),
                       dim3(256), 0, mg->stream, arena, packed_dev, L0.cells_gh);
}
template <typename T>
void ghosted_out(mf3d_mg *mg, const T *arena, double *packed_dev) {
    Level &L0 = mg->lv[0];
    hipLaunchKernelGGL((k_convert<double, T>), dim3((unsigned)std::min<long long>((L0.cells_gh + 255) / 256, 16384)),
                       dim3(256), 0, mg->stream, packed_dev, arena, L0.cells_gh);
}

int build(mf3d_mg *mg) {
    const mf3d_mg_cfg &c = mg->cfg;
    const int L = c.levels, nb = c.nblocks;
    const bool mixed = c.precision == MF3D_MIXED;
    const size_t w = mixed ? 4 : 8;
    mg->lv.resize(L);
    mg->cells = 0;
    for (int b = 0; b < nb; ++b) mg->cells += (long long)mg->blk[b].n[0] * mg->blk[b].n[1] * mg->blk[b].n[2];
    for (int l = 0; l < L; ++l) {
        Level &V = mg->lv[l];
        long long off = 0, poff = 0;
        V.hb.resize(nb);
        for (int b = 0; b < nb; ++b) {
            const mf3d_block_desc &B = mg->blk[b];
            BlockLevel &g = V.hb[b];
            g.n0 = dim_at(B, 0, l); g.n1 = dim_at(B, 1, l); g.n2 = dim_at(B, 2, l);
            g.neu = 0;
            for (int f = 0; f < 6; ++f)
                if (is_neumann(B, f)) g.neu |= (1 << f);
            g.sy = g.n0 + 2;
            g.sz = g.sy * (g.n1 + 2);
            g.off = off;
            g.poff = poff;
            off += g.sz * (g.n2 + 2);
            poff += (long long)g.n0 * g.n1 * g.n2;
            // constant coefficients (mgsolver.for:51-57,184-189)
            const double sc = (double)(1 << l);
            const double ndx = B.h[0] * sc, ndy = B.h[1] * sc, ndz = B.h[2] * sc;
            g.c0 = 1.0 / (ndx * ndx);
            g.c1 = 1.0 / (ndy * ndy);
            g.c2 = 1.0 / (ndz * ndz);
            g.f0 = (float)g.c0; g.f1 = (float)g.c1; g.f2 = (float)g.c2;
            g.lsm = c.variable_coef ? 1 : 0;
            g.h0 = B.h[0]; g.h1 = B.h[1]; g.h2 = B.h[2];
            g.gx = B.h[1] * B.h[2] / B.h[0];
            g.gy = B.h[0] * B.h[2] / B.h[1];
            g.gz = B.h[0] * B.h[1] / B.h[2];
        }
        V.cells_gh = off;
        int rc = upload(mg, &V.db, V.hb);
        if (rc) return rc;
        auto trb = build_tiles(V.hb, kRbTileX, false), tin = build_tiles(V.hb, kTileX, false),
             tgh = build_tiles(V.hb, kTileX, true);
        V.n_rb = (int)trb.size(); V.n_in = (int)tin.size(); V.n_gh = (int)tgh.size();
        if ((rc = upload(mg, &V.t_rb, trb)) || (rc = upload(mg, &V.t_in, tin)) || (rc = upload(mg, &V.t_gh, tgh)))
            return rc;
        auto rf = build_regions(mg->blk, l, false, &V.max_full);
        auto rn = build_regions(mg->blk, l, true, &V.max_neu);
        V.n_full = (int)rf.size(); V.n_neu = (int)rn.size();
        if ((rc = upload(mg, &V.r_full, rf)) || (rc = upload(mg, &V.r_neu, rn))) return rc;
        if ((rc = dev_alloc(mg, &V.phi, V.cells_gh * w)) || (rc = dev_alloc(mg, &V.rhs, V.cells_gh * w)))
            return rc;
        if (c.variable_coef) {
            if ((rc = dev_alloc(mg, (void **)&V.dcx, V.cells_gh * 8)) ||
                (rc = dev_alloc(mg, (void **)&V.dcy, V.cells_gh * 8)) ||
                (rc = dev_alloc(mg, (void **)&V.dcz, V.cells_gh * 8)) ||
                (rc = dev_alloc(mg, (void **)&V.dens, V.cells_gh * 8)))
                return rc;
            if (mixed) {
                if ((rc = dev_alloc(mg, &V.cx, V.cells_gh * 4)) || (rc = dev_alloc(mg, &V.cy, V.cells_gh * 4)) ||
                    (rc = dev_alloc(mg, &V.cz, V.cells_gh * 4)))
                    return rc;
            } else {
                V.cx = V.dcx; V.cy = V.dcy; V.cz = V.dcz;
            }
        }
    }
    Level &L0 = mg->lv[0];
    int rc;
    if ((rc = dev_alloc(mg, (void **)&mg->p_a, L0.cells_gh * 8)) ||
        (rc = dev_alloc(mg, (void **)&mg->p_b, L0.cells_gh * 8)) ||
        (rc = dev_alloc(mg, (void **)&mg->f_g, L0.cells_gh * 8)))
        return rc;
    mg->nparts = L0.n_gh;
    if ((rc = dev_alloc(mg, (void **)&mg->partials, sizeof(double) * 2 * (size_t)mg->nparts)) ||
        (rc = dev_alloc(mg, (void **)&mg->st, sizeof(SolveState))) ||
        (rc = dev_alloc(mg, (void **)&mg->d_hist, sizeof(double) * (c.maxcy + 1))) ||
        (rc = dev_alloc(mg, (void **)&mg->d_rmax, sizeof(double) * (c.maxcy + 1))))
        return rc;
    mg->io_len = std::max(L0.cells_gh, mg->cells);
    if ((rc = dev_alloc(mg, (void **)&mg->d_io, sizeof(double) * mg->io_len))) return rc;
    HIPCHK(hipHostMalloc((void **)&mg->h_io, sizeof(double) * mg->io_len, hipHostMallocDefault));
    // level fields start at zero (coeff zeroes phi, mgsolver.for:60-67)
    for (int l = 0; l < L; ++l) {
        Level &V = mg->lv[l];
        if (mixed) { fill<float>(mg, (float *)V.phi, V.cells_gh); fill<float>(mg, (float *)V.rhs, V.cells_gh); }
        else { fill<double>(mg, (double *)V.phi, V.cells_gh); fill<double>(mg, (double *)V.rhs, V.cells_gh); }
        if (c.variable_coef) {
            fill<double>(mg, V.dcx, V.cells_gh); fill<double>(mg, V.dcy, V.cells_gh);
            fill<double>(mg, V.dcz, V.cells_gh); fill<double>(mg, V.dens, V.cells_gh);
        }
    }
    HIPCHK(hipStreamSynchronize(mg->stream));
    HIPCHK(hipGetLastError());
    return 0;
}

// coeff, L_LSM branch, from level-0 densities already in lv[0].dens.
int build_lsm(mf3d_mg *mg) {
    const int L = mg->cfg.levels;
    for (int l = 1; l < L; ++l) {
        Level &C = mg->lv[l], &F = mg->lv[l - 1];
        hipLaunchKernelGGL(k_dens_restrict, dim3(C.n_in), dim3(kTileX, kTileY), 0, mg->stream, F.dens, F.db,
                           C.dens, C.db, C.t_in);
    }
    for (int l = 0; l < L; ++l) {
        Level &V = mg->lv[l];
        hipLaunchKernelGGL(k_lsm_coef, dim3(V.n_gh), dim3(kTileX, kTileY), 0, mg->stream, V.dens, V.dcx, V.dcy,
                           V.dcz, V.db, V.t_gh);
        if (mg->cfg.precision == MF3D_MIXED) {
            const unsigned gb = (unsigned)std::min<long long>((V.cells_gh + 255) / 256, 16384);
            hipLaunchKernelGGL((k_convert<float, double>), dim3(gb), dim3(256), 0, mg->stream, (float *)V.cx, V.dcx, V.cells_gh);
            hipLaunchKernelGGL((k_convert<float, double>), dim3(gb), dim3(256), 0, mg->stream, (float *)V.cy, V.dcy, V.cells_gh);
            hipLaunchKernelGGL((k_convert<float, double>), dim3(gb), dim3(256), 0, mg->stream, (float *)V.cz, V.dcz, V.cells_gh);
        }
    }
    HIPCHK(hipStreamSynchronize(mg->stream));
    HIPCHK(hipGetLastError());
    mg->density_set = true;
    return 0;
}

int set_device(mf3d_mg *mg) {
    HIPCHK(hipSetDevice(mg->cfg.device));
    return 0;
}

}  // namespace

// ===================================================================== C ABI
extern "C" {

int mf3d_mg_abi_version(void) { return MF3D_MG_ABI_VERSION; }

int64_t mf3d_mg_estimate_bytes(const mf3d_mg_cfg *cfg, const mf3d_block_desc *blocks) {
    if (!cfg || !blocks) return MF3D_EINVAL;
    if (!validate(*cfg, blocks).empty()) return MF3D_EINVAL;
    return estimate(*cfg, blocks);
}

int mf3d_mg_create(const mf3d_mg_cfg *cfg, const mf3d_block_desc *blocks, mf3d_mg **out) {
    if (!out) return MF3D_EINVAL;
    *out = nullptr;
    if (!cfg || !blocks) return MF3D_EINVAL;
    if (cfg->smoother != 3) return MF3D_ENOTSUP;
    std::string v = validate(*cfg, blocks);
    mf3d_mg *mg = new mf3d_mg();
    mg->cfg = *cfg;
    if (!v.empty()) {
        mg->err = v;
        *out = mg;  // handle returned so the caller can read the message
        return MF3D_EINVAL;
    }
    mg->blk.assign(blocks, blocks + cfg->nblocks);
    int rc = set_device(mg);
    if (rc) { *out = mg; return rc; }
    hipError_t e = hipStreamCreate(&mg->stream);
    if (e != hipSuccess) {
        mg->err = std::string("hipStreamCreate: ") + hipGetErrorString(e);
        *out = mg;
        return MF3D_EDEVICE;
    }
    rc = build(mg);
    *out = mg;
    return rc;
}

int mf3d_mg_set_density(mf3d_mg *mg, const double *const *dens_blocks) {
    if (!mg || !dens_blocks) return MF3D_EINVAL;
    if (!mg->cfg.variable_coef) { mg->err = "handle was created without variable_coef"; return MF3D_EINVAL; }
    int rc = set_device(mg);
    if (rc) return rc;
    host_to_packed(mg, dens_blocks, mg->h_io, 0);
    HIPCHK(hipMemcpyAsync(mg->d_io, mg->h_io, sizeof(double) * mg->cells, hipMemcpyHostToDevice, mg->stream));
    Level &L0 = mg->lv[0];
    hipLaunchKernelGGL(k_unpack<double>, dim3(L0.n_in), dim3(kTileX, kTileY), 0, mg->stream, mg->d_io, L0.dens,
                       L0.db, L0.t_in);
    return build_lsm(mg);
}

static int check_ready(mf3d_mg *mg) {
    if (mg->cfg.variable_coef && !mg->density_set) {
        mg->err = "variable-coefficient handle needs mf3d_mg_set_density first";
        return MF3D_EINVAL;
    }
    return 0;
}

int mf3d_mg_vcycle(mf3d_mg *mg, const double *const *rhs_blocks, double *const *pp_blocks) {
    if (!mg || !rhs_blocks || !pp_blocks) return MF3D_EINVAL;
    int rc = check_ready(mg);
    if (rc) return rc;
    for (int b = 0; b < mg->cfg.nblocks; ++b)
        if (mg->blk[b].ghost < 1) { mg->err = "vcycle needs caller arrays with ghost >= 1"; return MF3D_EINVAL; }
    if ((rc = set_device(mg))) return rc;
    Level &L0 = mg->lv[0];
    const bool mixed = mg->cfg.precision == MF3D_MIXED;
    // rhs (interior)
    host_to_packed(mg, rhs_blocks, mg->h_io, 0);
    HIPCHK(hipMemcpyAsync(mg->d_io, mg->h_io, sizeof(double) * mg->cells, hipMemcpyHostToDevice, mg->stream));
    if (mixed)
        hipLaunchKernelGGL(k_unpack<float>, dim3(L0.n_in), dim3(kTileX, kTileY), 0, mg->stream, mg->d_io,
                           (float *)L0.rhs, L0.db, L0.t_in);
    else
        hipLaunchKernelGGL(k_unpack<double>, dim3(L0.n_in), dim3(kTileX, kTileY), 0, mg->stream, mg->d_io,
                           (double *)L0.rhs, L0.db, L0.t_in);
    // pp including one ghost layer == the level-0 arena shape
    HIPCHK(hipStreamSynchronize(mg->stream));
    host_to_packed(mg, pp_blocks, mg->h_io, 1);
    HIPCHK(hipMemcpyAsync(mg->d_io, mg->h_io, sizeof(double) * L0.cells_gh, hipMemcpyHostToDevice, mg->stream));
    const bool var = mg->cfg.variable_coef != 0;
    if (mixed) {
        ghosted_in<float>(mg, mg->d_io, (float *)L0.phi);
        if (var) vcycle_host_t<float, true>(mg); else vcycle_host_t<float, false>(mg);
        ghosted_out<float>(mg, (const float *)L0.phi, mg->d_io);
    } else {
        ghosted_in<double>(mg, mg->d_io, (double *)L0.phi);
        if (var) vcycle_host_t<double, true>(mg); else vcycle_host_t<double, false>(mg);
        ghosted_out<double>(mg, (const double *)L0.phi, mg->d_io);
    }
    HIPCHK(hipMemcpyAsync(mg->h_io, mg->d_io, sizeof(double) * L0.cells_gh, hipMemcpyDeviceToHost, mg->stream));
    HIPCHK(hipStreamSynchronize(mg->stream));
    HIPCHK(hipGetLastError());
    packed_to_host(mg, mg->h_io, pp_blocks, 1);
    return 0;
}

static int solve_any(mf3d_mg *mg, const double *f_dev, double *p_dev, double *rmax_out, int32_t *iters_out,
                     double *hist) {
    const bool mixed = mg->cfg.precision == MF3D_MIXED, var = mg->cfg.variable_coef != 0;
    if (mixed) return var ? solve_dev<float, true>(mg, f_dev, p_dev, rmax_out, iters_out, hist)
                          : solve_dev<float, false>(mg, f_dev, p_dev, rmax_out, iters_out, hist);
    return var ? solve_dev<double, true>(mg, f_dev, p_dev, rmax_out, iters_out, hist)
               : solve_dev<double, false>(mg, f_dev, p_dev, rmax_out, iters_out, hist);
}

int mf3d_mg_solve(mf3d_mg *mg, const double *const *f_blocks, double *const *p_blocks, double *rmax_out,
                  int32_t *iters_out, double *hist) {
    if (!mg || !f_blocks || !p_blocks) return MF3D_EINVAL;
    int rc = check_ready(mg);
    if (rc) return rc;
    if ((rc = set_device(mg))) return rc;
    double *f_dev = nullptr;
    if ((rc = dev_alloc(mg, (void **)&f_dev, sizeof(double) * mg->cells))) return rc;
    host_to_packed(mg, f_blocks, mg->h_io, 0);
    HIPCHK(hipMemcpyAsync(f_dev, mg->h_io, sizeof(double) * mg->cells, hipMemcpyHostToDevice, mg->stream));
    rc = solve_any(mg, f_dev, mg->d_io, rmax_out, iters_out, hist);
    if (rc >= 0) {
        HIPCHK(hipMemcpyAsync(mg->h_io, mg->d_io, sizeof(double) * mg->cells, hipMemcpyDeviceToHost, mg->stream));
        HIPCHK(hipStreamSynchronize(mg->stream));
        packed_to_host(mg, mg->h_io, p_blocks, 0);
    }
    mg->allocs.erase(std::find(mg->allocs.begin(), mg->allocs.end(), (void *)f_dev));
    mg->bytes -= (long long)(((sizeof(double) * mg->cells) + 255) & ~size_t(255));
    (void)hipFree(f_dev);
    return rc;
}

int mf3d_mg_solve_device(mf3d_mg *mg, const double *f_dev, double *p_dev, void *stream, double *rmax_out,
                         int32_t *iters_out, double *hist) {
    if (!mg || !f_dev || !p_dev) return MF3D_EINVAL;
    int rc = check_ready(mg);
    if (rc) return rc;
    if ((rc = set_device(mg))) return rc;
    hipStream_t own = mg->stream;
    if (stream) mg->stream = (hipStream_t)stream;
    rc = solve_any(mg, f_dev, p_dev, rmax_out, iters_out, hist);
    mg->stream = own;
    return rc;
}

int mf3d_mg_time_kernel(mf3d_mg *mg, int kernel) {
    if (!mg || kernel < MF3D_K_OFF || kernel > MF3D_K_PROLONG) return MF3D_EINVAL;
    if (mg->stream) (void)hipStreamSynchronize(mg->stream);
    tk_flush(mg);
    mg->tk = kernel;
    mg->tk_launches = 0;
    mg->tk_ms = 0.0;
    return 0;
}

int mf3d_mg_kernel_time(mf3d_mg *mg, int64_t *launches, double *total_ms) {
    if (!mg) return MF3D_EINVAL;
    if (mg->stream) (void)hipStreamSynchronize(mg->stream);
    tk_flush(mg);
    if (launches) *launches = mg->tk_launches;
    if (total_ms) *total_ms = mg->tk_ms;
    return 0;
}

int64_t mf3d_mg_cells(const mf3d_mg *mg) { return mg ? mg->cells : MF3D_EINVAL; }
int64_t mf3d_mg_device_bytes(const mf3d_mg *mg) { return mg ? mg->bytes : MF3D_EINVAL; }

const char *mf3d_mg_last_error(const mf3d_mg *mg) {
    if (!mg) return "null handle";
    return mg->err.c_str();
}

void mf3d_mg_destroy(mf3d_mg *mg) {
    if (!mg) return;
    if (mg->stream) (void)hipStreamSynchronize(mg->stream);
    for (void *p : mg->allocs) (void)hipFree(p);
    if (mg->h_io) (void)hipHostFree(mg->h_io);
    for (hipEvent_t e : mg->ev) (void)hipEventDestroy(e);
    if (mg->stream) (void)hipStreamDestroy(mg->stream);
    delete mg;
}

}  // extern "C"
