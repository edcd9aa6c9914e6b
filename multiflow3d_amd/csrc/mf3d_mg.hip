// mf3d_mg.hip -- host side of the MI355X pressure-Poisson multigrid: the
// C ABI of include/mf3d_mg.h, device memory, the per-level block and tile
// tables, and the launch sequence of one V-cycle and of the outer
// defect-correction loop.
//
// V-cycle (mgkcyc, mgsolver.for:224-369 with kcycle = 1):
//   level 0: nu_pre sweeps; then down to the coarsest: restrict, nu_pre
//   sweeps (nu_pre + nu_post on the coarsest, which is only smoothed);
//   then up: prolong (with the Neumann copy), nu_post sweeps.
// One sweep = both colours in one kernel, then the ghost gather (mgbound +
// exchangepp + exchbc_mgpp, mgsolver.for:610-611).  No ghost refresh
// between the two colours, and none of the periodic or inter-block ghosts
// after prolongation: the reference refreshes neither.
//
// A solve captures one cycle (V-cycle + outer step + norm) in a hipGraph,
// two variants for the two directions of the pressure ping-pong, and
// replays it in chunks sized by the previous solve; the device raises a
// `done` flag at convergence, after which the remaining kernels of the
// chunk return at once.  The host synchronises once per chunk.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/mf3d_mg.h"
#include "mg_coarse.hpp"
#include "mg_common.hpp"
#include "mg_ghostmap.hpp"
#include "mg_refine.hpp"
#include "mg_outer.hpp"
#include "mg_setup.hpp"
#include "mg_sweep.hpp"
#include "mg_tail.hpp"
#include "mg_transfer.hpp"
#include "mg_levelset.hpp"
#include "mg_linesmooth.hpp"
#include "mg_velocity.hpp"

using namespace mf3d;

// Device memory of the allocation-free solve (mf3d_mg_solve_device_noalloc):
// one array per GPU, loaded with the library and held until the process
// ends.  Host code never uses its name as an address: each device's copy
// is found once with hipGetSymbolAddress (scratch_base) and that pointer
// is what kernels and copies receive.
static __device__ __align__(256) unsigned char scratch[MF3D_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kTileX = 64, kTileY = 4;

struct Level {
    std::vector<BlockLevel> hb;
    BlockLevel *db = nullptr;
    Tile *t_sw = nullptr, *t_in = nullptr, *t_row = nullptr, *t_oct = nullptr, *t_ch = nullptr;
    int n_sw = 0, n_in = 0, n_row = 0, n_oct = 0, n_ch = 0;
    // ADI smoother: dynamic LDS of the whole-sweep kernel (0: the level's
    // largest block does not fit, the per-pass kernels run), whether the
    // RHS is staged in LDS too, the largest block's ghosted cells
    size_t adi_lds = 0;
    bool adi_rhs = false;
    int adi_cells = 0, adi_sstride = 0;
    void *adi_apr[3] = {}, *adi_a[3] = {};  // the LDS kernel's Thomas coefficients per line axis
    int sw_ty = 4;  // sweep tile rows of the march; 0: 3-D tiles (k_rb3d)
    // the zero-start march (k_rbz) on its own chunks: the default depth
    // (the deepest that still gives ~1024 workgroups), which suits its
    // light steps better than the full sweep's tuned depth -- under
    // constant coefficients; with variable ones (coefficient loads make
    // its steps heavier) the full sweep's tuned tiles
    Tile *t_swz = nullptr;
    int n_swz = 0, swz_ty = 4;
    Region *r_full = nullptr;
    int n_full = 0, max_full = 0;
    // refined block grids: the level's ghosts as expressions (no regions)
    GhostExpr *gx = nullptr;
    int n_gx = 0;
    // restriction into this level: k_restrict tiles over blocks that coarsen
    // here, k_restrict_rep tiles over blocks that repeat their grid here
    Tile *t_rst = nullptr, *t_rrep = nullptr;
    int n_rst = 0, n_rrep = 0;
    // prolongation into this level from the next: k_prolong tiles (t_oct)
    // exclude blocks that repeat their grid at the next level, whose rows
    // k_prolong_rep takes
    Tile *t_prep = nullptr;
    int n_prep = 0;
    bool own_ghosts = false;  // lone block: the sweep writes the ghosts, no gather
    long long cells_gh = 0;  // arena length (all blocks, ghosted)
    void *phi[2] = {nullptr, nullptr};
    int cur = 0;  // which phi holds the current potential (host bookkeeping)
    void *rhs = nullptr;
    void *cx = nullptr, *cy = nullptr, *cz = nullptr;      // inner precision (variable coef)
    double *dcx = nullptr, *dcy = nullptr, *dcz = nullptr;  // FP64 (variable coef)
    double *dens = nullptr;
};

}  // namespace

struct mf3d_mg {
    mf3d_mg_cfg cfg{};
    std::vector<mf3d_block_desc> blk;
    std::vector<int> own;  // each block's own level count (refined blocks have more)
    bool refined = false;  // some block has rdiv > 1
    hipStream_t stream = nullptr;
    std::vector<Level> lv;
    double *pbuf[2] = {nullptr, nullptr};
    double *f_g = nullptr;
    // GCR outer iteration (krylov): directions (ping-pong), w = A d, FP64 residual
    double *kd[2] = {nullptr, nullptr};
    double *kw = nullptr, *kr = nullptr;
    double *partials = nullptr;
    int nparts = 0;
    SolveState *st = nullptr;
    double *d_hist = nullptr, *d_rmax = nullptr;
    double *d_io = nullptr;  // device staging for host transfers (packed)
    double *h_io = nullptr;  // pinned host staging
    long long io_len = 0;
    long long cells = 0;
    long long bytes = 0;
    bool density_set = false;
    std::vector<void *> allocs;
    std::string err;
    // captured cycles: graph g reads pbuf[g], writes pbuf[1-g]
    hipGraphExec_t cyc[2] = {nullptr, nullptr};
    // coarse tail: levels lt..L-1 run in LDS, one workgroup per block
    int lt = 0;
    BlockLevel *d_tail = nullptr;
    size_t tail_lds = 0;
    // outer step tiles (finest level)
    Tile *t_out = nullptr;
    int n_out = 0;
    bool graphs_ready = false;
    int guess = 0;  // cycles the previous solve needed
    // kernel timing hook: event pairs around each finest-level launch of
    // the selected kernel.  Eager launches record into ev_eager; cycle k is
    // replayed from its own captured graph tkg[k], which records into its
    // own events tkev[k], so timing keeps graph replay.
    struct EvList {
        std::vector<hipEvent_t> ev;
        size_t used = 0;
    };
    int tk = 0;
    EvList ev_eager;
    EvList *ev_rec = nullptr;  // where tk_mark records (nullptr: ev_eager)
    bool ev_fail = false;      // an event record failed (e.g. refused during capture)
    std::vector<hipGraphExec_t> tkg;
    std::vector<EvList> tkev;
    bool tk_eager = false;  // timed graphs unavailable: time eager launches
    int last_replay = 0;    // the last solve replayed captured cycles
    long long tk_launches = 0;
    double tk_ms = 0.0;
    // velocity form of newsolv_mg, built at its first call: per-block face
    // ranges and low-face sources, the composed ghost maps of exchange(1..3)
    // (u, v, w entries, cumulative counts vmap_c) and of exchange(4) +
    // pbound (vmap_p entries after them), their staging buffer, the arenas
    // of the host-array call, and the captured outer step with the field
    // pointers it was recorded with
    bool dens_ghosted = false;  // set_density received the density's ghost layer
    bool vel_ready = false;
    VelBlock *d_vb = nullptr;
    uint2 *d_vmap = nullptr;
    long long vmap_c[3] = {0, 0, 0};
    long long vmap_p = 0;
    double *d_vstage = nullptr;
    double *vf[8] = {};  // u*, v*, w*, u, v, w, p, su
    int vel_wg = 0;      // persistent grid of k_velmas (tiles spread evenly over resident workgroups)
    Tile *t_vel = nullptr;  // its tiles: 64 x 8 columns, chunks of at most VEL_TZ_MAX planes
    double *d_vpart = nullptr;  // its per-workgroup maxima
    double *d_vsink = nullptr;  // one slot per thread of its grid for stores a lane must not make
    int n_vel = 0;
    hipGraphExec_t vgraph = nullptr;
    const void *vkey[7] = {};
    // level-set density (heaviside), built at its first call: the composed
    // map of bound_lsm(18) + exchange(18), the per-block epsl table, and
    // the arenas of the host-array call (phi, mu)
    bool ls_ready = false;
    uint2 *d_lsmap = nullptr;
    long long n_lsmap = 0;
    LsBlock *d_lsb = nullptr;
    long long ls_max_cells = 0;
    double *d_lsio[2] = {nullptr, nullptr};
    // ADI smoother: a(m) scratch of the per-pass kernels (level-0 size,
    // shared by the levels whose blocks do not fit in LDS)
    void *d_adi = nullptr;
    // FP64 direct solve of the coarsest level: cell table, the inverse
    // operator (nc x nc, rebuilt with the coefficients), a pivot-column
    // scratch; cdirect set at build time
    bool cdirect = false;
    int nc = 0;
    // FP64 sweeps of the mixed form's coarsest level (c64): its potential
    // (two buffers) and RHS in FP64
    bool c64 = false;
    double *c64_phi[2] = {nullptr, nullptr}, *c64_rhs = nullptr;
    CCell *d_ccell = nullptr;
    // refined block grids: the coarsest operator, assembled on the host
    // (constant coefficients only, so factored once, at creation; released
    // after the inversion), row-major nc x nc
    std::vector<double> cmat;
    double *d_cinv = nullptr, *d_ccol = nullptr;
    // the zero-start sweep of the march levels: 1 a march over the RHS
    // alone (k_rbz), 0 the full sweep from a zero potential (k_rbs);
    // MF3D_ZERO_SWEEP picks one, both bitwise the same
    int zmode = 1;
    // levels on 3-D tiles fuse the prolongation into their first post-sweep
    // (k_rb3d<..., PRO>); MF3D_FUSE_PROLONG=0 keeps them apart, bitwise the same
    bool fuse_pro = true;
    // the velocity form's fields on the caller's ghost shell: G layers
    // (every block's `ghost`), per-block shell geometry, total length
    int G = 1;
    std::vector<BlockLevel> hw;
    BlockLevel *d_vw = nullptr;
    long long cells_w = 0;
    // allocation-free solve (mf3d_mg_solve_device_noalloc): every device
    // buffer is carved from this device's `scratch` array instead of
    // hipMalloc; no graphs, no tuner, no staging, the caller's stream
    char *arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
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

int dev_alloc(mf3d_mg *mg, void **p, size_t bytes) {
    bytes = (bytes + 255) & ~size_t(255);
    if (bytes == 0) bytes = 256;
    if (mg->arena) {  // carve from the scratch array (256-byte aligned)
        if (bytes > mg->arena_cap - mg->arena_used) {
            mg->err = "the solve needs more device memory than the " +
                      std::to_string(mg->arena_cap >> 20) + " MiB scratch of mf3d_mg_solve_device_noalloc";
            *p = nullptr;
            return MF3D_ENOMEM;
        }
        *p = mg->arena + mg->arena_used;
        mg->arena_used += bytes;
        mg->bytes += (long long)bytes;
        return 0;
    }
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
    if (!src.empty()) {  // on the handle's own stream: never the legacy stream
        HIPCHK(hipMemcpyAsync(*dst, src.data(), sizeof(X) * src.size(), hipMemcpyHostToDevice, mg->stream));
        HIPCHK(hipStreamSynchronize(mg->stream));
    }
    return 0;
}

inline int rdiv_of(const mf3d_block_desc &b) { return b.rdiv > 0 ? b.rdiv : 1; }
inline int lg2i(int v) {
    int r = 0;
    while ((1 << r) < v) ++r;
    return r;
}

// Each block's own level count (localparameters.for:33-56): levels is
// ngrd_gl = ngrid_input + log2(max rdiv), and a block of rdiv r runs
// ngrid_input + log2(r) levels of its own.
std::vector<int> own_levels(int nb, const mf3d_block_desc *b, int L) {
    int top = 0;
    for (int i = 0; i < nb; ++i) top = std::max(top, lg2i(rdiv_of(b[i])));
    std::vector<int> own(nb);
    for (int i = 0; i < nb; ++i) own[i] = L - top + lg2i(rdiv_of(b[i]));
    return own;
}

// Extent along d at level l of a block with `own` levels: beyond them it
// repeats its coarsest grid (mgsolver.for:26-34, gl = min(glevel, ngrid)).
inline int dim_at(const mf3d_block_desc &b, int d, int l, int own = 1 << 30) {
    return b.n[d] >> std::min(l, std::max(own, 1) - 1);
}

// rdv(dom, glevel) of localparameters.for:40-54 at level l (glevel l + 1).
int rdv_at(const mf3d_block_desc &b, int own, int lin, int l) {
    if (l + 1 > own) return 1;
    if (l + 1 <= lin) return rdiv_of(b);
    return rdv_at(b, own, lin, l - 1) / 2;
}

bool is_neumann(const mf3d_block_desc &b, int f) { return b.nbr[f] < 0 && b.bc[f] != 5; }

bool any_refined(int nb, const mf3d_block_desc *b) {
    for (int i = 0; i < nb; ++i)
        if (rdiv_of(b[i]) > 1) return true;
    return false;
}

// Ghost table of level l of a refined block grid (mg_refine.hpp): "" or why
// the level is refused.
std::string ghost_exprs(const std::vector<mf3d_block_desc> &blk, const std::vector<int> &own, int L, int l,
                        std::vector<GhostExpr> *out) {
    int top = 0;
    for (const mf3d_block_desc &B : blk) top = std::max(top, lg2i(rdiv_of(B)));
    std::vector<SymBlock> sb(blk.size());
    long long off = 0;
    for (size_t b = 0; b < blk.size(); ++b) {
        for (int d = 0; d < 3; ++d) sb[b].n[d] = dim_at(blk[b], d, l, own[b]);
        sb[b].off = off;
        sb[b].rdv = rdv_at(blk[b], own[b], L - top, l);
        sb[b].desc = &blk[b];
        off += (long long)(sb[b].n[0] + 2) * (sb[b].n[1] + 2) * (sb[b].n[2] + 2);
    }
    GhostSym sym(sb);
    sym.run();
    return sym.emit(out);
}

// What refined block grids this library runs: the exchange's 2:1 branches
// are face passes only (no block may touch another along an edge or at a
// corner alone: the reference's corner and edge messages have 2:1 branches
// of their own, exchangepp.for:433-1093), a periodic axis only as a lone
// block's wrap onto itself (exchbc_mgpp has no 2:1 branch), constant
// coefficients (the reference's LSM coeff indexes its level arrays past
// their bounds for blocks with fewer levels than the grid).
std::string refined_scope(const mf3d_mg_cfg &c, const mf3d_block_desc *b) {
    if (!any_refined(c.nblocks, b)) return "";
    if (c.variable_coef) return "refined blocks (rdiv > 1) with variable coefficients";
    for (int i = 0; i < c.nblocks; ++i)
        for (int f = 0; f < 6; ++f) {
            if (b[i].bc[f] == 5 && (b[i].per[f] >= 0 || b[i].nbr[f] >= 0 || b[i].nbr[f ^ 1] >= 0))
                return "refined blocks (rdiv > 1): a periodic axis only where a lone block wraps onto itself";
            const int o = b[i].nbr[f];
            if (o < 0) continue;
            const int ri = rdiv_of(b[i]), ro = rdiv_of(b[o]);
            if (ri != ro && ri != 2 * ro && ro != 2 * ri)
                return "refined blocks (rdiv > 1): face neighbours " + std::to_string(i) + " and " +
                       std::to_string(o) + " differ by more than a factor 2";
            for (int g = 0; g < 6; ++g)
                if (g / 2 != f / 2 && b[o].nbr[g] >= 0)
                    return "refined blocks (rdiv > 1): blocks that touch along an edge (block " + std::to_string(i) +
                           " and " + std::to_string(b[o].nbr[g]) + ")";
        }
    return "";
}

// Validate the block set: extents divisible by 2^(L-1), neighbours in range,
// symmetric and conforming (equal extents across a shared face).
std::string validate(const mf3d_mg_cfg &c, const mf3d_block_desc *b) {
    if (c.nblocks <= 0) return "nblocks must be positive";
    if (c.levels <= 0 || c.levels > 16) return "levels out of range";
    if (c.smoother < 1 || c.smoother > 3)
        return "unknown smoother (mg_itrsch: 1 lexicographic GS, 2 ADI-TDMA, 3 red-black GS)";
    if (c.nu_pre < 0 || c.nu_post < 0 || c.maxcy < 0) return "negative sweep or cycle count";
    if (c.precision != MF3D_FP64 && c.precision != MF3D_MIXED) return "unknown precision";
    if (c.coarse_op != MF3D_COARSE_REFERENCE && c.coarse_op != MF3D_COARSE_FACE) return "unknown coarse_op";
    if (c.krylov != MF3D_OUTER_DEFECT && c.krylov != MF3D_OUTER_GCR) return "unknown outer iteration (krylov)";
    if (c.coarse_solve < MF3D_COARSEST_DEFAULT || c.coarse_solve > MF3D_COARSEST_DIRECT)
        return "unknown coarsest-level solve (coarse_solve)";
    if (c.pl_ex < 0) return "negative pl_ex";
    if (c.differencing == 2 && c.variable_coef)
        return "differencing = 2 with a variable coefficient: the reference runs its level set with WENO "
               "differencing only (initial.for:133-135)";
    const std::vector<int> own = own_levels(c.nblocks, b, c.levels);
    for (int i = 0; i < c.nblocks; ++i) {
        if (b[i].rdiv < 0 || (b[i].rdiv & (b[i].rdiv - 1)))
            return "block " + std::to_string(i) + ": rdiv must be a power of 2";
        if (own[i] < 1) return "levels must exceed log2 of the largest rdiv";
    }
    for (int i = 0; i < c.nblocks; ++i) {
        const int div = 1 << (own[i] - 1);
        for (int d = 0; d < 3; ++d) {
            if (b[i].n[d] <= 0) return "block " + std::to_string(i) + ": empty extent";
            if (b[i].n[d] % div)
                return "block " + std::to_string(i) +
                       ": the given number not correct for mgsolver (extent not divisible by "
                       "2^(levels-1))";
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
                for (int e = 0; e < 3; ++e)  // equal physical extent: n x (rdiv of the other)
                    if (e != d && (long long)b[nb].n[e] * rdiv_of(b[i]) != (long long)b[i].n[e] * rdiv_of(b[nb]))
                        return "blocks " + std::to_string(i) + " and " + std::to_string(nb) +
                               " do not conform on their shared face";
            }
        }
    }
    return "";
}

// Ghost regions of one level: each ghost maps, axis by axis (z, then y,
// then x, the order in which the reference's face passes propagate edges
// and corners), to the interior cell it stands for -- across a face
// neighbour, across a periodic partner, onto the block itself for a lone
// periodic block, or reflected at a Neumann wall.
std::vector<Region> build_regions(const std::vector<mf3d_block_desc> &blk, int l, int *max_cnt) {
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
                    *max_cnt = std::max(*max_cnt, r.cnt[0] * r.cnt[1] * r.cnt[2]);
                    out.push_back(r);
                }
    }
    return out;
}

// Cell tiles: 64 x 4 cells of one plane (interior).
std::vector<Tile> cell_tiles(const std::vector<BlockLevel> &hb) {
    std::vector<Tile> t;
    for (int b = 0; b < (int)hb.size(); ++b) {
        const BlockLevel &g = hb[b];
        for (int z = 1; z <= g.n2; ++z)
            for (int y = 1; y <= g.n1; y += kTileY)
                for (int x = 1; x <= g.n0; x += kTileX) t.push_back(Tile{b, x, y, z, z});
    }
    return t;
}

// Row tiles: 4 full rows (ghosted range) of one ghosted plane.
std::vector<Tile> row_tiles(const std::vector<BlockLevel> &hb) {
    std::vector<Tile> t;
    for (int b = 0; b < (int)hb.size(); ++b) {
        const BlockLevel &g = hb[b];
        for (int z = 0; z <= g.n2 + 1; ++z)
            for (int y = 0; y <= g.n1 + 1; y += kTileY) t.push_back(Tile{b, 0, y, z, z});
    }
    return t;
}

// Chunks of kCoefChunk consecutive entries of each block's ghosted arena
// (x0 = the chunk's first entry within the block).
std::vector<Tile> chunk_tiles(const std::vector<BlockLevel> &hb) {
    std::vector<Tile> t;
    for (int b = 0; b < (int)hb.size(); ++b) {
        const BlockLevel &g = hb[b];
        const long long cells = g.sz * (g.n2 + 2);
        for (long long c = 0; c < cells; c += kCoefChunk) t.push_back(Tile{b, (int)c, 0, 0, 0});
    }
    return t;
}

// Octet tiles of the prolongation: 4 coarse-index rows m_y (fine rows
// 2m_y, 2m_y+1) of one coarse-index plane m_z, m = 0 .. n/2 on each axis.
std::vector<Tile> octet_tiles(const std::vector<BlockLevel> &hb) {
    std::vector<Tile> t;
    for (int b = 0; b < (int)hb.size(); ++b) {
        const BlockLevel &g = hb[b];
        for (int z = 0; 2 * z <= g.n2 + 1; ++z)
            for (int y = 0; 2 * y <= g.n1 + 1; y += kTileY) t.push_back(Tile{b, 0, y, z, z});
    }
    return t;
}

// Sweep tiles: 64 x ty columns, a chunk of tz planes (even, so chunks start
// at odd planes).  tz = 0: the deepest chunk (at most 32) that still gives
// the level ~1024 workgroups where it can.
std::vector<Tile> sweep_tiles(const std::vector<BlockLevel> &hb, int ty, int tz = 0) {
    long long cols = 0;
    int nzmax = 1;
    for (const BlockLevel &g : hb) {
        cols += (long long)((g.n0 + SW_TX - 1) / SW_TX) * ((g.n1 + ty - 1) / ty);
        nzmax = std::max(nzmax, g.n2);
    }
    if (tz <= 0) {
        tz = 32;
        while (tz > 2 && cols * ((nzmax + tz - 1) / tz) < 1024) tz >>= 1;
    }
    std::vector<Tile> t;
    for (int b = 0; b < (int)hb.size(); ++b) {
        const BlockLevel &g = hb[b];
        for (int z = 1; z <= g.n2; z += tz)
            for (int y = 1; y <= g.n1; y += ty)
                for (int x = 1; x <= g.n0; x += SW_TX)
                    t.push_back(Tile{b, x, y, z, std::min(g.n2, z + tz - 1)});
    }
    return t;
}

// 3-D sweep tiles (k_rb3d): R3_TX x R3_TY x R3_TZ cells, clipped to the block.
std::vector<Tile> tiles_3d(const std::vector<BlockLevel> &hb) {
    std::vector<Tile> t;
    for (int b = 0; b < (int)hb.size(); ++b) {
        const BlockLevel &g = hb[b];
        for (int z = 1; z <= g.n2; z += R3_TZ)
            for (int y = 1; y <= g.n1; y += R3_TY)
                for (int x = 1; x <= g.n0; x += R3_TX)
                    t.push_back(Tile{b, x, y, z, std::min(g.n2, z + R3_TZ - 1)});
    }
    return t;
}

// The dispatcher hands workgroup i to XCD i mod 8.  Reorder a spatially
// ordered tile list so that XCD k receives the k-th contiguous eighth:
// tiles that share halo planes then run under the same L2.  Results do not
// depend on the order.  MF3D_XCD_ORDER=0 keeps the spatial order.
std::vector<Tile> xcd_order(const std::vector<Tile> &t) {
    const char *env = std::getenv("MF3D_XCD_ORDER");
    if (env && env[0] == '0') return t;
    constexpr size_t X = 8;
    const size_t n = t.size(), q = n / X, r = n % X;
    std::vector<Tile> out(n);
    for (size_t i = 0; i < n; ++i) {
        const size_t k = i % X, j = i / X;
        out[i] = t[k * q + std::min(k, r) + j];
    }
    return out;
}

inline int inner_bytes(const mf3d_mg_cfg &c) { return c.precision == MF3D_MIXED ? 4 : 8; }

long long estimate(const mf3d_mg_cfg &c, const mf3d_block_desc *b) {
    const std::vector<int> own = own_levels(c.nblocks, b, c.levels);
    long long tot = 0;
    const long long w = inner_bytes(c);
    for (int l = 0; l < c.levels; ++l) {
        long long gh = 0;
        for (int i = 0; i < c.nblocks; ++i)
            gh += (long long)(dim_at(b[i], 0, l, own[i]) + 2) * (dim_at(b[i], 1, l, own[i]) + 2) * (dim_at(b[i], 2, l, own[i]) + 2);
        tot += gh * w * 3;                                     // phi (two) and rhs
        if (c.variable_coef) tot += gh * (w * 3 + 8 * 4);      // coefficients, FP64 copies, density
        if (c.smoother == 2) tot += gh * w * 6;                // ADI Thomas coefficients (LDS levels)
    }
    long long gh0 = 0, in0 = 0;
    for (int i = 0; i < c.nblocks; ++i) {
        gh0 += (long long)(b[i].n[0] + 2) * (b[i].n[1] + 2) * (b[i].n[2] + 2);
        in0 += (long long)b[i].n[0] * b[i].n[1] * b[i].n[2];
    }
    tot += gh0 * 8 * 3;             // p (two buffers) and f
    if (c.krylov == MF3D_OUTER_GCR) tot += gh0 * 8 * 4;  // d (two buffers), w, r
    if (c.smoother == 2) tot += gh0 * w;                 // ADI a(m) scratch (blocks too large for LDS)

    tot += std::max(gh0, in0) * 8;  // I/O staging
    if (c.levels >= 2 && c.precision == MF3D_MIXED) {  // FP64 coarsest potential (two) and RHS
        long long ghc = 0;
        for (int i = 0; i < c.nblocks; ++i)
            ghc += (long long)(dim_at(b[i], 0, c.levels - 1, own[i]) + 2) * (dim_at(b[i], 1, c.levels - 1, own[i]) + 2) *
                   (dim_at(b[i], 2, c.levels - 1, own[i]) + 2);
        tot += 3 * 8 * ghc;
    }
    const bool direct = c.levels >= 2 && c.coarse_solve == MF3D_COARSEST_DIRECT;
    if (direct) {  // the coarsest level's inverse operator
        long long nc = 0;
        for (int i = 0; i < c.nblocks; ++i)
            nc += (long long)dim_at(b[i], 0, c.levels - 1, own[i]) * dim_at(b[i], 1, c.levels - 1, own[i]) *
                  dim_at(b[i], 2, c.levels - 1, own[i]);
        if (nc <= MF3D_COARSEST_MAX_CELLS) tot += nc * nc * 8 + nc * 64;
    }
    return tot;
}

// ------------------------------------------------------------ timing hook
inline bool tk_on(const mf3d_mg *mg, int kernel, int level) { return mg->tk == kernel && level == 0; }

void tk_mark(mf3d_mg *mg) {
    mf3d_mg::EvList &E = mg->ev_rec ? *mg->ev_rec : mg->ev_eager;
    if (E.used == E.ev.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) {
            mg->ev_fail = true;
            return;
        }
        E.ev.push_back(e);
    }
    if (hipEventRecord(E.ev[E.used++], mg->stream) != hipSuccess) mg->ev_fail = true;
}

// Call after the stream is synchronised: fold the recorded pairs into the
// totals.  A captured list keeps its events (its graph records them again).
void tk_flush(mf3d_mg *mg, mf3d_mg::EvList &E, bool captured) {
    for (size_t i = 0; i + 1 < E.used; i += 2) {
        float ms = 0.f;
        const hipError_t e = hipEventElapsedTime(&ms, E.ev[i], E.ev[i + 1]);
        if (e == hipSuccess) {
            mg->tk_ms += ms;
            mg->tk_launches += 1;
        } else {
            (void)hipGetLastError();  // not a device error: the pair is just not counted
            if (captured && !mg->tk_eager) {
                std::fprintf(stderr, "mf3d_mg: no timing from events recorded by a graph (%s); "
                                     "timing eager launches from the next solve\n", hipGetErrorString(e));
                mg->tk_eager = true;  // the timed graphs are dropped at the next solve
            }
        }
    }
    if (!captured) E.used = 0;
}
void tk_flush(mf3d_mg *mg) { tk_flush(mg, mg->ev_eager, false); }

void drop_timed_graphs(mf3d_mg *mg) {
    for (hipGraphExec_t &g : mg->tkg)
        if (g) (void)hipGraphExecDestroy(g);
    mg->tkg.clear();
    for (mf3d_mg::EvList &E : mg->tkev)
        for (hipEvent_t e : E.ev) (void)hipEventDestroy(e);
    mg->tkev.clear();
}

// ---------------------------------------------------------------- launches
inline const int *done_ptr(mf3d_mg *mg) { return &mg->st->done; }

template <typename TD, typename TS>
void convert(mf3d_mg *mg, TD *dst, const TS *src, long long n);

template <typename T>
void fill(mf3d_mg *mg, T *a, long long n, const int *done = nullptr) {
    if (n <= 0) return;
    long long blocks = std::min<long long>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)blocks), dim3(256), 0, mg->stream, a, n, T(0), done);
}

// Every ghost of a level-l arena a (laid out as the level) from its
// interior: the region table, or a refined grid's expression table.
template <typename T>
void fill_ghosts(mf3d_mg *mg, const Level &L, T *a, const int *done) {
    if (L.n_gx) {
        const unsigned gb = (unsigned)std::min(4096, (L.n_gx + 255) / 256);
        hipLaunchKernelGGL(k_gather_expr<T>, dim3(gb), dim3(256), 0, mg->stream, a, L.gx, L.n_gx, done);
        return;
    }
    if (!L.n_full) return;
    const int gx = std::max(1, std::min(64, (L.max_full + 255) / 256));
    hipLaunchKernelGGL(k_gather<T>, dim3(gx, L.n_full), dim3(256), 0, mg->stream, a, L.db, L.r_full, done);
}

template <typename T>
void gather(mf3d_mg *mg, int l, const int *done) {
    Level &L = mg->lv[l];
    if (L.own_ghosts) return;
    fill_ghosts<T>(mg, L, (T *)L.phi[L.cur], done);
}

// The march's lag (planes between the two colours): 2 (one barrier a step)
// for the 8-row tiles' FP32 constant-coefficient sweeps, 1 (two barriers,
// one step fewer a chunk) elsewhere -- by measurement: 54.6 -> 49.4 us for
// the 256^3 FP32 sweep, while the FP64 one gains nothing (118 -> 120 us),
// the 4-row, 8-plane chunks of 128^3 lose (13.0 -> 14.2 us) and so does
// the variable-coefficient FP32 sweep at 384^3 (468 -> 539 us).
template <typename T, bool VAR>
constexpr int kRbsLag8 = sizeof(T) == 4 && !VAR ? 2 : 1;

// The mixed form's FP32 inner cycle multiplies by 1/ap; the FP64 form
// divides, as the reference does.
template <typename T, bool VAR, bool ZERO>
void sweep_launch(mf3d_mg *mg, Level &L, const int *done) {
    constexpr bool RCP = sizeof(T) == 4;
    const T *pin = (const T *)L.phi[L.cur];
    T *pout = (T *)L.phi[1 - L.cur];
    const uint32_t bytes = (uint32_t)(L.cells_gh * (long long)sizeof(T));
    const int own = L.own_ghosts ? 1 : 0;
    if (L.sw_ty == 0) {  // 3-D tiles
        hipLaunchKernelGGL((k_rb3d<T, VAR, ZERO, RCP>), dim3(L.n_sw), dim3(R3_NT), 0, mg->stream, pin, pout,
                           (const T *)L.rhs, (const T *)L.cx, (const T *)L.cy, (const T *)L.cz, bytes, own, L.db,
                           L.t_sw, done);
        return;
    }
    if (ZERO && mg->zmode == 1) {  // from zero: the march over the RHS alone
        if (L.swz_ty == 8)
            hipLaunchKernelGGL((k_rbz<T, VAR, RCP, 8>), dim3(L.n_swz), dim3(SW_TX, 8), 0, mg->stream, pout,
                               (const T *)L.rhs, (const T *)L.cx, (const T *)L.cy, (const T *)L.cz, bytes, own, L.db,
                               L.t_swz, done);
        else
            hipLaunchKernelGGL((k_rbz<T, VAR, RCP, 4>), dim3(L.n_swz), dim3(SW_TX, 4), 0, mg->stream, pout,
                               (const T *)L.rhs, (const T *)L.cx, (const T *)L.cy, (const T *)L.cz, bytes, own, L.db,
                               L.t_swz, done);
        return;
    }
    if (L.sw_ty == 8)
        hipLaunchKernelGGL((k_rbs<T, VAR, ZERO, RCP, 8, kRbsLag8<T, VAR>>), dim3(L.n_sw), dim3(SW_TX, 8), 0, mg->stream, pin, pout,
                           (const T *)L.rhs, (const T *)L.cx, (const T *)L.cy, (const T *)L.cz, bytes, own, L.db,
                           L.t_sw, done);
    else
        hipLaunchKernelGGL((k_rbs<T, VAR, ZERO, RCP, 4, 1>), dim3(L.n_sw), dim3(SW_TX, 4), 0, mg->stream, pin, pout,
                           (const T *)L.rhs, (const T *)L.cx, (const T *)L.cy, (const T *)L.cz, bytes, own, L.db,
                           L.t_sw, done);
}

// Threads of an ADI workgroup: the most lines one diagonal of the level holds.
unsigned adi_threads(const Level &V) {
    int lines = 1;  // the most lines one diagonal of any pass holds
    for (const BlockLevel &g : V.hb)
        lines = std::max({lines, std::min(g.n0, g.n2), std::min(g.n1, g.n2), std::min(g.n0, g.n1)});
    return (unsigned)std::min(256, (lines + 63) / 64 * 64);
}

// One in-place sweep of the lexicographic (1) or ADI (2) smoother on level
// l's current potential: a launch per hyperplane, or one launch per
// direction (j, i, k lines, mgsolver.for:433-535) with a workgroup per
// block.
template <typename T, bool VAR>
void sweep_inplace(mf3d_mg *mg, Level &L, const int *done) {
    int m0 = 1, m1 = 1, m2 = 1, nb = mg->cfg.nblocks;
    for (const BlockLevel &g : L.hb) {
        m0 = std::max(m0, g.n0);
        m1 = std::max(m1, g.n1);
        m2 = std::max(m2, g.n2);
    }
    T *p = (T *)L.phi[L.cur];
    const T *rhs = (const T *)L.rhs, *cx = (const T *)L.cx, *cy = (const T *)L.cy, *cz = (const T *)L.cz;
    if (mg->cfg.smoother == 1) {
        constexpr bool RCP = sizeof(T) == 4;
        const dim3 grid((unsigned)((m1 + 15) / 16), (unsigned)((m2 + 15) / 16), (unsigned)nb);
        int smax = 0;
        for (const BlockLevel &g : L.hb) smax = std::max(smax, g.n0 + g.n1 + g.n2);
        for (int s = 3; s <= smax; ++s)
            hipLaunchKernelGGL((k_lex<T, VAR, RCP>), grid, dim3(256), 0, mg->stream, p, rhs, cx, cy, cz, L.db, s,
                               done);
        return;
    }
    const unsigned nt = adi_threads(L);
    if (L.adi_lds) {  // a 256-thread workgroup per block
        const AdiCoef<T> ac{{(const T *)L.adi_apr[0], (const T *)L.adi_apr[1], (const T *)L.adi_apr[2]},
                            {(const T *)L.adi_a[0], (const T *)L.adi_a[1], (const T *)L.adi_a[2]}};
        if (L.adi_rhs)
            hipLaunchKernelGGL((k_adi_lds<T, VAR, true>), dim3(nb), dim3(256), L.adi_lds, mg->stream, p, rhs, cx, cy,
                               cz, ac, L.db, L.adi_cells, L.adi_sstride, done);
        else
            hipLaunchKernelGGL((k_adi_lds<T, VAR, false>), dim3(nb), dim3(256), L.adi_lds, mg->stream, p, rhs, cx, cy,
                               cz, ac, L.db, L.adi_cells, L.adi_sstride, done);
        return;
    }
    T *sa = (T *)mg->d_adi;
    hipLaunchKernelGGL((k_adi<T, VAR, 1>), dim3(nb), dim3(nt), 0, mg->stream, p, sa, rhs, cx, cy, cz, L.db, done);
    hipLaunchKernelGGL((k_adi<T, VAR, 0>), dim3(nb), dim3(nt), 0, mg->stream, p, sa, rhs, cx, cy, cz, L.db, done);
    hipLaunchKernelGGL((k_adi<T, VAR, 2>), dim3(nb), dim3(nt), 0, mg->stream, p, sa, rhs, cx, cy, cz, L.db, done);
}

// `sweeps` sweeps at level l; the first treats the potential as zero when
// zero_first (after restriction, or at the start of a cycle from pp = 0).
template <typename T, bool VAR>
void relax(mf3d_mg *mg, int l, int sweeps, bool zero_first, const int *done) {
    Level &L = mg->lv[l];
    if (mg->cfg.smoother != 3) {  // in place, from an explicit zero
        if (zero_first) fill<T>(mg, (T *)L.phi[L.cur], L.cells_gh, done);
        for (int s = 0; s < sweeps; ++s) {
            const bool tk = tk_on(mg, MF3D_K_SMOOTH, l);
            if (tk) tk_mark(mg);
            sweep_inplace<T, VAR>(mg, L, done);
            if (tk) tk_mark(mg);
            gather<T>(mg, l, done);
        }
        return;
    }
    if (sweeps == 0 && zero_first) {  // nothing will overwrite the zero the reference leaves
        fill<T>(mg, (T *)L.phi[L.cur], L.cells_gh, done);
        return;
    }
    for (int s = 0; s < sweeps; ++s) {
        const bool tk = tk_on(mg, MF3D_K_SMOOTH, l);
        if (tk) tk_mark(mg);
        if (zero_first && s == 0) sweep_launch<T, VAR, true>(mg, L, done);
        else sweep_launch<T, VAR, false>(mg, L, done);
        if (tk) tk_mark(mg);
        L.cur = 1 - L.cur;
        gather<T>(mg, l, done);
    }
}

template <typename T, bool VAR>
void restrict_to(mf3d_mg *mg, int lc, const int *done) {
    Level &C = mg->lv[lc], &F = mg->lv[lc - 1];
    const bool tk = tk_on(mg, MF3D_K_RESTRICT, lc - 1);
    if (tk) tk_mark(mg);
    if (C.n_rst)
        hipLaunchKernelGGL((k_restrict<T, VAR>), dim3(C.n_rst), dim3(kTileX, kTileY), 0, mg->stream,
                           (const T *)F.phi[F.cur], (const T *)F.rhs, (const T *)F.cx, (const T *)F.cy,
                           (const T *)F.cz, F.db, (T *)C.rhs, C.db, C.t_rst, done);
    if (C.n_rrep)  // blocks repeating their grid at lc (refined grids)
        hipLaunchKernelGGL((k_restrict_rep<T, VAR>), dim3(C.n_rrep), dim3(kTileX, kTileY), 0, mg->stream,
                           (const T *)F.phi[F.cur], (const T *)F.rhs, (const T *)F.cx, (const T *)F.cy,
                           (const T *)F.cz, F.db, (T *)C.rhs, C.db, C.t_rrep, done);
    if (tk) tk_mark(mg);
}

template <typename T>
void prolong_to(mf3d_mg *mg, int lf, const int *done) {
    Level &C = mg->lv[lf + 1], &F = mg->lv[lf];
    const bool tk = tk_on(mg, MF3D_K_PROLONG, lf);
    if (tk) tk_mark(mg);
    if (F.n_oct)
        hipLaunchKernelGGL(k_prolong<T>, dim3(F.n_oct), dim3(kTileX, kTileY), 0, mg->stream,
                           (const T *)F.phi[F.cur], (T *)F.phi[1 - F.cur], F.db, (const T *)C.phi[C.cur], C.db,
                           F.t_oct, done);
    if (F.n_prep)  // blocks whose next level repeats their grid (refined grids)
        hipLaunchKernelGGL(k_prolong_rep<T>, dim3(F.n_prep), dim3(kTileX, kTileY), 0, mg->stream,
                           (const T *)F.phi[F.cur], (T *)F.phi[1 - F.cur], F.db, (const T *)C.phi[C.cur], C.db,
                           F.t_prep, done);
    if (tk) tk_mark(mg);
    F.cur = 1 - F.cur;
}

// The reference's sweeps of the coarsest level in FP64 for the mixed form
// (per-level path): widen the level's RHS, run the sweeps on an FP64 copy
// of the level (same tiles, ghosts by the sweeps' own images or the
// gather), narrow the potential, ghosts included.
template <bool VAR>
void coarse_smooth64(mf3d_mg *mg, int sweeps, const int *done) {
    Level &C = mg->lv[mg->cfg.levels - 1];
    Level S = C;  // the FP64 twin: same tables and tiles
    S.phi[0] = mg->c64_phi[0];
    S.phi[1] = mg->c64_phi[1];
    S.rhs = mg->c64_rhs;
    S.cx = C.dcx;
    S.cy = C.dcy;
    S.cz = C.dcz;
    S.cur = 0;
    convert<double, float>(mg, mg->c64_rhs, (const float *)C.rhs, C.cells_gh);
    if (sweeps == 0) fill<double>(mg, (double *)S.phi[0], S.cells_gh, done);
    for (int s = 0; s < sweeps; ++s) {
        if (s == 0) sweep_launch<double, VAR, true>(mg, S, done);
        else sweep_launch<double, VAR, false>(mg, S, done);
        S.cur = 1 - S.cur;
        if (!S.own_ghosts) fill_ghosts<double>(mg, S, (double *)S.phi[S.cur], done);
    }
    convert<float, double>(mg, (float *)C.phi[1 - C.cur], (const double *)S.phi[S.cur], C.cells_gh);
    C.cur = 1 - C.cur;
}

// The coarsest level's FP64 direct solve: phi = Minv * rhs on every block's
// interior, then the ghost fill from the level's region table (lone blocks
// included, whose sweeps otherwise store their own ghost images).
template <typename T>
void coarse_direct(mf3d_mg *mg, const int *done) {
    Level &C = mg->lv[mg->cfg.levels - 1];
    T *out = (T *)C.phi[1 - C.cur];
    const unsigned gb = (unsigned)std::min(512, (mg->nc + 3) / 4);
    hipLaunchKernelGGL(k_coarse_solve<T>, dim3(gb), dim3(256), sizeof(double) * (size_t)mg->nc, mg->stream,
                       (const double *)mg->d_cinv, mg->nc, (const CCell *)mg->d_ccell, C.db, (const T *)C.rhs, out,
                       done);
    C.cur = 1 - C.cur;
    fill_ghosts<T>(mg, C, out, done);
}

// Levels lt..L-1 in one launch (k_tail); level lt's RHS is in place and its
// potential starts from zero (or from phi_in when given).
template <typename T, bool VAR>
void tail(mf3d_mg *mg, bool from_zero, const int *done) {
    const int lt = mg->lt, L = mg->cfg.levels;
    Level &V = mg->lv[lt];
    TailArgs<T> co{};
    for (int l = lt; l < L; ++l) {
        co.cx[l - lt] = (const T *)mg->lv[l].cx;
        co.cy[l - lt] = (const T *)mg->lv[l].cy;
        co.cz[l - lt] = (const T *)mg->lv[l].cz;
    }
    hipLaunchKernelGGL((k_tail<T, VAR>), dim3(mg->cfg.nblocks), dim3(TAIL_NT), mg->tail_lds, mg->stream, mg->d_tail,
                       mg->cfg.nblocks, L - lt, (const T *)V.rhs, from_zero ? nullptr : (const T *)V.phi[V.cur],
                       (T *)V.phi[1 - V.cur], co, mg->cfg.nu_pre, mg->cfg.nu_post,
                       mg->cdirect ? (const double *)mg->d_cinv : nullptr, mg->nc,
                       TailC64<T>{mg->c64 ? 1 : 0, mg->lv[L - 1].dcx, mg->lv[L - 1].dcy, mg->lv[L - 1].dcz}, done);
    V.cur = 1 - V.cur;
}

// The prolongation into level l fused with its first post-sweep: the level
// sweeps on 3-D tiles, every block takes the trilinear correction (none
// repeats the coarse grid) and the red-black smoother has a post-sweep.
bool fused_prolong(const mf3d_mg *mg, int l) {
    const Level &F = mg->lv[l];
    return mg->fuse_pro && F.sw_ty == 0 && F.n_prep == 0 && F.n_oct > 0 && mg->cfg.nu_post >= 1 &&
           mg->cfg.smoother == 3;
}

template <typename T, bool VAR>
void prolong_sweep(mf3d_mg *mg, int l, const int *done) {
    Level &C = mg->lv[l + 1], &F = mg->lv[l];
    constexpr bool RCP = sizeof(T) == 4;
    const uint32_t bytes = (uint32_t)(F.cells_gh * (long long)sizeof(T));
    const bool tk = tk_on(mg, MF3D_K_SMOOTH, l);
    if (tk) tk_mark(mg);
    hipLaunchKernelGGL((k_rb3d<T, VAR, false, RCP, true>), dim3(F.n_sw), dim3(R3_NT), 0, mg->stream,
                       (const T *)F.phi[F.cur], (T *)F.phi[1 - F.cur], (const T *)F.rhs, (const T *)F.cx,
                       (const T *)F.cy, (const T *)F.cz, bytes, F.own_ghosts ? 1 : 0, F.db, F.t_sw, done,
                       (const T *)C.phi[C.cur], C.db);
    if (tk) tk_mark(mg);
    F.cur = 1 - F.cur;
    gather<T>(mg, l, done);
}

template <typename T, bool VAR>
void vcycle_dev(mf3d_mg *mg, bool zero_start, const int *done) {
    const int L = mg->cfg.levels, lt = mg->lt;
    if (lt == 0) {
        tail<T, VAR>(mg, zero_start, done);
        return;
    }
    relax<T, VAR>(mg, 0, mg->cfg.nu_pre, zero_start, done);
    int post0 = mg->cfg.nu_post;  // level 0's post-sweeps left after a fused prolongation
    if (L > 1) {
        const int ldown = std::min(lt, L);
        for (int l = 1; l < ldown; ++l) {
            restrict_to<T, VAR>(mg, l, done);
            if (l == L - 1 && mg->cdirect) coarse_direct<T>(mg, done);
            else if (l == L - 1 && mg->c64 && sizeof(T) == 4)
                coarse_smooth64<VAR>(mg, mg->cfg.nu_pre + mg->cfg.nu_post, done);
            else relax<T, VAR>(mg, l, mg->cfg.nu_pre + (l == L - 1 ? mg->cfg.nu_post : 0), true, done);
        }
        if (lt < L) {
            restrict_to<T, VAR>(mg, lt, done);
            tail<T, VAR>(mg, true, done);
        }
        for (int l = ldown - 1; l >= 0; --l) {
            if (l == L - 1) continue;
            int post = mg->cfg.nu_post;
            if (fused_prolong(mg, l)) {
                prolong_sweep<T, VAR>(mg, l, done);
                --post;
            } else {
                prolong_to<T>(mg, l, done);
            }
            if (l > 0) relax<T, VAR>(mg, l, post, false, done);
            else post0 = post;
        }
    }
    relax<T, VAR>(mg, 0, post0, false, done);
}

// Outer step after the V-cycle of cycle k (from = k mod 2; k = 0: the
// initial residual, first = true).  Defect correction: p += e, r = f - A p.
// GCR(1): beta from (A z, w_prev), d = z - beta d_prev, w = A d, alpha,
// p += alpha d, r = f - A p.
template <typename T, bool VAR>
void outer_step(mf3d_mg *mg, int from, const int *done, bool first = false) {
    Level &L0 = mg->lv[0];
    const double *cx = VAR ? L0.dcx : nullptr, *cy = VAR ? L0.dcy : nullptr, *cz = VAR ? L0.dcz : nullptr;
    const bool tk = tk_on(mg, MF3D_K_OUTER, 0);
    if (mg->cfg.krylov == MF3D_OUTER_GCR) {
        const T *z = (const T *)L0.phi[L0.cur];
        if (!first) {
            hipLaunchKernelGGL((k_krylov_lds<2, T, T, VAR, 8>), dim3(mg->n_out), dim3(SW_TX, 8), 0, mg->stream, z,
                               nullptr, nullptr, nullptr, mg->kw, nullptr, nullptr, cx, cy, cz, mg->st, L0.db,
                               mg->t_out, mg->partials, done);
            hipLaunchKernelGGL(k_gcr_scalar<true>, dim3(1), dim3(1024), 0, mg->stream, mg->partials, mg->nparts,
                               mg->st);
            hipLaunchKernelGGL((k_krylov_lds<1, T, T, VAR, 8>), dim3(mg->n_out), dim3(SW_TX, 8), 0, mg->stream, z,
                               mg->kd[from], mg->kd[1 - from], nullptr, mg->kr, mg->kw, nullptr, cx, cy, cz, mg->st,
                               L0.db, mg->t_out, mg->partials, done);
            hipLaunchKernelGGL(k_gcr_scalar<false>, dim3(1), dim3(1024), 0, mg->stream, mg->partials, mg->nparts,
                               mg->st);
        }
        if (mg->refined) {  // p_new, its exchange (2:1 ghosts), then r = f - A p_new
            const unsigned gb = (unsigned)std::min<long long>(16384, (L0.cells_gh + 255) / 256);
            hipLaunchKernelGGL(k_paxpy, dim3(gb), dim3(256), 0, mg->stream, mg->pbuf[1 - from], mg->pbuf[from],
                               mg->kd[1 - from], mg->st, L0.cells_gh, done);
            fill_ghosts<double>(mg, L0, mg->pbuf[1 - from], done);
            if (tk) tk_mark(mg);
            hipLaunchKernelGGL((k_krylov_lds<3, double, T, VAR, 8>), dim3(mg->n_out), dim3(SW_TX, 8), 0,
                               mg->stream, mg->pbuf[1 - from], nullptr, nullptr, mg->f_g, nullptr, mg->kr,
                               (T *)L0.rhs, cx, cy, cz, mg->st, L0.db, mg->t_out, mg->partials, done);
            if (tk) tk_mark(mg);
        } else {
            if (tk) tk_mark(mg);
            hipLaunchKernelGGL((k_krylov_lds<0, double, T, VAR, 8>), dim3(mg->n_out), dim3(SW_TX, 8), 0,
                               mg->stream, mg->pbuf[from], mg->kd[1 - from], mg->pbuf[1 - from], mg->f_g, nullptr,
                               mg->kr, (T *)L0.rhs, cx, cy, cz, mg->st, L0.db, mg->t_out, mg->partials, done);
            if (tk) tk_mark(mg);
        }
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, mg->stream, mg->partials, mg->nparts, mg->st,
                           mg->d_hist, mg->d_rmax, mg->cfg.min_cycles);
        return;
    }
    if (mg->refined) {  // p_new, its exchange, then r = f - A p_new
        const unsigned gb = (unsigned)std::min<long long>(16384, (L0.cells_gh + 255) / 256);
        hipLaunchKernelGGL(k_padd<T>, dim3(gb), dim3(256), 0, mg->stream, mg->pbuf[1 - from], mg->pbuf[from],
                           (const T *)L0.phi[L0.cur], L0.cells_gh, done);
        fill_ghosts<double>(mg, L0, mg->pbuf[1 - from], done);
        if (tk) tk_mark(mg);
        hipLaunchKernelGGL((k_outer_lds<T, VAR, 8, false>), dim3(mg->n_out), dim3(SW_TX, 8), 0, mg->stream,
                           mg->pbuf[1 - from], nullptr, (const T *)nullptr, mg->f_g, (T *)L0.rhs, cx, cy, cz, L0.db,
                           mg->t_out, mg->partials, done);
        if (tk) tk_mark(mg);
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, mg->stream, mg->partials, mg->nparts, mg->st,
                           mg->d_hist, mg->d_rmax, mg->cfg.min_cycles);
        return;
    }
    if (tk) tk_mark(mg);
    hipLaunchKernelGGL((k_outer_lds<T, VAR, 8>), dim3(mg->n_out), dim3(SW_TX, 8), 0, mg->stream, mg->pbuf[from],
                       mg->pbuf[1 - from], (const T *)L0.phi[L0.cur], mg->f_g, (T *)L0.rhs, cx, cy, cz, L0.db,
                       mg->t_out, mg->partials, done);
    if (tk) tk_mark(mg);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, mg->stream, mg->partials, mg->nparts, mg->st,
                       mg->d_hist, mg->d_rmax, mg->cfg.min_cycles);
}

// Record one cycle reading pbuf[g] into a graph.
template <typename T, bool VAR>
int capture_cycles(mf3d_mg *mg) {
    for (int g = 0; g < 2; ++g) {
        for (Level &V : mg->lv) V.cur = 0;
        HIPCHK(hipStreamBeginCapture(mg->stream, hipStreamCaptureModeRelaxed));
        vcycle_dev<T, VAR>(mg, true, done_ptr(mg));
        outer_step<T, VAR>(mg, g, done_ptr(mg));
        hipGraph_t graph = nullptr;
        HIPCHK(hipStreamEndCapture(mg->stream, &graph));
        hipError_t e = hipGraphInstantiate(&mg->cyc[g], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            mg->err = std::string("hipGraphInstantiate: ") + hipGetErrorString(e);
            return MF3D_EDEVICE;
        }
    }
    mg->graphs_ready = true;
    return 0;
}

// Record cycle k (reading pbuf[k % 2]) with the timing hook's events into
// its own graph.  Returns 1 when the runtime refuses to capture the event
// records; the hook then times eager launches instead.
template <typename T, bool VAR>
int capture_timed(mf3d_mg *mg, int k) {
    if ((int)mg->tkg.size() <= k) {
        mg->tkg.resize(k + 1, nullptr);
        mg->tkev.resize(k + 1);
    }
    if (mg->tkg[k]) return 0;
    for (Level &V : mg->lv) V.cur = 0;
    mg->tkev[k].used = 0;
    // events are created before the capture, none while it runs
    for (size_t want = 2 * (size_t)(mg->cfg.nu_pre + mg->cfg.nu_post + 2); mg->tkev[k].ev.size() < want;) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        mg->tkev[k].ev.push_back(e);
    }
    mg->ev_rec = &mg->tkev[k];
    mg->ev_fail = false;
    HIPCHK(hipStreamBeginCapture(mg->stream, hipStreamCaptureModeRelaxed));
    vcycle_dev<T, VAR>(mg, true, done_ptr(mg));
    outer_step<T, VAR>(mg, k % 2, done_ptr(mg));
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(mg->stream, &graph);
    mg->ev_rec = nullptr;
    if (e != hipSuccess || mg->ev_fail || !graph) {
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        return 1;
    }
    e = hipGraphInstantiate(&mg->tkg[k], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        mg->tkg[k] = nullptr;
        (void)hipGetLastError();
        return 1;
    }
    return 0;
}

template <typename T, bool VAR>
int solve_dev(mf3d_mg *mg, const double *f_packed, double *p_packed, double *rmax_out, int32_t *iters_out,
              double *hist) {
    Level &L0 = mg->lv[0];
    const int maxcy = mg->cfg.maxcy;
    int chunk = mg->guess > 0 ? mg->guess : 4;
    // the allocation-free solve launches its kernels one by one
    // (instantiating a graph may allocate)
    const bool use_graph = mg->tk == MF3D_K_OFF && !mg->arena;
    if (use_graph && !mg->graphs_ready) {
        int rc = capture_cycles<T, VAR>(mg);
        if (rc) return rc;
    }
    // timed graphs for the first chunk, captured while the stream is idle
    if (mg->tk_eager) drop_timed_graphs(mg);
    bool timed = mg->tk != MF3D_K_OFF && !mg->tk_eager;
    for (int k = 1; timed && k <= std::min(chunk, maxcy); ++k)
        if (capture_timed<T, VAR>(mg, k)) {
            drop_timed_graphs(mg);
            mg->tk_eager = true;
            timed = false;
        }
    mg->last_replay = use_graph || timed;
    hipLaunchKernelGGL(k_reset_state, dim3(1), dim3(1), 0, mg->stream, mg->st, mg->cfg.eps);
    fill<double>(mg, mg->pbuf[0], L0.cells_gh);
    fill<T>(mg, (T *)L0.phi[L0.cur], L0.cells_gh);
    if (mg->cfg.krylov == MF3D_OUTER_GCR) {
        fill<double>(mg, mg->kd[0], L0.cells_gh);
        fill<double>(mg, mg->kd[1], L0.cells_gh);
        fill<double>(mg, mg->kw, L0.cells_gh);
    }
    hipLaunchKernelGGL(k_unpack<double>, dim3(L0.n_in), dim3(kTileX, kTileY), 0, mg->stream, f_packed, mg->f_g,
                       L0.db, L0.t_in);
    // step 0: r_0 = f - A*0 from pbuf[0] into pbuf[1]
    outer_step<T, VAR>(mg, 0, nullptr, true);
    int launched = 0;
    SolveState hs{};
    hs.done = 0;
    while (launched < maxcy) {
        const int n = std::min(chunk, maxcy - launched);
        // later chunks follow a synchronisation: the stream is idle here
        for (int c = 0; timed && c < n; ++c)
            if (capture_timed<T, VAR>(mg, launched + c + 1)) {
                mg->err = "capture of the timing hook's events failed after earlier captures succeeded";
                return MF3D_EDEVICE;
            }
        for (int c = 0; c < n; ++c) {
            const int k = launched + c + 1;  // cycle k reads pbuf[k % 2]
            if (use_graph) {
                HIPCHK(hipGraphLaunch(mg->cyc[k % 2], mg->stream));
            } else if (timed) {
                HIPCHK(hipGraphLaunch(mg->tkg[k], mg->stream));
            } else {
                for (Level &V : mg->lv) V.cur = 0;
                vcycle_dev<T, VAR>(mg, true, done_ptr(mg));
                outer_step<T, VAR>(mg, k % 2, done_ptr(mg));
            }
        }
        launched += n;
        HIPCHK(hipMemcpyAsync(&hs, mg->st, sizeof(SolveState), hipMemcpyDeviceToHost, mg->stream));
        HIPCHK(hipStreamSynchronize(mg->stream));
        tk_flush(mg);
        // cycles after convergence skip their work: their events are not counted
        const int ran = hs.done ? hs.last : launched;
        for (int k = launched - n + 1; timed && k <= ran; ++k) tk_flush(mg, mg->tkev[k], true);
        if (hs.done) break;
        chunk = 1;
    }
    const int k = hs.last;  // cycles completed; p_k sits in pbuf[(k + 1) % 2]
    if (k < 0) {
        mg->err = "no outer step recorded";
        return MF3D_EDEVICE;
    }
    mg->guess = k;
    hipLaunchKernelGGL(k_pack<double>, dim3(L0.n_in), dim3(kTileX, kTileY), 0, mg->stream, mg->pbuf[(k + 1) % 2],
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
    return hs.done ? MF3D_OK : MF3D_NOT_CONVERGED;
}

// Host arrays in the caller's layout <-> packed staging (ring = ghost
// layers kept around the interior).
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
                const double *row = src[b] + (size_t)s + (size_t)ld0 * ((size_t)(j + s) + (size_t)ld1 * (k + s));
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
                double *row = dst[b] + (size_t)s + (size_t)ld0 * ((size_t)(j + s) + (size_t)ld1 * (k + s));
                std::memcpy(row, src + o, sizeof(double) * e0);
                o += e0;
            }
    }
}

template <typename TD, typename TS>
void convert(mf3d_mg *mg, TD *dst, const TS *src, long long n) {
    const unsigned gb = (unsigned)std::min<long long>((n + 255) / 256, 16384);
    hipLaunchKernelGGL((k_convert<TD, TS>), dim3(gb), dim3(256), 0, mg->stream, dst, src, n);
}

// mgkcyc: the caller's pp (one ghost layer, same shape as the level-0 arena)
// in and out, rhs into the level-0 RHS.
template <typename T, bool VAR>
int vcycle_host_t(mf3d_mg *mg) {
    Level &L0 = mg->lv[0];
    for (Level &V : mg->lv) V.cur = 0;
    convert<T, double>(mg, (T *)L0.phi[0], mg->d_io, L0.cells_gh);
    vcycle_dev<T, VAR>(mg, false, nullptr);
    convert<double, T>(mg, mg->d_io, (const T *)L0.phi[L0.cur], L0.cells_gh);
    return 0;
}

// The coarse tail runs in LDS when every block stands alone (no face
// neighbour, no periodic partner) and the levels from lt down fit.
constexpr size_t kTailLdsMax = 150 * 1024;

int choose_tail(mf3d_mg *mg) {
    const int L = mg->cfg.levels, nb = mg->cfg.nblocks;
    mg->lt = L;
    const char *env = std::getenv("MF3D_TAIL");
    if (env && env[0] == '0') return 0;
    if (mg->cfg.smoother != 3) return 0;  // the tail smooths red-black
    for (const mf3d_block_desc &B : mg->blk)
        for (int f = 0; f < 6; ++f)
            if (B.nbr[f] >= 0 || B.per[f] >= 0) return 0;
    if (mg->cdirect && nb > 1) return 0;  // the tail's direct solve covers one block
    const size_t w = inner_bytes(mg->cfg);
    for (int lt = 0; lt < L; ++lt) {
        size_t need = 0;
        for (const mf3d_block_desc &B : mg->blk) {
            size_t s = 0;
            for (int l = lt; l < L; ++l)
                s += 2 * w * (size_t)(dim_at(B, 0, l) + 2) * (dim_at(B, 1, l) + 2) * (dim_at(B, 2, l) + 2);
            need = std::max(need, s);
        }
        // the direct solve's FP64 staging of the coarsest RHS, or the FP64
        // potential and RHS of its sweeps, follows the level arrays
        if (mg->cdirect) need = (need + 7) / 8 * 8 + sizeof(double) * (size_t)mg->nc;
        else if (mg->c64) {
            size_t ghc = 0;
            for (const mf3d_block_desc &B : mg->blk)
                ghc = std::max(ghc, (size_t)(dim_at(B, 0, L - 1) + 2) * (dim_at(B, 1, L - 1) + 2) *
                                        (dim_at(B, 2, L - 1) + 2));
            need = (need + 7) / 8 * 8 + 2 * sizeof(double) * ghc;
        }
        if (need <= kTailLdsMax) {
            mg->lt = lt;
            mg->tail_lds = need;
            break;
        }
    }
    if (mg->lt == L) return 0;
    // allow the tail's dynamic LDS beyond the default 64 KiB
    const hipFuncAttribute attr = hipFuncAttributeMaxDynamicSharedMemorySize;
    const int bytes = (int)mg->tail_lds;
    if (hipFuncSetAttribute((const void *)k_tail<double, false>, attr, bytes) != hipSuccess ||
        hipFuncSetAttribute((const void *)k_tail<double, true>, attr, bytes) != hipSuccess ||
        hipFuncSetAttribute((const void *)k_tail<float, false>, attr, bytes) != hipSuccess ||
        hipFuncSetAttribute((const void *)k_tail<float, true>, attr, bytes) != hipSuccess) {
        mg->err = "hipFuncSetAttribute (tail LDS) failed";
        return MF3D_EDEVICE;
    }
    std::vector<BlockLevel> tl;
    for (int l = mg->lt; l < L; ++l)
        for (int b = 0; b < nb; ++b) tl.push_back(mg->lv[l].hb[b]);
    return upload(mg, &mg->d_tail, tl);
}

template <typename T, bool VAR>
int autotune_sweeps(mf3d_mg *mg);
void shell_geometry(mf3d_mg *mg);

// Whether the coarsest level is solved directly (MF3D_COARSEST_*), and
// its cell count.  Called once the level tables exist, before the tail.
int decide_coarse(mf3d_mg *mg) {
    const mf3d_mg_cfg &c = mg->cfg;
    const int L = c.levels;
    mg->cdirect = false;
    mg->nc = 0;
    if (L < 2) return 0;  // one level: the reference's smoothing is the whole cycle
    // the reference's sweeps in FP64 on the coarsest level of the mixed
    // form (red-black smoother; the in-place smoothers keep the inner
    // precision there)
    mg->c64 = c.precision == MF3D_MIXED && c.coarse_solve == MF3D_COARSEST_DEFAULT && c.smoother == 3;
    if (c.coarse_solve != MF3D_COARSEST_DIRECT) return 0;
    long long nc = 0;
    for (const BlockLevel &g : mg->lv[L - 1].hb) nc += (long long)g.n0 * g.n1 * g.n2;
    if (nc > MF3D_COARSEST_MAX_CELLS) {
        mg->err = "the coarsest level has " + std::to_string(nc) + " cells, more than the " +
                  std::to_string(MF3D_COARSEST_MAX_CELLS) + " of its direct solve: use more levels";
        return MF3D_EINVAL;
    }
    mg->cdirect = true;
    mg->nc = (int)nc;
    return 0;
}

// The coarsest level's cell table: every cell's block and index, its six
// neighbours as cells (a ghost neighbour resolved through the level's
// region table to the cell the exchange copies into it), and one pinned
// cell per group of blocks connected through neighbours or wraps (the
// operator's constants); then the matrix buffers.
int coarse_setup_refined(mf3d_mg *mg);

int coarse_setup(mf3d_mg *mg) {
    if (mg->refined) return coarse_setup_refined(mg);
    const int L = mg->cfg.levels, nb = mg->cfg.nblocks;
    const Level &C = mg->lv[L - 1];
    int mx = 0;
    const std::vector<Region> regs = build_regions(mg->blk, L - 1, &mx);
    std::vector<int> base(nb + 1, 0);
    for (int b = 0; b < nb; ++b) base[b + 1] = base[b] + C.hb[b].n0 * C.hb[b].n1 * C.hb[b].n2;
    auto idx = [&](int b, int i, int j, int k) {
        const BlockLevel &g = C.hb[b];
        return base[b] + (i - 1) + g.n0 * ((j - 1) + g.n1 * (k - 1));
    };
    auto resolve = [&](int b, int i, int j, int k) -> int {
        const BlockLevel &g = C.hb[b];
        if (i >= 1 && i <= g.n0 && j >= 1 && j <= g.n1 && k >= 1 && k <= g.n2) return idx(b, i, j, k);
        const int v[3] = {i, j, k};
        for (const Region &r : regs) {
            if (r.dst_b != b) continue;
            bool in = true;
            for (int a = 0; a < 3; ++a) in = in && v[a] >= r.lo[a] && v[a] < r.lo[a] + r.cnt[a];
            if (!in) continue;
            int s[3];
            for (int a = 0; a < 3; ++a) s[a] = r.map[a] < 0 ? v[a] : r.map[a];
            return idx(r.src_b, s[0], s[1], s[2]);
        }
        return -1;
    };
    const int nc = mg->nc;
    std::vector<CCell> cells(nc);
    std::vector<int> parent(nc);
    for (int q = 0; q < nc; ++q) parent[q] = q;
    std::function<int(int)> root = [&](int q) { return parent[q] == q ? q : parent[q] = root(parent[q]); };
    for (int b = 0; b < nb; ++b) {
        const BlockLevel &g = C.hb[b];
        for (int k = 1; k <= g.n2; ++k)
            for (int j = 1; j <= g.n1; ++j)
                for (int i = 1; i <= g.n0; ++i) {
                    CCell &e = cells[idx(b, i, j, k)];
                    e.b = b; e.i = i; e.j = j; e.k = k; e.pin = 0;
                    const int nbq[6] = {resolve(b, i - 1, j, k), resolve(b, i + 1, j, k), resolve(b, i, j - 1, k),
                                        resolve(b, i, j + 1, k), resolve(b, i, j, k - 1), resolve(b, i, j, k + 1)};
                    for (int d = 0; d < 6; ++d) {
                        if (nbq[d] < 0) {
                            mg->err = "coarsest level: a ghost with no region (internal)";
                            return MF3D_EDEVICE;
                        }
                        e.nb[d] = nbq[d];
                        const int ra = root(idx(b, i, j, k)), rb = root(nbq[d]);
                        if (ra != rb) parent[ra] = rb;
                    }
                }
    }
    std::vector<int> last(nc, -1);  // the last cell of each connected group is pinned
    for (int q = 0; q < nc; ++q) last[root(q)] = q;
    for (int q = 0; q < nc; ++q)
        if (last[q] >= 0) cells[last[q]].pin = 1;
    int rc;
    if ((rc = upload(mg, &mg->d_ccell, cells))) return rc;
    if ((rc = dev_alloc(mg, (void **)&mg->d_cinv, sizeof(double) * (size_t)nc * nc)) ||
        (rc = dev_alloc(mg, (void **)&mg->d_ccol, sizeof(double) * (size_t)nc)))
        return rc;
    return 0;
}

// Refined block grids (constant coefficients): the coarsest operator
// assembled on the host.  A ghost neighbour is the expression of interior
// cells the level's exchange writes into it (its ghost table): a copy, the
// 8-cell average (0.125 each) or the 2:1 extrapolation ((2 s0 + 3 s1) -
// s2) / 4 (weights 0.5, 0.75, -0.25), so its coefficient spreads over those
// cells with those weights.  Pinned cells as in coarse_setup: the last cell
// of each group of cells connected through the operator.
int coarse_setup_refined(mf3d_mg *mg) {
    const int L = mg->cfg.levels, nb = mg->cfg.nblocks;
    const Level &C = mg->lv[L - 1];
    std::vector<GhostExpr> gx;
    const std::string why = ghost_exprs(mg->blk, mg->own, L, L - 1, &gx);
    if (!why.empty()) {
        mg->err = why;
        return MF3D_ENOTSUP;
    }
    const long long arena = C.cells_gh;
    std::vector<int> expr_of((size_t)arena, -1), cell_of((size_t)arena, -1);
    for (size_t e = 0; e < gx.size(); ++e) expr_of[(size_t)gx[e].dst] = (int)e;
    const int nc = mg->nc;
    std::vector<CCell> cells(nc);
    {
        int q = 0;
        for (int b = 0; b < nb; ++b) {
            const BlockLevel &g = C.hb[b];
            for (int k = 1; k <= g.n2; ++k)
                for (int j = 1; j <= g.n1; ++j)
                    for (int i = 1; i <= g.n0; ++i, ++q) {
                        cell_of[(size_t)(g.off + i + g.sy * j + g.sz * k)] = q;
                        CCell &e = cells[q];
                        e.b = b; e.i = i; e.j = j; e.k = k; e.pin = 0;
                        for (int d = 0; d < 6; ++d) e.nb[d] = q;
                    }
        }
        if (q != nc) {
            mg->err = "coarsest level: cell count (internal)";
            return MF3D_EDEVICE;
        }
    }
    struct Entry { int r, c; double w; };
    std::vector<Entry> ent;
    std::vector<int> parent(nc);
    for (int q = 0; q < nc; ++q) parent[q] = q;
    std::function<int(int)> root = [&](int q) { return parent[q] == q ? q : parent[q] = root(parent[q]); };
    for (int r = 0; r < nc; ++r) {
        const CCell &e = cells[r];
        const BlockLevel &g = C.hb[e.b];
        // coefs' constant form in FP64
        const double a[6] = {(e.i == 1 && (g.neu & kNeuW)) ? 0.0 : g.c0, (e.i == g.n0 && (g.neu & kNeuE)) ? 0.0 : g.c0,
                             (e.j == 1 && (g.neu & kNeuS)) ? 0.0 : g.c1, (e.j == g.n1 && (g.neu & kNeuN)) ? 0.0 : g.c1,
                             (e.k == 1 && (g.neu & kNeuB)) ? 0.0 : g.c2, (e.k == g.n2 && (g.neu & kNeuT)) ? 0.0 : g.c2};
        double s = a[0] + a[1];
        s = s + a[2];
        s = s + a[3];
        s = s + a[4];
        s = s + a[5];
        ent.push_back({r, r, -s});
        const long long q = g.off + e.i + g.sy * e.j + g.sz * e.k;
        const long long nq[6] = {q - 1, q + 1, q - g.sy, q + g.sy, q - g.sz, q + g.sz};
        for (int d = 0; d < 6; ++d) {
            if (a[d] == 0.0) continue;
            auto add = [&](long long o, double w) -> bool {
                const int c = cell_of[(size_t)o];
                if (c < 0) return false;
                ent.push_back({r, c, w});
                const int ra = root(r), rb = root(c);
                if (ra != rb) parent[ra] = rb;
                return true;
            };
            bool ok;
            if (cell_of[(size_t)nq[d]] >= 0) {
                ok = add(nq[d], a[d]);
            } else {
                const int x = expr_of[(size_t)nq[d]];
                if (x < 0) {
                    ok = false;
                } else {
                    const GhostExpr &ge = gx[(size_t)x];
                    if (ge.kind == 0) {
                        ok = add(ge.src[0], a[d]);
                    } else if (ge.kind == 1) {
                        ok = true;
                        for (int m = 0; m < 8; ++m) ok = add(ge.src[m], 0.125 * a[d]) && ok;
                    } else {
                        ok = add(ge.src[0], 0.5 * a[d]) && add(ge.src[1], 0.75 * a[d]) && add(ge.src[2], -0.25 * a[d]);
                    }
                }
            }
            if (!ok) {
                mg->err = "coarsest level: a ghost with no interior expression (internal)";
                return MF3D_EDEVICE;
            }
        }
    }
    std::vector<int> last(nc, -1);
    for (int q = 0; q < nc; ++q) last[root(q)] = q;
    for (int q = 0; q < nc; ++q)
        if (last[q] >= 0) cells[last[q]].pin = 1;
    mg->cmat.assign((size_t)nc * nc, 0.0);
    for (int q = 0; q < nc; ++q)
        if (cells[q].pin) mg->cmat[(size_t)q * nc + q] = 1.0;
    for (const Entry &x : ent)
        if (!cells[x.r].pin && !cells[x.c].pin) mg->cmat[(size_t)x.r * nc + x.c] += x.w;
    int rc;
    if ((rc = upload(mg, &mg->d_ccell, cells))) return rc;
    if ((rc = dev_alloc(mg, (void **)&mg->d_cinv, sizeof(double) * (size_t)nc * nc)) ||
        (rc = dev_alloc(mg, (void **)&mg->d_ccol, sizeof(double) * (size_t)nc)))
        return rc;
    return 0;
}

// Assemble the coarsest operator from the current coefficients (FP64: the
// constant 1/h^2 or the FP64 face coefficients; on refined grids the
// host-assembled matrix) and invert it in place.
template <bool VAR>
int coarse_factor(mf3d_mg *mg) {
    if (!mg->cdirect) return 0;
    const Level &C = mg->lv[mg->cfg.levels - 1];
    const int nc = mg->nc;
    if (mg->refined) {
        if (mg->cmat.size() != (size_t)nc * nc) {
            mg->err = "refined grid: the coarsest operator is factored once, at creation";
            return MF3D_EINVAL;
        }
        HIPCHK(hipMemcpyAsync(mg->d_cinv, mg->cmat.data(), sizeof(double) * (size_t)nc * nc, hipMemcpyHostToDevice,
                              mg->stream));
    } else {
        fill<double>(mg, mg->d_cinv, (long long)nc * nc);
        hipLaunchKernelGGL(k_coarse_assemble<VAR>, dim3((nc + 255) / 256), dim3(256), 0, mg->stream, mg->d_cinv,
                           nc, (const CCell *)mg->d_ccell, C.db, (const double *)C.dcx, (const double *)C.dcy,
                           (const double *)C.dcz);
    }
    if (nc <= 512) {
        hipLaunchKernelGGL(k_gj_small, dim3(1), dim3(1024), 0, mg->stream, mg->d_cinv, nc, mg->d_ccol);
    } else {
        const unsigned g1 = (unsigned)((nc + 255) / 256);
        for (int k = 0; k < nc; ++k) {
            hipLaunchKernelGGL(k_gj_col, dim3(g1), dim3(256), 0, mg->stream, (const double *)mg->d_cinv, nc, k,
                               mg->d_ccol);
            hipLaunchKernelGGL(k_gj_row, dim3(g1), dim3(256), 0, mg->stream, mg->d_cinv, nc, k,
                               (const double *)mg->d_ccol);
            hipLaunchKernelGGL(k_gj_update, dim3(g1, nc), dim3(256), 0, mg->stream, mg->d_cinv, nc, k,
                               (const double *)mg->d_ccol);
        }
    }
    hipLaunchKernelGGL(k_coarse_unpin, dim3((nc + 255) / 256, nc), dim3(256), 0, mg->stream, mg->d_cinv, nc,
                       (const CCell *)mg->d_ccell);
    HIPCHK(hipStreamSynchronize(mg->stream));
    HIPCHK(hipGetLastError());
    if (mg->refined) std::vector<double>().swap(mg->cmat);  // the copy has completed
    return 0;
}

// ADI smoother set-up: per level, the whole-sweep LDS kernel where the
// level's largest block (its ghosted potential and the per-diagonal
// scratch, with the RHS if that fits too) fits in kAdiLdsMax bytes.
constexpr size_t kAdiLdsMax = 150 * 1024;

int adi_setup(mf3d_mg *mg) {
    const size_t w = inner_bytes(mg->cfg);
    const char *env = std::getenv("MF3D_ADI_LDS");
    const bool allow = !(env && env[0] == '0');
    size_t most = 0;
    for (Level &V : mg->lv) {
        size_t cells = 0, diag = 1;  // the largest ghosted box; the most cells one diagonal holds
        for (const BlockLevel &g : V.hb) {
            cells = std::max(cells, (size_t)(g.sz * (g.n2 + 2)));
            const size_t n0 = g.n0, n1 = g.n1, n2 = g.n2;
            diag = std::max({diag, std::min(n0, n2) * n1, std::min(n1, n2) * n0, std::min(n0, n1) * n2});
        }
        V.adi_cells = (int)cells;
        V.adi_sstride = (int)diag;
        V.adi_lds = 0;
        V.adi_rhs = false;
        const bool fits = allow && diag <= 1024;  // the next diagonal's prefetch: 4 entries per thread
        if (fits && (2 * cells + 5 * diag) * w <= kAdiLdsMax) {
            V.adi_lds = (2 * cells + 5 * diag) * w;
            V.adi_rhs = true;
        } else if (fits && (cells + 5 * diag) * w <= kAdiLdsMax) {
            V.adi_lds = (cells + 5 * diag) * w;
        }
        most = std::max(most, V.adi_lds);
    }
    if (most > 64 * 1024) {
        const hipFuncAttribute attr = hipFuncAttributeMaxDynamicSharedMemorySize;
        const int bytes = (int)most;
        const void *fns[8] = {(const void *)k_adi_lds<double, false, false>, (const void *)k_adi_lds<double, false, true>,
                              (const void *)k_adi_lds<double, true, false>,  (const void *)k_adi_lds<double, true, true>,
                              (const void *)k_adi_lds<float, false, false>,  (const void *)k_adi_lds<float, false, true>,
                              (const void *)k_adi_lds<float, true, false>,   (const void *)k_adi_lds<float, true, true>};
        for (const void *f : fns)
            if (hipFuncSetAttribute(f, attr, bytes) != hipSuccess) {
                mg->err = "hipFuncSetAttribute (ADI LDS) failed";
                return MF3D_EDEVICE;
            }
    }
    return 0;
}

// The ADI Thomas coefficients of the levels whose sweeps run in LDS, from
// the (inner-precision) operator; after every coefficient build.
template <typename T, bool VAR>
int adi_coef_build(mf3d_mg *mg) {
    for (Level &V : mg->lv) {
        if (!V.adi_lds) continue;
        int most = 1;
        for (const BlockLevel &g : V.hb) most = std::max({most, g.n0 * g.n2, g.n1 * g.n2, g.n0 * g.n1});
        const dim3 grid((unsigned)((most + 255) / 256), (unsigned)mg->cfg.nblocks);
        const T *cx = (const T *)V.cx, *cy = (const T *)V.cy, *cz = (const T *)V.cz;
        hipLaunchKernelGGL((k_adi_coef<T, VAR, 0>), grid, dim3(256), 0, mg->stream, (T *)V.adi_apr[0], (T *)V.adi_a[0],
                           cx, cy, cz, V.db);
        hipLaunchKernelGGL((k_adi_coef<T, VAR, 1>), grid, dim3(256), 0, mg->stream, (T *)V.adi_apr[1], (T *)V.adi_a[1],
                           cx, cy, cz, V.db);
        hipLaunchKernelGGL((k_adi_coef<T, VAR, 2>), grid, dim3(256), 0, mg->stream, (T *)V.adi_apr[2], (T *)V.adi_a[2],
                           cx, cy, cz, V.db);
    }
    HIPCHK(hipStreamSynchronize(mg->stream));
    HIPCHK(hipGetLastError());
    return 0;
}

int build(mf3d_mg *mg) {
    const mf3d_mg_cfg &c = mg->cfg;
    {
        const char *e = std::getenv("MF3D_FUSE_PROLONG");
        if (e && std::strcmp(e, "0") != 0 && std::strcmp(e, "1") != 0) {
            mg->err = "MF3D_FUSE_PROLONG must be 0 or 1";
            return MF3D_EINVAL;
        }
        mg->fuse_pro = !(e && e[0] == '0');
    }
    {
        const char *e = std::getenv("MF3D_ZERO_SWEEP");
        mg->zmode = e ? std::atoi(e) : 1;
        if (mg->zmode < 0 || mg->zmode > 1) {
            mg->err = "MF3D_ZERO_SWEEP must be 0 or 1";
            return MF3D_EINVAL;
        }
    }
    const int L = c.levels, nb = c.nblocks;
    const bool mixed = c.precision == MF3D_MIXED;
    const size_t w = mixed ? 4 : 8;
    mg->own = own_levels(nb, mg->blk.data(), L);
    mg->refined = any_refined(nb, mg->blk.data());
    // level l of block b repeats level l-1's grid (the block has fewer levels)
    auto repeats = [&](int b, int l) { return l >= mg->own[b]; };
    mg->lv.resize(L);
    mg->cells = 0;
    for (int b = 0; b < nb; ++b) mg->cells += (long long)mg->blk[b].n[0] * mg->blk[b].n[1] * mg->blk[b].n[2];
    int rc;
    for (int l = 0; l < L; ++l) {
        Level &V = mg->lv[l];
        long long off = 0, poff = 0;
        V.hb.resize(nb);
        int n0max = 0, n1max = 0;
        for (int b = 0; b < nb; ++b) {
            const mf3d_block_desc &B = mg->blk[b];
            BlockLevel &g = V.hb[b];
            g.n0 = dim_at(B, 0, l, mg->own[b]); g.n1 = dim_at(B, 1, l, mg->own[b]); g.n2 = dim_at(B, 2, l, mg->own[b]);
            n0max = std::max(n0max, g.n0);
            n1max = std::max(n1max, g.n1);
            g.neu = 0;
            for (int f = 0; f < 6; ++f)
                if (is_neumann(B, f)) g.neu |= (1 << f);
            g.sy = g.n0 + 2;
            g.sz = g.sy * (g.n1 + 2);
            g.off = off;
            g.poff = poff;
            off += g.sz * (g.n2 + 2);
            poff += (long long)g.n0 * g.n1 * g.n2;
            // constant coefficients (mgsolver.for:51-57,184-189), the block's
            // own grid level gl = min(glevel, ngrid)
            const double sc = (double)(1 << std::min(l, mg->own[b] - 1));
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
        if (off * (long long)inner_bytes(c) >= (1LL << 32)) {
            mg->err = "a level holds more than 4 GiB of one field (the sweep addresses arenas with 32-bit "
                      "byte offsets): use fewer cells per solver";
            return MF3D_ENOTSUP;
        }
        // the finest level marches 8-row tiles; coarser levels have few
        // columns and use 4-row tiles, which keeps their z chunks deeper
        V.sw_ty = (l == 0 && n1max >= 8) ? 8 : 4;
        if ((rc = upload(mg, &V.db, V.hb))) return rc;
        auto tsw = xcd_order(sweep_tiles(V.hb, V.sw_ty)), tin = xcd_order(cell_tiles(V.hb)),
             trow = xcd_order(row_tiles(V.hb));
        auto pick = [](const std::vector<Tile> &t, auto keep) {
            std::vector<Tile> o;
            for (const Tile &x : t)
                if (keep(x.b)) o.push_back(x);
            return o;
        };
        // restriction into l: coarsening blocks / blocks repeating their grid
        auto trst = pick(tin, [&](int b) { return !repeats(b, l); });
        auto trrep = pick(tin, [&](int b) { return repeats(b, l); });
        // prolongation from l+1: blocks that repeat at l+1 take injection
        auto toct = xcd_order(pick(octet_tiles(V.hb), [&](int b) { return !(l + 1 < L && repeats(b, l + 1)); }));
        auto tprep = pick(trow, [&](int b) { return l + 1 < L && repeats(b, l + 1); });
        V.n_rst = (int)trst.size(); V.n_rrep = (int)trrep.size(); V.n_prep = (int)tprep.size();
        if ((rc = upload(mg, &V.t_rst, trst)) || (rc = upload(mg, &V.t_rrep, trrep)) ||
            (rc = upload(mg, &V.t_prep, tprep)))
            return rc;
        auto tch = xcd_order(chunk_tiles(V.hb));
        V.n_ch = (int)tch.size();
        if ((rc = upload(mg, &V.t_ch, tch))) return rc;
        V.n_sw = (int)tsw.size(); V.n_in = (int)tin.size(); V.n_row = (int)trow.size();
        V.n_oct = (int)toct.size();
        if ((rc = upload(mg, &V.t_oct, toct))) return rc;
        V.swz_ty = V.sw_ty;
        V.n_swz = (int)tsw.size();
        if ((rc = upload(mg, &V.t_swz, tsw))) return rc;
        if ((rc = upload(mg, &V.t_sw, tsw)) || (rc = upload(mg, &V.t_in, tin)) || (rc = upload(mg, &V.t_row, trow)))
            return rc;
        std::vector<Region> rf;
        if (mg->refined) {  // every ghost as an expression over interior cells
            std::vector<GhostExpr> gx;
            const std::string why = ghost_exprs(mg->blk, mg->own, L, l, &gx);
            if (!why.empty()) {
                mg->err = "level " + std::to_string(l) + ": " + why;
                return MF3D_ENOTSUP;
            }
            V.n_gx = (int)gx.size();
            if ((rc = upload(mg, &V.gx, gx))) return rc;
        } else {
            rf = build_regions(mg->blk, l, &V.max_full);
        }
        V.n_full = (int)rf.size();
        // lone blocks (no neighbour, no periodic partner): each sweep stores the
        // ghost images itself and no gather runs.  MF3D_OWN_GHOSTS=0: gather.
        {
            bool lone = true;
            for (const mf3d_block_desc &B : mg->blk) {
                for (int f = 0; f < 6; ++f) lone = lone && B.nbr[f] < 0 && B.per[f] < 0;
                for (int d = 0; d < 3; ++d)
                    lone = lone && dim_at(B, d, l) >= 2 && is_neumann(B, 2 * d) == is_neumann(B, 2 * d + 1);
            }
            const char *env = std::getenv("MF3D_OWN_GHOSTS");
            V.own_ghosts = lone && c.smoother == 3 && !(env && env[0] == '0');
        }
        if ((rc = upload(mg, &V.r_full, rf))) return rc;
        if ((rc = dev_alloc(mg, &V.phi[0], V.cells_gh * w)) || (rc = dev_alloc(mg, &V.phi[1], V.cells_gh * w)) ||
            (rc = dev_alloc(mg, &V.rhs, V.cells_gh * w)))
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
    {
        auto tout = xcd_order(sweep_tiles(L0.hb, 8));
        mg->n_out = (int)tout.size();
        if ((rc = upload(mg, &mg->t_out, tout))) return rc;
    }
    if ((rc = decide_coarse(mg)) || (rc = choose_tail(mg))) return rc;
    if (mg->cdirect && (rc = coarse_setup(mg))) return rc;
    if (mg->c64) {
        const long long ghc = mg->lv[L - 1].cells_gh;
        if ((rc = dev_alloc(mg, (void **)&mg->c64_phi[0], 8 * ghc)) ||
            (rc = dev_alloc(mg, (void **)&mg->c64_phi[1], 8 * ghc)) ||
            (rc = dev_alloc(mg, (void **)&mg->c64_rhs, 8 * ghc)))
            return rc;
    }
    if ((rc = dev_alloc(mg, (void **)&mg->pbuf[0], L0.cells_gh * 8)) ||
        (rc = dev_alloc(mg, (void **)&mg->pbuf[1], L0.cells_gh * 8)) ||
        (rc = dev_alloc(mg, (void **)&mg->f_g, L0.cells_gh * 8)))
        return rc;
    if (c.krylov == MF3D_OUTER_GCR &&
        ((rc = dev_alloc(mg, (void **)&mg->kd[0], L0.cells_gh * 8)) ||
         (rc = dev_alloc(mg, (void **)&mg->kd[1], L0.cells_gh * 8)) ||
         (rc = dev_alloc(mg, (void **)&mg->kw, L0.cells_gh * 8)) ||
         (rc = dev_alloc(mg, (void **)&mg->kr, L0.cells_gh * 8))))
        return rc;
    mg->nparts = mg->n_out;
    if ((rc = dev_alloc(mg, (void **)&mg->partials, sizeof(double) * 2 * (size_t)mg->nparts)) ||
        (rc = dev_alloc(mg, (void **)&mg->st, sizeof(SolveState))) ||
        (rc = dev_alloc(mg, (void **)&mg->d_hist, sizeof(double) * (c.maxcy + 1))) ||
        (rc = dev_alloc(mg, (void **)&mg->d_rmax, sizeof(double) * (c.maxcy + 1))))
        return rc;
    shell_geometry(mg);
    if (!mg->arena) {  // host-array entries' staging (the allocation-free solve has none)
        mg->io_len = std::max({L0.cells_gh, mg->cells, mg->cells_w});
        if ((rc = dev_alloc(mg, (void **)&mg->d_io, sizeof(double) * mg->io_len))) return rc;
        HIPCHK(hipHostMalloc((void **)&mg->h_io, sizeof(double) * mg->io_len, hipHostMallocDefault));
    }
    // every field starts at zero (coeff zeroes phi, mgsolver.for:60-67)
    for (int l = 0; l < L; ++l) {
        Level &V = mg->lv[l];
        if (mixed) {
            fill<float>(mg, (float *)V.phi[0], V.cells_gh); fill<float>(mg, (float *)V.phi[1], V.cells_gh);
            fill<float>(mg, (float *)V.rhs, V.cells_gh);
        } else {
            fill<double>(mg, (double *)V.phi[0], V.cells_gh); fill<double>(mg, (double *)V.phi[1], V.cells_gh);
            fill<double>(mg, (double *)V.rhs, V.cells_gh);
        }
        if (c.variable_coef) {
            fill<double>(mg, V.dcx, V.cells_gh); fill<double>(mg, V.dcy, V.cells_gh);
            fill<double>(mg, V.dcz, V.cells_gh); fill<double>(mg, V.dens, V.cells_gh);
        }
    }
    fill<double>(mg, mg->pbuf[1], L0.cells_gh);
    fill<double>(mg, mg->f_g, L0.cells_gh);
    hipLaunchKernelGGL(k_reset_state, dim3(1), dim3(1), 0, mg->stream, mg->st, mg->cfg.eps);
    HIPCHK(hipStreamSynchronize(mg->stream));
    HIPCHK(hipGetLastError());
    // constant coefficients are final now; variable ones at each coeff
    if (!c.variable_coef && (rc = coarse_factor<false>(mg))) return rc;
    if (c.smoother == 2) {
        if ((rc = adi_setup(mg))) return rc;
        bool global = false;  // a level runs the per-pass kernels: their a(m) scratch
        for (Level &V : mg->lv) {
            global = global || !V.adi_lds;
            if (V.adi_lds)
                for (int a = 0; a < 3; ++a)
                    if ((rc = dev_alloc(mg, &V.adi_apr[a], V.cells_gh * w)) ||
                        (rc = dev_alloc(mg, &V.adi_a[a], V.cells_gh * w)))
                        return rc;
        }
        if (global && (rc = dev_alloc(mg, &mg->d_adi, L0.cells_gh * w))) return rc;
        // constant coefficients are final now; variable ones at each coeff
        if (!c.variable_coef && (rc = mixed ? adi_coef_build<float, false>(mg) : adi_coef_build<double, false>(mg)))
            return rc;
    }
    if (c.smoother != 3) return 0;  // the tiles tuned below belong to the red-black sweep
    if (mg->arena) return 0;        // the tuner allocates: default tilings
    if ((rc = mixed ? (c.variable_coef ? autotune_sweeps<float, true>(mg) : autotune_sweeps<float, false>(mg))
                    : (c.variable_coef ? autotune_sweeps<double, true>(mg) : autotune_sweeps<double, false>(mg))))
        return rc;
    return 0;
}

// Tile autotuner.  On the finest level the march is timed over chunk
// depths tz from 2 to 64: each candidate is launched once to warm up and
// then timed over 3 launches (12 on levels under 2^20 cells) with HIP
// events on the handle's stream, on the level's own (zeroed) arrays, which
// are zeroed again afterwards; of the depths within 3 % of the fastest the
// deepest is kept (fewer halo planes read).  Coarser levels keep the
// default depth (the deepest chunk that still gives about 1024
// workgroups): inside a cycle it beat the depth such timing picked (128^3:
// 13 us at 8-plane chunks against 17 at 16), since the timing repeats one
// launch on cache-resident data.  Levels of at most 2^18 cells take the
// one-pass 3-D tile kernel (k_rb3d) instead.  The tiling never changes a
// result.  MF3D_AUTOTUNE=0 keeps the default depths everywhere;
// MF3D_SWEEP_TZ=<even depth> uses that march depth on every level without
// timing (to pin the tiling of a profile); MF3D_SWEEP_3D=1 uses the 3-D
// tiles on every level, =0 never.
// The finest level's depth is remembered per process, keyed by precision,
// coefficient form, tile rows and the level's block extents, so later
// handles of the same shape (the serving pool's) skip the timing.
std::mutex g_tune_mu;
std::map<std::string, int> g_tune_tz;

template <typename T, bool VAR>
int autotune_sweeps(mf3d_mg *mg) {
    const char *env = std::getenv("MF3D_AUTOTUNE");
    if (env && env[0] == '0') return 0;
    int pin_tz = 0;
    if (const char *p = std::getenv("MF3D_SWEEP_TZ")) {
        pin_tz = std::atoi(p);
        if (pin_tz < 2 || (pin_tz & 1)) {
            mg->err = "MF3D_SWEEP_TZ must be an even depth of at least 2";
            return MF3D_EINVAL;
        }
    }
    int force3d = -1;  // MF3D_SWEEP_3D: 1 every level, 0 never, unset timed
    if (const char *p = std::getenv("MF3D_SWEEP_3D")) force3d = p[0] == '1' ? 1 : 0;
    auto key_of = [&](const Level &V) {
        std::string k = std::to_string(sizeof(T)) + (VAR ? "v" : "c") + std::to_string(V.sw_ty);
        for (const BlockLevel &g : V.hb)
            k += ":" + std::to_string(g.n0) + "," + std::to_string(g.n1) + "," + std::to_string(g.n2);
        return k;
    };
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    // tiles of a choice: a march depth, or kTz3d for the 3-D tiles
    constexpr int kTz3d = -1;
    // Levels of at most this many cells take the 3-D tiles, larger ones the
    // march, by measurement inside the cycle: there the 3-D tiles sweep a
    // 64^3 or 32^3 level in about 6 us against 8-9 us for the best march,
    // and lose at 128^3 (15 us against 13) and 256^3.
    constexpr long long k3dCells = 1LL << 18;
    auto tiles_of = [&](const Level &V, int ty, int tz) {
        return xcd_order(tz == kTz3d ? tiles_3d(V.hb) : sweep_tiles(V.hb, ty, tz));
    };
    // 1 warm-up + 3 timed sweeps over a candidate tiling (12 on levels under
    // 2^20 cells, whose few-microsecond sweeps time noisily) from the tile
    // buffer tmp (allocated once per tuning pass, large enough for every
    // candidate); ms, 1e30 on failure
    auto time_tiles = [&](Level &V, const std::vector<Tile> &t, int ty, Tile *tmp) -> float {
        const int reps = V.cells_gh < (1LL << 20) ? 12 : 3;
        Tile *keep = V.t_sw;
        const int keep_n = V.n_sw, keep_ty = V.sw_ty;
        V.t_sw = tmp;
        V.n_sw = (int)t.size();
        V.sw_ty = ty;
        float ms = 1e30f;
        if (hipMemcpyAsync(tmp, t.data(), sizeof(Tile) * t.size(), hipMemcpyHostToDevice, mg->stream) ==
            hipSuccess) {
            sweep_launch<T, VAR, false>(mg, V, nullptr);
            (void)hipEventRecord(e0, mg->stream);
            for (int r = 0; r < reps; ++r) sweep_launch<T, VAR, false>(mg, V, nullptr);
            (void)hipEventRecord(e1, mg->stream);
            if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                ms = 1e30f;
        }
        V.t_sw = keep;
        V.n_sw = keep_n;
        V.sw_ty = keep_ty;
        return ms;
    };
    int rc = 0;
    for (int l = 0; l < std::min(mg->lt, (int)mg->lv.size()) && !rc; ++l) {
        Level &V = mg->lv[l];
        const int ty = V.sw_ty;  // the march's tile rows on this level
        const std::string key = key_of(V);
        int choice = 0;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_tune_tz.find(key);
            if (it != g_tune_tz.end()) choice = it->second;
        }
        long long cells = 0;
        for (const BlockLevel &g : V.hb) cells += (long long)g.n0 * g.n1 * g.n2;
        if (force3d == 1) choice = kTz3d;
        else if (pin_tz) choice = pin_tz;
        else if (force3d != 0 && cells <= k3dCells) choice = kTz3d;
        else if (l > 0) continue;  // coarser march levels keep the default depth
        if (!choice) {
            int nzmax = 1;
            for (const BlockLevel &g : V.hb) nzmax = std::max(nzmax, g.n2);
            // powers of two, and depths whose step count (tz + 2, or tz + 4
            // at lag 2) is a whole number of the march's 10- (lag 1) or
            // 6-step (lag 2) unrolled blocks
            static const int kTz[] = {2, 4, 8, 14, 16, 18, 20, 26, 28, 32, 38, 44, 48, 50, 56, 58, 62, 64};
            std::vector<std::pair<int, std::vector<Tile>>> cand;
            size_t most = 1;
            for (int tz : kTz) {
                if (tz > std::max(2, nzmax)) break;
                cand.emplace_back(tz, tiles_of(V, ty, tz));
                most = std::max(most, cand.back().second.size());
            }
            Tile *tmp = nullptr;
            if (hipMalloc((void **)&tmp, sizeof(Tile) * most) != hipSuccess) {
                (void)hipGetLastError();
                cand.clear();  // untimed: the default depth stays
            }
            std::vector<std::pair<int, float>> timed;  // (depth, ms)
            for (const auto &c : cand) timed.emplace_back(c.first, time_tiles(V, c.second, ty, tmp));
            if (tmp) {
                (void)hipStreamSynchronize(mg->stream);
                (void)hipFree(tmp);
            }
            float best = 1e30f;
            for (const auto &e : timed) best = std::min(best, e.second);
            for (const auto &e : timed)
                if (e.second <= 1.03f * best && best < 1e29f) choice = std::max(choice, e.first);
            if (choice) {
                std::lock_guard<std::mutex> lk(g_tune_mu);
                g_tune_tz[key] = choice;
            }
            fill<T>(mg, (T *)V.phi[0], V.cells_gh);
            fill<T>(mg, (T *)V.phi[1], V.cells_gh);
        }
        if (!choice) continue;
        const auto t = tiles_of(V, ty, choice);
        Tile *dst = nullptr;
        if ((rc = upload(mg, &dst, t))) break;
        V.t_sw = dst;
        V.n_sw = (int)t.size();
        V.sw_ty = choice == kTz3d ? 0 : ty;
        if (VAR && choice != kTz3d) {  // the zero-start march: the same tiles (384^3: 717 against 754 us
            V.t_swz = dst;             // on its own default-depth chunks)
            V.n_swz = V.n_sw;
            V.swz_ty = ty;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(mg->stream));
    HIPCHK(hipGetLastError());
    return 0;
}

// coeff, L_LSM branch, from level-0 densities already in lv[0].dens.
int build_lsm(mf3d_mg *mg) {
    const int L = mg->cfg.levels;
    const bool face = mg->cfg.coarse_op == MF3D_COARSE_FACE;
    for (int l = 1; l < L && !face; ++l) {
        Level &C = mg->lv[l], &F = mg->lv[l - 1];
        hipLaunchKernelGGL(k_dens_restrict, dim3(C.n_in), dim3(kTileX, kTileY), 0, mg->stream, F.dens, F.db,
                           C.dens, C.db, C.t_in);
    }
    for (int l = 0; l < L; ++l) {
        Level &V = mg->lv[l];
        if (face && l > 0) {
            const Level &F = mg->lv[l - 1];
            hipLaunchKernelGGL(k_face_restrict, dim3(V.n_row), dim3(kTileX, kTileY), 0, mg->stream, F.dcx, F.dcy,
                               F.dcz, F.db, V.dcx, V.dcy, V.dcz, V.db, V.t_row);
        } else if (mg->cfg.precision == MF3D_MIXED) {
            hipLaunchKernelGGL(k_lsm_coef<true>, dim3(V.n_ch), dim3(256), 0, mg->stream, V.dens, V.dcx, V.dcy,
                               V.dcz, (float *)V.cx, (float *)V.cy, (float *)V.cz, V.db, V.t_ch);
            continue;  // single-precision copies written by the same kernel
        } else {
            hipLaunchKernelGGL(k_lsm_coef<false>, dim3(V.n_ch), dim3(256), 0, mg->stream, V.dens, V.dcx, V.dcy,
                               V.dcz, (float *)nullptr, (float *)nullptr, (float *)nullptr, V.db, V.t_ch);
        }
        if (mg->cfg.precision == MF3D_MIXED) {
            convert<float, double>(mg, (float *)V.cx, V.dcx, V.cells_gh);
            convert<float, double>(mg, (float *)V.cy, V.dcy, V.cells_gh);
            convert<float, double>(mg, (float *)V.cz, V.dcz, V.cells_gh);
        }
    }
    HIPCHK(hipStreamSynchronize(mg->stream));
    HIPCHK(hipGetLastError());
    mg->density_set = true;
    if (int rc = coarse_factor<true>(mg)) return rc;  // the coarsest inverse follows the operator
    if (mg->cfg.smoother == 2) {  // the ADI Thomas coefficients follow the operator
        const int rc = mg->cfg.precision == MF3D_MIXED ? adi_coef_build<float, true>(mg)
                                                       : adi_coef_build<double, true>(mg);
        if (rc) return rc;
    }
    return 0;
}

// At every entry point: select the device and drop an error left in this
// thread's HIP state by unrelated earlier calls, so HIPCHK(hipGetLastError())
// below reports only this call's own failures.
// This device's copy of `scratch`, looked up once per device (with the
// device current) and kept for the process.
std::mutex g_scratch_mu;
std::map<int, void *> g_scratch_ptr;

int scratch_base(int dev, void **out, std::string *err) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it != g_scratch_ptr.end()) {
        *out = it->second;
        return 0;
    }
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return MF3D_EDEVICE;
    }
    g_scratch_ptr[dev] = ptr;
    *out = ptr;
    return 0;
}

int set_device(mf3d_mg *mg) {
    (void)hipGetLastError();
    HIPCHK(hipSetDevice(mg->cfg.device));
    return 0;
}

int check_ready(mf3d_mg *mg) {
    if (mg->cfg.variable_coef && !mg->density_set) {
        mg->err = "variable-coefficient handle needs mf3d_mg_set_density first";
        return MF3D_EINVAL;
    }
    return 0;
}

int solve_any(mf3d_mg *mg, const double *f_dev, double *p_dev, double *rmax_out, int32_t *iters_out,
              double *hist) {
    const bool mixed = mg->cfg.precision == MF3D_MIXED, var = mg->cfg.variable_coef != 0;
    if (mixed)
        return var ? solve_dev<float, true>(mg, f_dev, p_dev, rmax_out, iters_out, hist)
                   : solve_dev<float, false>(mg, f_dev, p_dev, rmax_out, iters_out, hist);
    return var ? solve_dev<double, true>(mg, f_dev, p_dev, rmax_out, iters_out, hist)
               : solve_dev<double, false>(mg, f_dev, p_dev, rmax_out, iters_out, hist);
}

void drop_graphs(mf3d_mg *mg) {
    for (hipGraphExec_t &g : mg->cyc)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    mg->graphs_ready = false;
    drop_timed_graphs(mg);
    if (mg->vgraph) {
        (void)hipGraphExecDestroy(mg->vgraph);
        mg->vgraph = nullptr;
    }
}

// ----------------------------------------------------- velocity form (host)
//
// exchange(1..4), pbound and bound_lsm only copy values; their sequences
// are run once on origins (mg_ghostmap.hpp) and applied as gather maps.

// bound_lsm (bounds_lsm.for:30-97) with one ghost layer: on each physical
// non-periodic face, in the order E, W, B, T, S, N, the whole ghosted
// ghost plane takes the plane inside it; the north loop starts at ipl = 0
// (bounds_lsm.for:91), so the last interior row first takes row n-1.
void shell_bound_lsm(const std::vector<mf3d_block_desc> &blk, ShellOrigins &S) {
    constexpr int order[6] = {1, 0, 4, 5, 2, 3};
    for (int b = 0; b < (int)blk.size(); ++b) {
        const BlockLevel &g = S.hb[b];
        for (const int f : order) {
            if (!is_neumann(blk[b], f)) continue;
            const int a = f / 2, n = shell_ext(g, a);
            auto copy = [&](int dst, int src) {
                shell_plane(g, S.G, a, [&](int p, int q) {
                    int d[3], s2[3];
                    shell_ijk(a, dst, p, q, d);
                    shell_ijk(a, src, p, q, s2);
                    S.set(b, d[0], d[1], d[2], S.get(b, s2[0], s2[1], s2[2]));
                });
            };
            if (f % 2 == 0) copy(0, 1);
            else {
                if (f == 3) copy(n, n - 1);
                copy(n + 1, n);
            }
        }
    }
}

// The velocity form's shell geometry: every block's array with G = ghost
// layers, blocks one after another; off addresses coordinate (0,0,0).
void shell_geometry(mf3d_mg *mg) {
    const Level &L0 = mg->lv[0];
    mg->G = std::max(1, mg->blk[0].ghost);
    const int G = mg->G;
    mg->hw = L0.hb;
    long long base = 0;
    for (size_t b = 0; b < mg->hw.size(); ++b) {
        BlockLevel &g = mg->hw[b];
        g.sy = g.n0 + 2 * G;
        g.sz = g.sy * (g.n1 + 2 * G);
        g.off = base + (G - 1) * (1 + g.sy + g.sz);
        base += g.sz * (g.n2 + 2 * G);
    }
    mg->cells_w = base;
}

int build_vel(mf3d_mg *mg) {
    if (mg->vel_ready) return 0;
    const Level &L0 = mg->lv[0];
    const int nb = mg->cfg.nblocks, G = mg->G;
    for (const mf3d_block_desc &B : mg->blk)
        if (std::max(1, B.ghost) != G) {
            mg->err = "velocity form: every block needs the same ghost width";
            return MF3D_EINVAL;
        }
    if (G < mg->cfg.pl_ex + 1) {
        mg->err = "velocity form: pl_ex = " + std::to_string(mg->cfg.pl_ex) + " needs ghost >= " +
                  std::to_string(mg->cfg.pl_ex + 1);
        return MF3D_EINVAL;
    }
    if (mg->cfg.differencing == 2 && G < 2) {
        mg->err = "velocity form: the 4th-order calmas (differencing = 2) reads two ghost layers: ghost >= 2";
        return MF3D_EINVAL;
    }
    if (mg->cells_w >= (1LL << 32) || L0.cells_gh >= (1LL << 32)) {
        mg->err = "velocity form: a field holds 2^32 entries or more";
        return MF3D_ENOTSUP;
    }
    std::vector<VelBlock> vb(nb);
    for (int b = 0; b < nb; ++b) {
        const mf3d_block_desc &B = mg->blk[b];
        for (int a = 0; a < 3; ++a) {
            vb[b].ie[a] = shell_end(B, a, a);
            const int f = 2 * a;
            if (B.nbr[f] >= 0) vb[b].lo_src[a] = B.nbr[f];
            else if (B.bc[f] == 5) vb[b].lo_src[a] = B.nbr[f + 1] < 0 ? b : B.per[f];
            else vb[b].lo_src[a] = -1;
        }
    }
    const bool periodic = std::any_of(mg->blk.begin(), mg->blk.end(), [](const mf3d_block_desc &B) {
        return std::any_of(B.bc, B.bc + 6, [](int c) { return c == 5; });
    });
    std::vector<uint2> ent;
    bool ok = true;
    for (int comp = 0; comp < 4; ++comp) {  // u, v, w (pl_ex extra layers), then p (cell field, one layer)
        ShellOrigins S(mg->hw, G, (size_t)mg->cells_w);
        const int c = comp < 3 ? comp : -1, nly = comp < 3 ? mg->cfg.pl_ex : 0;
        if (periodic) shell_exchange_bc(mg->blk, S, c, nly);
        ok = shell_exchange_smlvl(mg->blk, S, c, nly) && ok;
        if (comp == 3) shell_pbound(mg->blk, S);
        shell_entries(S, ent);
        if (comp < 3) mg->vmap_c[comp] = (long long)ent.size();
    }
    if (!ok) {
        mg->err = "velocity form: edge or corner neighbours are ambiguous (block grid not conforming)";
        return MF3D_ENOTSUP;
    }
    mg->vmap_p = (long long)ent.size() - mg->vmap_c[2];
    int rc;
    {
        auto tv = xcd_order(sweep_tiles(L0.hb, 8, VEL_TZ_MAX));
        mg->n_vel = (int)tv.size();
        if ((rc = upload(mg, &mg->t_vel, tv))) return rc;
    }
    if ((rc = upload(mg, &mg->d_vb, vb)) || (rc = upload(mg, &mg->d_vmap, ent)) ||
        (rc = upload(mg, &mg->d_vw, mg->hw)) ||
        (rc = dev_alloc(mg, (void **)&mg->d_vstage, sizeof(double) * std::max<size_t>(ent.size(), 1))))
        return rc;
    mg->vel_ready = true;
    return 0;
}

void map_launch(mf3d_mg *mg, const uint2 *ent, MapArgs m, const int *done) {
    const long long n = m.c[2];
    if (n <= 0) return;
    const unsigned gb = (unsigned)std::min<long long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_map_stage, dim3(gb), dim3(256), 0, mg->stream, m, ent, mg->d_vstage, done);
    hipLaunchKernelGGL(k_map_scatter, dim3(gb), dim3(256), 0, mg->stream, m, ent, mg->d_vstage, done);
}

// One outer step of the velocity form after a V-cycle left pp in the
// level-0 potential (update = true; newsolv_mg.for:38-55), or the step
// before the loop (update = false; newsolv_mg.for:16-17): p += pp and
// exchange(4) + pbound, then calvel + calmas and exchange(1..3), then rmax.
template <typename T, bool VAR>
void vel_step(mf3d_mg *mg, const VelFields &F, const int *done, bool update) {
    Level &L0 = mg->lv[0];
    if (update) {
        if (mg->G == 1) {  // shell and level-0 layouts equal
            const unsigned gb = (unsigned)std::min<long long>((L0.cells_gh + 255) / 256, 16384);
            hipLaunchKernelGGL(k_vel_pupdate<T>, dim3(gb), dim3(256), 0, mg->stream, (double *)F.p,
                               (const T *)L0.phi[L0.cur], L0.cells_gh, done);
        } else {
            hipLaunchKernelGGL(k_vel_pupdate_shell<T>, dim3(L0.n_row), dim3(kTileX, kTileY), 0, mg->stream,
                               (double *)F.p, (const T *)L0.phi[L0.cur], L0.db, mg->d_vw, L0.t_row, done);
        }
        const long long np = mg->vmap_p;
        map_launch(mg, mg->d_vmap + mg->vmap_c[2], MapArgs{{(double *)F.p, (double *)F.p, (double *)F.p}, {np, np, np}},
                   done);
    }
    const bool tk = tk_on(mg, MF3D_K_OUTER, 0);
    if (tk) tk_mark(mg);
    hipLaunchKernelGGL((k_velmas<T, VAR, false, 8>), dim3(mg->vel_wg), dim3(SW_TX, 8), 0, mg->stream, F,
                       (T *)L0.rhs, (double *)nullptr, mg->d_vsink, mg->st, L0.db, mg->d_vw, mg->d_vb, mg->t_vel,
                       mg->n_vel, mg->d_vpart, done);
    if (tk) tk_mark(mg);
    map_launch(mg, mg->d_vmap, MapArgs{{F.u, F.v, F.w}, {mg->vmap_c[0], mg->vmap_c[1], mg->vmap_c[2]}}, done);
    if (mg->cfg.differencing == 2)  // 4th order: on the exchanged faces; replaces the RHS and rmax above
        hipLaunchKernelGGL((k_calmas4<T, false>), dim3(mg->vel_wg), dim3(kTileX, kTileY), 0, mg->stream, F,
                           (T *)L0.rhs, (double *)nullptr, mg->st, L0.db, mg->d_vw, L0.t_in, L0.n_in, mg->d_vpart,
                           done);
    hipLaunchKernelGGL(k_vel_finalize, dim3(1), dim3(1024), 0, mg->stream, mg->d_vpart, mg->vel_wg, mg->st,
                       mg->d_rmax);
}

// Workgroups of k_velmas resident at once on the device (its persistent
// grid), at most one per tile.
template <typename T, bool VAR>
int vel_grid(mf3d_mg *mg) {
    if (mg->vel_wg) return 0;
    int per_cu = 0, cus = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_velmas<T, VAR, false, 8>, SW_TX * 8, 0));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, mg->cfg.device));
    const int cap = std::max(1, per_cu) * std::max(1, cus);
    const int rounds = (mg->n_vel + cap - 1) / cap;
    mg->vel_wg = std::max(1, (mg->n_vel + rounds - 1) / rounds);  // every workgroup walks `rounds` tiles or one fewer
    int rc = dev_alloc(mg, (void **)&mg->d_vpart, sizeof(double) * (size_t)mg->vel_wg);
    if (!rc) rc = dev_alloc(mg, (void **)&mg->d_vsink, sizeof(double) * (size_t)mg->vel_wg * SW_TX * 8);
    return rc;
}

template <typename T, bool VAR>
int newsolv_dev(mf3d_mg *mg, const VelFields &F, double dt, double alfapr, double eps, double *su,
                double *rmax_out, int32_t *iters_out, double *hist) {
    if (mg->refined) {
        mg->err = "the velocity form on refined blocks (rdiv > 1) needs the 2:1 interpolations of "
                  "exchangep/u/v/w, which this library does not have";
        return MF3D_ENOTSUP;
    }
    int rc = build_vel(mg);
    if (rc || (rc = vel_grid<T, VAR>(mg))) return rc;
    const int maxcy = mg->cfg.maxcy;
    const bool use_graph = mg->tk == MF3D_K_OFF;
    const void *key[7] = {F.us, F.vs, F.ws, F.u, F.v, F.w, F.p};
    if (use_graph && (!mg->vgraph || std::memcmp(key, mg->vkey, sizeof(key)) != 0)) {
        if (mg->vgraph) {
            (void)hipGraphExecDestroy(mg->vgraph);
            mg->vgraph = nullptr;
        }
        for (Level &V : mg->lv) V.cur = 0;
        HIPCHK(hipStreamBeginCapture(mg->stream, hipStreamCaptureModeRelaxed));
        vcycle_dev<T, VAR>(mg, true, done_ptr(mg));
        vel_step<T, VAR>(mg, F, done_ptr(mg), true);
        hipGraph_t graph = nullptr;
        HIPCHK(hipStreamEndCapture(mg->stream, &graph));
        hipError_t e = hipGraphInstantiate(&mg->vgraph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            mg->vgraph = nullptr;
            mg->err = std::string("hipGraphInstantiate: ") + hipGetErrorString(e);
            return MF3D_EDEVICE;
        }
        std::memcpy(mg->vkey, key, sizeof(key));
    }
    mg->last_replay = use_graph;
    hipLaunchKernelGGL(k_vel_reset, dim3(1), dim3(1), 0, mg->stream, mg->st, eps, dt * alfapr, dt);
    vel_step<T, VAR>(mg, F, nullptr, false);  // calvel + calmas on the caller's p
    int chunk = mg->guess > 0 ? mg->guess : 4, launched = 0;
    SolveState hs{};
    while (launched < maxcy) {
        const int n = std::min(chunk, maxcy - launched);
        for (int c = 0; c < n; ++c) {
            if (use_graph) {
                HIPCHK(hipGraphLaunch(mg->vgraph, mg->stream));
            } else {
                for (Level &V : mg->lv) V.cur = 0;
                vcycle_dev<T, VAR>(mg, true, done_ptr(mg));
                vel_step<T, VAR>(mg, F, done_ptr(mg), true);
            }
        }
        launched += n;
        HIPCHK(hipMemcpyAsync(&hs, mg->st, sizeof(SolveState), hipMemcpyDeviceToHost, mg->stream));
        HIPCHK(hipStreamSynchronize(mg->stream));
        tk_flush(mg);
        if (hs.done) break;
        chunk = 1;
    }
    const int k = hs.last;
    if (k < 0) {
        mg->err = "no outer step recorded";
        return MF3D_EDEVICE;
    }
    mg->guess = std::max(k, 1);
    if (su) {  // the last calmas once more, its su in FP64 (faces rewritten unchanged)
        Level &L0 = mg->lv[0];
        if (mg->cfg.differencing == 2)
            hipLaunchKernelGGL((k_calmas4<double, true>), dim3(mg->vel_wg), dim3(kTileX, kTileY), 0, mg->stream,
                               F, (double *)nullptr, su, mg->st, L0.db, mg->d_vw, L0.t_in, L0.n_in,
                               (double *)nullptr, (const int *)nullptr);
        else
            hipLaunchKernelGGL((k_velmas<double, VAR, true, 8>), dim3(mg->vel_wg), dim3(SW_TX, 8), 0, mg->stream,
                               F, (double *)nullptr, su, mg->d_vsink, mg->st, L0.db, mg->d_vw, mg->d_vb, mg->t_vel,
                               mg->n_vel, (double *)nullptr, (const int *)nullptr);
    }
    std::vector<double> rm(k + 1);
    HIPCHK(hipMemcpyAsync(rm.data(), mg->d_rmax, sizeof(double) * (k + 1), hipMemcpyDeviceToHost, mg->stream));
    HIPCHK(hipStreamSynchronize(mg->stream));
    HIPCHK(hipGetLastError());
    if (hist) std::memcpy(hist, rm.data(), sizeof(double) * (k + 1));
    if (iters_out) *iters_out = k;
    if (rmax_out) *rmax_out = rm[k];
    if (!(rm[k] <= 100.0)) {  // newsolv_mg.for:64 (a NaN counts as diverged here)
        mg->err = "BIG RMAX (rmax > 100): diverged";
        return MF3D_DIVERGED;
    }
    return hs.done ? MF3D_OK : MF3D_NOT_CONVERGED;
}

int newsolv_any(mf3d_mg *mg, const VelFields &F, double dt, double alfapr, double eps, double *su,
                double *rmax_out, int32_t *iters_out, double *hist) {
    if (mg->cfg.variable_coef && !mg->dens_ghosted) {
        mg->err = "velocity form under LSM needs the density with its ghost layer (mf3d_mg_set_density "
                  "with ghost >= 1)";
        return MF3D_EINVAL;
    }
    const bool mixed = mg->cfg.precision == MF3D_MIXED, var = mg->cfg.variable_coef != 0;
    if (mixed)
        return var ? newsolv_dev<float, true>(mg, F, dt, alfapr, eps, su, rmax_out, iters_out, hist)
                   : newsolv_dev<float, false>(mg, F, dt, alfapr, eps, su, rmax_out, iters_out, hist);
    return var ? newsolv_dev<double, true>(mg, F, dt, alfapr, eps, su, rmax_out, iters_out, hist)
               : newsolv_dev<double, false>(mg, F, dt, alfapr, eps, su, rmax_out, iters_out, hist);
}

// ------------------------------------------------ level-set density (host)

// The map of bound_lsm(18) + exchange(18) over the level-0 arena and the
// per-block epsl table (lsm.for:1475).
int build_ls(mf3d_mg *mg) {
    if (mg->ls_ready) return 0;
    const Level &L0 = mg->lv[0];
    const int nb = mg->cfg.nblocks;
    if (L0.cells_gh >= (1LL << 32)) {
        mg->err = "heaviside: a level-0 field holds 2^32 entries or more";
        return MF3D_ENOTSUP;
    }
    const bool periodic = std::any_of(mg->blk.begin(), mg->blk.end(), [](const mf3d_block_desc &B) {
        return std::any_of(B.bc, B.bc + 6, [](int c) { return c == 5; });
    });
    ShellOrigins S(L0.hb, 1, (size_t)L0.cells_gh);  // the one-layer level-0 geometry
    shell_bound_lsm(mg->blk, S);
    if (periodic) shell_exchange_bc(mg->blk, S, -1, 0);
    if (!shell_exchange_smlvl(mg->blk, S, -1, 0)) {
        mg->err = "heaviside: edge or corner neighbours are ambiguous (block grid not conforming)";
        return MF3D_ENOTSUP;
    }
    std::vector<uint2> ent;
    shell_entries(S, ent);
    std::vector<LsBlock> lb(nb);
    long long mx = 0;
    for (int b = 0; b < nb; ++b) {
        const BlockLevel &g = L0.hb[b];
        const double *h = mg->blk[b].h;
        lb[b].off = g.off;
        lb[b].cells = (long long)(g.n0 + 2) * (g.n1 + 2) * (g.n2 + 2);
        lb[b].epsl = 2 * std::max(std::max(h[0], h[1]), h[2]);
        mx = std::max(mx, lb[b].cells);
    }
    int rc;
    if ((rc = upload(mg, &mg->d_lsb, lb)) || (!ent.empty() && (rc = upload(mg, &mg->d_lsmap, ent)))) return rc;
    mg->n_lsmap = (long long)ent.size();
    mg->ls_max_cells = mx;
    mg->ls_ready = true;
    return 0;
}

// heaviside + bound_lsm(18/19) into the level-0 density (and the caller's
// copies), then coeff from it.
int heaviside_dev(mf3d_mg *mg, const double *phi, const LsProps &pr, double *dens_out, double *mu) {
    int rc = build_ls(mg);
    if (rc) return rc;
    Level &L0 = mg->lv[0];
    const int nb = mg->cfg.nblocks;
    const bool tk = tk_on(mg, MF3D_K_HEAVISIDE, 0);
    if (tk) tk_mark(mg);
    hipLaunchKernelGGL(k_heaviside, dim3((unsigned)((mg->ls_max_cells + 1023) / 1024), (unsigned)nb), dim3(256), 0,
                       mg->stream, phi, L0.dens, dens_out, mu, pr, mg->d_lsb);
    if (tk) tk_mark(mg);
    if (mg->n_lsmap > 0) {
        const unsigned gb = (unsigned)std::min<long long>((mg->n_lsmap + 255) / 256, 4096);
        hipLaunchKernelGGL(k_heaviside_map, dim3(gb), dim3(256), 0, mg->stream, phi, mg->d_lsmap, mg->n_lsmap,
                           L0.dens, dens_out, mu, pr, mg->d_lsb, nb);
    }
    HIPCHK(hipGetLastError());
    mg->dens_ghosted = true;
    return build_lsm(mg);
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
    std::string v = validate(*cfg, blocks);
    mf3d_mg *mg = new mf3d_mg();
    mg->cfg = *cfg;
    *out = mg;  // returned even on failure so the caller can read the message
    if (!v.empty()) {
        mg->err = v;
        return MF3D_EINVAL;
    }
    v = refined_scope(*cfg, blocks);
    if (!v.empty()) {
        mg->err = v;
        return MF3D_ENOTSUP;
    }
    mg->blk.assign(blocks, blocks + cfg->nblocks);
    int rc = set_device(mg);
    if (rc) return rc;
    // non-blocking: the handle's work never waits on (or makes the legacy
    // stream wait on) other threads' streams, some of which may be capturing
    hipError_t e = hipStreamCreateWithFlags(&mg->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        mg->err = std::string("hipStreamCreateWithFlags: ") + hipGetErrorString(e);
        return MF3D_EDEVICE;
    }
    return build(mg);
}

int mf3d_mg_set_density(mf3d_mg *mg, const double *const *dens_blocks) {
    if (!mg || !dens_blocks) return MF3D_EINVAL;
    if (!mg->cfg.variable_coef) {
        mg->err = "handle was created without variable_coef";
        return MF3D_EINVAL;
    }
    int rc = set_device(mg);
    if (rc) return rc;
    Level &L0 = mg->lv[0];
    // with a ghost layer in the caller's arrays the level-0 density keeps it
    // (the velocity form's calvel reads rho across block faces); the
    // coefficient hierarchy reads interior densities only
    mg->dens_ghosted = std::all_of(mg->blk.begin(), mg->blk.end(),
                                   [](const mf3d_block_desc &B) { return B.ghost >= 1; });
    if (mg->dens_ghosted) {
        host_to_packed(mg, dens_blocks, mg->h_io, 1);
        HIPCHK(hipMemcpyAsync(L0.dens, mg->h_io, sizeof(double) * L0.cells_gh, hipMemcpyHostToDevice, mg->stream));
    } else {
        host_to_packed(mg, dens_blocks, mg->h_io, 0);
        HIPCHK(hipMemcpyAsync(mg->d_io, mg->h_io, sizeof(double) * mg->cells, hipMemcpyHostToDevice, mg->stream));
        hipLaunchKernelGGL(k_unpack<double>, dim3(L0.n_in), dim3(kTileX, kTileY), 0, mg->stream, mg->d_io, L0.dens,
                           L0.db, L0.t_in);
    }
    return build_lsm(mg);
}

int mf3d_mg_set_density_device(mf3d_mg *mg, const double *dens, void *stream) {
    if (!mg || !dens) return MF3D_EINVAL;
    if (!mg->cfg.variable_coef) {
        mg->err = "handle was created without variable_coef";
        return MF3D_EINVAL;
    }
    int rc = set_device(mg);
    if (rc) return rc;
    Level &L0 = mg->lv[0];
    hipStream_t own = mg->stream;
    if (stream) mg->stream = (hipStream_t)stream;
    hipError_t e = hipMemcpyAsync(L0.dens, dens, sizeof(double) * L0.cells_gh, hipMemcpyDeviceToDevice, mg->stream);
    if (e == hipSuccess) {
        mg->dens_ghosted = true;
        rc = build_lsm(mg);
    } else {
        mg->err = std::string("hipMemcpyAsync: ") + hipGetErrorString(e);
        rc = MF3D_EDEVICE;
    }
    mg->stream = own;
    return rc;
}

int mf3d_mg_heaviside_device(mf3d_mg *mg, const double *phi, double densl, double densg, double mul, double mug,
                             double *dens, double *mu, void *stream) {
    if (!mg || !phi) return MF3D_EINVAL;
    if (!mg->cfg.variable_coef) {
        mg->err = "handle was created without variable_coef";
        return MF3D_EINVAL;
    }
    int rc = set_device(mg);
    if (rc) return rc;
    hipStream_t own = mg->stream;
    if (stream) mg->stream = (hipStream_t)stream;
    rc = heaviside_dev(mg, phi, LsProps{densl, densg, mul, mug}, dens, mu);
    mg->stream = own;
    return rc;
}

int mf3d_mg_heaviside(mf3d_mg *mg, const double *const *phi_blocks, double densl, double densg, double mul,
                      double mug, double *const *dens_blocks, double *const *mu_blocks) {
    if (!mg || !phi_blocks) return MF3D_EINVAL;
    if (!mg->cfg.variable_coef) {
        mg->err = "handle was created without variable_coef";
        return MF3D_EINVAL;
    }
    for (int b = 0; b < mg->cfg.nblocks; ++b)
        if (mg->blk[b].ghost < 1) {
            mg->err = "heaviside needs caller arrays with ghost >= 1";
            return MF3D_EINVAL;
        }
    int rc = set_device(mg);
    if (rc) return rc;
    Level &L0 = mg->lv[0];
    const size_t bytes = sizeof(double) * (size_t)L0.cells_gh;
    for (int i = 0; i < 2; ++i)
        if (!mg->d_lsio[i] && (i == 0 || mu_blocks) && (rc = dev_alloc(mg, (void **)&mg->d_lsio[i], bytes)))
            return rc;
    host_to_packed(mg, phi_blocks, mg->h_io, 1);
    HIPCHK(hipMemcpyAsync(mg->d_lsio[0], mg->h_io, bytes, hipMemcpyHostToDevice, mg->stream));
    HIPCHK(hipStreamSynchronize(mg->stream));
    if ((rc = heaviside_dev(mg, mg->d_lsio[0], LsProps{densl, densg, mul, mug}, nullptr,
                            mu_blocks ? mg->d_lsio[1] : nullptr)))
        return rc;
    const double *src[2] = {L0.dens, mg->d_lsio[1]};
    double *const *dst[2] = {dens_blocks, mu_blocks};
    for (int i = 0; i < 2; ++i) {
        if (!dst[i]) continue;
        HIPCHK(hipMemcpyAsync(mg->h_io, src[i], bytes, hipMemcpyDeviceToHost, mg->stream));
        HIPCHK(hipStreamSynchronize(mg->stream));
        packed_to_host(mg, mg->h_io, dst[i], 1);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

int mf3d_mg_vcycle(mf3d_mg *mg, const double *const *rhs_blocks, double *const *pp_blocks) {
    if (!mg || !rhs_blocks || !pp_blocks) return MF3D_EINVAL;
    int rc = check_ready(mg);
    if (rc) return rc;
    for (int b = 0; b < mg->cfg.nblocks; ++b)
        if (mg->blk[b].ghost < 1) {
            mg->err = "vcycle needs caller arrays with ghost >= 1";
            return MF3D_EINVAL;
        }
    if ((rc = set_device(mg))) return rc;
    Level &L0 = mg->lv[0];
    const bool mixed = mg->cfg.precision == MF3D_MIXED, var = mg->cfg.variable_coef != 0;
    host_to_packed(mg, rhs_blocks, mg->h_io, 0);
    HIPCHK(hipMemcpyAsync(mg->d_io, mg->h_io, sizeof(double) * mg->cells, hipMemcpyHostToDevice, mg->stream));
    if (mixed)
        hipLaunchKernelGGL(k_unpack<float>, dim3(L0.n_in), dim3(kTileX, kTileY), 0, mg->stream, mg->d_io,
                           (float *)L0.rhs, L0.db, L0.t_in);
    else
        hipLaunchKernelGGL(k_unpack<double>, dim3(L0.n_in), dim3(kTileX, kTileY), 0, mg->stream, mg->d_io,
                           (double *)L0.rhs, L0.db, L0.t_in);
    HIPCHK(hipStreamSynchronize(mg->stream));
    host_to_packed(mg, pp_blocks, mg->h_io, 1);
    HIPCHK(hipMemcpyAsync(mg->d_io, mg->h_io, sizeof(double) * L0.cells_gh, hipMemcpyHostToDevice, mg->stream));
    if (mixed) rc = var ? vcycle_host_t<float, true>(mg) : vcycle_host_t<float, false>(mg);
    else rc = var ? vcycle_host_t<double, true>(mg) : vcycle_host_t<double, false>(mg);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(mg->h_io, mg->d_io, sizeof(double) * L0.cells_gh, hipMemcpyDeviceToHost, mg->stream));
    HIPCHK(hipStreamSynchronize(mg->stream));
    HIPCHK(hipGetLastError());
    packed_to_host(mg, mg->h_io, pp_blocks, 1);
    return 0;
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
    (void)hipStreamSynchronize(mg->stream);
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

int mf3d_mg_solve_device_noalloc(const mf3d_mg_cfg *cfg, const mf3d_block_desc *blocks, const double *f_dev,
                                 double *p_dev, void *stream, double *rmax_out, int32_t *iters_out, double *hist,
                                 char *err, size_t err_len) {
    mf3d_mg tmp;  // host-side state of this call only
    auto fail = [&](int rc) {
        if (err && err_len) std::snprintf(err, err_len, "%s", tmp.err.c_str());
        return rc;
    };
    if (err && err_len) err[0] = '\0';
    if (!cfg || !blocks || !f_dev || !p_dev) {
        tmp.err = "null argument";
        return fail(MF3D_EINVAL);
    }
    std::string v = validate(*cfg, blocks);
    if (!v.empty()) {
        tmp.err = v;
        return fail(MF3D_EINVAL);
    }
    if (any_refined(cfg->nblocks, blocks)) {
        tmp.err = "mf3d_mg_solve_device_noalloc takes unrefined blocks only (rdiv 1)";
        return fail(MF3D_ENOTSUP);
    }
    if (cfg->variable_coef) {
        tmp.err = "mf3d_mg_solve_device_noalloc solves constant-coefficient problems only (a density "
                  "needs a handle: mf3d_mg_set_density)";
        return fail(MF3D_ENOTSUP);
    }
    // checked before any device work: the estimate less the host entries'
    // staging (which this solve does not hold), plus room for the tables
    long long gh0 = 0, in0 = 0;
    for (int i = 0; i < cfg->nblocks; ++i) {
        gh0 += (long long)(blocks[i].n[0] + 2) * (blocks[i].n[1] + 2) * (blocks[i].n[2] + 2);
        in0 += (long long)blocks[i].n[0] * blocks[i].n[1] * blocks[i].n[2];
    }
    const long long need = estimate(*cfg, blocks) - std::max(gh0, in0) * 8 + (1LL << 20);
    if (need > (long long)MF3D_NOALLOC_SCRATCH_BYTES) {
        tmp.err = "the solve needs about " + std::to_string(need >> 20) + " MiB of device memory, more than the " +
                  std::to_string((long long)MF3D_NOALLOC_SCRATCH_BYTES >> 20) +
                  " MiB scratch of mf3d_mg_solve_device_noalloc";
        return fail(MF3D_ENOMEM);
    }
    tmp.cfg = *cfg;
    tmp.blk.assign(blocks, blocks + cfg->nblocks);
    int rc = set_device(&tmp);
    if (rc) return fail(rc);
    void *base = nullptr;
    if ((rc = scratch_base(cfg->device, &base, &tmp.err))) return fail(rc);
    tmp.arena = (char *)base;
    tmp.arena_cap = MF3D_NOALLOC_SCRATCH_BYTES;
    tmp.stream = (hipStream_t)stream;  // NULL: the null stream
    if ((rc = build(&tmp))) return fail(rc);
    rc = solve_any(&tmp, f_dev, p_dev, rmax_out, iters_out, hist);
    (void)hipStreamSynchronize(tmp.stream);
    tmp.stream = nullptr;  // not ours to destroy
    return rc < 0 ? fail(rc) : rc;
}

int mf3d_mg_newsolv_device(mf3d_mg *mg, const double *ustar, const double *vstar, const double *wstar, double *u,
                           double *v, double *w, double *p, double dt, double alfapr, double eps, double *su,
                           void *stream, double *rmax_out, int32_t *iters_out, double *rmax_hist) {
    if (!mg || !ustar || !vstar || !wstar || !u || !v || !w || !p) return MF3D_EINVAL;
    int rc = check_ready(mg);
    if (rc) return rc;
    if ((rc = set_device(mg))) return rc;
    const VelFields F{ustar, vstar, wstar, u, v, w, p, mg->cfg.variable_coef ? mg->lv[0].dens : nullptr};
    hipStream_t own = mg->stream;
    if (stream) mg->stream = (hipStream_t)stream;
    rc = newsolv_any(mg, F, dt, alfapr, eps, su, rmax_out, iters_out, rmax_hist);
    mg->stream = own;
    return rc;
}

int mf3d_mg_newsolv(mf3d_mg *mg, const double *const *ustar, const double *const *vstar, const double *const *wstar,
                    double *const *u, double *const *v, double *const *w, double *const *p, double dt,
                    double alfapr, double eps, double *const *su, double *rmax_out, int32_t *iters_out,
                    double *rmax_hist) {
    if (!mg || !ustar || !vstar || !wstar || !u || !v || !w || !p) return MF3D_EINVAL;
    int rc = check_ready(mg);
    if (rc) return rc;
    for (int b = 0; b < mg->cfg.nblocks; ++b)
        if (mg->blk[b].ghost < 1) {
            mg->err = "newsolv needs caller arrays with ghost >= 1";
            return MF3D_EINVAL;
        }
    if ((rc = set_device(mg))) return rc;
    Level &L0 = mg->lv[0];
    const int G = mg->G;  // the whole caller shell travels
    const size_t bytes = sizeof(double) * (size_t)mg->cells_w;
    for (int i = 0; i < 8; ++i)
        if (!mg->vf[i] && (i < 7 || su) && (rc = dev_alloc(mg, (void **)&mg->vf[i], bytes))) return rc;
    const double *const *in[7] = {ustar, vstar, wstar, u, v, w, p};
    for (int i = 0; i < 7; ++i) {
        host_to_packed(mg, in[i], mg->h_io, G);
        HIPCHK(hipMemcpyAsync(mg->vf[i], mg->h_io, bytes, hipMemcpyHostToDevice, mg->stream));
        HIPCHK(hipStreamSynchronize(mg->stream));
    }
    const VelFields F{mg->vf[0], mg->vf[1], mg->vf[2], mg->vf[3], mg->vf[4], mg->vf[5], mg->vf[6],
                      mg->cfg.variable_coef ? L0.dens : nullptr};
    rc = newsolv_any(mg, F, dt, alfapr, eps, su ? mg->vf[7] : nullptr, rmax_out, iters_out, rmax_hist);
    if (rc < 0) return rc;
    double *const *out[4] = {u, v, w, p};
    for (int i = 0; i < 4; ++i) {
        HIPCHK(hipMemcpyAsync(mg->h_io, mg->vf[3 + i], bytes, hipMemcpyDeviceToHost, mg->stream));
        HIPCHK(hipStreamSynchronize(mg->stream));
        packed_to_host(mg, mg->h_io, out[i], G);
    }
    if (su) {
        hipLaunchKernelGGL(k_pack<double>, dim3(L0.n_in), dim3(kTileX, kTileY), 0, mg->stream, mg->vf[7], mg->d_io,
                           mg->d_vw, L0.t_in);
        HIPCHK(hipMemcpyAsync(mg->h_io, mg->d_io, sizeof(double) * mg->cells, hipMemcpyDeviceToHost, mg->stream));
        HIPCHK(hipStreamSynchronize(mg->stream));
        packed_to_host(mg, mg->h_io, su, 0);
    }
    HIPCHK(hipGetLastError());
    return rc;
}

int64_t mf3d_mg_ghosted_cells(const mf3d_mg *mg) { return mg ? mg->lv[0].cells_gh : MF3D_EINVAL; }

int64_t mf3d_mg_shell_cells(const mf3d_mg *mg) { return mg ? mg->cells_w : MF3D_EINVAL; }

int mf3d_mg_ghost_table(const mf3d_mg_cfg *cfg, const mf3d_block_desc *blocks, int32_t level, int32_t *out,
                        int64_t cap, int64_t *count, char *err, size_t err_len) {
    auto fail = [&](int rc, const std::string &why) {
        if (err && err_len) std::snprintf(err, err_len, "%s", why.c_str());
        return rc;
    };
    if (err && err_len) err[0] = '\0';
    if (!cfg || !blocks || !count || (cap > 0 && !out)) return fail(MF3D_EINVAL, "null argument");
    std::string v = validate(*cfg, blocks);
    if (!v.empty()) return fail(MF3D_EINVAL, v);
    v = refined_scope(*cfg, blocks);
    if (!v.empty()) return fail(MF3D_ENOTSUP, v);
    if (level < 0 || level >= cfg->levels) return fail(MF3D_EINVAL, "level out of range");
    std::vector<mf3d_block_desc> blk(blocks, blocks + cfg->nblocks);
    std::vector<GhostExpr> gx;
    v = ghost_exprs(blk, own_levels(cfg->nblocks, blocks, cfg->levels), cfg->levels, level, &gx);
    if (!v.empty()) return fail(MF3D_ENOTSUP, v);
    *count = (int64_t)gx.size();
    for (int64_t i = 0; i < std::min<int64_t>(cap, *count); ++i) {
        const GhostExpr &e = gx[(size_t)i];
        out[10 * i] = e.dst;
        out[10 * i + 1] = e.kind;
        for (int q = 0; q < 8; ++q) out[10 * i + 2 + q] = e.src[q];
    }
    return MF3D_OK;
}

int mf3d_mg_time_kernel(mf3d_mg *mg, int kernel) {
    if (!mg || kernel < MF3D_K_OFF || kernel > MF3D_K_HEAVISIDE) return MF3D_EINVAL;
    if (mg->stream) (void)hipStreamSynchronize(mg->stream);
    tk_flush(mg);
    if (kernel != mg->tk) drop_timed_graphs(mg);
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

int mf3d_mg_set_eps(mf3d_mg *mg, double eps) {
    if (!mg || !(eps >= 0.0) || !std::isfinite(eps)) return MF3D_EINVAL;
    mg->cfg.eps = eps;  // read by the next solve (captured cycles take it from device memory)
    return MF3D_OK;
}

int mf3d_mg_graph_replay(const mf3d_mg *mg) { return mg ? mg->last_replay : MF3D_EINVAL; }

int64_t mf3d_mg_cells(const mf3d_mg *mg) { return mg ? mg->cells : MF3D_EINVAL; }
int64_t mf3d_mg_device_bytes(const mf3d_mg *mg) { return mg ? mg->bytes : MF3D_EINVAL; }

const char *mf3d_mg_last_error(const mf3d_mg *mg) {
    if (!mg) return "null handle";
    return mg->err.c_str();
}

void mf3d_mg_destroy(mf3d_mg *mg) {
    if (!mg) return;
    if (mg->stream) (void)hipStreamSynchronize(mg->stream);
    drop_graphs(mg);
    for (void *p : mg->allocs) (void)hipFree(p);
    if (mg->h_io) (void)hipHostFree(mg->h_io);
    for (hipEvent_t e : mg->ev_eager.ev) (void)hipEventDestroy(e);
    if (mg->stream) (void)hipStreamDestroy(mg->stream);
    delete mg;
}

}  // extern "C"

// host-side reader of the reference's checkpoint files
#include "mf3d_tecbin.hpp"
