/*
 * mf3d_mg.h -- C ABI of the MI355X (gfx950) pressure-Poisson multigrid that
 * drops in for MultiFlow3D's FP64 solver path:
 *
 *   subroutine coeff       mgsolver.for:2-222       -> mf3d_mg_create / mf3d_mg_set_density
 *                                                     (and mf3d_mg_set_density_device)
 *   subroutine heaviside   lsm.for:1456-1509        -> mf3d_mg_heaviside(_device), which
 *                                                     rebuilds coeff's hierarchy as well
 *   subroutine tecbin      post.for:496-548         -> mf3d_tecbin_read (the checkpoint the
 *                                                     reference's restart reads back,
 *                                                     initial.for:689-760)
 *   subroutine mgkcyc      mgsolver.for:224-369     -> mf3d_mg_vcycle
 *   subroutine newsolv_mg  newsolv_mg.for:2-102     -> mf3d_mg_newsolv (the reference's own
 *                                                       velocity form, with calvel, calmas,
 *                                                       pbound and exchange(1..4),
 *                                                       press.for:2-200, exchange.for:2-2499)
 *                                                     and mf3d_mg_solve (Poisson form)
 *
 * The reference routines take no arguments: their inputs are module state
 * (module_vars.for:11-15, module_multidata.for:15-120).  Here the same state
 * is passed explicitly: one mf3d_block_desc per dom(ib), and the MG scalars
 * (ngrd_gl, mg_itrsch, irestr, iproln, maxcy, eps, L_LSM) in mf3d_mg_cfg.
 *
 * Arrays are plain pointers to FP64 data in the caller's layout:
 * column-major (i fastest), dims ld[0..2], interior starting at index
 * `ghost` on every axis -- the reference's dom(ib)%p / %pp / %su / %sup /
 * %dens with ld = ttc_* and ghost = pl (initial.for:199-204).
 *
 * Return codes: 0 converged / success, 1 not converged within maxcy
 * (newsolv_mg.for:60-61 prints and continues), 2 diverged (rmax > 100,
 * newsolv_mg.for:64-83 stops), < 0 an argument, setup or device error
 * (text via mf3d_mg_last_error).
 *
 * A handle owns its device memory and one HIP stream; it is not thread-safe.
 */
#ifndef MF3D_MG_H
#define MF3D_MG_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MF3D_MG_ABI_VERSION 7

enum {
    MF3D_OK = 0,
    MF3D_NOT_CONVERGED = 1,
    MF3D_DIVERGED = 2,
    MF3D_EINVAL = -22,
    MF3D_ENOTSUP = -95,
    MF3D_ENOMEM = -12,
    MF3D_EDEVICE = -5,
    MF3D_ENOENT = -2,   /* checkpoint reader: the file cannot be opened */
    MF3D_EFORMAT = -74  /* checkpoint reader: the file is not in the expected format */
};

enum { MF3D_FP64 = 0, MF3D_MIXED = 1 };

/* Coarse operators of the variable-coefficient form (mf3d_mg_cfg.coarse_op):
 * the reference's density hierarchy (8-cell density average, face
 * coefficients from the finest dx, dy, dz on every level, mgsolver.for:69-181),
 * or face-coefficient restriction (each coarse face coefficient is the sum of
 * the four fine face coefficients it covers / 16, i.e. the mean flux
 * coefficient rescaled to the coarse spacing).  No counterpart of the second
 * in the reference: with a 1000:1 density ratio the first stalls. */
enum { MF3D_COARSE_REFERENCE = 0, MF3D_COARSE_FACE = 1 };

/* Outer iteration (mf3d_mg_cfg.krylov): the reference's defect correction
 * p += V(f - A p) (newsolv_mg.for:23-58), or GCR(1) with the V-cycle as
 * preconditioner (residual-minimising, one direction kept, valid for the
 * reference's non-symmetric multi-block density operator; no counterpart in
 * the reference). */
enum { MF3D_OUTER_DEFECT = 0, MF3D_OUTER_GCR = 1 };

/* Coarsest level (mf3d_mg_cfg.coarse_solve).  The reference solves it by
 * irestr + iproln smoothing sweeps (mgsolver.for:224-369).
 * MF3D_COARSEST_DEFAULT: those sweeps, computed in FP64 in both precisions
 * (the mixed form widens the coarse RHS, sweeps in FP64 and narrows the
 * result), so the mixed form follows the FP64 form's path.
 * MF3D_COARSEST_INNER: the sweeps in the inner precision (FP32 in mixed).
 * MF3D_COARSEST_DIRECT: an exact FP64 solve (a dense inverse of the
 * coarsest operator, rebuilt with the coefficients, applied every cycle);
 * it removes the smooth error the sweeps leave, so a solve stops at a
 * different (more accurate) iterate than the reference's path.  It needs
 * levels >= 2 and at most MF3D_COARSEST_MAX_CELLS coarsest cells. */
enum { MF3D_COARSEST_DEFAULT = 0, MF3D_COARSEST_INNER = 1, MF3D_COARSEST_DIRECT = 2 };
#define MF3D_COARSEST_MAX_CELLS 4096

typedef struct mf3d_mg mf3d_mg;

/* One block: the fields of dom(ib) the multigrid reads
 * (module_multidata.for:16-28,36,116; alloc_dom.for:456-616,972-1150).
 * Faces are ordered W,E,S,N,B,T.  Edge and corner neighbours (the
 * corprev, cornext, edgprev and edgnext fields) are not passed: on a
 * conforming block grid they follow from the face neighbours. */
typedef struct {
    int32_t n[3];    /* interior cells along i, j, k (iep-isp+1, ...) */
    int32_t ghost;   /* ghost layers in the caller's arrays (pl) */
    int32_t ld[3];   /* caller array dims (ttc_i, ttc_j, ttc_k); 0 -> n + 2*ghost */
    double  h[3];    /* dx, dy, dz */
    int32_t nbr[6];  /* iprev, inext, jprev, jnext, kprev, knext: block index or -1 */
    int32_t per[6];  /* per_ip, per_in, per_jp, per_jn, per_kp, per_kn: block index or -1 */
    int32_t bc[6];   /* bc_west .. bc_top; 5 = periodic, any other code = Neumann for p */
    int32_t rdiv;    /* rdiv(dom_id(ib)) (infodom.cin, localparameters.for:26-56): a power of
                        2, 0 = 1.  n and h are the block's own (refined) extents and
                        spacing; the block runs levels - log2(max rdiv) + log2(rdiv) levels
                        of its own and repeats its coarsest grid beyond them */
} mf3d_block_desc;

typedef struct {
    int32_t nblocks;       /* nbp */
    int32_t levels;        /* ngrd_gl; every extent must divide by 2^(levels-1) */
    int32_t smoother;      /* mg_itrsch (mgsolver.for:403-611): 1 lexicographic Gauss-Seidel,
                              2 ADI line relaxation with TDMA, 3 red-black Gauss-Seidel */
    int32_t nu_pre;        /* irestr */
    int32_t nu_post;       /* iproln */
    int32_t maxcy;         /* maximum outer cycles */
    double  eps;           /* stop when ||f - A p||_2 / ||f||_2 <= eps */
    int32_t min_cycles;    /* newsolv_mg requires iter > 1, i.e. 2 */
    int32_t precision;     /* MF3D_FP64 (bitwise the reference) or MF3D_MIXED */
    int32_t variable_coef; /* 0: 1/h^2 coefficients; 1: LSM density coefficients */
    int32_t device;        /* HIP device ordinal */
    int32_t coarse_op;     /* MF3D_COARSE_*; variable-coefficient form only */
    int32_t krylov;        /* MF3D_OUTER_* */
    int32_t coarse_solve;  /* MF3D_COARSEST_* */
    int32_t differencing;  /* the reference's differencing (control.cin): 2 = 4th-order calmas
                              (press.for:13-36, needs ghost >= 2), any other value 2nd order */
    int32_t pl_ex;         /* extra ghost layers exchange(1..3) refreshes for u, v, w
                              (exchange.for:14-23); needs ghost >= pl_ex + 1 */
} mf3d_mg_cfg;

/* coeff sizing + constant coefficients (mgsolver.for:2-222, localparameters.for:200-237). */
int mf3d_mg_create(const mf3d_mg_cfg *cfg, const mf3d_block_desc *blocks, mf3d_mg **out);

/* coeff, L_LSM branch (mgsolver.for:69-181): cell density dom(ib)%dens of
 * every block, caller layout.  Builds the whole level hierarchy on device. */
int mf3d_mg_set_density(mf3d_mg *mg, const double *const *dens_blocks);

/* The same from a device-resident density: every block's ghosted array
 * ((n0+2)(n1+2)(n2+2) values, i fastest) concatenated in block order,
 * mf3d_mg_ghosted_cells values (the layout of mf3d_mg_newsolv_device
 * when ghost is 1); the ghost layer is kept for the velocity form.  Runs on `stream` (NULL =
 * the handle's own) and returns when the hierarchy is built. */
int mf3d_mg_set_density_device(mf3d_mg *mg, const double *dens, void *stream);

/* heaviside (lsm.for:1456-1509) followed by coeff (mgsolver.for:69-181),
 * the density update of the reference's LSM time step (flosol.for:114-117)
 * and the coefficient rebuild of newsolv_mg.for:12-14, without leaving
 * the device.  Per block epsl = 2*max(dx, dy, dz); over every cell of the
 * one-layer ghosted box
 *   h = 0 (phi < -epsl), 1 (phi > epsl),
 *       0.5*(1 + phi/epsl + 1/pi*sin(pi*phi/epsl)) otherwise, pi = 3.14159265359
 *   dens = densg + (densl-densg)*h,  mu = mug + (mul-mug)*h,
 * then bound_lsm(18/19) (bounds_lsm.for:2-104: zero-gradient ghosts on the
 * physical non-periodic faces, E, W, B, T, S, N, the north loop also
 * rewriting the last interior row from the one below it) and
 * exchange(18/19).  With a periodic face the reference's exchange_bc
 * selects no array for the density (exchange_bc.for:105,422); the path it
 * defines for the viscosity (op 19) is applied to both.  phi must hold its
 * ghost layer as the reference's bound_lsm(14) leaves it.  The handle's
 * level-0 density (with its ghost layer) and coefficient hierarchy are
 * replaced, as by mf3d_mg_set_density.  Needs variable_coef.
 *
 * Device form: phi, dens and mu are one-layer ghosted device arrays
 * (mf3d_mg_ghosted_cells values); dens and mu may be NULL (not written).  Runs on
 * `stream` (NULL = the handle's own) and returns when the hierarchy is
 * built.  Host form: caller arrays (ghost >= 1); dens and mu (may be NULL)
 * are written over the interior and the first ghost layer.  sin is the
 * device library's: in the transition band |phi| <= epsl a value may
 * differ from the reference's (host libm) in the last place. */
int mf3d_mg_heaviside_device(mf3d_mg *mg, const double *phi, double densl, double densg, double mul,
                             double mug, double *dens, double *mu, void *stream);
int mf3d_mg_heaviside(mf3d_mg *mg, const double *const *phi_blocks, double densl, double densg,
                      double mul, double mug, double *const *dens_blocks, double *const *mu_blocks);

/* mgkcyc (mgsolver.for:224-369): one V-cycle on rhs (su, or sup under LSM);
 * pp is read and written including one ghost layer around the interior,
 * so the caller arrays need ghost >= 1. */
int mf3d_mg_vcycle(mf3d_mg *mg, const double *const *rhs_blocks, double *const *pp_blocks);

/* newsolv_mg in Poisson form: p = 0, then repeat { one V-cycle on
 * r = f - A p; p += e } until ||r||/||f|| <= eps (and at least min_cycles
 * cycles) or maxcy cycles.  rmax_out gets the reference's own absolute
 * criterion max|r|*dx*dy*dz (max|r| under LSM, whose RHS carries the cell
 * volume); hist (maxcy+1 doubles, may be NULL) gets ||r_k||/||f|| for
 * k = 0..iters.  p is written over the interior only. */
int mf3d_mg_solve(mf3d_mg *mg, const double *const *f_blocks, double *const *p_blocks,
                  double *rmax_out, int32_t *iters_out, double *hist);

/* Device-resident solve: f_dev and p_dev are device pointers to all blocks'
 * interiors concatenated in block order (i fastest within a block).  Runs
 * on `stream` (a hipStream_t; NULL = the handle's own stream) and returns
 * when the solve is complete.  No host transfer of fields. */
int mf3d_mg_solve_device(mf3d_mg *mg, const double *f_dev, double *p_dev, void *stream,
                         double *rmax_out, int32_t *iters_out, double *hist);

/* Allocation-free form of mf3d_mg_solve_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  One call builds the level hierarchy of cfg/blocks in a
 * fixed device array of MF3D_NOALLOC_SCRATCH_BYTES that the library holds
 * on each GPU it runs on until the process ends, runs the solve with
 * kernels launched one by one on `stream` (NULL = the null stream), and
 * returns when it is complete; it calls no hipMalloc, hipFree or other
 * allocating routine, and it needs no handle.  f_dev, p_dev, rmax_out,
 * iters_out and hist are as in mf3d_mg_solve_device (results are bitwise
 * those of a handle created with the same cfg and blocks, with the default
 * sweep tiling, MF3D_AUTOTUNE=0).  Constant coefficients only
 * (variable_coef = 0, else MF3D_ENOTSUP).  A problem whose hierarchy does
 * not fit the array returns MF3D_ENOMEM before any device work.  The array
 * is shared by every call on a GPU and keeps the last call's contents
 * (each call clears what it reads); calls that overlap on one GPU (other
 * threads, other streams) race and must be serialised by the caller.
 * err (may be NULL) receives up to err_len bytes of text on failure. */
#define MF3D_NOALLOC_SCRATCH_BYTES (32u << 20)
int mf3d_mg_solve_device_noalloc(const mf3d_mg_cfg *cfg, const mf3d_block_desc *blocks, const double *f_dev,
                                 double *p_dev, void *stream, double *rmax_out, int32_t *iters_out, double *hist,
                                 char *err, size_t err_len);

/* newsolv_mg (newsolv_mg.for:2-102) in the reference's velocity form, bitwise
 * the reference in FP64:
 *
 *   calvel; calmas                      (press.for:139-200, 2-75)
 *   do iter = 1, maxcy
 *     mgkcyc on su (sup under LSM)      (mgsolver.for:224-369)
 *     p += pp over the ghost layer      (newsolv_mg.for:38-51)
 *     exchange(4); pbound               (exchange.for:2-341, press.for:77-137)
 *     calvel (u = u* - dt*alfapr*grad p [/rho_face]; exchange(1..3)); calmas
 *     exit when rmax < eps and iter > 1 (newsolv_mg.for:55)
 *
 * ustar, vstar, wstar (dom(ib)%ustar ...) are read; u, v, w and p are read
 * and written in place: calvel writes the faces isu..ieu (a face on a
 * physical non-periodic wall keeps the caller's value) and the exchanges
 * and pbound write the ghost shell exactly as the reference's copy
 * sequence does on arrays with `ghost` = pl layers (every block the same
 * width): exchange(1..3) over the layers ly = 0..pl_ex, exchange(4) over
 * ly = 0, whole ghosted planes across periodic faces, (pl+1)-deep corner
 * and edge cubes (exchange.for:2-341, exchange_bc.for:2-841).  With
 * differencing = 2, calmas is the 4th-order branch (press.for:13-36) and
 * needs ghost >= 2.  Caller arrays need ghost >= pl_ex + 1; a face array's
 * index i is the face east of (north of, above) cell i.  eps is the
 * reference's absolute criterion on
 * rmax = max|div u|*dx*dy*dz (the eps of control.cin), not mf3d_mg_cfg.eps.
 * Under LSM (variable_coef) mf3d_mg_set_density must have received the
 * cell density with its ghost layer; newsolv_mg.for:12-14 rebuilds the
 * coefficients every call, so call it before each solve whose density
 * changed.  su (may be NULL) gets the last calmas right-hand side su (sup
 * under LSM) over the interior.  rmax_hist (maxcy+1 doubles, may be NULL)
 * gets rmax before the first cycle and after each cycle.  Return codes as
 * mf3d_mg_solve (rmax_out = the last rmax).  The mixed precision runs
 * the V-cycle in FP32; p, u, v, w and su stay FP64. */
int mf3d_mg_newsolv(mf3d_mg *mg, const double *const *ustar, const double *const *vstar,
                    const double *const *wstar, double *const *u, double *const *v, double *const *w,
                    double *const *p, double dt, double alfapr, double eps, double *const *su,
                    double *rmax_out, int32_t *iters_out, double *rmax_hist);

/* The same with device-resident fields: device pointers to every block's
 * array with its `ghost` layers ((n0+2g)(n1+2g)(n2+2g) values, i fastest)
 * concatenated in block order, mf3d_mg_shell_cells values each (su
 * likewise, interior written; may be NULL).  Runs on `stream` (NULL = the
 * handle's own) and returns when the solve is complete.  The captured
 * outer step is kept for the last set of field pointers. */
int mf3d_mg_newsolv_device(mf3d_mg *mg, const double *ustar, const double *vstar, const double *wstar,
                           double *u, double *v, double *w, double *p, double dt, double alfapr,
                           double eps, double *su, void *stream, double *rmax_out, int32_t *iters_out,
                           double *rmax_hist);

/* Length of the one-layer ghosted device fields ((n0+2)(n1+2)(n2+2) per
 * block) of mf3d_mg_set_density_device and mf3d_mg_heaviside_device. */
int64_t mf3d_mg_ghosted_cells(const mf3d_mg *mg);

/* Length of the device fields of mf3d_mg_newsolv_device: every block's
 * array with the blocks' `ghost` layers (= mf3d_mg_ghosted_cells when
 * ghost is 1). */
int64_t mf3d_mg_shell_cells(const mf3d_mg *mg);

/* Interior cells of all blocks (length of the device-resident vectors). */
int64_t mf3d_mg_cells(const mf3d_mg *mg);

/* Device bytes held by the handle (for HBM-budgeted admission). */
int64_t mf3d_mg_device_bytes(const mf3d_mg *mg);

/* Bytes a handle with this configuration would hold, without creating it. */
int64_t mf3d_mg_estimate_bytes(const mf3d_mg_cfg *cfg, const mf3d_block_desc *blocks);

/* The ghost fill of one level -- exchangepp (with its 2:1 branches on
 * refined grids), mgbound and exchbc_mgpp -- as a table of expressions over
 * interior cells, the form the library applies on refined block grids
 * (no counterpart in the reference; for host-side checks of the exchange).
 * Host only, no device work.  Entries are 10 int32 each: destination, kind
 * (0 copy of src0; 1 0.125 x (src0 + ... + src7) summed in that order;
 * 2 ((2 src0 + 3 src1) - src2) / 4), src0 .. src7, as element offsets into
 * the level's arena (every block's (n0+2)(n1+2)(n2+2) array, i fastest,
 * concatenated in block order).  Writes at most cap entries to out (may be
 * NULL when cap is 0) and the table length to *count; MF3D_EINVAL /
 * MF3D_ENOTSUP as mf3d_mg_create would return for the configuration, with
 * the reason in err (err_len bytes, may be NULL). */
int mf3d_mg_ghost_table(const mf3d_mg_cfg *cfg, const mf3d_block_desc *blocks, int32_t level, int32_t *out,
                        int64_t cap, int64_t *count, char *err, size_t err_len);

/* Measurement hook (no counterpart in the reference; for benchmarks):
 * while enabled, every finest-level launch of the chosen kernel is
 * bracketed by HIP events on the stream it runs on.  Each cycle index is
 * then replayed from a captured graph of its own that records its own
 * events, so timing keeps graph replay (where the runtime refuses to
 * capture event records, the cycles are launched eagerly instead; see
 * mf3d_mg_graph_replay).  Cycles skipped after convergence are not
 * counted.  mf3d_mg_time_kernel selects the kernel (MF3D_K_*; 0 = off)
 * and resets the counters; mf3d_mg_kernel_time returns the launches seen
 * and their summed time. */
enum {
    MF3D_K_OFF = 0,
    MF3D_K_SMOOTH = 1,
    MF3D_K_OUTER = 2,
    MF3D_K_RESTRICT = 3,
    MF3D_K_PROLONG = 4,
    MF3D_K_HEAVISIDE = 5
};
int mf3d_mg_time_kernel(mf3d_mg *mg, int kernel);
int mf3d_mg_kernel_time(mf3d_mg *mg, int64_t *launches, double *total_ms);

/* Stop tolerance of the following solves (the eps of mf3d_mg_cfg), so one
 * resident handle serves requests of different tolerances.  No
 * counterpart in the reference, whose eps is a global of control.cin. */
int mf3d_mg_set_eps(mf3d_mg *mg, double eps);

/* 1 if the last solve replayed captured cycle graphs, 0 if it launched its
 * kernels one by one (no counterpart in the reference; for benchmarks). */
int mf3d_mg_graph_replay(const mf3d_mg *mg);

/* ------------------------------------------------------------------------
 * The reference's checkpoint, to start the solver from a real run's state.
 *
 * tecbin####.bin: one file per block, written by tecbin (post.for:496-548)
 * and read at restart by initial.for:708-748.  Fortran unformatted
 * sequential (little-endian, 4-byte record markers): the extents, pl, the
 * wall flags, then one record per cell (i fastest, then j, then k) of 19
 * FP64 values: x(i), y(j), z(k) and the 16 fields p, pm, ppm, vis, u, um,
 * uum, v, vm, vvm, w, wm, wwm, uvm, uwm, vwm. */
typedef struct {
    int32_t tt[3];  /* ttc_i, ttc_j, ttc_k: extents with the pl ghost layers */
    int32_t pl;     /* ghost layers */
    int32_t ind[3]; /* inind, jnind, knind: -1 where the block ends at a physical
                       non-periodic east / north / top face, else 0 */
} mf3d_tecbin_hdr;

/* The header alone. */
int mf3d_tecbin_header(const char *path, mf3d_tecbin_hdr *hdr);

/* The whole file.  fields (may be NULL): 16 pointers in the order above,
 * each to ttc_i*ttc_j*ttc_k doubles in the reference's own array layout
 * (column-major, i fastest); a NULL entry skips that field.  x, y, z (may
 * be NULL): ttc_i, ttc_j, ttc_k doubles.  hdr (may be NULL) receives the
 * header.  Returns MF3D_ENOENT or MF3D_EFORMAT (text via
 * mf3d_io_last_error) on an unreadable or malformed file. */
int mf3d_tecbin_read(const char *path, mf3d_tecbin_hdr *hdr, double *const *fields, double *x, double *y,
                     double *z);

/* final_ctime.dat (written flosol.for:296-299, read initial.for:692-694):
 * format '(i8,3F15.6)' over ntime, ctime, forcn, qstpn, count, ntav1_count,
 * ntav2_count, i.e. two lines. */
typedef struct {
    double ctime, forcn, qstpn, ntav1_count, ntav2_count;
    int32_t ntime, count;
} mf3d_final_ctime;
int mf3d_final_ctime_read(const char *path, mf3d_final_ctime *out);

/* Text of the last checkpoint-reader failure in this thread. */
const char *mf3d_io_last_error(void);

const char *mf3d_mg_last_error(const mf3d_mg *mg);
void mf3d_mg_destroy(mf3d_mg *mg);
int mf3d_mg_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif
