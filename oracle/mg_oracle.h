/*
 * mg_oracle.h -- CPU FP64 restatement of MultiFlow3D's pressure multigrid
 * (newsolv_mg / mgkcyc / coeff).  TEST INFRASTRUCTURE ONLY: what the parity
 * tests compare against, and the CPU baseline.  Nothing in the product library links it.
 *
 * Levels are 0-based here (level 0 = finest = the reference's glevel 1).
 * Arrays handed in and out are concatenated over blocks in block order; a
 * block's interior is stored i fastest, then j, then k.  "Ghosted" arrays
 * carry one ghost layer on every side: (n0+2)*(n1+2)*(n2+2) values.
 */
#ifndef MF3D_MG_ORACLE_H
#define MF3D_MG_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* One block: the MG-relevant fields of the reference's type multidom
 * (module_multidata.for:16-28,36,116). Face order is W,E,S,N,B,T
 * (iprev,inext,jprev,jnext,kprev,knext). */
typedef struct {
    int32_t n[3];     /* interior cells along i, j, k at the finest level */
    double  h[3];     /* dx, dy, dz */
    int32_t nbr[6];   /* face neighbour block index, -1 if none */
    int32_t per[6];   /* periodic partner block index (per_ip..per_kn), -1 if none */
    int32_t bc[6];    /* boundary codes; 5 = periodic, anything else Neumann for p */
    int32_t rdiv;     /* refinement ratio rdiv (infodom.cin), a power of 2; 0 = 1.
                         n is at the block's own resolution, h its own spacing */
} orc_block;

typedef struct orc_mg orc_mg;

/* levels: ngrd_gl = ngrid_input + log2(max rdiv).  Returns NULL for a block
 * set outside the covered scope (extents not divisible by 2^(own levels - 1);
 * refined blocks with lsm, a periodic face, face neighbours more than a
 * factor 2 apart in rdiv, or an edge neighbour).
 * lsm: 0 constant coefficients; 1 the reference's density hierarchy
 * (mgsolver.for:69-181); 2 the same finest level with coarse operators by
 * face-coefficient restriction (the library's MF3D_COARSE_FACE). */
orc_mg *orc_create(int nblocks, const orc_block *blocks, int levels, int lsm,
                   int nu_pre, int nu_post);
void    orc_destroy(orc_mg *mg);
/* Smoother (mg_itrsch): 1 lexicographic GS, 2 ADI-TDMA, 3 red-black GS
 * (the default).  Returns 0, or -1 for another value. */
int     orc_set_smoother(orc_mg *mg, int itrsch);
/* Level-0 cell density (interior, concatenated).  Only for lsm handles. */
int     orc_set_density(orc_mg *mg, const double *dens);
/* Coefficient hierarchy (mgsolver.for:2-222). */
void    orc_coeff(orc_mg *mg);
/* One V-cycle (mgsolver.for:224-369): rhs is interior, pp is ghosted in/out. */
void    orc_vcycle(orc_mg *mg, const double *rhs, double *pp);
/* Poisson-form outer loop: p = 0; repeat { V-cycle on r = f - A p; p += e }
 * until ||r||/||f|| <= tol (and at least min_cycles cycles) or maxcy cycles.
 * p is interior in/out (overwritten with the result, started from 0).
 * hist_rho / hist_rmax receive maxcy+1 entries (entry 0 is the start).
 * Returns the number of cycles run. */
int     orc_solve(orc_mg *mg, const double *f, double *p, double tol, int maxcy,
                  int min_cycles, double *hist_rho, double *hist_rmax);
/* The same problem with GCR(1) around the V-cycle (the library's
 * MF3D_OUTER_GCR): beta = (A z, w_prev)/(w_prev, w_prev), d = z - beta d_prev,
 * w = A d, alpha = (r, w)/(w, w), p += alpha d, r = f - A p. */
int     orc_solve_gcr(orc_mg *mg, const double *f, double *p, double tol, int maxcy,
                      int min_cycles, double *hist_rho, double *hist_rmax);

/* Refined blocks: write NaN into the ghost rows of a 2:1 fine-side ghost
 * face that no sender fills (the reference rewrites them from stale receive
 * buffer entries), so a result that depends on them shows NaN. */
void    orc_set_poison(orc_mg *mg, int on);

/* Unit-level access for kernel tests. */
int     orc_dims(const orc_mg *mg, int block, int level, int32_t out_n[3]);
double *orc_phi(orc_mg *mg, int block, int level);            /* ghosted */
double *orc_rhs(orc_mg *mg, int block, int level);            /* interior */
double *orc_cof(orc_mg *mg, int block, int level, int which); /* 0..6 = aw ae as an ab at ap */
double *orc_dens(orc_mg *mg, int block, int level);           /* interior, lsm only */
void    orc_relax(orc_mg *mg, int level, int sweeps);
void    orc_restrict(orc_mg *mg, int coarse_level);
void    orc_correct(orc_mg *mg, int fine_level);
void    orc_exchange(orc_mg *mg, int level);
void    orc_bound(orc_mg *mg, int level);
/* r = rhs - A phi at a level (interior, concatenated). */
void    orc_residual(orc_mg *mg, int level, double *r);

#ifdef __cplusplus
}
#endif
#endif
