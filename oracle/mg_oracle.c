/*
 * mg_oracle.c -- CPU FP64 restatement of MultiFlow3D's pressure multigrid.
 *
 * TEST INFRASTRUCTURE ONLY.  The parity tests compare the HIP library
 * against this file, and it is the CPU baseline timed beside it; the product
 * never links it.
 *
 * It restates, in its own data layout (one ghosted array per block and level,
 * i fastest), the arithmetic and the order of operations of:
 *   coeff    mgsolver.for:2-222     coefficient hierarchy (constant and LSM)
 *   mgkcyc   mgsolver.for:224-369   V-cycle driven by the K-cycle counter
 *   mgrelax  mgsolver.for:371-617   lexicographic Gauss-Seidel (mg_itrsch = 1),
 *                                   ADI line relaxation with TDMA (2),
 *                                   red-black Gauss-Seidel (3)
 *   mgrestr  mgsolver.for:619-761   residual + 8-cell average restriction
 *   mgcorr   mgsolver.for:763-905   trilinear prolongation + correction
 *   mgbound  mgsolver.for:907-992   Neumann ghost copy
 *   exchangepp / exchbc_mgpp        exchangepp.for:43-430,3997-4763,4768-5383
 *                                   (same-resolution, same-rank branches)
 * Every floating-point expression keeps the reference's operand order
 * (Fortran evaluates a+b+c left to right; the reference is built without
 * FMA contraction), so results are bitwise comparable with the compiled
 * reference.  Build with -ffp-contract=off.
 *
 * Built with -fopenmp (liborc_mg_omp.so, the multi-core CPU baseline) the
 * red-black colour passes, residuals, restriction, prolongation and the
 * outer update run their plane loops in parallel.  Each of those loops
 * writes disjoint cells from values no other iteration writes, and every
 * reduction stays serial, so the results are bitwise those of the serial
 * build.
 *
 * Refined blocks (rdiv > 1, localparameters.for:26-56): a block with
 * rdiv = 2^r has r more levels of its own; at a global level beyond its own
 * count it repeats its coarsest grid, and its restriction and prolongation
 * become injection (mgsolver.for:672-680,719-732,807-819).  Across a face
 * whose two blocks differ in rdv at a level, the face pass of exchangepp
 * sends 0.125 x the 8 fine cells a coarse ghost covers (fine to coarse,
 * exchangepp.for:163-172,1110-1120) and 2 x the coarse cell each of the 4
 * fine ghosts over it lies against (coarse to fine, :173-188); the coarse
 * side stores the interior rows of its ghost face, the fine side
 * (received + 3 phi_1 - phi_2) / 4 (:4046-4053,4257-4262).  The reference's
 * fine side also rewrites the edge rows of that ghost face, from receive
 * buffer entries no sender filled (left over from earlier exchanges); here
 * they are left as they are (orc_set_poison writes NaN there instead, to
 * show whether a result depends on them).  Scope: constant coefficients,
 * a periodic axis only where a block wraps onto itself (no partner block:
 * exchbc_mgpp has no 2:1 branch), face neighbours within a factor 2 in
 * rdiv, and no two
 * blocks that touch along an edge or at a corner only (the reference's
 * corner and edge messages, exchangepp.for:433-1093, have 2:1 branches of
 * their own).
 *
 * Scope notes (what is not restated): the explicit corner/edge messages of
 * exchangepp.for:433-1093: on a conforming block grid they carry the same
 * values the three full-face passes already deliver (each face pass copies
 * whole ghosted planes, so x -> y -> z propagates edges and corners).
 * Multi-block fixtures from the compiled reference pin this.
 */
#include "mg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { FW = 0, FE = 1, FS = 2, FN = 3, FB = 4, FT = 5 };

struct orc_mg {
    int nb, L, lsm, nu_pre, nu_post, periodic;
    int lin;            /* ngrid_input: levels of an rdiv = 1 block */
    int poison;         /* write NaN into the 2:1 ghost rows no sender fills */
    int refined;        /* some block has rdiv > 1 */
    int *ngrid;         /* [b] the block's own level count (dom(ib)%ngrid) */
    int itrsch;         /* mg_itrsch: 1 lexicographic GS, 2 ADI-TDMA, 3 red-black GS */
    double *ta, *tc;    /* TDMA recurrence arrays a(), c() of mgsolver.for:381 */
    orc_block *blk;
    int32_t (*dims)[3]; /* [b*L + l] interior dims */
    double **phi;       /* [b*L + l] ghosted */
    double **rhs;       /* [b*L + l] interior */
    double **cof;       /* [b*L + l] 7 interior arrays: aw ae as an ab at ap */
    double **dens;      /* [b*L + l] interior (lsm only) */
    double **bufm;      /* [b] staging for the minus-side ghost plane */
    double **bufp;      /* [b] staging for the plus-side ghost plane */
    double **pf;        /* [b] ghosted level-0 pressure used by orc_solve */
};

static inline size_t gsz(const int32_t *n) {
    return (size_t)(n[0] + 2) * (size_t)(n[1] + 2) * (size_t)(n[2] + 2);
}
static inline size_t isz(const int32_t *n) {
    return (size_t)n[0] * (size_t)n[1] * (size_t)n[2];
}
/* ghosted index, coordinates 0..n+1 */
static inline size_t gi(const int32_t *n, int i, int j, int k) {
    return (size_t)i + (size_t)(n[0] + 2) * ((size_t)j + (size_t)(n[1] + 2) * (size_t)k);
}
/* interior index, coordinates 1..n */
static inline size_t ci(const int32_t *n, int i, int j, int k) {
    return (size_t)(i - 1) + (size_t)n[0] * ((size_t)(j - 1) + (size_t)n[1] * (size_t)(k - 1));
}

#define DIMS(mg, b, l) ((mg)->dims[(size_t)(b) * (mg)->L + (l)])
#define SLOT(mg, b, l) ((size_t)(b) * (mg)->L + (l))

static int ilog2(int v) {
    int r = 0;
    while ((1 << r) < v) ++r;
    return r;
}
static int rdiv_of(const orc_block *b) { return b->rdiv > 0 ? b->rdiv : 1; }

/* The grid level l of block b runs on (0-based: min(l, ngrid - 1)). */
static int grid_level(const orc_mg *mg, int b, int l) {
    return l < mg->ngrid[b] ? l : mg->ngrid[b] - 1;
}
/* Level l of block b repeats level l-1's grid (glevel > dom(ib)%ngrid). */
static int repeats(const orc_mg *mg, int b, int l) { return l >= mg->ngrid[b]; }
/* rdv(dom, glevel) of localparameters.for:40-54, glevel = l + 1. */
static int rdv_at(const orc_mg *mg, int b, int l) {
    if (l + 1 > mg->ngrid[b]) return 1;
    if (l + 1 <= mg->lin) return rdiv_of(&mg->blk[b]);
    return rdv_at(mg, b, l - 1) / 2;
}

/* Refined block sets this restatement covers (see the header). */
static int refined_ok(int nb, const orc_block *blk) {
    for (int b = 0; b < nb; ++b) {
        int r = rdiv_of(&blk[b]);
        if (r & (r - 1)) return 0;
        for (int f = 0; f < 6; ++f) {
            /* a periodic face only as a lone block's wrap onto itself */
            if (blk[b].bc[f] == 5 && (blk[b].per[f] >= 0 || blk[b].nbr[f ^ 1] >= 0 || blk[b].nbr[f] >= 0))
                return 0;
            const int o = blk[b].nbr[f];
            if (o < 0) continue;
            const int ro = rdiv_of(&blk[o]);
            if (ro != r && ro != 2 * r && 2 * ro != r) return 0;
            for (int g = 0; g < 6; ++g) /* no edge neighbour (through two faces) */
                if (g / 2 != f / 2 && blk[o].nbr[g] >= 0) return 0;
        }
    }
    return 1;
}

orc_mg *orc_create(int nblocks, const orc_block *blocks, int levels, int lsm,
                   int nu_pre, int nu_post) {
    if (nblocks <= 0 || levels <= 0) return NULL;
    int maxr = 1, refined = 0;
    for (int b = 0; b < nblocks; ++b) {
        const int r = rdiv_of(&blocks[b]);
        if (r > maxr) maxr = r;
        if (r > 1) refined = 1;
    }
    if (refined && (lsm || !refined_ok(nblocks, blocks))) return NULL;
    const int lin = levels - ilog2(maxr); /* ngrd_gl = ngrid_input + log2(max rdiv) */
    if (lin < 1) return NULL;
    for (int b = 0; b < nblocks; ++b) {
        const int own = lin + ilog2(rdiv_of(&blocks[b]));
        for (int d = 0; d < 3; ++d) {
            int n = blocks[b].n[d];
            if (n <= 0 || (n % (1 << (own - 1))) != 0) return NULL;
        }
    }
    orc_mg *mg = (orc_mg *)calloc(1, sizeof(orc_mg));
    mg->nb = nblocks;
    mg->L = levels;
    mg->lin = lin;
    mg->refined = refined;
    mg->ngrid = malloc(sizeof(int) * nblocks);
    for (int b = 0; b < nblocks; ++b) mg->ngrid[b] = lin + ilog2(rdiv_of(&blocks[b]));
    mg->lsm = lsm;
    mg->nu_pre = nu_pre;
    mg->nu_post = nu_post;
    mg->blk = (orc_block *)malloc(sizeof(orc_block) * nblocks);
    memcpy(mg->blk, blocks, sizeof(orc_block) * nblocks);
    size_t ns = (size_t)nblocks * levels;
    mg->dims = malloc(sizeof(*mg->dims) * ns);
    mg->phi = calloc(ns, sizeof(double *));
    mg->rhs = calloc(ns, sizeof(double *));
    mg->cof = calloc(ns, sizeof(double *));
    mg->dens = calloc(ns, sizeof(double *));
    mg->bufm = calloc(nblocks, sizeof(double *));
    mg->bufp = calloc(nblocks, sizeof(double *));
    mg->pf = calloc(nblocks, sizeof(double *));
    mg->periodic = 0;
    mg->itrsch = 3;
    size_t face_max = 0, line_max = 0;
    for (int b = 0; b < nblocks; ++b) {
        for (int f = 0; f < 6; ++f)
            if (blocks[b].bc[f] == 5) mg->periodic = 1; /* alloc_dom.for:1156-1163 */
        for (int l = 0; l < levels; ++l) {
            int32_t *n = DIMS(mg, b, l);
            for (int d = 0; d < 3; ++d) n[d] = blocks[b].n[d] >> grid_level(mg, b, l);
            mg->phi[SLOT(mg, b, l)] = calloc(gsz(n), sizeof(double));
            mg->rhs[SLOT(mg, b, l)] = calloc(isz(n), sizeof(double));
            mg->cof[SLOT(mg, b, l)] = calloc(7 * isz(n), sizeof(double));
            if (lsm) mg->dens[SLOT(mg, b, l)] = calloc(isz(n), sizeof(double));
        }
        const int32_t *n0 = DIMS(mg, b, 0);
        size_t f01 = (size_t)(n0[0] + 2) * (n0[1] + 2), f02 = (size_t)(n0[0] + 2) * (n0[2] + 2),
               f12 = (size_t)(n0[1] + 2) * (n0[2] + 2);
        size_t fm = f01 > f02 ? f01 : f02;
        fm = fm > f12 ? fm : f12;
        if (fm > face_max) face_max = fm;
        mg->pf[b] = calloc(gsz(n0), sizeof(double));
        for (int d = 0; d < 3; ++d)
            if ((size_t)n0[d] + 2 > line_max) line_max = (size_t)n0[d] + 2;
    }
    mg->ta = calloc(line_max, sizeof(double));
    mg->tc = calloc(line_max, sizeof(double));
    for (int b = 0; b < nblocks; ++b) {
        mg->bufm[b] = calloc(face_max, sizeof(double));
        mg->bufp[b] = calloc(face_max, sizeof(double));
    }
    return mg;
}

void orc_destroy(orc_mg *mg) {
    if (!mg) return;
    free(mg->ta);
    free(mg->tc);
    size_t ns = (size_t)mg->nb * mg->L;
    for (size_t s = 0; s < ns; ++s) {
        free(mg->phi[s]);
        free(mg->rhs[s]);
        free(mg->cof[s]);
        free(mg->dens[s]);
    }
    for (int b = 0; b < mg->nb; ++b) {
        free(mg->bufm[b]);
        free(mg->bufp[b]);
        free(mg->pf[b]);
    }
    free(mg->phi); free(mg->rhs); free(mg->cof); free(mg->dens);
    free(mg->bufm); free(mg->bufp); free(mg->pf);
    free(mg->dims); free(mg->blk); free(mg->ngrid); free(mg);
}

void orc_set_poison(orc_mg *mg, int on) { mg->poison = on != 0; }

int orc_dims(const orc_mg *mg, int block, int level, int32_t out_n[3]) {
    if (block < 0 || block >= mg->nb || level < 0 || level >= mg->L) return -1;
    const int32_t *n = mg->dims[(size_t)block * mg->L + level];
    out_n[0] = n[0]; out_n[1] = n[1]; out_n[2] = n[2];
    return 0;
}
double *orc_phi(orc_mg *mg, int b, int l) { return mg->phi[SLOT(mg, b, l)]; }
double *orc_rhs(orc_mg *mg, int b, int l) { return mg->rhs[SLOT(mg, b, l)]; }
double *orc_cof(orc_mg *mg, int b, int l, int which) {
    return mg->cof[SLOT(mg, b, l)] + (size_t)which * isz(DIMS(mg, b, l));
}
double *orc_dens(orc_mg *mg, int b, int l) { return mg->dens[SLOT(mg, b, l)]; }

int orc_set_density(orc_mg *mg, const double *dens) {
    if (!mg->lsm) return -1;
    size_t off = 0;
    for (int b = 0; b < mg->nb; ++b) {
        size_t n = isz(DIMS(mg, b, 0));
        memcpy(mg->dens[SLOT(mg, b, 0)], dens + off, n * sizeof(double));
        off += n;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* coeff (mgsolver.for:2-222)                                               */
/* ------------------------------------------------------------------------ */
void orc_coeff(orc_mg *mg) {
    for (int b = 0; b < mg->nb; ++b) {
        const orc_block *bk = &mg->blk[b];
        for (int l = 0; l < mg->L; ++l) {
            const int32_t *n = DIMS(mg, b, l);
            size_t N = isz(n);
            memset(mg->phi[SLOT(mg, b, l)], 0, gsz(n) * sizeof(double)); /* :60-67 */
            double *c = mg->cof[SLOT(mg, b, l)];
            double *aw = c, *ae = c + N, *as = c + 2 * N, *an = c + 3 * N, *ab = c + 4 * N,
                   *at = c + 5 * N, *ap = c + 6 * N;
            double *rh = mg->rhs[SLOT(mg, b, l)];
            double *dl = mg->lsm ? mg->dens[SLOT(mg, b, l)] : NULL;
            if (mg->lsm && l > 0) { /* :81-102, 8-cell average in the ij1..ij8 order */
                const int32_t *nf = DIMS(mg, b, l - 1);
                const double *df = mg->dens[SLOT(mg, b, l - 1)];
                for (int k = 1; k <= n[2]; ++k)
                    for (int j = 1; j <= n[1]; ++j)
                        for (int i = 1; i <= n[0]; ++i) {
                            int fi = 2 * i - 1, fj = 2 * j - 1, fk = 2 * k - 1;
                            double s = df[ci(nf, fi, fj, fk)];
                            s = s + df[ci(nf, fi, fj + 1, fk)];
                            s = s + df[ci(nf, fi + 1, fj, fk)];
                            s = s + df[ci(nf, fi + 1, fj + 1, fk)];
                            s = s + df[ci(nf, fi, fj, fk + 1)];
                            s = s + df[ci(nf, fi, fj + 1, fk + 1)];
                            s = s + df[ci(nf, fi + 1, fj, fk + 1)];
                            s = s + df[ci(nf, fi + 1, fj + 1, fk + 1)];
                            dl[ci(n, i, j, k)] = 0.125 * s;
                        }
            }
            /* constant branch :184-189: 1/h_l^2 with h_l = h * 2^gl, gl the
             * block's own grid level (:41-53) */
            const int gl = grid_level(mg, b, l);
            double ndx = bk->h[0] * (double)(1 << gl), ndy = bk->h[1] * (double)(1 << gl),
                   ndz = bk->h[2] * (double)(1 << gl);
            double cxx = 1.0 / (ndx * ndx), cyy = 1.0 / (ndy * ndy), czz = 1.0 / (ndz * ndz);
            /* LSM branch :170-181 uses the finest dx, dy, dz on every level */
            double gx = bk->h[1] * bk->h[2] / bk->h[0];
            double gy = bk->h[0] * bk->h[2] / bk->h[1];
            double gz = bk->h[0] * bk->h[1] / bk->h[2];
            for (int k = 1; k <= n[2]; ++k)
                for (int j = 1; j <= n[1]; ++j)
                    for (int i = 1; i <= n[0]; ++i) {
                        size_t q = ci(n, i, j, k);
                        if (mg->lsm) {
                            double d0 = dl[q];
                            double dim = (i == 1) ? d0 : 0.5 * (d0 + dl[q - 1]);
                            double dip = (i == n[0]) ? d0 : 0.5 * (d0 + dl[q + 1]);
                            double djm = (j == 1) ? d0 : 0.5 * (d0 + dl[q - n[0]]);
                            double djp = (j == n[1]) ? d0 : 0.5 * (d0 + dl[q + n[0]]);
                            size_t sk = (size_t)n[0] * n[1];
                            double dkm = (k == 1) ? d0 : 0.5 * (d0 + dl[q - sk]);
                            double dkp = (k == n[2]) ? d0 : 0.5 * (d0 + dl[q + sk]);
                            aw[q] = gx / dim;
                            ae[q] = gx / dip;
                            as[q] = gy / djm;
                            an[q] = gy / djp;
                            ab[q] = gz / dkm;
                            at[q] = gz / dkp;
                        } else {
                            aw[q] = cxx; ae[q] = cxx;
                            as[q] = cyy; an[q] = cyy;
                            ab[q] = czz; at[q] = czz;
                        }
                        /* :192-203: zero the coefficient on a physical non-periodic face */
                        if (bk->nbr[FW] < 0 && bk->bc[FW] != 5 && i == 1) aw[q] = 0.0;
                        if (bk->nbr[FE] < 0 && bk->bc[FE] != 5 && i == n[0]) ae[q] = 0.0;
                        if (bk->nbr[FS] < 0 && bk->bc[FS] != 5 && j == 1) as[q] = 0.0;
                        if (bk->nbr[FN] < 0 && bk->bc[FN] != 5 && j == n[1]) an[q] = 0.0;
                        if (bk->nbr[FB] < 0 && bk->bc[FB] != 5 && k == 1) ab[q] = 0.0;
                        if (bk->nbr[FT] < 0 && bk->bc[FT] != 5 && k == n[2]) at[q] = 0.0;
                        if (mg->lsm == 2 && l > 0) {
                            /* face-coefficient restriction (the library's
                             * MF3D_COARSE_FACE; not in the reference): each coarse
                             * face = sum of the 4 fine faces it covers / 16, the
                             * lower tangential axis fastest */
                            const int32_t *nf = DIMS(mg, b, l - 1);
                            const double *cf = mg->cof[SLOT(mg, b, l - 1)];
                            const size_t NF = isz(nf);
                            const int fi = 2 * i - 1, fj = 2 * j - 1, fk = 2 * k - 1;
#define CF(w, a, bb, c) cf[(size_t)(w) * NF + ci(nf, a, bb, c)]
                            aw[q] = (CF(0, fi, fj, fk) + CF(0, fi, fj + 1, fk) + CF(0, fi, fj, fk + 1) +
                                     CF(0, fi, fj + 1, fk + 1)) / 16.0;
                            ae[q] = (CF(1, fi + 1, fj, fk) + CF(1, fi + 1, fj + 1, fk) + CF(1, fi + 1, fj, fk + 1) +
                                     CF(1, fi + 1, fj + 1, fk + 1)) / 16.0;
                            as[q] = (CF(2, fi, fj, fk) + CF(2, fi + 1, fj, fk) + CF(2, fi, fj, fk + 1) +
                                     CF(2, fi + 1, fj, fk + 1)) / 16.0;
                            an[q] = (CF(3, fi, fj + 1, fk) + CF(3, fi + 1, fj + 1, fk) + CF(3, fi, fj + 1, fk + 1) +
                                     CF(3, fi + 1, fj + 1, fk + 1)) / 16.0;
                            ab[q] = (CF(4, fi, fj, fk) + CF(4, fi + 1, fj, fk) + CF(4, fi, fj + 1, fk) +
                                     CF(4, fi + 1, fj + 1, fk)) / 16.0;
                            at[q] = (CF(5, fi, fj, fk + 1) + CF(5, fi + 1, fj, fk + 1) + CF(5, fi, fj + 1, fk + 1) +
                                     CF(5, fi + 1, fj + 1, fk + 1)) / 16.0;
#undef CF
                        }
                        /* :205-208 */
                        ap[q] = -1.0 * (aw[q] + ae[q] + as[q] + an[q] + ab[q] + at[q]);
                        rh[q] = 0.0; /* :210 */
                    }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* ghost planes                                                             */
/* ------------------------------------------------------------------------ */
/* Copy the full ghosted plane coord_d == c of a ghosted array to/from buf. */
static void plane_get(const double *a, const int32_t *n, int d, int c, double *buf) {
    size_t m = 0;
    if (d == 0) {
        for (int k = 0; k <= n[2] + 1; ++k)
            for (int j = 0; j <= n[1] + 1; ++j) buf[m++] = a[gi(n, c, j, k)];
    } else if (d == 1) {
        for (int k = 0; k <= n[2] + 1; ++k)
            for (int i = 0; i <= n[0] + 1; ++i) buf[m++] = a[gi(n, i, c, k)];
    } else {
        for (int j = 0; j <= n[1] + 1; ++j)
            for (int i = 0; i <= n[0] + 1; ++i) buf[m++] = a[gi(n, i, j, c)];
    }
}
static void plane_set(double *a, const int32_t *n, int d, int c, const double *buf) {
    size_t m = 0;
    if (d == 0) {
        for (int k = 0; k <= n[2] + 1; ++k)
            for (int j = 0; j <= n[1] + 1; ++j) a[gi(n, c, j, k)] = buf[m++];
    } else if (d == 1) {
        for (int k = 0; k <= n[2] + 1; ++k)
            for (int i = 0; i <= n[0] + 1; ++i) a[gi(n, i, c, k)] = buf[m++];
    } else {
        for (int j = 0; j <= n[1] + 1; ++j)
            for (int i = 0; i <= n[0] + 1; ++i) a[gi(n, i, j, c)] = buf[m++];
    }
}
static void plane_copy(double *a, const int32_t *n, int d, int dst, int src) {
    if (d == 0) {
        for (int k = 0; k <= n[2] + 1; ++k)
            for (int j = 0; j <= n[1] + 1; ++j) a[gi(n, dst, j, k)] = a[gi(n, src, j, k)];
    } else if (d == 1) {
        for (int k = 0; k <= n[2] + 1; ++k)
            for (int i = 0; i <= n[0] + 1; ++i) a[gi(n, i, dst, k)] = a[gi(n, i, src, k)];
    } else {
        for (int j = 0; j <= n[1] + 1; ++j)
            for (int i = 0; i <= n[0] + 1; ++i) a[gi(n, i, j, dst)] = a[gi(n, i, j, src)];
    }
}

/* mgbound (mgsolver.for:907-992): Neumann copy, faces in order W,E,S,N,B,T. */
static void bound_fields(orc_mg *mg, int l, double **f) {
    for (int b = 0; b < mg->nb; ++b) {
        const orc_block *bk = &mg->blk[b];
        const int32_t *n = DIMS(mg, b, l);
        for (int d = 0; d < 3; ++d) {
            if (bk->nbr[2 * d] < 0 && bk->bc[2 * d] != 5) plane_copy(f[b], n, d, 0, 1);
            if (bk->nbr[2 * d + 1] < 0 && bk->bc[2 * d + 1] != 5)
                plane_copy(f[b], n, d, n[d] + 1, n[d]);
        }
    }
}

/* Ghosted index of the cell with normal coordinate c on axis d and
 * tangential coordinates (t1, t2): (j, k) for d = 0, (i, k) for 1, (i, j)
 * for 2 -- the plane_get order, which also indexes the staging buffers. */
static inline size_t face_gi(const int32_t *n, int d, int c, int t1, int t2) {
    return d == 0 ? gi(n, c, t1, t2) : (d == 1 ? gi(n, t1, c, t2) : gi(n, t1, t2, c));
}
static inline size_t face_m(const int32_t *n, int d, int t1, int t2) {
    return (size_t)t1 + (size_t)(n[d == 0 ? 1 : 0] + 2) * (size_t)t2;
}

/* The 8 fine cells of a 2:1 fine-to-coarse ghost in the summation order of
 * exchangepp.for:163-172 (x), :210-220 (y), :245-255 (z): offsets along
 * (normal, t1, t2) from the low corner. */
static const int avg8_order[3][8][3] = {
    {{0, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}, {1, 0, 0}, {1, 1, 0}, {1, 0, 1}, {1, 1, 1}},
    {{0, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}, {1, 0, 0}, {1, 1, 0}, {1, 0, 1}, {1, 1, 1}},
    {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 1, 1}, {1, 0, 0}, {1, 0, 1}, {1, 1, 0}, {1, 1, 1}}};

/* Pack of block b (sender) for its face neighbour o on side hi (0: o is the
 * prev neighbour, so b's first interior planes go out; 1: next) into buf,
 * laid out on o's face. */
static void pack_face(orc_mg *mg, int l, double **f, int b, int o, int d, int hi, double *buf) {
    const int32_t *n = DIMS(mg, b, l), *no = DIMS(mg, o, l);
    const int rb = rdv_at(mg, b, l), ro = rdv_at(mg, o, l);
    if (rb == ro) { /* :151-161 */
        plane_get(f[b], n, d, hi ? n[d] : 1, buf);
        return;
    }
    const int a1 = d == 0 ? 1 : 0, a2 = d == 2 ? 1 : 2;
    if (rb > ro) { /* fine to coarse: 0.125 x 8 cells, o's interior rows */
        const int c0 = hi ? n[d] - 1 : 1;
        for (int t2 = 1; t2 <= no[a2]; ++t2)
            for (int t1 = 1; t1 <= no[a1]; ++t1) {
                double sum = 0.0;
                for (int q = 0; q < 8; ++q) {
                    const int *w = avg8_order[d][q];
                    const double v = f[b][face_gi(n, d, c0 + w[0], 2 * t1 - 1 + w[1], 2 * t2 - 1 + w[2])];
                    sum = q == 0 ? v : sum + v;
                }
                buf[face_m(no, d, t1, t2)] = 0.125 * sum;
            }
        return;
    }
    /* coarse to fine: 2 x the face cell to the 4 fine positions over it */
    const int c = hi ? n[d] : 1;
    for (int t2 = 1; t2 <= n[a2]; ++t2)
        for (int t1 = 1; t1 <= n[a1]; ++t1) {
            const double v = 2.0 * f[b][face_gi(n, d, c, t1, t2)];
            for (int e2 = 0; e2 < 2; ++e2)
                for (int e1 = 0; e1 < 2; ++e1) buf[face_m(no, d, 2 * t1 - 1 + e1, 2 * t2 - 1 + e2)] = v;
        }
}

/* Unpack into block b's ghost plane on side hi from neighbour o's buf. */
static void unpack_face(orc_mg *mg, int l, double **f, int b, int o, int d, int hi, const double *buf) {
    const int32_t *n = DIMS(mg, b, l);
    const int rb = rdv_at(mg, b, l), ro = rdv_at(mg, o, l);
    const int g = hi ? n[d] + 1 : 0;
    if (rb == ro) { /* :4038-4044 */
        plane_set(f[b], n, d, g, buf);
        return;
    }
    const int a1 = d == 0 ? 1 : 0, a2 = d == 2 ? 1 : 2;
    const int c1 = hi ? n[d] : 1, c2 = hi ? n[d] - 1 : 2;
    for (int t2 = 0; t2 <= n[a2] + 1; ++t2)
        for (int t1 = 0; t1 <= n[a1] + 1; ++t1) {
            const int inner = t1 >= 1 && t1 <= n[a1] && t2 >= 1 && t2 <= n[a2];
            const size_t q = face_gi(n, d, g, t1, t2);
            if (!inner) { /* rows no sender fills */
                if (rb > ro && mg->poison) f[b][q] = NAN;
                continue;
            }
            const double r = buf[face_m(n, d, t1, t2)];
            if (rb < ro) /* coarse side: :4054-4061 */
                f[b][q] = r;
            else /* fine side: :4046-4053 */
                f[b][q] = ((r + 3.0 * f[b][face_gi(n, d, c1, t1, t2)]) - f[b][face_gi(n, d, c2, t1, t2)]) / 4.0;
        }
}

/* exchangepp (same-rank branches) followed by mgbound and exchbc_mgpp;
 * f[b] is block b's ghosted array at level l. */
static void exchange_fields(orc_mg *mg, int l, double **f) {
    /* face passes x, y, z: all blocks pack, then all blocks unpack
     * (exchangepp.for:43-430,1094-1240 pack, 3997-4345 unpack) */
    for (int d = 0; d < 3; ++d) {
        for (int b = 0; b < mg->nb; ++b) {
            const orc_block *bk = &mg->blk[b];
            int prv = bk->nbr[2 * d], nxt = bk->nbr[2 * d + 1];
            if (prv >= 0) pack_face(mg, l, f, b, prv, d, 0, mg->bufp[prv]);
            if (nxt >= 0) pack_face(mg, l, f, b, nxt, d, 1, mg->bufm[nxt]);
        }
        for (int b = 0; b < mg->nb; ++b) {
            const orc_block *bk = &mg->blk[b];
            if (bk->nbr[2 * d] >= 0) unpack_face(mg, l, f, b, bk->nbr[2 * d], d, 0, mg->bufm[b]);
            if (bk->nbr[2 * d + 1] >= 0) unpack_face(mg, l, f, b, bk->nbr[2 * d + 1], d, 1, mg->bufp[b]);
        }
    }
    bound_fields(mg, l, f); /* exchangepp.for:4761 */
    if (!mg->periodic) return;
    /* exchbc_mgpp (exchangepp.for:4768-5383), directions x, y, z */
    for (int d = 0; d < 3; ++d) {
        for (int b = 0; b < mg->nb; ++b) {
            const orc_block *bk = &mg->blk[b];
            if (!(bk->bc[2 * d] == 5 || bk->bc[2 * d + 1] == 5)) continue;
            const int32_t *n = DIMS(mg, b, l);
            int prv = bk->nbr[2 * d], nxt = bk->nbr[2 * d + 1];
            if (prv < 0) {
                if (nxt < 0) plane_copy(f[b], n, d, 0, n[d]);
                else if (bk->per[2 * d] >= 0) plane_get(f[b], n, d, 1, mg->bufp[bk->per[2 * d]]);
            }
            if (nxt < 0) {
                if (prv < 0) plane_copy(f[b], n, d, n[d] + 1, 1);
                else if (bk->per[2 * d + 1] >= 0)
                    plane_get(f[b], n, d, n[d], mg->bufm[bk->per[2 * d + 1]]);
            }
        }
        for (int b = 0; b < mg->nb; ++b) {
            const orc_block *bk = &mg->blk[b];
            if (!(bk->bc[2 * d] == 5 || bk->bc[2 * d + 1] == 5)) continue;
            const int32_t *n = DIMS(mg, b, l);
            int prv = bk->nbr[2 * d], nxt = bk->nbr[2 * d + 1];
            if (prv < 0 && nxt >= 0 && bk->per[2 * d] >= 0) plane_set(f[b], n, d, 0, mg->bufm[b]);
            if (nxt < 0 && prv >= 0 && bk->per[2 * d + 1] >= 0)
                plane_set(f[b], n, d, n[d] + 1, mg->bufp[b]);
        }
    }
}

static double **level_phis(orc_mg *mg, int l, double **tmp) {
    for (int b = 0; b < mg->nb; ++b) tmp[b] = mg->phi[SLOT(mg, b, l)];
    return tmp;
}

void orc_exchange(orc_mg *mg, int l) {
    double **f = malloc(sizeof(double *) * mg->nb);
    exchange_fields(mg, l, level_phis(mg, l, f));
    free(f);
}

void orc_bound(orc_mg *mg, int l) {
    double **f = malloc(sizeof(double *) * mg->nb);
    bound_fields(mg, l, level_phis(mg, l, f));
    free(f);
}

/* ------------------------------------------------------------------------ */
/* mgrelax, mg_itrsch = 3 (mgsolver.for:547-611)                            */
/* ------------------------------------------------------------------------ */
static void rb_colour(orc_mg *mg, int b, int l, int parity) {
    const int32_t *n = DIMS(mg, b, l);
    size_t N = isz(n);
    double *p = mg->phi[SLOT(mg, b, l)];
    const double *c = mg->cof[SLOT(mg, b, l)];
    const double *rh = mg->rhs[SLOT(mg, b, l)];
    size_t sy = (size_t)(n[0] + 2), sz = sy * (size_t)(n[1] + 2);
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= n[2]; ++k)
        for (int j = 1; j <= n[1]; ++j)
            for (int i = 1; i <= n[0]; ++i) {
                /* first colour: mod(i+j+k,2)==0 on the reference's 1-based
                 * ghosted indices == (i+j+k) odd on interior indices here */
                if (((i + j + k) & 1) != parity) continue;
                size_t q = ci(n, i, j, k), g = gi(n, i, j, k);
                double s = c[q] * p[g - 1];
                s = s + c[N + q] * p[g + 1];
                s = s + c[2 * N + q] * p[g - sy];
                s = s + c[3 * N + q] * p[g + sy];
                s = s + c[4 * N + q] * p[g - sz];
                s = s + c[5 * N + q] * p[g + sz];
                p[g] = (rh[q] - s) / c[6 * N + q];
            }
}

/* mgrelax, mg_itrsch = 1 (mgsolver.for:403-428): Gauss-Seidel in the loop
 * order k, j, i (i innermost), updating in place, so the W, S and B
 * neighbours are this sweep's values and E, N, T the previous ones. */
static void lex_block(orc_mg *mg, int b, int l) {
    const int32_t *n = DIMS(mg, b, l);
    size_t N = isz(n);
    double *p = mg->phi[SLOT(mg, b, l)];
    const double *c = mg->cof[SLOT(mg, b, l)];
    const double *rh = mg->rhs[SLOT(mg, b, l)];
    size_t sy = (size_t)(n[0] + 2), sz = sy * (size_t)(n[1] + 2);
    for (int k = 1; k <= n[2]; ++k)
        for (int j = 1; j <= n[1]; ++j)
            for (int i = 1; i <= n[0]; ++i) {
                size_t q = ci(n, i, j, k), g = gi(n, i, j, k);
                double s = c[q] * p[g - 1];
                s = s + c[N + q] * p[g + 1];
                s = s + c[2 * N + q] * p[g - sy];
                s = s + c[3 * N + q] * p[g + sy];
                s = s + c[4 * N + q] * p[g - sz];
                s = s + c[5 * N + q] * p[g + sz];
                p[g] = (rh[q] - s) / c[6 * N + q];
            }
}

/* mgrelax, mg_itrsch = 2 (mgsolver.for:430-545): TDMA along j lines (lines
 * in the order k, i), then along i lines (k, j), then along k lines (i, j),
 * each line solved in place with the Thomas algorithm; the other
 * directions' neighbours are read as they stand, so lines earlier in the
 * pass contribute their new values.  a(1) = 0 and c(1) = the ghost before
 * the line; apr = 1/(ap - a_lo*a(m-1)), a(m) = a_hi*apr,
 * c(m) = (su - the four cross terms in the reference's order - a_lo*c(m-1))*apr,
 * then phi(m) = c(m) - a(m)*phi(m+1) from the ghost after the line down. */
static void adi_block(orc_mg *mg, int b, int l) {
    const int32_t *n = DIMS(mg, b, l);
    size_t N = isz(n);
    double *p = mg->phi[SLOT(mg, b, l)];
    const double *c = mg->cof[SLOT(mg, b, l)];
    const double *rh = mg->rhs[SLOT(mg, b, l)];
    double *a = mg->ta, *cc = mg->tc;
    const double *aw = c, *ae = c + N, *as = c + 2 * N, *an = c + 3 * N, *ab = c + 4 * N, *at = c + 5 * N,
                 *ap = c + 6 * N;
    size_t sy = (size_t)(n[0] + 2), sz = sy * (size_t)(n[1] + 2);
    a[0] = 0.0;
    /* j lines (:433-463) */
    for (int k = 1; k <= n[2]; ++k)
        for (int i = 1; i <= n[0]; ++i) {
            cc[0] = p[gi(n, i, 0, k)];
            for (int j = 1; j <= n[1]; ++j) {
                size_t q = ci(n, i, j, k), g = gi(n, i, j, k);
                double apr = 1.0 / (ap[q] - as[q] * a[j - 1]);
                a[j] = an[q] * apr;
                double t = rh[q] - ae[q] * p[g + 1];
                t = t - aw[q] * p[g - 1];
                t = t - at[q] * p[g + sz];
                t = t - ab[q] * p[g - sz];
                t = t - as[q] * cc[j - 1];
                cc[j] = t * apr;
            }
            for (int j = n[1]; j >= 1; --j) {
                size_t g = gi(n, i, j, k);
                p[g] = cc[j] - a[j] * p[g + sy];
            }
        }
    /* i lines (:466-499) */
    for (int k = 1; k <= n[2]; ++k)
        for (int j = 1; j <= n[1]; ++j) {
            cc[0] = p[gi(n, 0, j, k)];
            for (int i = 1; i <= n[0]; ++i) {
                size_t q = ci(n, i, j, k), g = gi(n, i, j, k);
                double apr = 1.0 / (ap[q] - aw[q] * a[i - 1]);
                a[i] = ae[q] * apr;
                double t = rh[q] - an[q] * p[g + sy];
                t = t - as[q] * p[g - sy];
                t = t - at[q] * p[g + sz];
                t = t - ab[q] * p[g - sz];
                t = t - aw[q] * cc[i - 1];
                cc[i] = t * apr;
            }
            for (int i = n[0]; i >= 1; --i) {
                size_t g = gi(n, i, j, k);
                p[g] = cc[i] - a[i] * p[g + 1];
            }
        }
    /* k lines (:502-535) */
    for (int i = 1; i <= n[0]; ++i)
        for (int j = 1; j <= n[1]; ++j) {
            cc[0] = p[gi(n, i, j, 0)];
            for (int k = 1; k <= n[2]; ++k) {
                size_t q = ci(n, i, j, k), g = gi(n, i, j, k);
                double apr = 1.0 / (ap[q] - ab[q] * a[k - 1]);
                a[k] = at[q] * apr;
                double t = rh[q] - an[q] * p[g + sy];
                t = t - as[q] * p[g - sy];
                t = t - ae[q] * p[g + 1];
                t = t - aw[q] * p[g - 1];
                t = t - ab[q] * cc[k - 1];
                cc[k] = t * apr;
            }
            for (int k = n[2]; k >= 1; --k) {
                size_t g = gi(n, i, j, k);
                p[g] = cc[k] - a[k] * p[g + sz];
            }
        }
}

int orc_set_smoother(orc_mg *mg, int itrsch) {
    if (itrsch < 1 || itrsch > 3) return -1;
    mg->itrsch = itrsch;
    return 0;
}

void orc_relax(orc_mg *mg, int l, int sweeps) {
    double **f = malloc(sizeof(double *) * mg->nb);
    level_phis(mg, l, f);
    for (int it = 0; it < sweeps; ++it) {
        for (int b = 0; b < mg->nb; ++b) {
            if (mg->itrsch == 1) {
                lex_block(mg, b, l);
            } else if (mg->itrsch == 2) {
                adi_block(mg, b, l);
            } else {
                rb_colour(mg, b, l, 1);
                rb_colour(mg, b, l, 0);
            }
        }
        bound_fields(mg, l, f);   /* mgsolver.for:610 */
        exchange_fields(mg, l, f); /* mgsolver.for:611 */
    }
    free(f);
}

/* r = rhs - (aw W + ae E + as S + an N + ab B + at T + ap P)  (:702-709) */
static void residual_block(orc_mg *mg, int b, int l, const double *p, const double *rh, double *r) {
    const int32_t *n = DIMS(mg, b, l);
    size_t N = isz(n);
    const double *c = mg->cof[SLOT(mg, b, l)];
    size_t sy = (size_t)(n[0] + 2), sz = sy * (size_t)(n[1] + 2);
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= n[2]; ++k)
        for (int j = 1; j <= n[1]; ++j)
            for (int i = 1; i <= n[0]; ++i) {
                size_t q = ci(n, i, j, k), g = gi(n, i, j, k);
                double s = c[q] * p[g - 1];
                s = s + c[N + q] * p[g + 1];
                s = s + c[2 * N + q] * p[g - sy];
                s = s + c[3 * N + q] * p[g + sy];
                s = s + c[4 * N + q] * p[g - sz];
                s = s + c[5 * N + q] * p[g + sz];
                s = s + c[6 * N + q] * p[g];
                r[q] = rh[q] - s;
            }
}

void orc_residual(orc_mg *mg, int l, double *r) {
    size_t off = 0;
    for (int b = 0; b < mg->nb; ++b) {
        residual_block(mg, b, l, mg->phi[SLOT(mg, b, l)], mg->rhs[SLOT(mg, b, l)], r + off);
        off += isz(DIMS(mg, b, l));
    }
}

/* mgrestr (mgsolver.for:619-761) */
void orc_restrict(orc_mg *mg, int lc) {
    int lf = lc - 1;
    for (int b = 0; b < mg->nb; ++b) {
        const int32_t *nc = DIMS(mg, b, lc), *nf = DIMS(mg, b, lf);
        memset(mg->phi[SLOT(mg, b, lc)], 0, gsz(nc) * sizeof(double)); /* :650-657 */
        double *rc = mg->rhs[SLOT(mg, b, lc)];
        memset(rc, 0, isz(nc) * sizeof(double)); /* :659-667 */
        double *resf = malloc(isz(nf) * sizeof(double));
        residual_block(mg, b, lf, mg->phi[SLOT(mg, b, lf)], mg->rhs[SLOT(mg, b, lf)], resf);
        if (repeats(mg, b, lc)) { /* injection on the repeated grid (:719-732) */
            memcpy(rc, resf, isz(nc) * sizeof(double));
            free(resf);
            continue;
        }
#pragma omp parallel for schedule(static)
        for (int k = 1; k <= nc[2]; ++k)
            for (int j = 1; j <= nc[1]; ++j)
                for (int i = 1; i <= nc[0]; ++i) {
                    int x0 = 2 * i - 1, y0 = 2 * j - 1, z0 = 2 * k - 1;
                    double s = resf[ci(nf, x0, y0, z0)]; /* order of :743-747 */
                    s = s + resf[ci(nf, x0 + 1, y0, z0)];
                    s = s + resf[ci(nf, x0, y0 + 1, z0)];
                    s = s + resf[ci(nf, x0 + 1, y0 + 1, z0)];
                    s = s + resf[ci(nf, x0, y0, z0 + 1)];
                    s = s + resf[ci(nf, x0 + 1, y0, z0 + 1)];
                    s = s + resf[ci(nf, x0, y0 + 1, z0 + 1)];
                    s = s + resf[ci(nf, x0 + 1, y0 + 1, z0 + 1)];
                    rc[ci(nc, i, j, k)] = 0.125 * s;
                }
        free(resf);
    }
}

/* mgcorr (mgsolver.for:763-905): cell-centred trilinear prolongation added
 * to every fine point including the ghost layer, then mgbound only. */
void orc_correct(orc_mg *mg, int lf) {
    int lc = lf + 1;
    /* coarse stencil terms T0..T7 as offsets (di,dj,dk), in the reference's
     * summation order: c, c+j, c+k, c+jk, c+i, c+ij, c+ik, c+ijk */
    static const int toff[8][3] = {{0, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1},
                                   {1, 0, 0}, {1, 1, 0}, {1, 0, 1}, {1, 1, 1}};
    for (int b = 0; b < mg->nb; ++b) {
        const int32_t *nc = DIMS(mg, b, lc), *nf = DIMS(mg, b, lf);
        const double *pc = mg->phi[SLOT(mg, b, lc)];
        double *pfi = mg->phi[SLOT(mg, b, lf)];
        if (repeats(mg, b, lc)) { /* injection, ghosts included (:807-819) */
            for (size_t g = 0; g < gsz(nf); ++g) pfi[g] = pfi[g] + pc[g];
            continue;
        }
        /* coarse plane kc writes fine planes 2kc and 2kc+1 only */
#pragma omp parallel for schedule(static)
        for (int kc = 0; kc <= nc[2]; ++kc)
            for (int ic = 0; ic <= nc[0]; ++ic)
                for (int jc = 0; jc <= nc[1]; ++jc) {
                    double cv[8];
                    for (int t = 0; t < 8; ++t)
                        cv[t] = pc[gi(nc, ic + toff[t][0], jc + toff[t][1], kc + toff[t][2])];
                    for (int ch = 0; ch < 8; ++ch) {
                        int fi = (ch >> 2) & 1, fj = (ch >> 1) & 1, fk = ch & 1;
                        double s = 0.0;
                        for (int t = 0; t < 8; ++t) {
                            double w = (toff[t][0] == fi ? 3.0 : 1.0) * (toff[t][1] == fj ? 3.0 : 1.0) *
                                       (toff[t][2] == fk ? 3.0 : 1.0);
                            s = (t == 0) ? w * cv[0] : s + w * cv[t];
                        }
                        size_t g = gi(nf, 2 * ic + fi, 2 * jc + fj, 2 * kc + fk);
                        pfi[g] = pfi[g] + s / 64.0;
                    }
                }
        (void)nf;
    }
    orc_bound(mg, lf); /* mgsolver.for:902 */
}

/* ------------------------------------------------------------------------ */
/* mgkcyc (mgsolver.for:224-369), kcycle = 1                                */
/* ------------------------------------------------------------------------ */
void orc_vcycle(orc_mg *mg, const double *rhs, double *pp) {
    size_t oi = 0, og = 0;
    for (int b = 0; b < mg->nb; ++b) { /* :236-265 load */
        const int32_t *n = DIMS(mg, b, 0);
        memcpy(mg->phi[SLOT(mg, b, 0)], pp + og, gsz(n) * sizeof(double));
        memcpy(mg->rhs[SLOT(mg, b, 0)], rhs + oi, isz(n) * sizeof(double));
        oi += isz(n);
        og += gsz(n);
    }
    const int L = mg->L;
    orc_relax(mg, 0, mg->nu_pre); /* :273 */
    if (L > 1) {
        orc_restrict(mg, 1); /* :281 */
        int *kount = calloc(L, sizeof(int));
        int l = 1, nrel = mg->nu_pre;
        while (l != 0) { /* label 10 */
            kount[l] += 1;
            orc_relax(mg, l, nrel);
            if (kount[l] == 2) {
                orc_correct(mg, l - 1);
                kount[l] = 0;
                l -= 1;
                nrel = mg->nu_post;
            } else if (l < L - 1) {
                orc_restrict(mg, l + 1);
                l += 1;
                nrel = mg->nu_pre;
            } else {
                orc_relax(mg, l, mg->nu_post); /* :332 */
                orc_correct(mg, l - 1);
                nrel = mg->nu_post;
                l = L - 2;
            }
        }
        free(kount);
    }
    orc_relax(mg, 0, mg->nu_post); /* :349 */
    og = 0;
    for (int b = 0; b < mg->nb; ++b) { /* :351-364 store */
        const int32_t *n = DIMS(mg, b, 0);
        memcpy(pp + og, mg->phi[SLOT(mg, b, 0)], gsz(n) * sizeof(double));
        og += gsz(n);
    }
}

/* ------------------------------------------------------------------------ */
/* Poisson-form outer loop (newsolv_mg.for:19-58 with calvel+calmas replaced
 * by the identity div(u* - dt grad p)/dt = f - A p)                          */
/* ------------------------------------------------------------------------ */
static void norms(orc_mg *mg, const double *r, double *l2, double *rmax) {
    double s = 0.0, m = 0.0;
    size_t off = 0;
    for (int b = 0; b < mg->nb; ++b) {
        const int32_t *n = DIMS(mg, b, 0);
        const orc_block *bk = &mg->blk[b];
        size_t N = isz(n);
        for (size_t q = 0; q < N; ++q) {
            double v = r[off + q];
            s = s + v * v;
            /* press.for:60-62: rmax = max |div * dx*dy*dz|; the LSM RHS
             * already carries the cell volume (press.for:51-53) */
            double a = mg->lsm ? fabs(v) : fabs(v * bk->h[0] * bk->h[1] * bk->h[2]);
            if (a > m) m = a;
        }
        off += N;
    }
    *l2 = sqrt(s);
    *rmax = m;
}

int orc_solve(orc_mg *mg, const double *f, double *p, double tol, int maxcy, int min_cycles,
              double *hist_rho, double *hist_rmax) {
    size_t ntot = 0, gtot = 0;
    for (int b = 0; b < mg->nb; ++b) {
        ntot += isz(DIMS(mg, b, 0));
        gtot += gsz(DIMS(mg, b, 0));
    }
    double *r = malloc(ntot * sizeof(double));
    double *pp = malloc(gtot * sizeof(double));
    for (int b = 0; b < mg->nb; ++b) memset(mg->pf[b], 0, gsz(DIMS(mg, b, 0)) * sizeof(double));
    size_t off = 0;
    for (int b = 0; b < mg->nb; ++b) {
        residual_block(mg, b, 0, mg->pf[b], f + off, r + off);
        off += isz(DIMS(mg, b, 0));
    }
    double nf, rm, nr;
    norms(mg, f, &nf, &rm);
    norms(mg, r, &nr, &rm);
    if (nf == 0.0) nf = 1.0;
    hist_rho[0] = nr / nf;
    hist_rmax[0] = rm;
    int k = 0;
    for (k = 1; k <= maxcy; ++k) {
        memset(pp, 0, gtot * sizeof(double)); /* newsolv_mg.for:19-21,44 */
        orc_vcycle(mg, r, pp);
        size_t og = 0;
        for (int b = 0; b < mg->nb; ++b) { /* newsolv_mg.for:38-48, ppref == 0 */
            const int32_t *n = DIMS(mg, b, 0);
#pragma omp parallel for schedule(static)
            for (int kk = 1; kk <= n[2]; ++kk)
                for (int j = 1; j <= n[1]; ++j)
                    for (int i = 1; i <= n[0]; ++i) {
                        size_t g = gi(n, i, j, kk);
                        mg->pf[b][g] = mg->pf[b][g] + pp[og + g];
                    }
            og += gsz(n);
        }
        exchange_fields(mg, 0, mg->pf); /* exchange(4) + pbound */
        off = 0;
        for (int b = 0; b < mg->nb; ++b) {
            residual_block(mg, b, 0, mg->pf[b], f + off, r + off);
            off += isz(DIMS(mg, b, 0));
        }
        norms(mg, r, &nr, &rm);
        hist_rho[k] = nr / nf;
        hist_rmax[k] = rm;
        if (hist_rho[k] <= tol && k >= min_cycles) break;
    }
    if (k > maxcy) k = maxcy;
    off = 0;
    for (int b = 0; b < mg->nb; ++b) {
        const int32_t *n = DIMS(mg, b, 0);
        for (int kk = 1; kk <= n[2]; ++kk)
            for (int j = 1; j <= n[1]; ++j)
                for (int i = 1; i <= n[0]; ++i) p[off + ci(n, i, j, kk)] = mg->pf[b][gi(n, i, j, kk)];
        off += isz(n);
    }
    free(r);
    free(pp);
    return k;
}

/* ------------------------------------------------------------------------ */
/* GCR(1) with the V-cycle as preconditioner: the library's krylov =         */
/* MF3D_OUTER_GCR outer loop (not in the reference).  Sums run in plain      */
/* order, so results agree with the library's tree sums to rounding only.   */
/* ------------------------------------------------------------------------ */
static double dot_int(orc_mg *mg, const double *a_i, const double *b_i) {
    double s = 0.0;
    for (size_t q = 0, n = 0; n < (size_t)mg->nb; ++n) {
        const size_t m = isz(DIMS(mg, (int)n, 0));
        for (size_t e = 0; e < m; ++e, ++q) s += a_i[q] * b_i[q];
    }
    return s;
}

/* out = A u on level-0 interiors (u ghosted with mapped ghosts) */
static void apply_a(orc_mg *mg, const double *u_g, const double *zero, double *out) {
    size_t og = 0, oi = 0;
    for (int b = 0; b < mg->nb; ++b) {
        residual_block(mg, b, 0, u_g + og, zero + oi, out + oi); /* 0 - A u */
        const size_t m = isz(DIMS(mg, b, 0));
        for (size_t q = 0; q < m; ++q) out[oi + q] = -out[oi + q];
        og += gsz(DIMS(mg, b, 0));
        oi += m;
    }
}

int orc_solve_gcr(orc_mg *mg, const double *f, double *p, double tol, int maxcy, int min_cycles,
                  double *hist_rho, double *hist_rmax) {
    size_t ntot = 0, gtot = 0;
    for (int b = 0; b < mg->nb; ++b) {
        ntot += isz(DIMS(mg, b, 0));
        gtot += gsz(DIMS(mg, b, 0));
    }
    double *r = malloc(ntot * sizeof(double)), *w = calloc(ntot, sizeof(double));
    double *q = malloc(ntot * sizeof(double)), *zero = calloc(ntot, sizeof(double));
    double *z = malloc(gtot * sizeof(double)), *d = calloc(gtot, sizeof(double));
    for (int b = 0; b < mg->nb; ++b) memset(mg->pf[b], 0, gsz(DIMS(mg, b, 0)) * sizeof(double));
    size_t off = 0;
    for (int b = 0; b < mg->nb; ++b) {
        residual_block(mg, b, 0, mg->pf[b], f + off, r + off);
        off += isz(DIMS(mg, b, 0));
    }
    double nf, rm, nr;
    norms(mg, f, &nf, &rm);
    norms(mg, r, &nr, &rm);
    if (nf == 0.0) nf = 1.0;
    hist_rho[0] = nr / nf;
    hist_rmax[0] = rm;
    double gamma = 1.0;
    int k = 0;
    for (k = 1; k <= maxcy; ++k) {
        memset(z, 0, gtot * sizeof(double));
        orc_vcycle(mg, r, z);
        apply_a(mg, z, zero, q);
        const double beta = k > 1 && gamma != 0.0 ? dot_int(mg, q, w) / gamma : 0.0;
        for (size_t g = 0; g < gtot; ++g) d[g] = z[g] + (-beta) * d[g];
        apply_a(mg, d, zero, w);
        gamma = dot_int(mg, w, w);
        const double alpha = gamma != 0.0 ? dot_int(mg, r, w) / gamma : 0.0;
        size_t og = 0;
        for (int b = 0; b < mg->nb; ++b) {
            const int32_t *n = DIMS(mg, b, 0);
            for (int kk = 1; kk <= n[2]; ++kk)
                for (int j = 1; j <= n[1]; ++j)
                    for (int i = 1; i <= n[0]; ++i) {
                        size_t g = gi(n, i, j, kk);
                        mg->pf[b][g] = mg->pf[b][g] + alpha * d[og + g];
                    }
            og += gsz(n);
        }
        exchange_fields(mg, 0, mg->pf);
        off = 0;
        for (int b = 0; b < mg->nb; ++b) {
            residual_block(mg, b, 0, mg->pf[b], f + off, r + off);
            off += isz(DIMS(mg, b, 0));
        }
        norms(mg, r, &nr, &rm);
        hist_rho[k] = nr / nf;
        hist_rmax[k] = rm;
        if (hist_rho[k] <= tol && k >= min_cycles) break;
    }
    if (k > maxcy) k = maxcy;
    off = 0;
    for (int b = 0; b < mg->nb; ++b) {
        const int32_t *n = DIMS(mg, b, 0);
        for (int kk = 1; kk <= n[2]; ++kk)
            for (int j = 1; j <= n[1]; ++j)
                for (int i = 1; i <= n[0]; ++i) p[off + ci(n, i, j, kk)] = mg->pf[b][gi(n, i, j, kk)];
        off += isz(n);
    }
    free(r);
    free(w);
    free(q);
    free(zero);
    free(z);
    free(d);
    return k;
}
