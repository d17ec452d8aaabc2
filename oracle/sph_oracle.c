/*
 * sph_oracle.c -- CPU restatement of the reference's per-step particle loop.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this file; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it, as the checker.  It follows the reference's *serial* FP64 path
 * (nthreads == 1; the OpenMP variants accumulate in FP32 slabs, e.g.
 * m_particles.f90:2441) operation by operation:
 *   - pairs are enumerated and scattered in the reference's list order;
 *   - Fortran default-real literals are widened from FP32 exactly as the
 *     reference does (pi = 3.14159265358979323846 without a kind suffix is the
 *     FP32 value 3.1415927410125732, 2./3. is FP32 0.6666666865348816, ...);
 *   - expression evaluation order is Fortran left-to-right; x**2, x**3, x**4,
 *     x**5 with integer exponents are repeated squaring, x**r with a real
 *     exponent is pow().
 * Build with -O2 -ffp-contract=off (no FMA contraction), see oracle/Makefile.
 *
 * The reference itself is Fortran whose live compile set needs stand-ins
 * for an un-vendored library (modules `constants`, `declarations`, `ifport`,
 * externals get_unit/scount/trim_pvalu) and source edits (!DEC$ blocks), so
 * it is not built here.  This restatement is pinned instead by closed forms,
 * invariants and brute force (tests/test_oracle_*.py).
 *
 * Parity status per function is listed in the header of each block.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "sph_oracle.h"

/* FP32 literals of the reference, widened to double. */
#define PI_F   ((double)3.14159265358979323846f)   /* m_particles.f90:1483, param.f90:121 */
#define C23_F  ((double)(2.0f / 3.0f))               /* 2./3. m_particles.f90:1508; 2.e0/3.e0 :2902 */
#define EPS_YIELD_F ((double)1.0e-5f)                /* plastic_or_not :3492 */
#define EPSP_SCALE_F ((double)0.000005f)             /* plastic_flow_rule3 :3756 */

static double sq(double a) { return a * a; }
static double cube(double a) { return (a * a) * a; }           /* a**3 */
static double pow4(double a) { double t = a * a; return t * t; } /* a**4 */
static double pow5(double a) { return pow4(a) * a; }           /* a**5 */

/* ------------------------------------------------------------------------ */
/* get_scale_k, m_particles.f90:915-934                                     */
/* ------------------------------------------------------------------------ */
int oracle_scale_k(int skf)
{
    switch (skf) {
    case 1: return 2;
    case 2: return 3;
    case 3: return 3;
    case 4: return 2;
    default: return -1;
    }
}

/* ------------------------------------------------------------------------ */
/* kernel, m_particles.f90:1469-1602 (dim = 2).                              */
/* Pinned by: closed forms at sample q, unit integral, finite-difference      */
/* gradient, symmetry (tests/test_oracle_kernel.py).                         */
/* ------------------------------------------------------------------------ */
void oracle_kernel(int skf, double r, const double dx[2], double hsml,
                   double *w, double dwdx[2])
{
    double q = r / hsml, factor;
    int d;
    *w = 0.0;
    dwdx[0] = 0.0; dwdx[1] = 0.0;
    if (skf == 1) {
        factor = 15.0 / (7.0 * PI_F * hsml * hsml);
        if (q >= 0.0 && q <= 1.0) {
            *w = factor * (C23_F - q * q + cube(q) / 2.0);
            for (d = 0; d < 2; d++)
                dwdx[d] = factor * (-2.0 + 1.5 * q) / sq(hsml) * dx[d];
        } else if (q > 1.0 && q <= 2.0) {
            *w = factor * 1.0 / 6.0 * cube(2.0 - q);
            for (d = 0; d < 2; d++)
                dwdx[d] = -(factor * 1.0 / 6.0 * 3.0 * sq(2.0 - q) / hsml * (dx[d] / r));
        }
    } else if (skf == 2) {
        factor = 1.0 / (sq(hsml) * pow(PI_F, 1.0));
        if (q >= 0.0 && q <= 3.0) {
            *w = factor * exp(-q * q);
            for (d = 0; d < 2; d++)
                dwdx[d] = *w * (-2.0 * dx[d] / hsml / hsml);
        }
    } else if (skf == 3) {
        factor = 7.0 / (478.0 * PI_F * hsml * hsml);
        if (q >= 0.0 && q <= 1.0) {
            *w = factor * (pow5(3.0 - q) - 6.0 * pow5(2.0 - q) + 15.0 * pow5(1.0 - q));
            for (d = 0; d < 2; d++)
                dwdx[d] = factor * ((-120.0 + 120.0 * q - 50.0 * sq(q)) / sq(hsml) * dx[d]);
        } else if (q > 1.0 && q <= 2.0) {
            *w = factor * (pow5(3.0 - q) - 6.0 * pow5(2.0 - q));
            for (d = 0; d < 2; d++)
                dwdx[d] = factor * (-(5.0 * pow4(3.0 - q)) + 30.0 * pow4(2.0 - q)) / hsml * (dx[d] / r);
        } else if (q > 2.0 && q <= 3.0) {
            *w = factor * pow5(3.0 - q);
            for (d = 0; d < 2; d++)
                dwdx[d] = factor * (-(5.0 * pow4(3.0 - q))) / hsml * (dx[d] / r);
        }
    } else if (skf == 4) {
        factor = 7.0 / (4.0 * PI_F * hsml * hsml);
        if (q >= 0.0 && q <= 2.0) {
            *w = factor * (pow4(1.0 - q / 2.0) * (1.0 + 2.0 * q));
            for (d = 0; d < 2; d++)
                dwdx[d] = factor * (-5.0 + 15.0 * q / 2.0 - 15.0 * sq(q) / 4.0 + 5.0 * cube(q) / 8.0)
                          / sq(hsml) * dx[d];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* direct_find, m_particles.f90:938-1001.                                    */
/* Pinned by: numpy brute force with the identical predicate                 */
/* (tests/test_oracle_pairs.py).                                             */
/* ------------------------------------------------------------------------ */
#ifdef _OPENMP
/* Timing build (cpu_baseline) only: the pair searches split over threads by
 * i, as the reference's threaded search splits its particles
 * (link_list_omp, m_particles.f90:1269-1443; get_niac_start_end :4219-4251).
 * Chunks of consecutive i collect their pairs in the serial visiting order
 * and are concatenated in order, so the result is the serial search's pair
 * list; the checker build (no OpenMP) runs the serial loops below. */
static int pair_gen = 1;   /* bumped by every pair search: invalidates the pass schedules */

/* Pairs found for a range of i, in the serial order. */
typedef struct {
    int *pi, *pj;
    double *w, *dw;
    int n, cap;
} pair_buf;

static void pbuf_add(pair_buf *B, int i, int j, double w, const double dw[2])
{
    if (B->n == B->cap) {
        B->cap = B->cap ? 2 * B->cap : 1024;
        B->pi = (int *)realloc(B->pi, sizeof(int) * (size_t)B->cap);
        B->pj = (int *)realloc(B->pj, sizeof(int) * (size_t)B->cap);
        B->w = (double *)realloc(B->w, sizeof(double) * (size_t)B->cap);
        B->dw = (double *)realloc(B->dw, 2 * sizeof(double) * (size_t)B->cap);
    }
    B->pi[B->n] = i + 1;
    B->pj[B->n] = j + 1;
    B->w[B->n] = w;
    B->dw[2 * B->n] = dw[0];
    B->dw[2 * B->n + 1] = dw[1];
    B->n++;
}

typedef void (*pair_visit_fn)(const sphg_params *P, const sphg_phase_view *a, const sphg_phase_view *b, int i,
                              int scale_k, const void *aux, pair_buf *out);

/* The i range [0, na) in many chunks, taken dynamically; each chunk's pairs
 * are kept in order, and the chunks are concatenated in order. */
static int omp_pairs(const sphg_params *P, const sphg_phase_view *a, const sphg_phase_view *b, int na,
                     int scale_k, const void *aux, pair_visit_fn visit, sphg_phase_view *out)
{
    enum { NCH = 512 };
    pair_buf *bufs = (pair_buf *)calloc(NCH, sizeof(pair_buf));
    long long total = 0;
    int c, rc = SPHG_OK;
#pragma omp parallel for schedule(dynamic, 1)
    for (c = 0; c < NCH; c++) {
        int i, i0 = (int)((long long)na * c / NCH), i1 = (int)((long long)na * (c + 1) / NCH);
        for (i = i0; i < i1; i++) visit(P, a, b, i, scale_k, aux, &bufs[c]);
    }
    for (c = 0; c < NCH; c++) total += bufs[c].n;
    if (total > out->max_interaction) {
        out->niac = out->max_interaction;
        rc = SPHG_ERR_TOO_MANY_INTERACTIONS;
    } else {
        long long o = 0;
        for (c = 0; c < NCH; c++) {
            pair_buf *B = &bufs[c];
            if (B->n) {
                memcpy(out->pair_i + o, B->pi, sizeof(int) * (size_t)B->n);
                memcpy(out->pair_j + o, B->pj, sizeof(int) * (size_t)B->n);
                memcpy(out->w + o, B->w, sizeof(double) * (size_t)B->n);
                memcpy(out->dwdx + 2 * o, B->dw, 2 * sizeof(double) * (size_t)B->n);
            }
            o += B->n;
        }
        out->niac = (int)total;
    }
    for (c = 0; c < NCH; c++) { free(bufs[c].pi); free(bufs[c].pj); free(bufs[c].w); free(bufs[c].dw); }
    free(bufs);
    return rc;
}

/* partners j > i of direct_find for one i */
static void visit_direct(const sphg_params *P, const sphg_phase_view *ph, const sphg_phase_view *b, int i,
                         int scale_k, const void *aux, pair_buf *out)
{
    int n = ph->ntotal + ph->nvirt, j;
    (void)b; (void)aux;
    for (j = i + 1; j < n; j++) {
        double dxiac[2], driac, mhsml;
        dxiac[0] = ph->x[2 * i] - ph->x[2 * j];
        driac = dxiac[0] * dxiac[0];
        dxiac[1] = ph->x[2 * i + 1] - ph->x[2 * j + 1];
        driac = driac + dxiac[1] * dxiac[1];
        mhsml = (ph->hsml[i] + ph->hsml[j]) / 2.0;
        if (sqrt(driac) < scale_k * mhsml) {
            double ww, tdw[2];
            oracle_kernel(P->skf, sqrt(driac), dxiac, mhsml, &ww, tdw);
            pbuf_add(out, i, j, ww, tdw);
        }
    }
}

/* all soil partners j of water particle i (direct_find_2) */
static void visit_cross(const sphg_params *P, const sphg_phase_view *wa, const sphg_phase_view *so, int i,
                        int scale_k, const void *aux, pair_buf *out)
{
    int ns = so->ntotal + so->nvirt, j;
    (void)aux;
    for (j = 0; j < ns; j++) {
        double dxiac[2], driac, mhsml;
        dxiac[0] = wa->x[2 * i] - so->x[2 * j];
        driac = dxiac[0] * dxiac[0];
        dxiac[1] = wa->x[2 * i + 1] - so->x[2 * j + 1];
        driac = driac + dxiac[1] * dxiac[1];
        mhsml = (wa->hsml[i] + so->hsml[j]) / 2.0;
        if (sqrt(driac) < scale_k * mhsml) {
            double ww, tdw[2];
            oracle_kernel(P->skf, sqrt(driac), dxiac, mhsml, &ww, tdw);
            pbuf_add(out, i, j, ww, tdw);
        }
    }
}

/* countiac of the pairs of a list (both ends) */
static void count_ends(const sphg_phase_view *lst, int *ci, int ni, int *cj, int nj)
{
    int k;
    for (k = 0; k < ni; k++) ci[k] = 0;
    if (cj && cj != ci) for (k = 0; k < nj; k++) cj[k] = 0;
    for (k = 0; k < lst->niac; k++) {
        ci[lst->pair_i[k] - 1]++;
        (cj ? cj : ci)[lst->pair_j[k] - 1]++;
    }
}
#endif

int oracle_direct_find(const sphg_params *P, sphg_phase_view *ph)
{
    int n = ph->ntotal + ph->nvirt, i, j, niac = 0;
    int scale_k = oracle_scale_k(P->skf);
    if (scale_k < 0) return SPHG_ERR_BAD_KERNEL;
#ifdef _OPENMP
    {
        int rc;
        pair_gen++;
        rc = omp_pairs(P, ph, ph, n > 0 ? n - 1 : 0, scale_k, NULL, visit_direct, ph);
        if (rc == SPHG_OK) count_ends(ph, ph->countiac, n, NULL, 0);
        return rc;
    }
#endif
    for (i = 0; i < n; i++) ph->countiac[i] = 0;
    for (i = 0; i < n - 1; i++) {
        for (j = i + 1; j < n; j++) {
            double dxiac[2], driac, mhsml;
            dxiac[0] = ph->x[2 * i] - ph->x[2 * j];
            driac = dxiac[0] * dxiac[0];
            dxiac[1] = ph->x[2 * i + 1] - ph->x[2 * j + 1];
            driac = driac + dxiac[1] * dxiac[1];
            mhsml = (ph->hsml[i] + ph->hsml[j]) / 2.0;
            if (sqrt(driac) < scale_k * mhsml) {
                double r, tdw[2];
                if (niac >= ph->max_interaction) { ph->niac = niac; return SPHG_ERR_TOO_MANY_INTERACTIONS; }
                ph->pair_i[niac] = i + 1;
                ph->pair_j[niac] = j + 1;
                r = sqrt(driac);
                ph->countiac[i]++;
                ph->countiac[j]++;
                oracle_kernel(P->skf, r, dxiac, mhsml, &ph->w[niac], tdw);
                ph->dwdx[2 * niac] = tdw[0];
                ph->dwdx[2 * niac + 1] = tdw[1];
                niac++;
            }
        }
    }
    ph->niac = niac;
    return SPHG_OK;
}

/* ------------------------------------------------------------------------ */
/* direct_find_2, m_particles.f90:1004-1062: water i x soil j, all pairs     */
/* (virtual x virtual and r = 0 included), into the water pair arrays.        */
/* ------------------------------------------------------------------------ */
int oracle_direct_find_2(const sphg_params *P, sphg_phase_view *wa, sphg_phase_view *so)
{
    int nw = wa->ntotal + wa->nvirt, ns = so->ntotal + so->nvirt, i, j, niac = 0;
    int scale_k = oracle_scale_k(P->skf);
    if (scale_k < 0) return SPHG_ERR_BAD_KERNEL;
#ifdef _OPENMP
    {
        int rc;
        pair_gen++;
        rc = omp_pairs(P, wa, so, nw, scale_k, NULL, visit_cross, wa);
        if (rc == SPHG_OK) count_ends(wa, wa->countiac, nw, so->countiac, ns);
        return rc;
    }
#endif
    for (i = 0; i < nw; i++) wa->countiac[i] = 0;
    for (j = 0; j < ns; j++) so->countiac[j] = 0;
    for (i = 0; i < nw; i++) {
        for (j = 0; j < ns; j++) {
            double dxiac[2], driac, mhsml;
            dxiac[0] = wa->x[2 * i] - so->x[2 * j];
            driac = dxiac[0] * dxiac[0];
            dxiac[1] = wa->x[2 * i + 1] - so->x[2 * j + 1];
            driac = driac + dxiac[1] * dxiac[1];
            mhsml = (wa->hsml[i] + so->hsml[j]) / 2.0;
            if (sqrt(driac) < scale_k * mhsml) {
                double r, tdw[2];
                if (niac >= wa->max_interaction) { wa->niac = niac; return SPHG_ERR_TOO_MANY_INTERACTIONS; }
                wa->pair_i[niac] = i + 1;
                wa->pair_j[niac] = j + 1;
                r = sqrt(driac);
                wa->countiac[i]++;
                so->countiac[j]++;
                oracle_kernel(P->skf, r, dxiac, mhsml, &wa->w[niac], tdw);
                wa->dwdx[2 * niac] = tdw[0];
                wa->dwdx[2 * niac + 1] = tdw[1];
                niac++;
            }
        }
    }
    wa->niac = niac;
    return SPHG_OK;
}

/* ------------------------------------------------------------------------ */
/* init_grid / grid_geom / link_list, m_particles.f90:1069-1265 (dim = 2).    */
/* Pinned by: equality with direct_find when the cell size >= 2h, and by a    */
/* brute force that adds the reference's cell-reach restriction.             */
/* ------------------------------------------------------------------------ */
void oracle_init_grid(const sphg_params *P, int ntotal, double hsml,
                      int ngridx[2], int ghsmlx[2], double mingridx[2], double dgeomx[2])
{
    const double nppg = 3.0;
    double maxgridx[2];
    int d;
    maxgridx[0] = P->x_maxgeom; mingridx[0] = P->x_mingeom;
    maxgridx[1] = P->y_maxgeom; mingridx[1] = P->y_mingeom;
    dgeomx[0] = maxgridx[0] - mingridx[0];
    dgeomx[1] = maxgridx[1] - mingridx[1];
    ngridx[0] = (int)sqrt(ntotal * dgeomx[0] / (dgeomx[1] * nppg)) + 1;
    if (ngridx[0] > P->maxngx) ngridx[0] = P->maxngx;
    ngridx[1] = (int)(ngridx[0] * dgeomx[1] / dgeomx[0]) + 1;
    if (ngridx[1] > P->maxngy) ngridx[1] = P->maxngy;
    for (d = 0; d < 2; d++)
        ghsmlx[d] = (int)((double)(float)ngridx[d] * hsml / dgeomx[d]) + 1;
}

/* grid_geom: 1-based cell of a position, or 0 when out of the box. */
int oracle_grid_geom(const sphg_params *P, const int ngridx[2], const double mingridx[2],
                     const double dgeomx[2], const double x[2], int gcell[2])
{
    double maxg[2] = {P->x_maxgeom, P->y_maxgeom};
    int d;
    for (d = 0; d < 2; d++) {
        if (x[d] > maxg[d] || x[d] < mingridx[d]) return SPHG_ERR_OUT_OF_RANGE;
        gcell[d] = (int)((double)(float)ngridx[d] / dgeomx[d] * (x[d] - mingridx[d]) + 1.0);
    }
    return SPHG_OK;
}

#ifdef _OPENMP
struct ll_aux {
    const int *ngridx, *ghsmlx, *grid, *celldata, *xg;
    double hsml;
};

/* partners of i in link_list's cell walk, in its order */
static void visit_link(const sphg_params *P, const sphg_phase_view *ph, const sphg_phase_view *b, int i,
                       int scale_k, const void *aux, pair_buf *out)
{
    const struct ll_aux *A = (const struct ll_aux *)aux;
    int mn[2], mx[2], xc, yc, d, j;
    (void)b;
    for (d = 0; d < 2; d++) {
        mn[d] = A->xg[2 * i + d] - A->ghsmlx[d]; if (mn[d] < 1) mn[d] = 1;
        mx[d] = A->xg[2 * i + d] + A->ghsmlx[d]; if (mx[d] > A->ngridx[d]) mx[d] = A->ngridx[d];
    }
    for (yc = mn[1]; yc <= mx[1]; yc++) {
        for (xc = mn[0]; xc <= mx[0]; xc++) {
            j = A->grid[(yc - 1) * A->ngridx[0] + (xc - 1)];
            while (j > i + 1) {
                int jj = j - 1;
                double dx[2], dr;
                dx[0] = ph->x[2 * i] - ph->x[2 * jj];
                dr = dx[0] * dx[0];
                dx[1] = ph->x[2 * i + 1] - ph->x[2 * jj + 1];
                dr = dr + dx[1] * dx[1];
                if (sqrt(dr) < scale_k * A->hsml) {
                    double ww, tdw[2];
                    oracle_kernel(P->skf, sqrt(dr), dx, A->hsml, &ww, tdw);
                    pbuf_add(out, i, jj, ww, tdw);
                }
                j = A->celldata[jj];
            }
        }
    }
}
#endif

int oracle_link_list(const sphg_params *P, sphg_phase_view *ph)
{
    int n = ph->ntotal + ph->nvirt, i, j, niac = 0, rc;
    int scale_k = oracle_scale_k(P->skf);
    int ngridx[2], ghsmlx[2], *grid, *celldata, *xg;
    double mingridx[2], dgeomx[2], hsml;
    if (scale_k < 0) return SPHG_ERR_BAD_KERNEL;
    hsml = ph->hsml[0];
    for (i = 0; i < n; i++) ph->countiac[i] = 0;
    oracle_init_grid(P, n, hsml, ngridx, ghsmlx, mingridx, dgeomx);
    grid = (int *)calloc((size_t)ngridx[0] * ngridx[1], sizeof(int));
    celldata = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    xg = (int *)malloc(sizeof(int) * 2 * (size_t)(n > 0 ? n : 1));
    for (i = 0; i < n; i++) {
        int g[2];
        rc = oracle_grid_geom(P, ngridx, mingridx, dgeomx, &ph->x[2 * i], g);
        if (rc) { free(grid); free(celldata); free(xg); return rc; }
        xg[2 * i] = g[0]; xg[2 * i + 1] = g[1];
        /* 1-based particle numbers in the chains; 0 terminates */
        celldata[i] = grid[(g[1] - 1) * ngridx[0] + (g[0] - 1)];
        grid[(g[1] - 1) * ngridx[0] + (g[0] - 1)] = i + 1;
    }
#ifdef _OPENMP
    {
        struct ll_aux A;
        A.ngridx = ngridx; A.ghsmlx = ghsmlx; A.grid = grid; A.celldata = celldata; A.xg = xg; A.hsml = hsml;
        pair_gen++;
        rc = omp_pairs(P, ph, ph, n > 0 ? n - 1 : 0, scale_k, &A, visit_link, ph);
        if (rc == SPHG_OK) count_ends(ph, ph->countiac, n, NULL, 0);
        free(grid); free(celldata); free(xg);
        return rc;
    }
#endif
    for (i = 0; i < n - 1; i++) {
        int mn[2], mx[2], xc, yc, d;
        for (d = 0; d < 2; d++) {
            mn[d] = xg[2 * i + d] - ghsmlx[d]; if (mn[d] < 1) mn[d] = 1;
            mx[d] = xg[2 * i + d] + ghsmlx[d]; if (mx[d] > ngridx[d]) mx[d] = ngridx[d];
        }
        for (yc = mn[1]; yc <= mx[1]; yc++) {
            for (xc = mn[0]; xc <= mx[0]; xc++) {
                j = grid[(yc - 1) * ngridx[0] + (xc - 1)];
                while (j > i + 1) {
                    int jj = j - 1;
                    double dx[2], dr;
                    dx[0] = ph->x[2 * i] - ph->x[2 * jj];
                    dr = dx[0] * dx[0];
                    dx[1] = ph->x[2 * i + 1] - ph->x[2 * jj + 1];
                    dr = dr + dx[1] * dx[1];
                    if (sqrt(dr) < scale_k * hsml) {
                        double r, tdw[2];
                        if (niac >= ph->max_interaction) {
                            free(grid); free(celldata); free(xg);
                            ph->niac = niac;
                            return SPHG_ERR_TOO_MANY_INTERACTIONS;
                        }
                        ph->pair_i[niac] = i + 1;
                        ph->pair_j[niac] = jj + 1;
                        r = sqrt(dr);
                        ph->countiac[i]++;
                        ph->countiac[jj]++;
                        oracle_kernel(P->skf, r, dx, hsml, &ph->w[niac], tdw);
                        ph->dwdx[2 * niac] = tdw[0];
                        ph->dwdx[2 * niac + 1] = tdw[1];
                        niac++;
                    }
                    j = celldata[jj];
                }
            }
        }
    }
    ph->niac = niac;
    free(grid); free(celldata); free(xg);
    return SPHG_OK;
}

int oracle_find_pairs(const sphg_params *P, sphg_phase_view *ph)
{
    if (P->nnps == 1) return oracle_direct_find(P, ph);
    if (P->nnps == 2) return oracle_link_list(P, ph);
    return SPHG_ERR_UNSUPPORTED;
}

/* ------------------------------------------------------------------------ */
/* Pair loops.  All scatter over the phase's current pair list.              */
/*                                                                           */
/* PAIR_LOOP(lst, nj, body) runs body for every pair k of the list lst with  */
/* i = pair_i(k) - 1, j = pair_j(k) - 1 and side = 3: both ends are updated, */
/* pair after pair in list order, as the reference's serial loops do (the    */
/* checker build).  The threaded timing build (_OPENMP) splits the pairs     */
/* over threads by owner instead, the data-race-free form of the reference's */
/* even pair split with per-thread slabs (get_niac_start_end,                */
/* m_particles.f90:4219-4251): body runs with side = 1 over chunks of        */
/* consecutive i (the lists are i-major) updating only i, then with side = 2 */
/* over chunks of a j-sorted order updating only j.  nj: particles the       */
/* j ends index.                                                             */
/* ------------------------------------------------------------------------ */
#ifdef _OPENMP
#include <omp.h>
typedef struct {
    const int *pair_i, *pair_j;
    int niac, nj, nt, gen;
    int *kj;        /* pairs sorted by j (stable) */
    int ib[65], jb[65];
} pair_sched;

static pair_sched sched_cache[4];

static const pair_sched *sched_get(const sphg_phase_view *lst, int nj)
{
    int c, t, k, *cnt;
    pair_sched *S = NULL;
    for (c = 0; c < 4; c++) {
        pair_sched *q = &sched_cache[c];
        if (q->gen == pair_gen && q->pair_i == lst->pair_i && q->niac == lst->niac && q->nj == nj) return q;
    }
    for (c = 0; c < 4 && !S; c++)
        if (sched_cache[c].gen != pair_gen) S = &sched_cache[c];
    if (!S) S = &sched_cache[0];
    free(S->kj);
    S->pair_i = lst->pair_i; S->pair_j = lst->pair_j; S->niac = lst->niac; S->nj = nj; S->gen = pair_gen;
    S->nt = omp_get_max_threads();
    if (S->nt > 64) S->nt = 64;
    S->kj = (int *)malloc(sizeof(int) * (size_t)(S->niac > 0 ? S->niac : 1));
    cnt = (int *)calloc((size_t)nj + 1, sizeof(int));
    for (k = 0; k < S->niac; k++) cnt[lst->pair_j[k]]++;          /* pair_j is 1-based */
    for (k = 1; k <= nj; k++) cnt[k] += cnt[k - 1];
    for (k = 0; k < S->niac; k++) S->kj[cnt[lst->pair_j[k] - 1]++] = k;
    free(cnt);
    /* chunk bounds on particle boundaries of each order */
    S->ib[0] = 0; S->jb[0] = 0;
    for (t = 1; t < S->nt; t++) {
        int a = (int)((long long)S->niac * t / S->nt), b = a;
        if (a < S->ib[t - 1]) a = S->ib[t - 1];
        while (a > 0 && a < S->niac && lst->pair_i[a] == lst->pair_i[a - 1]) a++;
        S->ib[t] = a;
        if (b < S->jb[t - 1]) b = S->jb[t - 1];
        while (b > 0 && b < S->niac && lst->pair_j[S->kj[b]] == lst->pair_j[S->kj[b - 1]]) b++;
        S->jb[t] = b;
    }
    S->ib[S->nt] = S->niac; S->jb[S->nt] = S->niac;
    return S;
}

#define PAIR_LOOP(lst, nj, ...)                                                          \
    do {                                                                                 \
        const pair_sched *S_ = sched_get((lst), (nj));                                   \
        int t_;                                                                          \
        _Pragma("omp parallel for schedule(static, 1) private(i, j, k)")                 \
        for (t_ = 0; t_ < S_->nt; t_++) {                                                \
            const int side = 1;                                                          \
            int q_;                                                                      \
            for (q_ = S_->ib[t_]; q_ < S_->ib[t_ + 1]; q_++) {                           \
                k = q_; i = (lst)->pair_i[k] - 1; j = (lst)->pair_j[k] - 1;              \
                __VA_ARGS__                                                              \
            }                                                                            \
        }                                                                                \
        _Pragma("omp parallel for schedule(static, 1) private(i, j, k)")                 \
        for (t_ = 0; t_ < S_->nt; t_++) {                                                \
            const int side = 2;                                                          \
            int q_;                                                                      \
            for (q_ = S_->jb[t_]; q_ < S_->jb[t_ + 1]; q_++) {                           \
                k = S_->kj[q_]; i = (lst)->pair_i[k] - 1; j = (lst)->pair_j[k] - 1;      \
                __VA_ARGS__                                                              \
            }                                                                            \
        }                                                                                \
    } while (0)
#else
#define PAIR_LOOP(lst, nj, ...)                                                          \
    do {                                                                                 \
        const int side = 3;                                                              \
        for (k = 0; k < (lst)->niac; k++) {                                              \
            i = (lst)->pair_i[k] - 1; j = (lst)->pair_j[k] - 1;                          \
            __VA_ARGS__                                                                  \
        }                                                                                \
    } while (0)
#endif
#define I_END (side & 1)
#define J_END (side & 2)

#define PI_(k) (ph->pair_i[k] - 1)
#define PJ_(k) (ph->pair_j[k] - 1)

/* con_density, m_particles.f90:2402-2431 */
void oracle_con_density(const sphg_params *P, sphg_phase_view *ph)
{
    int n = ph->ntotal + ph->nvirt, i, j, k;
    (void)P;
    for (i = 0; i < n; i++) ph->drho[i] = 0.0;
    PAIR_LOOP(ph, n, {
        double dvx[2], vcc;
        dvx[0] = ph->vx[2 * i] - ph->vx[2 * j];
        dvx[1] = ph->vx[2 * i + 1] - ph->vx[2 * j + 1];
        vcc = dvx[0] * ph->dwdx[2 * k];
        vcc = vcc + dvx[1] * ph->dwdx[2 * k + 1];
        if (I_END) ph->drho[i] = ph->drho[i] + ph->mass[j] * vcc;
        if (J_END) ph->drho[j] = ph->drho[j] + ph->mass[i] * vcc;
    });
}

/* sum_density, m_particles.f90:2243-2298 */
void oracle_sum_density(const sphg_params *P, sphg_phase_view *ph)
{
    int n = ph->ntotal + ph->nvirt, i, j, k;
    double *wi = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double hv[2] = {0.0, 0.0}, r = 0.0, selfdens, dw[2];
    for (i = 0; i < n; i++) {
        oracle_kernel(P->skf, r, hv, ph->hsml[i], &selfdens, dw);
        wi[i] = selfdens * ph->mass[i] / ph->rho[i];
    }
    PAIR_LOOP(ph, n, {
        if (I_END) wi[i] = wi[i] + ph->mass[j] / ph->rho[j] * ph->w[k];
        if (J_END) wi[j] = wi[j] + ph->mass[i] / ph->rho[i] * ph->w[k];
    });
    for (i = 0; i < n; i++) {
        oracle_kernel(P->skf, r, hv, ph->hsml[i], &selfdens, dw);
        ph->rho[i] = selfdens * ph->mass[i];
    }
    PAIR_LOOP(ph, n, {
        if (I_END) ph->rho[i] = ph->rho[i] + ph->mass[j] * ph->w[k];
        if (J_END) ph->rho[j] = ph->rho[j] + ph->mass[i] * ph->w[k];
    });
    for (i = 0; i < n; i++) ph->rho[i] = ph->rho[i] / wi[i];
    free(wi);
}

/* delta_sph, m_particles.f90:2658-2695: df += delta-SPH diffusion of f. */
void oracle_delta_sph(const sphg_params *P, sphg_phase_view *ph, const double *f, double *df)
{
    int i, j, k;
    double delta = P->delta;
    PAIR_LOOP(ph, ph->ntotal + ph->nvirt, {
        double dx[2], rr, muv, h;
        rr = 0.0;
        dx[0] = ph->x[2 * i] - ph->x[2 * j];
        rr = rr + dx[0] * dx[0];
        dx[1] = ph->x[2 * i + 1] - ph->x[2 * j + 1];
        rr = rr + dx[1] * dx[1];
        muv = 2.0 * (f[i] - f[j]) / rr;
        h = 0.0;
        h = h + dx[0] * muv * ph->dwdx[2 * k];
        h = h + dx[1] * muv * ph->dwdx[2 * k + 1];
        if (I_END) df[i] = df[i] + delta * ph->hsml[i] * ph->c[i] * ph->mass[j] * h / ph->rho[j];
        if (J_END) df[j] = df[j] - delta * ph->hsml[j] * ph->c[j] * ph->mass[i] * h / ph->rho[i];
    });
}

/* shear_strain_rate, m_particles.f90:2855-2932 (dim = 2) */
void oracle_shear_strain_rate(const sphg_params *P, sphg_phase_view *ph)
{
    int n = ph->ntotal + ph->nvirt, i, j, k;
    (void)P;
    for (i = 0; i < n; i++) { ph->txx[i] = 0.0; ph->tyy[i] = 0.0; ph->txy[i] = 0.0; }
    PAIR_LOOP(ph, n, {
        double dvx[2], hxx, hxy, hyy, d1, d2;
        dvx[0] = ph->vx[2 * j] - ph->vx[2 * i];
        dvx[1] = ph->vx[2 * j + 1] - ph->vx[2 * i + 1];
        d1 = ph->dwdx[2 * k]; d2 = ph->dwdx[2 * k + 1];
        hxx = 2.0 * dvx[0] * d1 - dvx[1] * d2;
        hxy = dvx[0] * d2 + dvx[1] * d1;
        hyy = 2.0 * dvx[1] * d2 - dvx[0] * d1;
        hxx = C23_F * hxx;
        hyy = C23_F * hyy;
        if (I_END) ph->txx[i] = ph->txx[i] + ph->mass[j] * hxx / ph->rho[j];
        if (J_END) ph->txx[j] = ph->txx[j] + ph->mass[i] * hxx / ph->rho[i];
        if (I_END) ph->txy[i] = ph->txy[i] + ph->mass[j] * hxy / ph->rho[j];
        if (J_END) ph->txy[j] = ph->txy[j] + ph->mass[i] * hxy / ph->rho[i];
        if (I_END) ph->tyy[i] = ph->tyy[i] + ph->mass[j] * hyy / ph->rho[j];
        if (J_END) ph->tyy[j] = ph->tyy[j] + ph->mass[i] * hyy / ph->rho[i];
    });
}

/* velocity_divergence, m_particles.f90:1996-2037 */
void oracle_velocity_divergence(const sphg_params *P, sphg_phase_view *ph)
{
    int n = ph->ntotal + ph->nvirt, i, j, k;
    (void)P;
    for (i = 0; i < n; i++) ph->vcc[i] = 0.0;
    PAIR_LOOP(ph, n, {
        double dvx[2], hvcc;
        dvx[0] = ph->vx[2 * j] - ph->vx[2 * i];
        dvx[1] = ph->vx[2 * j + 1] - ph->vx[2 * i + 1];
        hvcc = dvx[0] * ph->dwdx[2 * k];
        hvcc = hvcc + dvx[1] * ph->dwdx[2 * k + 1];
        if (I_END) ph->vcc[i] = ph->vcc[i] + ph->mass[j] * hvcc / ph->rho[j];
        if (J_END) ph->vcc[j] = ph->vcc[j] + ph->mass[i] * hvcc / ph->rho[i];
    });
}

/* pressure, m_particles.f90:2133-2175 */
void oracle_pressure(const sphg_params *P, sphg_phase_view *ph)
{
    int n = ph->ntotal + ph->nvirt, i;
    if (ph->material == SPHG_WATER) {
        for (i = 0; i < n; i++)
            ph->p[i] = P->water_b * (pow(ph->rho[i] / (P->water_rho0 * ph->vof[i]), P->water_gamma) - 1.0);
        for (i = 0; i < n; i++)
            ph->c[i] = P->water_c * pow(ph->rho[i] / (P->water_rho0 * ph->vof[i]), 3.0);
    } else {
        for (i = 0; i < n; i++) ph->c[i] = P->soil_c;
    }
}

/* newtonian_fluid, m_particles.f90:2484-2500 (sigma = mu * t) */
void oracle_newtonian_fluid(const sphg_params *P, sphg_phase_view *ph)
{
    int n = ph->ntotal + ph->nvirt, i;
    for (i = 0; i < n; i++) {
        ph->sxx[i] = P->water_viscosity * ph->txx[i];
        ph->syy[i] = P->water_viscosity * ph->tyy[i];
        ph->sxy[i] = P->water_viscosity * ph->txy[i];
    }
}

static double dp_alpha1(double phi)
{
    /* tan(phi)/sqrt(9.+12.*tan(phi)**2.), m_particles.f90:3337 */
    double t = tan(phi);
    return t / sqrt(9.0 + 12.0 * (t * t));
}

static double dp_sqrt_j2(double sxx, double sxy, double syy)
{
    /* m_particles.f90:3362-3363 */
    double J2 = sxx * sxx + 2.0 * (sxy * sxy) + syy * syy + (sxx + syy) * (sxx + syy);
    return sqrt(J2 / 2.0) + 1.0e-6;
}

/* drucker_prager_failure_criterion, m_particles.f90:3323-3381 */
void oracle_drucker_prager(const sphg_params *P, sphg_phase_view *ph)
{
    int n = ph->ntotal + ph->nvirt, i, k = 0;
    double alpha1 = dp_alpha1(P->soil_phi);
    for (i = 0; i < n; i++) ph->fail[i] = 0;
    for (i = 0; i < n; i++) {
        double I1, J2;
        if (ph->p[i] < 0.0) ph->p[i] = 0.0;
        I1 = 3.0 * ph->p[i];
        J2 = dp_sqrt_j2(ph->sxx[i], ph->sxy[i], ph->syy[i]);
        if (J2 > alpha1 * I1) {
            double skale = alpha1 * I1 / J2;
            k++;
            ph->fail[i] = 1;
            ph->sxx[i] = skale * ph->sxx[i];
            ph->sxy[i] = skale * ph->sxy[i];
            ph->syy[i] = skale * ph->syy[i];
        }
    }
    ph->nfail = k;
}

/* mohr_coulomb_failure_criterion, m_particles.f90:3220-3266 */
void oracle_mohr_coulomb(const sphg_params *P, sphg_phase_view *ph)
{
    int n = ph->ntotal + ph->nvirt, i, k = 0;
    double cohesion = P->soil_cohesion, phi = P->soil_phi;
    for (i = 0; i < n; i++) ph->fail[i] = 0;
    for (i = 0; i < n; i++) {
        double tmax, yield;
        tmax = sqrt(sq((ph->sxx[i] - ph->syy[i]) / 2.0) + sq(ph->sxy[i]));
        if (tmax < (double)1.e-6f) continue;
        yield = cohesion * cos(phi) + ph->p[i] * sin(phi);
        if (yield <= 0.0) {
            yield = 0.0;
            ph->p[i] = -cohesion * pow(tan(phi), -1.0);
            ph->rho[i] = P->soil_rho0;
        }
        if (tmax > yield) {
            double skale = yield / tmax;
            k++;
            ph->fail[i] = 1;
            ph->sxx[i] = skale * ph->sxx[i];
            ph->sxy[i] = skale * ph->sxy[i];
            ph->syy[i] = skale * ph->syy[i];
        }
    }
    ph->nfail = k;
}

/* plastic_or_not, m_particles.f90:3447-3505 */
void oracle_plastic_or_not(const sphg_params *P, sphg_phase_view *ph)
{
    int n = ph->ntotal + ph->nvirt, i, k = 0;
    double alpha1 = dp_alpha1(P->soil_phi);
    for (i = 0; i < n; i++) ph->fail[i] = 0;
    for (i = 0; i < n; i++) {
        double I1 = 3.0 * ph->p[i];
        double J2 = dp_sqrt_j2(ph->sxx[i], ph->sxy[i], ph->syy[i]);
        if (fabs(J2 - alpha1 * I1) < EPS_YIELD_F) { k++; ph->fail[i] = 1; }
    }
    ph->nfail = k;
}

/* plastic_flow_rule3, m_particles.f90:3700-3762 (failed points only) and
   plastic_flow_rule, :3571-3633 (every point; the same correction). */
static void flow_rule(const sphg_params *P, sphg_phase_view *ph, int failed_only)
{
    int n = ph->ntotal + ph->nvirt, i;
    double phi = P->soil_phi, K = P->soil_k, e = P->soil_E, niu = P->soil_niu;
    double alpha = dp_alpha1(phi);
    double G = e / (2.0 * (1.0 + niu));
    for (i = 0; i < n; i++) {
        double exx, exy, eyy, sde, J2, dlambda, sxx, sxy, syy;
        if (failed_only && ph->fail[i] != 1) continue;
        sxx = ph->sxx[i]; sxy = ph->sxy[i]; syy = ph->syy[i];
        exx = ph->txx[i] / 2.0 + ph->vcc[i] / 3.0;
        exy = ph->txy[i] / 2.0;
        eyy = ph->tyy[i] / 2.0 + ph->vcc[i] / 3.0;
        sde = sxx * exx + 2.0 * sxy * exy + syy * eyy;
        J2 = (sxx * sxx + 2.0 * (sxy * sxy) + syy * syy + (sxx + syy) * (sxx + syy)) / 2.0;
        J2 = J2 + 1.0e-10;
        dlambda = (3.0 * alpha * K * ph->vcc[i] + G / sqrt(J2) * sde) / (9.0 * (alpha * alpha) * K + G);
        if (dlambda < 0.0) continue;
        ph->dsxx[i] = ph->dsxx[i] - dlambda * (G / sqrt(J2) * sxx);
        ph->dsxy[i] = ph->dsxy[i] - dlambda * (G / sqrt(J2) * sxy);
        ph->dsyy[i] = ph->dsyy[i] - dlambda * (G / sqrt(J2) * syy);
        ph->dp[i] = ph->dp[i] + 3.0 * K * alpha * dlambda;
        ph->epsilon_p[i] = ph->epsilon_p[i] + dlambda * sxy / (2.0 * sqrt(J2)) * EPSP_SCALE_F;
    }
}

void oracle_plastic_flow_rule3(const sphg_params *P, sphg_phase_view *ph) { flow_rule(P, ph, 1); }

/* plastic_flow_rule2, m_particles.f90:3636-3697: the same dlambda at every
   point, without the sign test, into the return-mapping rates dsxx2 .. dp2
   (not accumulated; epsilon_p untouched).  Pinned: equal to plastic_flow_rule's
   increments where dlambda >= 0 (tests/test_oracle_physics.py). */
void oracle_plastic_flow_rule2(const sphg_params *P, sphg_phase_view *ph)
{
    int n = ph->ntotal + ph->nvirt, i;
    double phi = P->soil_phi, K = P->soil_k, e = P->soil_E, niu = P->soil_niu;
    double alpha = dp_alpha1(phi);
    double G = e / (2.0 * (1.0 + niu));
    for (i = 0; i < n; i++) {
        double exx, exy, eyy, sde, J2, dlambda, sxx = ph->sxx[i], sxy = ph->sxy[i], syy = ph->syy[i];
        exx = ph->txx[i] / 2.0 + ph->vcc[i] / 3.0;
        exy = ph->txy[i] / 2.0;
        eyy = ph->tyy[i] / 2.0 + ph->vcc[i] / 3.0;
        sde = sxx * exx + 2.0 * sxy * exy + syy * eyy;
        J2 = (sxx * sxx + 2.0 * (sxy * sxy) + syy * syy + (sxx + syy) * (sxx + syy)) / 2.0;
        J2 = J2 + 1.0e-10;
        dlambda = (3.0 * alpha * K * ph->vcc[i] + G / sqrt(J2) * sde) / (9.0 * (alpha * alpha) * K + G);
        ph->dsxx2[i] = -dlambda * (G / sqrt(J2) * sxx);
        ph->dsxy2[i] = -dlambda * (G / sqrt(J2) * sxy);
        ph->dsyy2[i] = -dlambda * (G / sqrt(J2) * syy);
        ph->dp2[i] = 3.0 * K * alpha * dlambda;
    }
}

/* Pinned: equals plastic_flow_rule3 bit for bit where every point is flagged,
   and acts where none is (tests/test_oracle_physics.py). */
void oracle_plastic_flow_rule(const sphg_params *P, sphg_phase_view *ph) { flow_rule(P, ph, 0); }

/* Jaumann_rate, m_particles.f90:3047-3119.  `wxy` is the spin; in the
   reference it aliases the shared txy buffer and is zeroed after Hooke's law
   has read txy, which this separate array reproduces. */
void oracle_jaumann_rate(const sphg_params *P, sphg_phase_view *ph)
{
    int n = ph->ntotal + ph->nvirt, i, j, k;
    double G = P->soil_E / (2.0 * (1.0 + P->soil_niu));
    for (i = 0; i < n; i++) {
        ph->dsxx[i] = ph->dsxx[i] + G * ph->txx[i];
        ph->dsxy[i] = ph->dsxy[i] + G * ph->txy[i];
        ph->dsyy[i] = ph->dsyy[i] + G * ph->tyy[i];
    }
    for (i = 0; i < n; i++) ph->dp[i] = ph->dp[i] - P->soil_k * ph->vcc[i];
    for (i = 0; i < n; i++) ph->wxy[i] = 0.0;
    PAIR_LOOP(ph, n, {
        double dvx[2], hxy;
        dvx[0] = ph->vx[2 * j] - ph->vx[2 * i];
        dvx[1] = ph->vx[2 * j + 1] - ph->vx[2 * i + 1];
        hxy = 0.5 * (dvx[0] * ph->dwdx[2 * k + 1] - dvx[1] * ph->dwdx[2 * k]);
        if (I_END) ph->wxy[i] = ph->wxy[i] + ph->mass[j] * hxy / ph->rho[j];
        if (J_END) ph->wxy[j] = ph->wxy[j] + ph->mass[i] * hxy / ph->rho[i];
    });
    for (i = 0; i < n; i++) {
        ph->dsxx[i] = ph->dsxx[i] + 2.0 * ph->sxy[i] * ph->wxy[i];
        ph->dsxy[i] = ph->dsxy[i] - (ph->sxx[i] - ph->syy[i]) * ph->wxy[i];
        ph->dsyy[i] = ph->dsyy[i] - 2.0 * ph->sxy[i] * ph->wxy[i];
    }
}

/* art_visc, m_particles.f90:2504-2564 */
void oracle_art_visc(const sphg_params *P, sphg_phase_view *ph)
{
    int i, j, k;
    double alpha = P->alpha, beta = P->beta, etq = P->etq;
    PAIR_LOOP(ph, ph->ntotal + ph->nvirt, {
        double dvx[2], dx, vr, rr, mhsml;
        int d;
        mhsml = (ph->hsml[i] + ph->hsml[j]) / 2.0;
        vr = 0.0; rr = 0.0;
        for (d = 0; d < 2; d++) {
            dvx[d] = ph->vx[2 * i + d] - ph->vx[2 * j + d];
            dx = ph->x[2 * i + d] - ph->x[2 * j + d];
            vr = vr + dvx[d] * dx;
            rr = rr + dx * dx;
        }
        if (vr < 0.0) {
            double muv, mc, mrho, piv;
            muv = mhsml * vr / (rr + mhsml * mhsml * etq * etq);
            mc = 0.5 * (ph->c[i] + ph->c[j]);
            mrho = 0.5 * (ph->rho[i] + ph->rho[j]);
            piv = (beta * muv - alpha * mc) * muv / mrho;
            for (d = 0; d < 2; d++) {
                double h = -piv * ph->dwdx[2 * k + d];
                if (I_END) ph->dvx[2 * i + d] = ph->dvx[2 * i + d] + ph->mass[j] * h;
                if (J_END) ph->dvx[2 * j + d] = ph->dvx[2 * j + d] - ph->mass[i] * h;
            }
        }
    });
}

/* repulsive_force, m_particles.f90:1605-1642 */
void oracle_repulsive_force(const sphg_params *P, sphg_phase_view *ph)
{
    int i, j, k;
    double rr0 = ph->dspp, dd = P->dd, p1 = P->p1, p2 = P->p2;
    PAIR_LOOP(ph, ph->ntotal + ph->nvirt, {
        double dx[2], rr, f;
        int d, ii;
        if (ph->itype[i] * ph->itype[j] > 0) continue;
        rr = 0.0;
        for (d = 0; d < 2; d++) {
            dx[d] = ph->x[2 * i + d] - ph->x[2 * j + d];
            rr = rr + dx[d] * dx[d];
        }
        rr = sqrt(rr);
        if (rr > rr0) continue;
        f = (pow(rr0 / rr, p1) - pow(rr0 / rr, p2)) / (rr * rr);
        ii = i;
        if (ph->itype[i] < 0) ii = j;
        if ((ii == i && I_END) || (ii == j && J_END))
            for (d = 0; d < 2; d++)
                ph->dvx[2 * ii + d] = ph->dvx[2 * ii + d] + dd * dx[d] * f;
    });
}

/* av_vel, m_particles.f90:2761-2793 */
void oracle_av_vel(const sphg_params *P, sphg_phase_view *ph)
{
    int n = ph->ntotal + ph->nvirt, i, j, k, d;
    double epsilon = P->epsilon;
    for (i = 0; i < 2 * n; i++) ph->av[i] = 0.0;
    PAIR_LOOP(ph, n, {
        double mrho;
        int e;
        mrho = (ph->rho[i] + ph->rho[j]) / 2.0;
        for (e = 0; e < 2; e++) {
            double dvx = ph->vx[2 * i + e] - ph->vx[2 * j + e];
            if (I_END) ph->av[2 * i + e] = ph->av[2 * i + e] - ph->mass[j] * dvx / mrho * ph->w[k];
            if (J_END) ph->av[2 * j + e] = ph->av[2 * j + e] + ph->mass[i] * dvx / mrho * ph->w[k];
        }
    });
    for (i = 0; i < ph->ntotal; i++)
        for (d = 0; d < 2; d++) ph->av[2 * i + d] = epsilon * ph->av[2 * i + d];
}

/* df, m_particles.f90:1711-1736: out_i = sum m_j/rho_j (f_i+f_j) dW/dx_d */
void oracle_df(sphg_phase_view *ph, const double *f, int d, double *out)
{
    int n = ph->ntotal + ph->nvirt, i, j, k;
    for (i = 0; i < n; i++) out[i] = 0.0;
    PAIR_LOOP(ph, n, {
        double fwx;
        fwx = (f[i] + f[j]) * ph->dwdx[2 * k + d];
        if (I_END) out[i] = out[i] + ph->mass[j] / ph->rho[j] * fwx;
        if (J_END) out[j] = out[j] - ph->mass[i] / ph->rho[i] * fwx;
    });
}

/* df2, m_particles.f90:1802-1831: (1/rho_i) sum m_j (f_j-f_i) dW/dx_d */
void oracle_df2(sphg_phase_view *ph, const double *f, int d, double *out)
{
    int n = ph->ntotal + ph->nvirt, i, j, k;
    for (i = 0; i < n; i++) out[i] = 0.0;
    PAIR_LOOP(ph, n, {
        double fwx;
        fwx = (f[j] - f[i]) * ph->dwdx[2 * k + d];
        if (I_END) out[i] = out[i] + ph->mass[j] * fwx;
        if (J_END) out[j] = out[j] + ph->mass[i] * fwx;
    });
    for (i = 0; i < n; i++) out[i] = out[i] / ph->rho[i];
}

/* df3, m_particles.f90:1898-1927: rho_i sum m_j (f_i/rho_i^2+f_j/rho_j^2) dW/dx_d */
void oracle_df3(sphg_phase_view *ph, const double *f, int d, double *out)
{
    int n = ph->ntotal + ph->nvirt, i, j, k;
    for (i = 0; i < n; i++) out[i] = 0.0;
    PAIR_LOOP(ph, n, {
        double fwx;
        fwx = ((f[i] / (ph->rho[i] * ph->rho[i])) + (f[j] / (ph->rho[j] * ph->rho[j]))) * ph->dwdx[2 * k + d];
        if (I_END) out[i] = out[i] + ph->mass[j] * fwx;
        if (J_END) out[j] = out[j] - ph->mass[i] * fwx;
    });
    for (i = 0; i < n; i++) out[i] = out[i] * ph->rho[i];
}
#undef PI_
#undef PJ_

/* ------------------------------------------------------------------------ */
/* Cross-phase routines over the cross pair list held by the water view.     */
/* ------------------------------------------------------------------------ */

/* darcy_law, m_particles.f90:3832-3874 */
void oracle_darcy_law(const sphg_params *P, sphg_phase_view *wa, sphg_phase_view *so)
{
    int i, j, k;
    double gravity = P->numeric_gravity, ks = P->soil_permeability;
    PAIR_LOOP(wa, so->ntotal + so->nvirt, {
        double rrw = wa->w[k] / (wa->rho[i] * so->rho[j]);
        double cf = wa->vof[i] * so->vof[j] * wa->rho[i] * (-gravity) / ks;
        int d;
        for (d = 0; d < 2; d++) {
            double sp = cf * (wa->vx[2 * i + d] - so->vx[2 * j + d]) * rrw;
            if (I_END) wa->dvx[2 * i + d] = wa->dvx[2 * i + d] - so->mass[j] * sp;
            if (J_END) so->dvx[2 * j + d] = so->dvx[2 * j + d] + wa->mass[i] * sp;
        }
    });
}

/* pore_water_pressure, m_particles.f90:3877-3909 */
void oracle_pore_water_pressure(const sphg_params *P, sphg_phase_view *wa, sphg_phase_view *so)
{
    int i, j, k;
    (void)P;
    PAIR_LOOP(wa, so->ntotal + so->nvirt, {
        double mprr = wa->mass[i] * wa->p[i] / (wa->rho[i] * so->rho[j]);
        int d;
        if (J_END)
            for (d = 0; d < 2; d++)
                so->dvx[2 * j + d] = so->dvx[2 * j + d] + mprr * wa->dwdx[2 * k + d] * so->vof[j];
    });
}

/* volume_fraction_soil, m_particles.f90:3970-3988 */
void oracle_volume_fraction_soil(const sphg_params *P, sphg_phase_view *so)
{
    int n = so->ntotal + so->nvirt, i;
    for (i = 0; i < n; i++) so->vof[i] = so->rho[i] / P->soil_rho0;
}

/* volume_fraction_water2, m_particles.f90:3940-3967 */
void oracle_volume_fraction_water2(const sphg_params *P, sphg_phase_view *wa, sphg_phase_view *so)
{
    int i, j, k;
    PAIR_LOOP(wa, so->ntotal + so->nvirt, {
        double dvx0 = wa->vx[2 * i] - so->vx[2 * j];
        double dvx1 = wa->vx[2 * i + 1] - so->vx[2 * j + 1];
        double tmp = dvx0 * wa->dwdx[2 * k] + dvx1 * wa->dwdx[2 * k + 1];
        if (I_END) wa->dvof[i] = wa->dvof[i] - so->mass[j] * tmp / P->soil_rho0;
    });
}

/* volume_fraction_water, m_particles.f90:3912-3937 */
void oracle_volume_fraction_water(const sphg_params *P, sphg_phase_view *wa, sphg_phase_view *so)
{
    int n = wa->ntotal + wa->nvirt, i, j, k;
    for (i = 0; i < n; i++) wa->vof2[i] = 0.0;
    PAIR_LOOP(wa, so->ntotal + so->nvirt, {
        if (I_END) wa->vof2[i] = wa->vof2[i] + so->mass[j] * wa->w[k];
    });
    for (i = 0; i < n; i++) wa->vof2[i] = 1.0 - wa->vof2[i] / P->soil_rho0;
}

/* ------------------------------------------------------------------------ */
/* single_step, declarations_sph.f90:1131-1435 (serial path).                */
/* Pinned beyond the per-routine checks by bit-exact mirror equivariance and  */
/* Galilean invariance of the composed two-phase step.                        */
/* ------------------------------------------------------------------------ */
static int phase_rhs(const sphg_params *P, sphg_phase_view *pl, int itimestep)
{
    int n = pl->ntotal + pl->nvirt, i, rc;
    double *t1, *t2, *t3, *fa, *fb;
    (void)itimestep;
    for (i = 0; i < 2 * n; i++) pl->dvx[i] = 0.0;
    for (i = 0; i < n; i++) pl->drho[i] = 0.0;
    if (pl->material == SPHG_WATER) { for (i = 0; i < n; i++) pl->dvof[i] = 0.0; }
    else for (i = 0; i < n; i++) { pl->dsxx[i] = 0.0; pl->dsxy[i] = 0.0; pl->dsyy[i] = 0.0; pl->dp[i] = 0.0; }

    rc = oracle_find_pairs(P, pl);
    if (rc) return rc;

    oracle_con_density(P, pl);
    if (P->artificial_density) oracle_delta_sph(P, pl, pl->rho, pl->drho);

    oracle_shear_strain_rate(P, pl);
    if (pl->material == SPHG_SOIL) oracle_velocity_divergence(P, pl);
    oracle_pressure(P, pl);

    if (pl->material == SPHG_WATER) {
        oracle_newtonian_fluid(P, pl);
    } else {
        if (P->yield_criterion == 1) oracle_mohr_coulomb(P, pl);
        else if (P->yield_criterion == 2) oracle_drucker_prager(P, pl);
    }

    t1 = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    t2 = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    t3 = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    fa = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    fb = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (pl->material == SPHG_WATER) {
        /* dvx = -vof*df(p) + df(vof*sxx,'x') + df(vof*sxy,'y'), :1257-1258 */
        for (i = 0; i < n; i++) { fa[i] = pl->vof[i] * pl->sxx[i]; fb[i] = pl->vof[i] * pl->sxy[i]; }
        oracle_df(pl, pl->p, 0, t1); oracle_df(pl, fa, 0, t2); oracle_df(pl, fb, 1, t3);
        for (i = 0; i < n; i++) pl->dvx[2 * i] = -(pl->vof[i] * t1[i]) + t2[i] + t3[i];
        for (i = 0; i < n; i++) { fa[i] = pl->vof[i] * pl->syy[i]; }
        oracle_df(pl, pl->p, 1, t1); oracle_df(pl, fb, 0, t2); oracle_df(pl, fa, 1, t3);
        for (i = 0; i < n; i++) pl->dvx[2 * i + 1] = -(pl->vof[i] * t1[i]) + t2[i] + t3[i];
    } else {
        /* dvx = -df3(vof*p) + df3(vof*sxx,'x') + df3(vof*sxy,'y'), :1270-1271 */
        double *fp = t1;
        double *g1 = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
        double *g2 = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
        double *g3 = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
        for (i = 0; i < n; i++) { fp[i] = pl->vof[i] * pl->p[i]; fa[i] = pl->vof[i] * pl->sxx[i]; fb[i] = pl->vof[i] * pl->sxy[i]; }
        oracle_df3(pl, fp, 0, g1); oracle_df3(pl, fa, 0, g2); oracle_df3(pl, fb, 1, g3);
        for (i = 0; i < n; i++) pl->dvx[2 * i] = -g1[i] + g2[i] + g3[i];
        for (i = 0; i < n; i++) fa[i] = pl->vof[i] * pl->syy[i];
        oracle_df3(pl, fp, 1, g1); oracle_df3(pl, fb, 0, g2); oracle_df3(pl, fa, 1, g3);
        for (i = 0; i < n; i++) pl->dvx[2 * i + 1] = -g1[i] + g2[i] + g3[i];
        free(g1); free(g2); free(g3);
    }
    for (i = 0; i < n; i++) if (pl->rho[i] > 0.0) pl->dvx[2 * i] = pl->dvx[2 * i] / pl->rho[i];
    for (i = 0; i < n; i++) if (pl->rho[i] > 0.0) pl->dvx[2 * i + 1] = pl->dvx[2 * i + 1] / pl->rho[i];
    free(t1); free(t2); free(t3); free(fa); free(fb);

    if (pl->material == SPHG_SOIL) {
        if (P->plasticity == 3) {
            oracle_plastic_or_not(P, pl);
            oracle_plastic_flow_rule3(P, pl);
        } else if (P->plasticity == 1) {
            oracle_plastic_flow_rule(P, pl);
        } else if (P->plasticity == 2) {
            oracle_plastic_flow_rule2(P, pl);
        } else if (P->plasticity != 0) {
            return SPHG_ERR_UNSUPPORTED;
        }
        oracle_jaumann_rate(P, pl);
    }
    if (P->visc_artificial) oracle_art_visc(P, pl);
    if (pl->material == SPHG_SOIL && P->soil_artificial_stress) {
        oracle_delta_sph(P, pl, pl->p, pl->dp);
        oracle_delta_sph(P, pl, pl->sxx, pl->dsxx);
        oracle_delta_sph(P, pl, pl->sxy, pl->dsxy);
        oracle_delta_sph(P, pl, pl->syy, pl->dsyy);
    }
    if (pl->material == SPHG_WATER && P->water_artificial_volume)
        oracle_delta_sph(P, pl, pl->vof, pl->dvof);
    oracle_repulsive_force(P, pl);
    for (i = 0; i < n; i++) pl->dvx[2 * i + 1] = pl->dvx[2 * i + 1] + P->gravity;
    if (P->average_velocity) oracle_av_vel(P, pl);
    return SPHG_OK;
}

int oracle_single_step(const sphg_params *P, sphg_phase_view *wa, sphg_phase_view *so, int itimestep)
{
    int rc, i;
    /* summation_density: single_step keeps con_density (its sum_density call
       is commented out, declarations_sph.f90:1190-1199) */
    rc = phase_rhs(P, wa, itimestep);
    if (rc) return rc;
    if (P->single_phase) return SPHG_OK;
    rc = phase_rhs(P, so, itimestep);
    if (rc) return rc;
    /* water/soil interaction, declarations_sph.f90:1393-1421 */
    rc = oracle_direct_find_2(P, wa, so);
    if (rc) return rc;
    oracle_darcy_law(P, wa, so);
    oracle_pore_water_pressure(P, wa, so);
    if (P->volume_fraction) {
        oracle_volume_fraction_soil(P, so);
        oracle_volume_fraction_water2(P, wa, so);
        oracle_volume_fraction_water(P, wa, so);
        if (P->volume_fraction_renorm && itimestep % 40 == 0) {
            int n = wa->ntotal + wa->nvirt;
            for (i = 0; i < n; i++) wa->rho[i] = wa->rho[i] / wa->vof[i];
            for (i = 0; i < n; i++) wa->vof[i] = wa->vof2[i];
            for (i = 0; i < n; i++) wa->rho[i] = wa->rho[i] * wa->vof[i];
        }
    }
    return SPHG_OK;
}

/* single_step_for_water, declarations_sph.f90:1439-1526 */
int oracle_single_step_for_water(const sphg_params *P, sphg_phase_view *pl, int itimestep)
{
    int n = pl->ntotal + pl->nvirt, i, rc;
    double *a, *b, *c3, *vxx, *vyy;
    (void)itimestep;
    for (i = 0; i < 2 * n; i++) pl->dvx[i] = 0.0;
    for (i = 0; i < n; i++) pl->drho[i] = 0.0;
    rc = oracle_find_pairs(P, pl);
    if (rc) return rc;
    oracle_sum_density(P, pl);
    a = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    b = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    c3 = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    vxx = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    vyy = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (i = 0; i < n; i++) { vxx[i] = pl->vx[2 * i]; vyy[i] = pl->vx[2 * i + 1]; }
    /* drho = -rho*(df2(vx,'x')+df2(vy,'y')), :1466 */
    oracle_df2(pl, vxx, 0, a); oracle_df2(pl, vyy, 1, b);
    for (i = 0; i < n; i++) pl->drho[i] = -(pl->rho[i] * (a[i] + b[i]));
    if (P->artificial_density) oracle_delta_sph(P, pl, pl->rho, pl->drho);
    /* where(rho>0) p = b*((rho/rho0)**gamma-1), :1478 */
    for (i = 0; i < n; i++)
        if (pl->rho[i] > 0.0) pl->p[i] = P->water_b * (pow(pl->rho[i] / P->water_rho0, P->water_gamma) - 1.0);
    /* txx = 2./3.*(2.0*df(vx,'x')-df(vy,'y')) etc., :1482-1484 */
    oracle_df(pl, vxx, 0, a); oracle_df(pl, vyy, 1, b);
    for (i = 0; i < n; i++) pl->txx[i] = C23_F * (2.0 * a[i] - b[i]);
    oracle_df(pl, vxx, 1, a); oracle_df(pl, vyy, 0, b);
    for (i = 0; i < n; i++) pl->txy[i] = a[i] + b[i];
    oracle_df(pl, vyy, 1, a); oracle_df(pl, vxx, 0, b);
    for (i = 0; i < n; i++) pl->tyy[i] = C23_F * (2.0 * a[i] - b[i]);
    for (i = 0; i < n; i++) {
        pl->sxx[i] = P->water_viscosity * pl->txx[i];
        pl->syy[i] = P->water_viscosity * pl->tyy[i];
        pl->sxy[i] = P->water_viscosity * pl->txy[i];
    }
    /* dvx = -df(p,'x') + df(sxx,'x') + df(sxy,'y'), :1494-1495 */
    oracle_df(pl, pl->p, 0, a); oracle_df(pl, pl->sxx, 0, b); oracle_df(pl, pl->sxy, 1, c3);
    for (i = 0; i < n; i++) pl->dvx[2 * i] = -a[i] + b[i] + c3[i];
    oracle_df(pl, pl->p, 1, a); oracle_df(pl, pl->sxy, 0, b); oracle_df(pl, pl->syy, 1, c3);
    for (i = 0; i < n; i++) pl->dvx[2 * i + 1] = -a[i] + b[i] + c3[i];
    for (i = 0; i < n; i++) if (pl->rho[i] > 0.0) pl->dvx[2 * i] = pl->dvx[2 * i] / pl->rho[i];
    for (i = 0; i < n; i++) if (pl->rho[i] > 0.0) pl->dvx[2 * i + 1] = pl->dvx[2 * i + 1] / pl->rho[i];
    free(a); free(b); free(c3); free(vxx); free(vyy);
    if (P->visc_artificial) oracle_art_visc(P, pl);
    for (i = 0; i < n; i++) pl->dvx[2 * i + 1] = pl->dvx[2 * i + 1] + P->gravity;
    if (P->average_velocity) oracle_av_vel(P, pl);
    return SPHG_OK;
}

/* single_step_for_soil, declarations_sph.f90:1530-1655 (serial path,
 * summation_density = .false.).  The same routines as the
 * soil phase of single_step, except the momentum: -df(p) + df(sxx) + df(sxy)
 * (the df form, no vof factors, :1596-1600) and no repulsive force.
 * Pinned: the df-form internal force with art_visc is pairwise
 * antisymmetric, so sum_i m_i (dvx_i - g) = 0 over all particles (to
 * rounding), and a uniform-stress lattice interior has no internal force
 * (tests/test_oracle_soil_only.py); the routines it composes are pinned in
 * tests/test_oracle_physics.py. */
int oracle_single_step_for_soil(const sphg_params *P, sphg_phase_view *pl, int itimestep)
{
    int n = pl->ntotal + pl->nvirt, i, rc;
    double *a, *b, *c3;
    (void)itimestep;
    if (P->plasticity < 0 || P->plasticity > 3) return SPHG_ERR_UNSUPPORTED;
    for (i = 0; i < 2 * n; i++) pl->dvx[i] = 0.0;
    for (i = 0; i < n; i++) {
        pl->drho[i] = 0.0; pl->dsxx[i] = 0.0; pl->dsxy[i] = 0.0; pl->dsyy[i] = 0.0; pl->dp[i] = 0.0;
    }
    rc = oracle_find_pairs(P, pl);
    if (rc) return rc;
    /* density approximation or change rate, :1558-1562 */
    if (P->summation_density) oracle_sum_density(P, pl);
    else oracle_con_density(P, pl);
    if (P->artificial_density) oracle_delta_sph(P, pl, pl->rho, pl->drho);
    oracle_shear_strain_rate(P, pl);
    oracle_velocity_divergence(P, pl);
    oracle_pressure(P, pl);
    if (P->yield_criterion == 1) oracle_mohr_coulomb(P, pl);
    else if (P->yield_criterion == 2) oracle_drucker_prager(P, pl);
    a = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    b = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    c3 = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    /* dvx = -df(p,'x') + df(sxx,'x') + df(sxy,'y'), :1596-1597 */
    oracle_df(pl, pl->p, 0, a); oracle_df(pl, pl->sxx, 0, b); oracle_df(pl, pl->sxy, 1, c3);
    for (i = 0; i < n; i++) pl->dvx[2 * i] = -a[i] + b[i] + c3[i];
    oracle_df(pl, pl->p, 1, a); oracle_df(pl, pl->sxy, 0, b); oracle_df(pl, pl->syy, 1, c3);
    for (i = 0; i < n; i++) pl->dvx[2 * i + 1] = -a[i] + b[i] + c3[i];
    for (i = 0; i < n; i++) if (pl->rho[i] > 0.0) pl->dvx[2 * i] = pl->dvx[2 * i] / pl->rho[i];
    for (i = 0; i < n; i++) if (pl->rho[i] > 0.0) pl->dvx[2 * i + 1] = pl->dvx[2 * i + 1] / pl->rho[i];
    free(a); free(b); free(c3);
    if (P->plasticity == 3) {
        oracle_plastic_or_not(P, pl);
        oracle_plastic_flow_rule3(P, pl);
    } else if (P->plasticity == 1) {
        oracle_plastic_flow_rule(P, pl);
    } else if (P->plasticity == 2) {
        oracle_plastic_flow_rule2(P, pl);
    }
    oracle_jaumann_rate(P, pl);
    if (P->visc_artificial) oracle_art_visc(P, pl);
    if (P->soil_artificial_stress) {
        oracle_delta_sph(P, pl, pl->p, pl->dp);
        oracle_delta_sph(P, pl, pl->sxx, pl->dsxx);
        oracle_delta_sph(P, pl, pl->sxy, pl->dsxy);
        oracle_delta_sph(P, pl, pl->syy, pl->dsyy);
    }
    for (i = 0; i < n; i++) pl->dvx[2 * i + 1] = pl->dvx[2 * i + 1] + P->gravity;
    if (P->average_velocity) oracle_av_vel(P, pl);
    return SPHG_OK;
}

/* ------------------------------------------------------------------------ */
/* time_integration, declarations_sph.f90:525-791.  first_half / first_step */
/* / second_half / return_mapping are the                                    */
/* routines contained there (:643-789).  Pinned by the free-fall recurrence  */
/* of an isolated particle and by closed forms of the stress levels under    */
/* constant rates, both stress_integration values and the return mapping     */
/* (tests/test_oracle_physics.py).                                           */
/* ------------------------------------------------------------------------ */
void oracle_first_half(const sphg_params *P, sphg_phase_view *pl)
{
    int n = pl->ntotal + pl->nvirt, i, d;
    double dt = P->dt;
    for (i = 0; i < n; i++) {
        if (!P->summation_density) {
            pl->rho_min[i] = pl->rho[i];
            pl->rho[i] = pl->rho[i] + (dt / 2.0) * pl->drho[i];
        }
        if (pl->material == SPHG_WATER && P->volume_fraction) {
            pl->vof_min[i] = pl->vof[i];
            pl->vof[i] = pl->vof[i] + (dt / 2.0) * pl->dvof[i];
        }
        if (pl->material == SPHG_SOIL && P->stress_integration == 2) {   /* :664-672 */
            pl->sxx_min[i] = pl->sxx[i];
            pl->sxy_min[i] = pl->sxy[i];
            pl->syy_min[i] = pl->syy[i];
            pl->p_min[i] = pl->p[i];
            pl->sxx[i] = pl->sxx[i] + (dt / 2.0) * pl->dsxx[i];
            pl->sxy[i] = pl->sxy[i] + (dt / 2.0) * pl->dsxy[i];
            pl->syy[i] = pl->syy[i] + (dt / 2.0) * pl->dsyy[i];
            pl->p[i] = pl->p[i] + (dt / 2.0) * pl->dp[i];
        }
        if (pl->itype[i] < 0) continue;
        for (d = 0; d < 2; d++) {
            pl->v_min[2 * i + d] = pl->vx[2 * i + d];
            pl->vx[2 * i + d] = pl->vx[2 * i + d] + (dt / 2.0) * pl->dvx[2 * i + d];
        }
    }
}

void oracle_first_step(const sphg_params *P, sphg_phase_view *pl)
{
    int n = pl->ntotal + pl->nvirt, i, d;
    double dt = P->dt;
    for (i = 0; i < n; i++) {
        if (!P->summation_density) pl->rho[i] = pl->rho[i] + (dt / 2.0) * pl->drho[i];
        if (pl->material == SPHG_WATER && P->volume_fraction)
            pl->vof[i] = pl->vof[i] + (dt / 2.0) * pl->dvof[i];
        if (pl->material == SPHG_SOIL && P->stress_integration == 2) {   /* :700-705 */
            pl->sxx[i] = pl->sxx[i] + (dt / 2.0) * pl->dsxx[i];
            pl->sxy[i] = pl->sxy[i] + (dt / 2.0) * pl->dsxy[i];
            pl->syy[i] = pl->syy[i] + (dt / 2.0) * pl->dsyy[i];
            pl->p[i] = pl->p[i] + (dt / 2.0) * pl->dp[i];
        } else if (pl->material == SPHG_SOIL) {
            pl->sxx[i] = pl->sxx[i] + dt * pl->dsxx[i];
            pl->sxy[i] = pl->sxy[i] + dt * pl->dsxy[i];
            pl->syy[i] = pl->syy[i] + dt * pl->dsyy[i];
            pl->p[i] = pl->p[i] + dt * pl->dp[i];
        }
        if (pl->itype[i] < 0) continue;
        for (d = 0; d < 2; d++) {
            pl->vx[2 * i + d] = pl->vx[2 * i + d] + (dt / 2.0) * pl->dvx[2 * i + d] + pl->av[2 * i + d];
            pl->x[2 * i + d] = pl->x[2 * i + d] + dt * pl->vx[2 * i + d];
        }
    }
}

void oracle_second_half(const sphg_params *P, sphg_phase_view *pl)
{
    int n = pl->ntotal + pl->nvirt, i, d;
    double dt = P->dt;
    for (i = 0; i < n; i++) {
        if (!P->summation_density) pl->rho[i] = pl->rho_min[i] + dt * pl->drho[i];
        if (pl->material == SPHG_WATER && P->volume_fraction)
            pl->vof[i] = pl->vof_min[i] + dt * pl->dvof[i];
        if (pl->material == SPHG_SOIL && P->stress_integration == 2) {   /* :744-749 */
            pl->sxx[i] = pl->sxx_min[i] + dt * pl->dsxx[i];
            pl->sxy[i] = pl->sxy_min[i] + dt * pl->dsxy[i];
            pl->syy[i] = pl->syy_min[i] + dt * pl->dsyy[i];
            pl->p[i] = pl->p_min[i] + dt * pl->dp[i];
        } else if (pl->material == SPHG_SOIL) {
            pl->sxx[i] = pl->sxx[i] + dt * pl->dsxx[i];
            pl->sxy[i] = pl->sxy[i] + dt * pl->dsxy[i];
            pl->syy[i] = pl->syy[i] + dt * pl->dsyy[i];
            pl->p[i] = pl->p[i] + dt * pl->dp[i];
        }
        if (pl->itype[i] < 0) continue;
        for (d = 0; d < 2; d++) {
            pl->vx[2 * i + d] = pl->v_min[2 * i + d] + dt * pl->dvx[2 * i + d] + pl->av[2 * i + d];
            pl->x[2 * i + d] = pl->x[2 * i + d] + dt * pl->vx[2 * i + d];
        }
    }
}

static int check_switches(const sphg_params *P)
{
    if (P->stress_integration != 1 && P->stress_integration != 2) return SPHG_ERR_UNSUPPORTED;
    return SPHG_OK;
}

/* return_mapping, declarations_sph.f90:764-789: points flagged by
   plastic_or_not take the plastic_flow_rule2 rates (stress_integration 1)
   or are recomputed from the _min levels (stress_integration 2). */
void oracle_return_mapping(const sphg_params *P, sphg_phase_view *pl)
{
    int n = pl->ntotal + pl->nvirt, i;
    double dt = P->dt;
    for (i = 0; i < n; i++) {
        if (pl->fail[i] != 1) continue;
        if (P->stress_integration == 1) {
            pl->sxx[i] = pl->sxx[i] + dt * pl->dsxx2[i];
            pl->sxy[i] = pl->sxy[i] + dt * pl->dsxy2[i];
            pl->syy[i] = pl->syy[i] + dt * pl->dsyy2[i];
            pl->p[i] = pl->p[i] + dt * pl->dp2[i];
        } else {
            pl->sxx[i] = pl->sxx_min[i] + dt * pl->dsxx[i];
            pl->sxy[i] = pl->sxy_min[i] + dt * pl->dsxy[i];
            pl->syy[i] = pl->syy_min[i] + dt * pl->dsyy[i];
            pl->p[i] = pl->p_min[i] + dt * pl->dp[i];
        }
    }
}

int oracle_time_integration(const sphg_params *P, sphg_phase_view *wa, sphg_phase_view *so,
                            int *itimestep, int nsteps)
{
    int it, rc = check_switches(P);
    if (rc) return rc;
    for (it = 0; it < nsteps; it++) {
        int step = ++(*itimestep);
        if (step != 1) {
            oracle_first_half(P, wa);
            oracle_first_half(P, so);
            oracle_drucker_prager(P, so);
        }
        rc = oracle_single_step(P, wa, so, step);
        if (rc) return rc;
        if (step == 1) {
            oracle_first_step(P, wa);
            oracle_first_step(P, so);
        } else {
            oracle_second_half(P, wa);
            oracle_second_half(P, so);
        }
        if (P->plasticity == 2) {   /* :576-579, :594-597 */
            oracle_plastic_or_not(P, so);
            oracle_return_mapping(P, so);
        }
        oracle_drucker_prager(P, so);
    }
    return SPHG_OK;
}

/* time_integration_for_soil, declarations_sph.f90:978-1127
 * (stress_integration = 1): first_half and a Drucker-Prager projection
 * (steps > 1), single_step_for_soil, first_step / second_half and the closing
 * projection -- the soil updates of time_integration without the water
 * phase.  Pinned by the free-fall recurrence of isolated particles
 * (tests/test_oracle_soil_only.py). */
int oracle_time_integration_for_soil(const sphg_params *P, sphg_phase_view *so, int *itimestep, int nsteps)
{
    int it, rc;
    if (P->stress_integration != 1 && P->stress_integration != 2) return SPHG_ERR_UNSUPPORTED;
    for (it = 0; it < nsteps; it++) {
        int step = ++(*itimestep);
        if (step != 1) {
            oracle_first_half(P, so);
            oracle_drucker_prager(P, so);
        }
        rc = oracle_single_step_for_soil(P, so, step);
        if (rc) return rc;
        if (step == 1) oracle_first_step(P, so);
        else oracle_second_half(P, so);
        oracle_drucker_prager(P, so);
    }
    return SPHG_OK;
}

/* time_integration_for_water, declarations_sph.f90:794-886 */
int oracle_time_integration_for_water(const sphg_params *P, sphg_phase_view *wa, int *itimestep, int nsteps)
{
    int it, i, d, rc;
    double dt = P->dt;
    for (it = 0; it < nsteps; it++) {
        int step = ++(*itimestep), n = wa->ntotal + wa->nvirt;
        if (step != 1) {
            for (i = 0; i < n; i++) {
                if (!P->summation_density) {   /* :814-817 */
                    wa->rho_min[i] = wa->rho[i];
                    wa->rho[i] = wa->rho[i] + (dt / 2.0) * wa->drho[i];
                }
                if (wa->itype[i] < 0) continue;
                for (d = 0; d < 2; d++) {
                    wa->v_min[2 * i + d] = wa->vx[2 * i + d];
                    wa->vx[2 * i + d] = wa->vx[2 * i + d] + (dt / 2.0) * wa->dvx[2 * i + d];
                }
            }
        }
        rc = oracle_single_step_for_water(P, wa, step);
        if (rc) return rc;
        for (i = 0; i < n; i++) {
            if (P->summation_density) { /* rho left alone, :835-856 */ }
            else if (step == 1) wa->rho[i] = wa->rho[i] + (dt / 2.0) * wa->drho[i];
            else wa->rho[i] = wa->rho_min[i] + dt * wa->drho[i];
            if (wa->itype[i] < 0) continue;
            for (d = 0; d < 2; d++) {
                if (step == 1)
                    wa->vx[2 * i + d] = wa->vx[2 * i + d] + (dt / 2.0) * wa->dvx[2 * i + d] + wa->av[2 * i + d];
                else
                    wa->vx[2 * i + d] = wa->v_min[2 * i + d] + dt * wa->dvx[2 * i + d] + wa->av[2 * i + d];
                wa->x[2 * i + d] = wa->x[2 * i + d] + dt * wa->vx[2 * i + d];
            }
        }
    }
    return SPHG_OK;
}

/* The update of time_integration_for_water_by_verlet after single_step_for_water
 * of step `step` (declarations_sph.f90:902-948): real particles (the first
 * ntotal) move by dt v + dt**2./2. dvx and advance their other velocity level
 * by 2 dt dvx (every 50th step: restart from v + dt dvx); every particle's
 * other density level advances by 2 dt drho (step 1 and every 50th step:
 * restart); then both levels swap (whole arrays, as the reference does).
 * Pinned against the scheme's closed form under constant dvx and drho: with
 * u_k = (v_k - v_0) / (dt a), u_k = 2 ceil(k/2) for k < 50 and k + 1 from the
 * first restart on; the same for the density; x by the sum of the velocities
 * (tests/test_oracle_verlet.py). */
void oracle_verlet_update(const sphg_params *P, sphg_phase_view *wa, double *lastvx, double *lastrho, int step)
{
    int n = wa->ntotal + wa->nvirt, nr = wa->ntotal, i, d;
    double dt = P->dt, h2 = pow(dt, 2.0) / 2.0;   /* dt**2./2. */
    for (i = 0; i < nr; i++)
        for (d = 0; d < 2; d++) {
            int k = 2 * i + d;
            wa->x[k] = wa->x[k] + dt * wa->vx[k] + h2 * wa->dvx[k];
            if (step % 50 != 0) lastvx[k] = lastvx[k] + 2.0 * dt * wa->dvx[k];
            else lastvx[k] = wa->vx[k] + dt * wa->dvx[k];
        }
    for (i = 0; i < 2 * n; i++) { double t = wa->vx[i]; wa->vx[i] = lastvx[i]; lastvx[i] = t; }
    for (i = 0; i < n; i++) {
        if (step == 1) lastrho[i] = wa->rho[i] + dt * 2.0 * wa->drho[i];
        else if (step % 50 != 0) lastrho[i] = lastrho[i] + dt * 2.0 * wa->drho[i];
        else lastrho[i] = wa->rho[i] + dt * wa->drho[i];
    }
    for (i = 0; i < n; i++) { double t = wa->rho[i]; wa->rho[i] = lastrho[i]; lastrho[i] = t; }
}

/* time_integration_for_water_by_verlet, declarations_sph.f90:889-974
 * (integrate_scheme = 2).  lastvx / lastrho are the routine's local arrays:
 * they start as copies of vx and rho at every call (the reference allocates
 * and fills them per call).  Composition of oracle_single_step_for_water and
 * oracle_verlet_update, each pinned on its own. */
int oracle_time_integration_for_water_by_verlet(const sphg_params *P, sphg_phase_view *wa, int *itimestep,
                                                int nsteps)
{
    int n = wa->ntotal + wa->nvirt, it, rc = SPHG_OK;
    double *lastvx, *lastrho;
    if (P->summation_density) return SPHG_ERR_UNSUPPORTED;
    lastvx = (double *)malloc(sizeof(double) * 2 * (size_t)(n > 0 ? n : 1));
    lastrho = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    memcpy(lastvx, wa->vx, sizeof(double) * 2 * (size_t)n);
    memcpy(lastrho, wa->rho, sizeof(double) * (size_t)n);
    for (it = 0; it < nsteps; it++) {
        int step = ++(*itimestep);
        rc = oracle_single_step_for_water(P, wa, step);
        if (rc) break;
        oracle_verlet_update(P, wa, lastvx, lastrho, step);
    }
    free(lastvx);
    free(lastrho);
    return rc;
}
