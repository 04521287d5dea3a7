/*
 * lbm_oracle.c - FP64 CPU restatement of the reference time step
 * (R-IFFERY/LBM_Mutiwing, Fortran 90 + OpenACC).
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker of the HIP path and the CPU
 * baseline of bench.py.  The product never links or loads this file.
 *
 * Every routine follows one reference routine statement by statement, in the
 * same floating-point evaluation order (Fortran evaluates a*b*c as (a*b)*c
 * and the build uses no FMA contraction), and cites it.  Indices inside this
 * file are 1-based like the reference (IDX(i,j) converts), so each formula
 * reads as in the Fortran source.
 *
 * Deliberate differences from the reference:
 *   - cuBLAS DGEMM/DGEMV and cuSOLVER DGETRF/DGETRS (boundary_condition.f90:
 *     1618-1640) are replaced by plain loops: the products sum in ascending
 *     index order skipping exact-zero terms (identical to a naive dense
 *     product), the solve is textbook LU with partial pivoting.
 *   - the dead corner / top / bottom characteristic updates of
 *     boundary_condition1 (boundary_condition.f90:303-401, 599-649, 728-776)
 *     are not computed: their results are overwritten before any use
 *     (boundary_condition.f90:651-661, 778-788).
 *   - body_point / distanceq (airfoil.f90:2-189) are skipped: body is zeroed
 *     right after (solution.f90:71).
 *   - an IBM support box that reaches within 3 nodes of the lattice edge is
 *     an error (the reference indexes outside the arrays in that case).
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXW LBM_MAX_WINGS
#define PAI 3.1415926535897932384626433832795
/* fields of the rigid-body state, restart-file order (solution.f90:1663-1665) */
enum { REL_A, REL_A_PREV, REL_V, REL_V_PREV, REL_S, REL_S_PREV, REL_POS, REL_THETA_PREV, REL_THETA,
       REL_TEMP_THETA_PREV, REL_TEMP_THETA };

struct orc {
    lbm_config cfg;
    int nx, ny, W, P, SPW, nseg, npts, ncols, row_len;
    size_t np;
    double scale, ox, oy, uu, dt, niu, tau, s[9];
    double w[9];
    int e[9][2];
    double matm[9][9], minv[9][9];
    double cs;
    double *f, *f0, *rho, *u, *v, *u0, *v0, *pre, *g1, *g2;
    int *body;
    double *xb0, *yb0, *xb1, *yb1, *xb2, *yb2, *arc;
    int *box; /* per segment i1,i2,j1,j2 */
    int cap, adj_n;
    int *adj_i, *adj_j, *adj_w;
    double angle[MAXW], omega[MAXW], omgac[MAXW], sa[MAXW], sau[MAXW],
        saac[MAXW], sb[MAXW], sbv[MAXW], sbac[MAXW];
    /* body_rel_* in restart-file order (solution.f90:1663-1665) */
    double rel[LBM_BODY_REL_FIELDS][MAXW];
    lbm_dynamics dyn;
    double body_extra_v[MAXW], body_extra_s[MAXW], body_extra_pos[MAXW];
    int movenum, flag2, n;
    double *table;
    int table_rows;
    double *step_table;
    int step_count;
    double motion_dt, motion_end_time;
    int motion_loaded;
    double stream_start[LBM_MAX_MOTION_COLS], stream_target[LBM_MAX_MOTION_COLS];
    int stream_ready;
    double wing_force[MAXW][2], wing_moment[MAXW];
    double err;
    double *sysA, *sysB, *sysX;
    double *ibm_us, *ibm_vs; /* u* at the adjoint nodes of the last step */
    int have_sys;
    char msg[512];
};

static char g_create_msg[512];
static int g_threads = 1;

static int fail(orc_t *o, int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(o ? o->msg : g_create_msg, 512, fmt, ap);
    va_end(ap);
    return code;
}

#define IDX(i, j) ((size_t)((j)-1) * (size_t)o->nx + (size_t)((i)-1))
#define FK(a, i, j, k) (a)[(size_t)((k)-1) * o->np + IDX(i, j)]

/* ---------------------------------------------------------------- model */

/* D2Q9model (D2Q9model.f90:1-33): the weights are single-precision
 * quotients stored in double precision. */
void orc_weights(double w[9])
{
    w[0] = (double)(4.0f / 9.0f);
    for (int k = 1; k < 5; k++) w[k] = (double)(1.0f / 9.0f);
    for (int k = 5; k < 9; k++) w[k] = (double)(1.0f / 36.0f);
}

static const int E_TAB[9][2] = {{0, 0}, {1, 0}, {0, 1}, {-1, 0}, {0, -1},
                                {1, 1}, {-1, 1}, {-1, -1}, {1, -1}};

static const double MATM[9][9] = { /* solution.f90:17-25 */
    {1, 1, 1, 1, 1, 1, 1, 1, 1},
    {-4, -1, -1, -1, -1, 2, 2, 2, 2},
    {4, -2, -2, -2, -2, 1, 1, 1, 1},
    {0, 1, 0, -1, 0, 1, -1, -1, 1},
    {0, -2, 0, 2, 0, 1, -1, -1, 1},
    {0, 0, 1, 0, -1, 1, 1, -1, -1},
    {0, 0, -2, 0, 2, 1, 1, -1, -1},
    {0, 1, -1, 1, -1, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 1, -1, 1, -1}};

/* nizhen (solution.f90:1474-1529): Gauss-Jordan inverse with row pivoting,
 * reproduced operation by operation (its result differs from the closed form
 * M^T diag(1/9,...) in the last bit of a few entries). */
static void nizhen(const double AA[9][9], double b1[9][9])
{
    const int Nm = 9;
    double a[9][9];
    memcpy(a, AA, sizeof(a));
    for (int i = 0; i < Nm; i++)
        for (int j = 0; j < Nm; j++) b1[i][j] = (i == j) ? 1.0 : 0.0;
    for (int k = 0; k < Nm - 1; k++) {
        int ik = k;
        double temp = fabs(a[ik][k]);
        for (int i = k; i < Nm; i++)
            if (fabs(a[i][k]) > fabs(temp)) {
                temp = fabs(a[i][k]);
                ik = i;
            }
        for (int j = 0; j < Nm; j++) {
            temp = a[ik][j]; a[ik][j] = a[k][j]; a[k][j] = temp;
            temp = b1[ik][j]; b1[ik][j] = b1[k][j]; b1[k][j] = temp;
        }
        for (int i = 0; i < Nm; i++) {
            double mik = a[i][k] / a[k][k];
            temp = a[k][k];
            for (int j = k; j < Nm; j++)
                if (i != k) a[i][j] = a[i][j] - mik * a[k][j];
            for (int j = k; j < Nm; j++) a[k][j] = a[k][j] / temp;
            for (int j = 0; j < Nm; j++)
                if (i != k) b1[i][j] = b1[i][j] - mik * b1[k][j];
            for (int j = 0; j < Nm; j++) b1[k][j] = b1[k][j] / temp;
        }
    }
    double temp = a[Nm - 1][Nm - 1];
    for (int i = 0; i < Nm - 1; i++) {
        double mik = a[i][Nm - 1] / a[Nm - 1][Nm - 1];
        a[i][Nm - 1] = 0.0;
        for (int j = 0; j < Nm; j++) b1[i][j] = b1[i][j] - mik * b1[Nm - 1][j];
    }
    for (int j = 0; j < Nm; j++) b1[Nm - 1][j] = b1[Nm - 1][j] / temp;
}

void orc_mrt_matrices(double m[81], double minv[81])
{
    double mi[9][9];
    nizhen(MATM, mi);
    for (int i = 0; i < 9; i++)
        for (int j = 0; j < 9; j++) {
            m[i * 9 + j] = MATM[i][j];
            minv[i * 9 + j] = mi[i][j];
        }
}

/* matmulvoc (boundary_condition.f90:1738-1755): c = A b, sum in j order. */
static inline void matmulvoc(double c[9], const double a[9][9], const double b[9])
{
    for (int i = 0; i < 9; i++) {
        double sum = 0.0;
        for (int j = 0; j < 9; j++) sum = sum + a[i][j] * b[j];
        c[i] = sum;
    }
}

/* D2Q9feq (D2Q9model.f90:35-46). */
static double d2q9feq(const orc_t *o, double rh, double u1, double v1, int k)
{
    double eu = (double)o->e[k][0] * u1 + (double)o->e[k][1] * v1;
    double utwo = u1 * u1 + v1 * v1;
    return o->w[k] * rh * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * utwo);
}

/* smoothed delta kernel, smooth=1 branch (boundary_condition.f90:1581-1600).
 * disr is already an absolute value there. */
double orc_delta_kernel(double disr)
{
    if (fabs(disr) <= 0.5)
        return 0.375 + 0.03125 * PAI - 0.25 * disr * disr;
    else if (fabs(disr) <= 1.5 && fabs(disr) >= 0.5)
        return 0.25 + 0.125 * (1.0 - disr) * sqrt(-2.0 + 8.0 * disr - 4.0 * disr * disr) -
               0.125 * asin(sqrt(2.0) * (disr - 1.0));
    else if (fabs(disr) <= 2.5 && fabs(disr) >= 1.5)
        return 1.0625 - 0.015625 * PAI - 0.75 * disr + 0.125 * disr * disr +
               0.0625 * (disr - 2.0) * sqrt(-14.0 + 16.0 * disr - 4.0 * disr * disr) +
               0.0625 * asin(sqrt(2.0) * (disr - 2.0));
    return 0.0;
}

int orc_set_threads(int n)
{
    if (n < 1) n = 1;
    g_threads = n;
#ifdef _OPENMP
    omp_set_num_threads(n);
    return n;
#else
    return 1;
#endif
}

/* ------------------------------------------------------------- lifetime */

int orc_create(const lbm_config *cfg, orc_t **out)
{
    if (!cfg || !out) return fail(NULL, LBM_E_INVALID, "null argument");
    *out = NULL;
    if (cfg->nx < 8 || cfg->ny < 8)
        return fail(NULL, LBM_E_INVALID, "lattice must be at least 8x8");
    if (cfg->wing_count < 0 || cfg->wing_count > MAXW)
        return fail(NULL, LBM_E_INVALID, "wing_count out of range");
    if (cfg->wing_count > 0 && (cfg->points_per_wing < 3 || !cfg->xb0 || !cfg->yb0))
        return fail(NULL, LBM_E_INVALID, "marker arrays missing");
    if (!(cfg->scale > 0.0) || !(cfg->uu > 0.0) || !(cfg->re > 0.0))
        return fail(NULL, LBM_E_INVALID, "scale, uu and re must be positive");
    if (cfg->motion_source != LBM_MOTION_FILE && cfg->motion_source != LBM_MOTION_STREAM)
        return fail(NULL, LBM_E_INVALID, "motion_source must be 1 (file) or 2 (stream)");
    orc_t *o = (orc_t *)calloc(1, sizeof(orc_t));
    if (!o) return fail(NULL, LBM_E_INVALID, "out of host memory");
    o->cfg = *cfg;
    o->nx = cfg->nx;
    o->ny = cfg->ny;
    o->np = (size_t)o->nx * (size_t)o->ny;
    o->W = cfg->wing_count;
    o->P = cfg->wing_count > 0 ? cfg->points_per_wing : 0;
    o->SPW = o->P > 0 ? o->P - 1 : 0;
    o->nseg = o->W * o->SPW;
    o->npts = o->W * o->P;
    o->ncols = 1 + LBM_MOTION_COLS_PER_WING * o->W;
    o->row_len = 4 + 12 * o->W;
    o->scale = cfg->scale;
    o->ox = cfg->ox;
    o->oy = cfg->oy;
    o->uu = cfg->uu;
    o->dt = o->scale * o->uu;                        /* main.f90:375 */
    o->niu = o->uu * (1.0 / o->scale) / cfg->re;     /* ini_mesh.f90:30 */
    o->tau = 3.0 * o->niu + 0.5;                     /* ini_mesh.f90:31 */
    for (int k = 0; k < 9; k++) o->s[k] = cfg->s[k];
    if (!(o->s[7] > 0.0)) o->s[7] = 1.0 / o->tau;    /* ini_mesh.f90:66-67 */
    if (!(o->s[8] > 0.0)) o->s[8] = 1.0 / o->tau;
    orc_weights(o->w);
    memcpy(o->e, E_TAB, sizeof(E_TAB));
    memcpy(o->matm, MATM, sizeof(MATM));
    nizhen(MATM, o->minv);
    o->cs = (double)sqrtf(3.0f) / 3.0;                /* boundary_condition.f90:304 */
    size_t np = o->np;
    o->f = (double *)malloc(9 * np * sizeof(double));
    o->f0 = (double *)malloc(9 * np * sizeof(double));
    o->rho = (double *)malloc(np * sizeof(double));
    o->u = (double *)malloc(np * sizeof(double));
    o->v = (double *)malloc(np * sizeof(double));
    o->u0 = (double *)malloc(np * sizeof(double));
    o->v0 = (double *)malloc(np * sizeof(double));
    o->pre = (double *)malloc(np * sizeof(double));
    o->g1 = (double *)calloc(np, sizeof(double));
    o->g2 = (double *)calloc(np, sizeof(double));
    o->body = (int *)calloc(np, sizeof(int));
    int npts = o->npts > 0 ? o->npts : 1, nseg = o->nseg > 0 ? o->nseg : 1;
    o->xb0 = (double *)calloc(npts, sizeof(double));
    o->yb0 = (double *)calloc(npts, sizeof(double));
    o->xb1 = (double *)calloc(npts, sizeof(double));
    o->yb1 = (double *)calloc(npts, sizeof(double));
    o->xb2 = (double *)calloc(nseg, sizeof(double));
    o->yb2 = (double *)calloc(nseg, sizeof(double));
    o->arc = (double *)calloc(nseg, sizeof(double));
    o->box = (int *)calloc(4 * (size_t)nseg, sizeof(int));
    o->cap = 36 * npts; /* main.f90:131 */
    o->adj_i = (int *)calloc(o->cap, sizeof(int));
    o->adj_j = (int *)calloc(o->cap, sizeof(int));
    o->adj_w = (int *)calloc(o->cap, sizeof(int));
    o->sysA = (double *)calloc((size_t)nseg * nseg, sizeof(double));
    o->sysB = (double *)calloc(2 * (size_t)nseg, sizeof(double));
    o->sysX = (double *)calloc(2 * (size_t)nseg, sizeof(double));
    o->ibm_us = (double *)calloc(o->cap > 0 ? o->cap : 1, sizeof(double));
    o->ibm_vs = (double *)calloc(o->cap > 0 ? o->cap : 1, sizeof(double));
    if (!o->f || !o->f0 || !o->rho || !o->u || !o->v || !o->u0 || !o->v0 || !o->pre ||
        !o->g1 || !o->g2 || !o->body || !o->sysA) {
        orc_destroy(o);
        return fail(NULL, LBM_E_INVALID, "out of host memory");
    }
    /* read_boundary (airfoil.f90:207-242): arc lengths in lattice units and
     * the (xc-xp, yc-yp) shift of the marker coordinates. */
    for (int wing = 0; wing < o->W; wing++) {
        int ps = wing * o->P, ss = wing * o->SPW;
        for (int i = 0; i < o->P; i++) {
            o->xb0[ps + i] = cfg->xb0[ps + i];
            o->yb0[ps + i] = cfg->yb0[ps + i];
        }
        for (int i0 = 0; i0 < o->SPW; i0++) {
            double dx = o->xb0[ps + i0 + 1] - o->xb0[ps + i0];
            double dy = o->yb0[ps + i0 + 1] - o->yb0[ps + i0];
            o->arc[ss + i0] = sqrt(dx * dx + dy * dy) / o->scale;
        }
    }
    for (int i = 0; i < o->npts; i++) {
        o->xb0[i] = o->xb0[i] + (cfg->xc - cfg->xp);
        o->yb0[i] = o->yb0[i] + (cfg->yc - cfg->yp);
    }
    *out = o;
    return orc_reset(o);
}

int orc_destroy(orc_t *o)
{
    if (!o) return 0;
    free(o->f); free(o->f0); free(o->rho); free(o->u); free(o->v); free(o->u0);
    free(o->v0); free(o->pre); free(o->g1); free(o->g2); free(o->body);
    free(o->xb0); free(o->yb0); free(o->xb1); free(o->yb1); free(o->xb2);
    free(o->yb2); free(o->arc); free(o->box); free(o->adj_i); free(o->adj_j);
    free(o->adj_w); free(o->table); free(o->step_table); free(o->sysA);
    free(o->sysB); free(o->sysX); free(o->ibm_us); free(o->ibm_vs);
    free(o);
    return 0;
}

int orc_last_error(const orc_t *o, char *buf, size_t n)
{
    if (!buf || n == 0) return LBM_E_INVALID;
    snprintf(buf, n, "%s", o ? o->msg : g_create_msg);
    return 0;
}

/* --------------------------------------------------------------- motion */

/* update_body_extra_from_rel (mode.f90:473-495). */
static void update_body_extra(orc_t *o)
{
    double av = 0.0, as = 0.0, ap = 0.0;
    for (int w = 0; w < MAXW; w++) {
        o->body_extra_v[w] = 0.0;
        o->body_extra_s[w] = 0.0;
        o->body_extra_pos[w] = 0.0;
    }
    for (int w = 0; w < o->W; w++) {
        av = av + o->rel[REL_V][w];
        as = as + o->rel[REL_S][w];
        ap = ap + o->rel[REL_POS][w];
        o->body_extra_v[w] = av;
        o->body_extra_s[w] = as;
        o->body_extra_pos[w] = ap;
    }
}

/* apply_motion_row (mode.f90:180-218). */
static void apply_motion_row(orc_t *o, const double *row)
{
    for (int w = 0; w < MAXW; w++) {
        o->angle[w] = o->omega[w] = o->omgac[w] = 0.0;
        o->sa[w] = o->sau[w] = o->saac[w] = 0.0;
        o->sb[w] = o->sbv[w] = o->sbac[w] = 0.0;
    }
    for (int w = 0; w < o->W; w++) {
        int col = 1 + w * LBM_MOTION_COLS_PER_WING;
        o->angle[w] = row[col];
        o->omega[w] = row[col + 1];
        o->omgac[w] = row[col + 2];
        o->sa[w] = row[col + 3];
        o->sau[w] = row[col + 4];
        o->saac[w] = row[col + 5];
        o->sb[w] = row[col + 6];
        o->sbv[w] = row[col + 7];
        o->sbac[w] = row[col + 8];
        if (o->cfg.motion_source == LBM_MOTION_FILE) {
            double x_base = o->cfg.wing_x0 + (double)w * o->cfg.wing_dx;
            double y_base = o->cfg.wing_y0;
            o->sa[w] = o->sa[w] - x_base;
            o->sb[w] = o->sb[w] - y_base;
        }
    }
}

/* smooth_r / smooth_h / smooth_d (mode.f90:220-261). */
static double smooth_r(double x)
{
    if (x < 0.0) return 0.0;
    if (x <= 1.0) return 0.5 * x * x + (cos(2.0 * PAI * x) - 1.0) / (4.0 * PAI * PAI);
    return x - 0.5;
}
static double smooth_h(double x)
{
    if (x < 0.0) return 0.0;
    if (x <= 1.0) return x - sin(2.0 * PAI * x) / (2.0 * PAI);
    return 1.0;
}
static double smooth_d(double x)
{
    if (x >= 0.0 && x <= 1.0) return 1.0 - cos(2.0 * PAI * x);
    return 0.0;
}

/* interpolate_motion_triplet (mode.f90:263-285). */
static void interp_triplet(double s0, double v0, double v1, double dt, double lt,
                           double *s, double *v, double *a)
{
    double dv = v1 - v0;
    if (dt <= 1.0e-14) {
        *s = (v0 + dv) * lt + s0;
        *v = v0 + dv;
        *a = 0.0;
        return;
    }
    double x = lt / dt;
    *s = dv * dt * smooth_r(x) + v0 * lt + s0;
    *v = dv * smooth_h(x) + v0;
    *a = dv / dt * smooth_d(x);
}

/* build_stream_target_row (mode.f90:287-325). */
int orc_build_stream_target_row(int32_t W, const double *start, double target_time,
                                const double *delta, double *target)
{
    if (W < 0 || W > MAXW || !start || !delta || !target) return LBM_E_INVALID;
    int ncols = 1 + LBM_MOTION_COLS_PER_WING * W;
    double dt = target_time - start[0];
    if (dt < -1.0e-12) return LBM_E_INVALID;
    for (int c = 0; c < ncols; c++) target[c] = start[c];
    target[0] = target_time;
    double lt = dt > 0.0 ? dt : 0.0;
    for (int w = 0; w < W; w++) {
        int col = 1 + w * 9, dc = 1 + w * 3;
        interp_triplet(start[col], start[col + 1], start[col + 1] + delta[dc], dt, lt,
                       &target[col], &target[col + 1], &target[col + 2]);
        interp_triplet(start[col + 3], start[col + 4], start[col + 4] + delta[dc + 1], dt, lt,
                       &target[col + 3], &target[col + 4], &target[col + 5]);
        interp_triplet(start[col + 6], start[col + 7], start[col + 7] + delta[dc + 2], dt, lt,
                       &target[col + 6], &target[col + 7], &target[col + 8]);
    }
    return 0;
}

/* lookup_stream_motion_row (mode.f90:414-449). */
static void lookup_stream_row(const orc_t *o, double tt, double *row)
{
    const double tol = 1.0e-12;
    double t1 = o->stream_start[0], t2 = o->stream_target[0], dt = t2 - t1;
    if (tt <= t1 + tol) {
        for (int c = 0; c < o->ncols; c++) row[c] = o->stream_start[c];
        row[0] = tt;
        return;
    }
    double lt = tt - t1;
    if (lt < 0.0) lt = 0.0;
    double dmax = dt > 0.0 ? dt : 0.0;
    if (lt > dmax) lt = dmax;
    for (int c = 0; c < o->ncols; c++) row[c] = 0.0;
    row[0] = tt;
    for (int w = 0; w < o->W; w++) {
        int col = 1 + w * 9;
        const double *S = o->stream_start, *T = o->stream_target;
        interp_triplet(S[col], S[col + 1], T[col + 1], dt, lt, &row[col], &row[col + 1], &row[col + 2]);
        interp_triplet(S[col + 3], S[col + 4], T[col + 4], dt, lt, &row[col + 3], &row[col + 4], &row[col + 5]);
        interp_triplet(S[col + 6], S[col + 7], T[col + 7], dt, lt, &row[col + 6], &row[col + 7], &row[col + 8]);
    }
}

/* interpolate_motion_row + lookup_motion_row (mode.f90:327-386). */
static void lookup_motion_row(const orc_t *o, int *midx, double tt, double *row)
{
    const double tol = 1.0e-12;
    const int nc = o->ncols, N = o->table_rows;
    const double *T = o->table;
#define TAB(r, c) T[(size_t)((r)-1) * nc + (c)]
    if (tt <= TAB(1, 0) + tol) {
        *midx = 1;
        for (int c = 0; c < nc; c++) row[c] = TAB(1, c);
        row[0] = tt;
        return;
    }
    if (tt >= TAB(N, 0) - tol) {
        *midx = N;
        for (int c = 0; c < nc; c++) row[c] = TAB(N, c);
        row[0] = tt;
        return;
    }
    int m = *midx;
    if (m > N - 1) m = N - 1;
    if (m < 1) m = 1;
    while (m > 1 && tt <= TAB(m, 0) + tol) m--;
    while (m < N - 1 && tt > TAB(m + 1, 0) + tol) m++;
    *midx = m;
    double t1 = TAB(m, 0), t2 = TAB(m + 1, 0), dt = t2 - t1;
    if (fabs(dt) <= 1.0e-14) {
        for (int c = 0; c < nc; c++) row[c] = TAB(m + 1, c);
        row[0] = tt;
        return;
    }
    double ratio = (tt - t1) / dt;
    if (ratio < 0.0) ratio = 0.0;
    if (ratio > 1.0) ratio = 1.0;
    for (int c = 0; c < nc; c++) row[c] = TAB(m, c) + ratio * (TAB(m + 1, c) - TAB(m, c));
    row[0] = tt;
#undef TAB
}

/* load_motion_table validation + build_motion_step_table
 * (mode.f90:104-178, 388-412). */
int orc_set_motion_table(orc_t *o, const double *rows, int32_t nrows, int32_t ncols,
                         double *end_time)
{
    if (!o || !rows) return fail(o, LBM_E_INVALID, "null argument");
    if (ncols != o->ncols)
        return fail(o, LBM_E_INVALID, "motion table must have %d columns, got %d", o->ncols, ncols);
    if (nrows < 1) return fail(o, LBM_E_INVALID, "motion table is empty");
    for (int i = 1; i < nrows; i++)
        if (rows[(size_t)i * ncols] < rows[(size_t)(i - 1) * ncols])
            return fail(o, LBM_E_INVALID, "motion table time column must be non-decreasing");
    free(o->table);
    o->table = (double *)malloc((size_t)nrows * ncols * sizeof(double));
    memcpy(o->table, rows, (size_t)nrows * ncols * sizeof(double));
    o->table_rows = nrows;
    o->motion_end_time = rows[(size_t)(nrows - 1) * ncols];
    o->motion_dt = o->dt;
    o->step_count = (int)((o->motion_end_time + 1.0e-12) / o->motion_dt) + 1;
    free(o->step_table);
    o->step_table = (double *)malloc((size_t)o->step_count * ncols * sizeof(double));
    int midx = 1;
    for (int i = 1; i <= o->step_count; i++)
        lookup_motion_row(o, &midx, (double)(i - 1) * o->motion_dt,
                          &o->step_table[(size_t)(i - 1) * ncols]);
    o->motion_loaded = 1;
    if (end_time) *end_time = o->motion_end_time;
    return 0;
}

int orc_set_stream_segment(orc_t *o, const double *start, const double *target, int32_t ncols)
{
    if (!o || !start || !target) return fail(o, LBM_E_INVALID, "null argument");
    if (ncols != o->ncols)
        return fail(o, LBM_E_INVALID, "stream rows must have %d columns, got %d", o->ncols, ncols);
    for (int c = 0; c < LBM_MAX_MOTION_COLS; c++) {
        o->stream_start[c] = c < ncols ? start[c] : 0.0;
        o->stream_target[c] = c < ncols ? target[c] : 0.0;
    }
    o->stream_ready = 1;
    return 0;
}

/* mode (mode.f90:451-471) at the module time `time`. */
static int mode(orc_t *o, double time)
{
    double row[LBM_MAX_MOTION_COLS];
    if (o->cfg.motion_source == LBM_MOTION_STREAM) {
        if (!o->stream_ready) return fail(o, LBM_E_STATE, "stream motion not initialised");
        lookup_stream_row(o, time, row);
        apply_motion_row(o, row);
        return 0;
    }
    if (!o->motion_loaded) return fail(o, LBM_E_STATE, "motion table not loaded");
    int idx = (int)lround(time / o->motion_dt) + 1;
    if (idx > o->step_count) idx = o->step_count;
    if (idx < 1) idx = 1;
    apply_motion_row(o, &o->step_table[(size_t)(idx - 1) * o->ncols]);
    return 0;
}

/* ------------------------------------------------------------ geometry */

/* move_boundary (airfoil.f90:244-338). */
static int move_boundary(orc_t *o)
{
    const double scale = o->scale;
    for (int w = 0; w < o->W; w++) {
        int ps = w * o->P, ss = w * o->SPW;
        double ca = cos(o->angle[w]), sn = sin(o->angle[w]);
        for (int i0 = ps; i0 < ps + o->P; i0++) {
            o->xb1[i0] = ca * o->xb0[i0] + sn * o->yb0[i0] + o->cfg.wing_x0 +
                         (double)w * o->cfg.wing_dx + o->sa[w] + o->cfg.xc +
                         o->body_extra_s[w] / (1.0 / scale) - (double)o->movenum * scale;
            o->yb1[i0] = -sn * o->xb0[i0] + ca * o->yb0[i0] + o->cfg.wing_y0 + o->sb[w] + o->cfg.yc;
        }
        for (int i0 = ss; i0 < ss + o->SPW; i0++) {
            o->xb2[i0] = 0.5 * (o->xb1[ps + i0 - ss + 1] + o->xb1[ps + i0 - ss]);
            o->yb2[i0] = 0.5 * (o->yb1[ps + i0 - ss + 1] + o->yb1[ps + i0 - ss]);
        }
    }
    for (int i = 0; i < o->adj_n; i++) {
        size_t id = IDX(o->adj_i[i], o->adj_j[i]);
        o->body[id] = 0;
        o->g1[id] = 0.0;
        o->g2[id] = 0.0;
        o->adj_w[i] = 0;
    }
    o->adj_n = 0;
    const double ox = o->ox, oy = o->oy;
    for (int s = 0; s < o->nseg; s++) {
        int wing = s / o->SPW + 1;
        double X = o->xb2[s], Y = o->yb2[s];
        int i1 = (int)((X - ox) / scale + 1.0), i2 = i1;
        int j1 = (int)((Y - oy) / scale + 1.0), j2 = j1;
        for (int i = i1; i >= 1; i--)
            if (fabs((ox + (double)(i - 1) * scale - X) / scale) > 2.5) { i1 = i + 1; break; }
        for (int i = i2; i <= o->nx; i++)
            if (fabs((ox + (double)(i - 1) * scale - X) / scale) > 2.5) { i2 = i - 1; break; }
        for (int i = j2; i <= o->ny; i++)
            if (fabs((oy + (double)(i - 1) * scale - Y) / scale) > 2.5) { j2 = i - 1; break; }
        for (int i = j1; i >= 1; i--)
            if (fabs((oy + (double)(i - 1) * scale - Y) / scale) > 2.5) { j1 = i + 1; break; }
        if (i1 < 4 || j1 < 4 || i2 > o->nx - 3 || j2 > o->ny - 3)
            return fail(o, LBM_E_RANGE,
                        "IBM support of segment %d (wing %d) reaches the lattice edge", s + 1, wing);
        o->box[4 * s + 0] = i1;
        o->box[4 * s + 1] = i2;
        o->box[4 * s + 2] = j1;
        o->box[4 * s + 3] = j2;
        for (int i = i1; i <= i2; i++)
            for (int j = j1; j <= j2; j++) {
                size_t id = IDX(i, j);
                if (o->body[id] == 0) {
                    if (o->adj_n >= o->cap) return fail(o, LBM_E_RANGE, "adjoint list overflow");
                    o->body[id] = 1;
                    o->adj_i[o->adj_n] = i;
                    o->adj_j[o->adj_n] = j;
                    o->adj_w[o->adj_n] = wing;
                    o->adj_n++;
                }
            }
    }
    return 0;
}

/* --------------------------------------------------------------- lattice */

/* streaming (solution.f90:440-466): pull for interior nodes. */
static void streaming(orc_t *o)
{
    const int nx = o->nx, ny = o->ny;
    for (int k = 1; k <= 9; k++) {
        const int ex = o->e[k - 1][0], ey = o->e[k - 1][1];
#pragma omp parallel for schedule(static)
        for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nx - 1; i++)
                FK(o->f0, i, j, k) = FK(o->f, i - ex, j - ey, k);
    }
}

/* physical_variables (solution.f90:483-554): moments at all nodes, f <- f0
 * at interior nodes. */
static void physical_variables(orc_t *o)
{
    const int nx = o->nx, ny = o->ny;
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
            size_t id = IDX(i, j);
            o->u0[id] = o->u[id];
            o->v0[id] = o->v[id];
            double tr = 0.0, tu = 0.0, tv = 0.0;
            int ring = (i == 1) || (i == nx) || (j == 1) || (j == ny);
            for (int k = 1; k <= 9; k++) {
                double t = FK(o->f0, i, j, k);
                if (!ring) FK(o->f, i, j, k) = t;
                tr = tr + t;
                tu = tu + (double)o->e[k - 1][0] * t;
                tv = tv + (double)o->e[k - 1][1] * t;
            }
            o->rho[id] = tr;
            o->pre[id] = tr / 3.0;
            o->u[id] = tu / tr;
            o->v[id] = tv / tr;
        }
}

/* colission (solution.f90:585-655): MRT relaxation with Guo forcing. */
static void colission(orc_t *o)
{
    const int nx = o->nx, ny = o->ny;
    const double *s = o->s;
#pragma omp parallel for schedule(static)
    for (int j = 2; j <= ny - 1; j++)
        for (int i = 2; i <= nx - 1; i++) {
            size_t id = IDX(i, j);
            double temp[9], temp1[9], fp[9];
            for (int k = 1; k <= 9; k++) temp[k - 1] = FK(o->f, i, j, k);
            matmulvoc(temp1, o->matm, temp);
            double ut = o->u[id], vt = o->v[id], rt = o->rho[id];
            temp[0] = 0.0;
            temp[1] = -2.0 * rt + 3.0 * (rt * ut * rt * ut + rt * vt * rt * vt);
            temp[2] = rt - 3.0 * (rt * ut * rt * ut + rt * vt * rt * vt);
            temp[3] = 0.0;
            temp[4] = -rt * ut;
            temp[5] = 0.0;
            temp[6] = -rt * vt;
            temp[7] = (rt * ut * rt * ut - rt * vt * rt * vt);
            temp[8] = (rt * ut * rt * vt);
            double gfx = o->g1[id], gfy = o->g2[id];
            fp[0] = 0.0;
            fp[1] = 6.0 * (1.0 - 0.5 * s[1]) * (ut * gfx + vt * gfy);
            fp[2] = -6.0 * (1.0 - 0.5 * s[2]) * (ut * gfx + vt * gfy);
            fp[3] = gfx;
            fp[4] = -(1.0 - 0.5 * s[4]) * gfx;
            fp[5] = gfy;
            fp[6] = -(1.0 - 0.5 * s[6]) * gfy;
            fp[7] = 2.0 * (1.0 - 0.5 * s[7]) * (ut * gfx - vt * gfy);
            fp[8] = (1.0 - 0.5 * s[8]) * (ut * gfy + vt * gfx);
            for (int k = 0; k < 9; k++) temp[k] = temp1[k] + (temp[k] - temp1[k]) * s[k] + fp[k];
            matmulvoc(temp1, o->minv, temp);
            for (int k = 1; k <= 9; k++) FK(o->f, i, j, k) = temp1[k - 1];
        }
}

/* compressible moment-space equilibrium used by the far-field BC
 * (boundary_condition.f90:444-455). */
static inline void meq_bc(double rh, double uu, double vv, double t[9])
{
    double jx = rh * uu, jy = rh * vv;
    t[0] = rh;
    t[1] = -2.0 * rh + 3.0 * (jx * jx + jy * jy) / rh;
    t[2] = rh - 3.0 * (jx * jx + jy * jy) / rh;
    t[3] = jx;
    t[4] = -jx;
    t[5] = jy;
    t[6] = -jy;
    t[7] = (jx * jx - jy * jy) / rh;
    t[8] = (jx * jy) / rh;
}

/* f(node b) = M^-1 [ m_eq(b) + (1-s)(M f0(node n) - m_eq(n)) ]
 * (boundary_condition.f90:443-474). */
static void bc_extrapolate(orc_t *o, int ib, int jb, int in, int jn)
{
    double t1[9], t2[9], t3[9], tmp[9];
    size_t b = IDX(ib, jb), n = IDX(in, jn);
    meq_bc(o->rho[b], o->u[b], o->v[b], t1);
    meq_bc(o->rho[n], o->u[n], o->v[n], t2);
    for (int k = 1; k <= 9; k++) tmp[k - 1] = FK(o->f0, in, jn, k);
    matmulvoc(t3, o->matm, tmp);
    for (int k = 0; k < 9; k++) t3[k] = t1[k] + (1.0 - o->s[k]) * (t3[k] - t2[k]);
    matmulvoc(tmp, o->minv, t3);
    for (int k = 1; k <= 9; k++) FK(o->f, ib, jb, k) = tmp[k - 1];
}

/* boundary_condition1 (boundary_condition.f90:288-858), ref_freestream=1,
 * periodical_bc=0, single block. */
/* LODI update of one far-field node (boundary_condition.f90:413-483 left,
 * :495-567 right): r, u, v hold rho, u, v at the boundary node and at the
 * next two nodes inward; out receives the boundary node's updated rho, u, v.
 * The u update divides by the updated rho, as the reference does. */
static void lodi_node(double cs, int right, const double *r, const double *u, const double *v, double *out)
{
    double rhox, ux, vx, lx1, lx2, lx3;
    if (!right) {
        rhox = (-3.0 * r[0] + 4.0 * r[1] - r[2]) * 0.5;
        ux = (-3.0 * u[0] + 4.0 * u[1] - u[2]) * 0.5;
        vx = (-3.0 * v[0] + 4.0 * v[1] - v[2]) * 0.5;
        lx1 = (u[0] - cs) * (cs * cs * rhox - cs * r[0] * ux);
        lx2 = u[0] < 0.0 ? u[0] * vx : 0.0;
        lx3 = 0.0;
    } else {
        rhox = (3.0 * r[0] - 4.0 * r[1] + r[2]) * 0.5;
        ux = (3.0 * u[0] - 4.0 * u[1] + u[2]) * 0.5;
        vx = (3.0 * v[0] - 4.0 * v[1] + v[2]) * 0.5;
        lx1 = 0.0;
        lx2 = u[0] >= 0.0 ? u[0] * vx : 0.0;
        lx3 = (u[0] + cs) * (cs * cs * rhox + cs * r[0] * ux);
    }
    out[0] = r[0] - (lx1 + lx3) * 1.5;
    out[1] = u[0] + (lx1 - lx3) / (2.0 * cs * out[0]);
    out[2] = v[0] - lx2;
}

int orc_lodi_node(const orc_t *o, int32_t right, const double *r, const double *u, const double *v, double *out)
{
    if (!o || !r || !u || !v || !out) return LBM_E_INVALID;
    lodi_node(o->cs, right != 0, r, u, v, out);
    return 0;
}

static void boundary_condition1(orc_t *o)
{
    const int nx = o->nx, ny = o->ny;
    const double cs = o->cs;
    double *tr = (double *)malloc(sizeof(double) * (ny + 1));
    double *tu = (double *)malloc(sizeof(double) * (ny + 1));
    double *tv = (double *)malloc(sizeof(double) * (ny + 1));
    /* left (boundary_condition.f90:413-483) */
    for (int j = 1; j <= ny; j++) {
        tr[j] = o->rho[IDX(1, j)];
        tu[j] = o->u[IDX(1, j)];
        tv[j] = o->v[IDX(1, j)];
    }
    for (int j = 2; j <= ny - 1; j++) {
        const double r[3] = {tr[j], o->rho[IDX(2, j)], o->rho[IDX(3, j)]};
        const double uu[3] = {tu[j], o->u[IDX(2, j)], o->u[IDX(3, j)]};
        const double vv[3] = {tv[j], o->v[IDX(2, j)], o->v[IDX(3, j)]};
        double out[3];
        lodi_node(cs, 0, r, uu, vv, out);
        size_t id = IDX(1, j);
        o->rho[id] = out[0];
        o->u[id] = out[1];
        o->v[id] = out[2];
    }
    for (int j = 2; j <= ny - 1; j++) bc_extrapolate(o, 1, j, 2, j);
    /* right (boundary_condition.f90:495-567) */
    for (int j = 1; j <= ny; j++) {
        tr[j] = o->rho[IDX(nx, j)];
        tu[j] = o->u[IDX(nx, j)];
        tv[j] = o->v[IDX(nx, j)];
    }
    for (int j = 2; j <= ny - 1; j++) {
        const double r[3] = {tr[j], o->rho[IDX(nx - 1, j)], o->rho[IDX(nx - 2, j)]};
        const double uu[3] = {tu[j], o->u[IDX(nx - 1, j)], o->u[IDX(nx - 2, j)]};
        const double vv[3] = {tv[j], o->v[IDX(nx - 1, j)], o->v[IDX(nx - 2, j)]};
        double out[3];
        lodi_node(cs, 1, r, uu, vv, out);
        size_t id = IDX(nx, j);
        o->rho[id] = out[0];
        o->u[id] = out[1];
        o->v[id] = out[2];
    }
    for (int j = 2; j <= ny - 1; j++) bc_extrapolate(o, nx, j, nx - 1, j);
    /* top wall (boundary_condition.f90:657-723) */
    for (int i = 1; i <= nx; i++) {
        size_t id = IDX(i, ny);
        o->rho[id] = o->rho[IDX(i, ny - 1)];
        o->u[id] = 0.0;
        o->v[id] = 0.0;
        bc_extrapolate(o, i, ny, i, ny - 1);
    }
    /* bottom wall (boundary_condition.f90:784-850) */
    for (int i = 1; i <= nx; i++) {
        size_t id = IDX(i, 1);
        o->rho[id] = o->rho[IDX(i, 2)];
        o->u[id] = 0.0;
        o->v[id] = 0.0;
        bc_extrapolate(o, i, 1, i, 2);
    }
    free(tr);
    free(tu);
    free(tv);
}

/* -------------------------------------------------------------------- IBM */

typedef struct { int idx; double val; } spent;

static int cmp_spent(const void *a, const void *b)
{
    int x = ((const spent *)a)->idx, y = ((const spent *)b)->idx;
    return (x > y) - (x < y);
}

/* LU with partial pivoting (first maximal |pivot|) and a 2-column solve; the
 * math contract of cusolverDnDgetrf/getrs (boundary_condition.f90:1631-1638). */
static int lu_solve(double *a, int n, double *b, int nrhs)
{
    int *piv = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    for (int j = 0; j < n; j++) {
        int p = j;
        double mx = fabs(a[j + (size_t)j * n]);
        for (int i = j + 1; i < n; i++)
            if (fabs(a[i + (size_t)j * n]) > mx) { mx = fabs(a[i + (size_t)j * n]); p = i; }
        piv[j] = p;
        if (p != j)
            for (int c = 0; c < n; c++) {
                double t = a[j + (size_t)c * n];
                a[j + (size_t)c * n] = a[p + (size_t)c * n];
                a[p + (size_t)c * n] = t;
            }
        double d = a[j + (size_t)j * n];
        if (d == 0.0) { free(piv); return -1; }
        for (int i = j + 1; i < n; i++) a[i + (size_t)j * n] /= d;
        for (int c = j + 1; c < n; c++) {
            double ajc = a[j + (size_t)c * n];
            if (ajc != 0.0)
                for (int i = j + 1; i < n; i++) a[i + (size_t)c * n] -= a[i + (size_t)j * n] * ajc;
        }
    }
    for (int r = 0; r < nrhs; r++) {
        double *x = b + (size_t)r * n;
        for (int j = 0; j < n; j++)
            if (piv[j] != j) { double t = x[j]; x[j] = x[piv[j]]; x[piv[j]] = t; }
        for (int j = 0; j < n; j++)
            for (int i = j + 1; i < n; i++) x[i] -= a[i + (size_t)j * n] * x[j];
        for (int j = n - 1; j >= 0; j--) {
            x[j] /= a[j + (size_t)j * n];
            for (int i = 0; i < j; i++) x[i] -= a[i + (size_t)j * n] * x[j];
        }
    }
    free(piv);
    return 0;
}

/* immersed_boundary_Wu (boundary_condition.f90:1512-1692), smooth=1. */
static int immersed_boundary(orc_t *o)
{
    const int nseg = o->nseg, num = o->adj_n;
    const double scale = o->scale, UU = o->uu;
    for (int w = 0; w < MAXW; w++) {
        o->wing_force[w][0] = o->wing_force[w][1] = 0.0;
        o->wing_moment[w] = 0.0;
    }
    if (nseg == 0) return 0;
    double *ub = (double *)malloc(sizeof(double) * nseg);
    double *vb = (double *)malloc(sizeof(double) * nseg);
    for (int w = 0; w < o->W; w++) { /* :1540-1555 */
        int ss = w * o->SPW, ps = w * o->P;
        for (int i = ss; i < ss + o->SPW; i++) {
            double xs = o->xb0[ps + i - ss + 1] + o->xb0[ps + i - ss];
            double ys = o->yb0[ps + i - ss + 1] + o->yb0[ps + i - ss];
            double om = o->omega[w], an = o->angle[w];
            ub[i] = UU * (-om * sin(an) * 0.5 * xs + om * cos(an) * 0.5 * ys + o->sau[w]) +
                    o->body_extra_v[w];
            vb[i] = UU * (-om * cos(an) * 0.5 * xs - om * sin(an) * 0.5 * ys + o->sbv[w]);
        }
    }
    /* adjoint index of each node, to address xx1(k,i) by lattice position */
    int *amap = (int *)malloc(sizeof(int) * o->np);
    for (size_t q = 0; q < o->np; q++) amap[q] = -1;
    for (int i = 0; i < num; i++) amap[IDX(o->adj_i[i], o->adj_j[i])] = i;
    /* xx1(k,i) = phi_x*phi_y (:1558-1606): nonzeros lie in the support box
     * of segment k; kept sorted by adjoint index. */
    spent *rows = (spent *)malloc(sizeof(spent) * 36 * (size_t)nseg);
    int *rlen = (int *)calloc(nseg, sizeof(int));
    for (int k = 0; k < nseg; k++) {
        int i1 = o->box[4 * k], i2 = o->box[4 * k + 1], j1 = o->box[4 * k + 2], j2 = o->box[4 * k + 3];
        int c = 0;
        for (int i = i1; i <= i2; i++)
            for (int j = j1; j <= j2; j++) {
                double x0 = o->ox + scale * (double)(i - 1);
                double y0 = o->oy + scale * (double)(j - 1);
                double k1 = orc_delta_kernel(fabs((x0 - o->xb2[k]) / scale));
                double k2 = orc_delta_kernel(fabs((y0 - o->yb2[k]) / scale));
                if (c >= 36) continue;
                rows[36 * k + c].idx = amap[IDX(i, j)];
                rows[36 * k + c].val = k1 * k2;
                c++;
            }
        qsort(&rows[36 * k], c, sizeof(spent), cmp_spent);
        rlen[k] = c;
    }
    /* A = xx1 * xx2 (DGEMM, :1619): A(k,l) = sum_i xx1(k,i) * (xx1(l,i)*arc(l)) */
    double *A = o->sysA;
    memset(A, 0, sizeof(double) * (size_t)nseg * nseg);
    for (int k = 0; k < nseg; k++)
        for (int l = 0; l < nseg; l++) {
            const int *bk = &o->box[4 * k], *bl = &o->box[4 * l];
            if (bk[1] < bl[0] || bl[1] < bk[0] || bk[3] < bl[2] || bl[3] < bk[2]) continue;
            const spent *rk = &rows[36 * k], *rl = &rows[36 * l];
            int a = 0, b = 0;
            double sum = 0.0;
            while (a < rlen[k] && b < rlen[l]) {
                if (rk[a].idx < rl[b].idx) a++;
                else if (rk[a].idx > rl[b].idx) b++;
                else { sum = sum + rk[a].val * (rl[b].val * o->arc[l]); a++; b++; }
            }
            A[k + (size_t)l * nseg] = sum;
        }
    /* ustar (:1609-1616), bu/bv = xx1 * ustar (DGEMV, :1620-1621), comj (:1624-1629) */
    double *us = (double *)malloc(sizeof(double) * (num > 0 ? num : 1));
    double *vs = (double *)malloc(sizeof(double) * (num > 0 ? num : 1));
    for (int i = 0; i < num; i++) {
        size_t id = IDX(o->adj_i[i], o->adj_j[i]);
        us[i] = o->u[id];
        vs[i] = o->v[id];
        o->ibm_us[i] = us[i];
        o->ibm_vs[i] = vs[i];
    }
    double *B = o->sysB;
    for (int k = 0; k < nseg; k++) {
        double su = 0.0, sv = 0.0;
        for (int c = 0; c < rlen[k]; c++) {
            su = su + rows[36 * k + c].val * us[rows[36 * k + c].idx];
            sv = sv + rows[36 * k + c].val * vs[rows[36 * k + c].idx];
        }
        B[k] = ub[k] - su;
        B[nseg + k] = vb[k] - sv;
    }
    double *LU = (double *)malloc(sizeof(double) * (size_t)nseg * nseg);
    memcpy(LU, A, sizeof(double) * (size_t)nseg * nseg);
    double *X = o->sysX;
    memcpy(X, B, sizeof(double) * 2 * (size_t)nseg);
    int rc = lu_solve(LU, nseg, X, 2);
    free(LU);
    o->have_sys = 1;
    if (rc != 0) {
        free(ub); free(vb); free(amap); free(rows); free(rlen); free(us); free(vs);
        return fail(o, LBM_E_NUMERIC, "IBM matrix is singular");
    }
    /* glx/gly = xx2 * comj (DGEMV, :1639-1640): g(i) = sum_k xx1(k,i)*arc(k)*comj(k)
     * summed in ascending k */
    double *gx = (double *)calloc(num > 0 ? num : 1, sizeof(double));
    double *gy = (double *)calloc(num > 0 ? num : 1, sizeof(double));
    for (int k = 0; k < nseg; k++)
        for (int c = 0; c < rlen[k]; c++) {
            int i = rows[36 * k + c].idx;
            double x2 = rows[36 * k + c].val * o->arc[k];
            gx[i] = gx[i] + x2 * X[k];
            gy[i] = gy[i] + x2 * X[nseg + k];
        }
    /* spread (:1645-1659) and per-wing force / moment (:1662-1675) */
    for (int i = 0; i < num; i++) {
        size_t id = IDX(o->adj_i[i], o->adj_j[i]);
        o->u[id] = us[i] + gx[i];
        o->v[id] = vs[i] + gy[i];
        o->g1[id] = 2.0 * o->rho[id] * gx[i];
        o->g2[id] = 2.0 * o->rho[id] * gy[i];
    }
    for (int i = 0; i < num; i++) {
        int w = o->adj_w[i] - 1;
        size_t id = IDX(o->adj_i[i], o->adj_j[i]);
        double t1 = 2.0 * o->rho[id] * gx[i];
        double t2 = 2.0 * o->rho[id] * gy[i];
        o->wing_force[w][0] = o->wing_force[w][0] - t1;
        o->wing_force[w][1] = o->wing_force[w][1] - t2;
        double x0 = (o->ox + scale * (double)(o->adj_i[i] - 1) -
                     (o->cfg.xc + o->cfg.wing_x0 + (double)w * o->cfg.wing_dx + o->sa[w])) / scale -
                    o->body_extra_pos[w];
        double y0 = (o->oy + scale * (double)(o->adj_j[i] - 1) - (o->cfg.yc + o->cfg.wing_y0 + o->sb[w])) / scale;
        o->wing_moment[w] = o->wing_moment[w] - x0 * t2 + y0 * t1;
    }
    free(ub); free(vb); free(amap); free(rows); free(rlen); free(us); free(vs);
    free(gx); free(gy);
    if (o->dyn.enable && o->W > 0) { /* :1677-1719, then update_body_extra_from_rel (:1731) */
        double fx = 0.0, st[LBM_BODY_REL_FIELDS];
        for (int w = 0; w < o->W; w++) fx = fx + o->wing_force[w][0];
        for (int q = 0; q < LBM_BODY_REL_FIELDS; q++) st[q] = o->rel[q][0];
        orc_body_step(&o->dyn, scale, UU, fx, o->wing_moment[0], st);
        for (int q = 0; q < LBM_BODY_REL_FIELDS; q++) o->rel[q][0] = st[q];
        update_body_extra(o);
    }
    return 0;
}

/* ------------------------------------------------------------- the step */

/* erranalysis (solution.f90:77-130), i outer / j inner. */
static void erranalysis(orc_t *o)
{
    double t1 = 0.0, t2 = 0.0;
    for (int i = 2; i <= o->nx - 1; i++)
        for (int j = 2; j <= o->ny - 1; j++) {
            size_t id = IDX(i, j);
            double du = o->u[id] - o->u0[id], dv = o->v[id] - o->v0[id];
            t1 = t1 + du * du + dv * dv;
            t2 = t2 + o->u[id] * o->u[id] + o->v[id] * o->v[id];
        }
    o->err = sqrt(t1) / (sqrt(t2) + 1e-30);
}

/* move_mesh (move_mesh.f90:17-58 dr=+1, 171-214 dr=-1), single rank. */
static void move_mesh(orc_t *o, int dr)
{
    const int nx = o->nx, ny = o->ny;
    if (dr == 1) {
        for (int j = 1; j <= ny; j++)
            for (int k = 1; k <= nx - 1; k++) {
                for (int q = 1; q <= 9; q++) FK(o->f0, k, j, q) = FK(o->f, k + 1, j, q);
                o->u0[IDX(k, j)] = o->u[IDX(k + 1, j)];
                o->v0[IDX(k, j)] = o->v[IDX(k + 1, j)];
                o->pre[IDX(k, j)] = o->rho[IDX(k + 1, j)];
            }
        for (int j = 1; j <= ny; j++)
            for (int k = 1; k <= nx - 1; k++) {
                for (int q = 1; q <= 9; q++) FK(o->f, k, j, q) = FK(o->f0, k, j, q);
                o->u[IDX(k, j)] = o->u0[IDX(k, j)];
                o->v[IDX(k, j)] = o->v0[IDX(k, j)];
                o->rho[IDX(k, j)] = o->pre[IDX(k, j)];
            }
    } else {
        for (int j = 1; j <= ny; j++)
            for (int k = 2; k <= nx; k++) {
                for (int q = 1; q <= 9; q++) FK(o->f0, k, j, q) = FK(o->f, k - 1, j, q);
                o->u0[IDX(k, j)] = o->u[IDX(k - 1, j)];
                o->v0[IDX(k, j)] = o->v[IDX(k - 1, j)];
                o->pre[IDX(k, j)] = o->rho[IDX(k - 1, j)];
            }
        for (int j = 1; j <= ny; j++)
            for (int k = 2; k <= nx; k++) {
                for (int q = 1; q <= 9; q++) FK(o->f, k, j, q) = FK(o->f0, k, j, q);
                o->u[IDX(k, j)] = o->u0[IDX(k, j)];
                o->v[IDX(k, j)] = o->v0[IDX(k, j)];
                o->rho[IDX(k, j)] = o->pre[IDX(k, j)];
            }
    }
}

/* force_row assembly (solution.f90:330-378) and airfoil_motion row
 * (:390-401) of step n. */
static void force_rows(orc_t *o, double time, double *row, double *mrow)
{
    const double scale = o->scale, UU = o->uu;
    double F0 = 0.0, F1 = 0.0, Mt = 0.0;
    for (int w = 0; w < o->W; w++) {
        o->wing_force[w][0] = 2.0 * o->wing_force[w][0] / (UU * UU * (1.0 / scale));
        o->wing_force[w][1] = 2.0 * o->wing_force[w][1] / (UU * UU * (1.0 / scale));
        o->wing_moment[w] = -2.0 * o->wing_moment[w] / (UU * UU * (1.0 / (scale * scale)));
    }
    for (int w = 0; w < o->W; w++) {
        F0 = F0 + o->wing_force[w][0];
        F1 = F1 + o->wing_force[w][1];
    }
    for (int w = 0; w < o->W; w++) Mt = Mt + o->wing_moment[w];
    int col = 0;
    row[col++] = time;
    for (int w = 0; w < o->W; w++) {
        row[col++] = o->wing_force[w][0];
        row[col++] = o->wing_force[w][1];
        row[col++] = o->wing_moment[w];
    }
    row[col++] = F0;
    row[col++] = F1;
    row[col++] = Mt;
    for (int w = 0; w < o->W; w++) {
        row[col++] = o->cfg.wing_x0 + (double)w * o->cfg.wing_dx + o->sa[w] +
                     o->body_extra_s[w] / (1.0 / scale);
        row[col++] = o->sau[w];
        row[col++] = o->saac[w];
        row[col++] = o->cfg.wing_y0 + o->sb[w];
        row[col++] = o->sbv[w];
        row[col++] = o->sbac[w];
        row[col++] = o->angle[w];
        row[col++] = o->omega[w];
        row[col++] = o->omgac[w];
    }
    if (mrow) {
        int c = 0;
        mrow[c++] = time;
        for (int w = 0; w < o->W; w++) {
            mrow[c++] = 0.0;
            mrow[c++] = o->body_extra_v[w] / UU;
            mrow[c++] = o->cfg.wing_x0 + (double)w * o->cfg.wing_dx + o->sa[w] +
                        o->body_extra_s[w] / (1.0 / scale);
        }
    }
}

/* init (solution.f90:1-75) + read_boundary + the rigid-body reset of
 * reset_zmq_episode (driver.f90:205-258). */
int orc_reset(orc_t *o)
{
    if (!o) return LBM_E_INVALID;
    if (o->cfg.motion_source == LBM_MOTION_STREAM) { /* driver.f90:120-124 */
        for (int c = 0; c < LBM_MAX_MOTION_COLS; c++) o->stream_start[c] = o->stream_target[c] = 0.0;
        o->stream_ready = 1;
    }
    for (int w = 0; w < MAXW; w++) {
        for (int q = 0; q < LBM_BODY_REL_FIELDS; q++) o->rel[q][w] = 0.0;
        o->wing_force[w][0] = o->wing_force[w][1] = 0.0;
        o->wing_moment[w] = 0.0;
    }
    update_body_extra(o);
    const double UU = o->uu;
    double feq[9];
    for (int k = 0; k < 9; k++) feq[k] = d2q9feq(o, 1.0, UU, 0.0, k);
    for (size_t q = 0; q < o->np; q++) {
        o->u[q] = UU;
        o->u0[q] = UU;
        o->v[q] = 0.0;
        o->v0[q] = 0.0;
        o->rho[q] = 1.0;
        o->pre[q] = 1.0 / 3.0;
        o->g1[q] = 0.0;
        o->g2[q] = 0.0;
        o->body[q] = 0;
    }
    for (int k = 0; k < 9; k++)
        for (size_t q = 0; q < o->np; q++) {
            o->f[(size_t)k * o->np + q] = feq[k];
            o->f0[(size_t)k * o->np + q] = feq[k];
        }
    o->adj_n = 0;
    o->movenum = 0;
    o->flag2 = 1;
    o->n = 1;
    o->err = 0.0;
    o->have_sys = 0;
    return 0;
}

/* one pass of the run loop (main.f90:374-412 / driver.f90:369-408). */
static int one_step(orc_t *o, double *frow, double *mrow)
{
    int n = o->n;
    double time = (double)(n - 1) * o->dt;
    if (n >= 2) erranalysis(o);
    /* evolution (solution.f90:272-422) */
    int rc = o->W > 0 ? mode(o, time) : 0; /* no wings: the motion source is unused */
    if (rc) return rc;
    rc = move_boundary(o);
    if (rc) return rc;
    streaming(o);
    physical_variables(o);
    rc = immersed_boundary(o);
    if (rc) return rc;
    colission(o);
    force_rows(o, time, frow, mrow);
    boundary_condition1(o);
    /* move_mesh check (main.f90:393-412) */
    if (o->W > 0 && o->rel[REL_POS][0] >= 1.0) {
        move_mesh(o, 1);
        o->rel[REL_POS][0] = o->rel[REL_POS][0] - 1.0;
        update_body_extra(o);
        o->movenum = o->movenum + 1;
        o->flag2 = 3;
    } else if (o->W > 0 && o->rel[REL_POS][0] <= -1.0) {
        move_mesh(o, -1);
        o->rel[REL_POS][0] = o->rel[REL_POS][0] + 1.0;
        update_body_extra(o);
        o->movenum = o->movenum - 1;
        o->flag2 = 3;
    }
    if (o->flag2 > 0) o->flag2 = o->flag2 - 1;
    o->n = n + 1;
    return 0;
}

int orc_advance(orc_t *o, int32_t n_last, double *force_rows_out, double *motion_rows,
                double *row_sum, double *last_row, double *err)
{
    if (!o) return LBM_E_INVALID;
    const int rl = o->row_len, ml = 1 + 3 * o->W;
    double row[LBM_MAX_FORCE_ROW], mrow[1 + 3 * MAXW];
    if (row_sum)
        for (int c = 0; c < rl; c++) row_sum[c] = 0.0;
    int cnt = 0;
    while (o->n <= n_last) {
        int rc = one_step(o, row, mrow);
        if (rc) return rc;
        if (force_rows_out) memcpy(force_rows_out + (size_t)cnt * rl, row, sizeof(double) * rl);
        if (motion_rows) memcpy(motion_rows + (size_t)cnt * ml, mrow, sizeof(double) * ml);
        if (row_sum)
            for (int c = 0; c < rl; c++) row_sum[c] = row_sum[c] + row[c];
        if (last_row) memcpy(last_row, row, sizeof(double) * rl);
        cnt++;
    }
    if (err) *err = o->err;
    return 0;
}

int orc_get_state(orc_t *o, int32_t *n_next, int32_t *movenum, int32_t *flag2,
                  int32_t *row_len, int32_t *adjoint_n)
{
    if (!o) return LBM_E_INVALID;
    if (n_next) *n_next = o->n;
    if (movenum) *movenum = o->movenum;
    if (flag2) *flag2 = o->flag2;
    if (row_len) *row_len = o->row_len;
    if (adjoint_n) *adjoint_n = o->adj_n;
    return 0;
}

int orc_get_fields(orc_t *o, double *rho, double *u, double *v, double *pre)
{
    if (!o) return LBM_E_INVALID;
    size_t b = o->np * sizeof(double);
    if (rho) memcpy(rho, o->rho, b);
    if (u) memcpy(u, o->u, b);
    if (v) memcpy(v, o->v, b);
    if (pre)
        for (size_t q = 0; q < o->np; q++) pre[q] = o->rho[q] / 3.0;
    return 0;
}

int orc_get_populations(orc_t *o, double *f, double *f0)
{
    if (!o) return LBM_E_INVALID;
    if (f) memcpy(f, o->f, 9 * o->np * sizeof(double));
    if (f0) memcpy(f0, o->f0, 9 * o->np * sizeof(double));
    return 0;
}

/* Total mass: the sum over every node of its nine populations f (no
 * counterpart in the reference, which has no conservation check; the
 * lattice-wide FP64 reduction follows the pattern of erranalysis,
 * solution.f90:77-130).  Compensated (Neumaier) summation, so the value is
 * the exact sum to a few ulp whatever the lattice size. */
int orc_get_mass(orc_t *o, double *mass)
{
    if (!o || !mass) return LBM_E_INVALID;
    double s = 0.0, c = 0.0;
    for (size_t q = 0; q < o->np; q++)
        for (int k = 0; k < 9; k++) {
            double x = o->f[(size_t)k * o->np + q];
            double t = s + x;
            if (fabs(s) >= fabs(x))
                c += (s - t) + x;
            else
                c += (x - t) + s;
            s = t;
        }
    *mass = s + c;
    return 0;
}

int orc_set_populations(orc_t *o, const double *f, const double *f0)
{
    if (!o) return LBM_E_INVALID;
    if (f) memcpy(o->f, f, 9 * o->np * sizeof(double));
    if (f0) {
        const int nx = o->nx, ny = o->ny;
        for (int k = 1; k <= 9; k++)
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++)
                    if (i == 1 || i == nx || j == 1 || j == ny)
                        FK(o->f0, i, j, k) = f0[(size_t)(k - 1) * o->np + IDX(i, j)];
    }
    if (f)
        for (size_t q = 0; q < o->np; q++) {
            double tr = 0.0, tu = 0.0, tv = 0.0;
            for (int k = 0; k < 9; k++) {
                double t = o->f[(size_t)k * o->np + q];
                tr = tr + t;
                tu = tu + (double)o->e[k][0] * t;
                tv = tv + (double)o->e[k][1] * t;
            }
            o->rho[q] = tr;
            o->u[q] = tu / tr;
            o->v[q] = tv / tr;
        }
    return 0;
}

int orc_get_markers(orc_t *o, double *xb1, double *yb1, double *xb2, double *yb2)
{
    if (!o) return LBM_E_INVALID;
    if (xb1) memcpy(xb1, o->xb1, o->npts * sizeof(double));
    if (yb1) memcpy(yb1, o->yb1, o->npts * sizeof(double));
    if (xb2) memcpy(xb2, o->xb2, o->nseg * sizeof(double));
    if (yb2) memcpy(yb2, o->yb2, o->nseg * sizeof(double));
    return 0;
}

int orc_get_ibm(orc_t *o, int32_t capacity, int32_t *ij, int32_t *wing_id,
                double *wing_force, double *wing_moment)
{
    if (!o) return LBM_E_INVALID;
    int n = o->adj_n < capacity ? o->adj_n : capacity;
    for (int i = 0; i < n; i++) {
        if (ij) {
            ij[2 * i] = o->adj_i[i];
            ij[2 * i + 1] = o->adj_j[i];
        }
        if (wing_id) wing_id[i] = o->adj_w[i];
    }
    for (int w = 0; w < MAXW; w++) {
        if (wing_force) {
            wing_force[2 * w] = o->wing_force[w][0];
            wing_force[2 * w + 1] = o->wing_force[w][1];
        }
        if (wing_moment) wing_moment[w] = o->wing_moment[w];
    }
    return 0;
}

int orc_get_ibm_system(orc_t *o, double *a, double *rhs, double *sol)
{
    if (!o || !o->have_sys) return LBM_E_STATE;
    size_t n = (size_t)o->nseg;
    if (a) memcpy(a, o->sysA, n * n * sizeof(double));
    if (rhs) memcpy(rhs, o->sysB, 2 * n * sizeof(double));
    if (sol) memcpy(sol, o->sysX, 2 * n * sizeof(double));
    return 0;
}

/* Restart hook (readflowfield, solution.f90:1721-1851, main.f90:343-349):
 * next step, frame shift, flag2 = 1. */
int orc_set_step(orc_t *o, int32_t n_next, int32_t movenum)
{
    if (!o || n_next < 1) return LBM_E_INVALID;
    o->n = n_next;
    o->movenum = movenum;
    o->flag2 = 1;
    update_body_extra(o);
    return 0;
}

int orc_set_body_rel_pos(orc_t *o, int32_t wing, double pos)
{
    if (!o || wing < 1 || wing > o->W) return LBM_E_INVALID;
    o->rel[REL_POS][wing - 1] = pos;
    update_body_extra(o);
    return 0;
}

int orc_get_body_rel(orc_t *o, double *rel)
{
    if (!o || !rel) return LBM_E_INVALID;
    for (int q = 0; q < LBM_BODY_REL_FIELDS; q++)
        for (int w = 0; w < o->W; w++) rel[q * o->W + w] = o->rel[q][w];
    return 0;
}

/* readflowfield's body_rel_* (solution.f90:1721-1851) */
int orc_set_body_rel(orc_t *o, const double *rel)
{
    if (!o || !rel) return LBM_E_INVALID;
    for (int q = 0; q < LBM_BODY_REL_FIELDS; q++)
        for (int w = 0; w < o->W; w++) o->rel[q][w] = rel[q * o->W + w];
    update_body_extra(o);
    return 0;
}

int orc_set_dynamics(orc_t *o, const lbm_dynamics *d)
{
    if (!o || !d) return LBM_E_INVALID;
    if (d->enable && o->W == 0) return fail(o, LBM_E_INVALID, "rigid-body coupling needs at least one wing");
    if (d->enable && !(d->solid_den > 0.0 && (!d->closed_boundary || d->area > 0.0)))
        return fail(o, LBM_E_INVALID, "solid_den and area must be positive");
    o->dyn = *d;
    return 0;
}

/* rk4 (boundary_condition.f90:3168-3202): s1..a1 are wing 1's s_prev, s,
 * pos, v_prev, v, a; jd1..jsd its theta_prev, theta, temp_theta_prev,
 * temp_theta; mom the moment of wing 1.  The k terms are evaluated in the
 * reference's order, which reuses the same names for the pitch equation. */
static void rk4(double s1, double *s, double *s2, double v1, double *v, double a1, double jd1, double *jd,
                double jsd1, double *jsd, double p1, double p2, double alphad, double mom)
{
    double k11 = v1;
    double k21 = a1;
    double k12 = v1 + 0.5 * k21;
    double k22 = a1;
    double k13 = v1 + 0.5 * k22;
    double k23 = a1;
    double k14 = v1 + k23;
    double k24 = a1;
    *s = s1 + (k11 + 2.0 * k12 + 2.0 * k13 + k14) / 6.0;
    *v = v1 + (k21 + 2.0 * k22 + 2.0 * k23 + k24) / 6.0;
    *s2 = *s2 + (k11 + 2.0 * k12 + 2.0 * k13 + k14) / 6.0;
    k11 = jsd1;
    k21 = p1 * (jd1 + alphad) + p2 * mom;
    k12 = jsd1 + 0.5 * k21;
    k22 = p1 * (jd1 + alphad + 0.5 * k11) + p2 * mom;
    k13 = jsd1 + 0.5 * k22;
    k23 = p1 * (jd1 + alphad + 0.5 * k12) + p2 * mom;
    k14 = jsd1 + k23;
    k24 = p1 * (jd1 + alphad + k13) + p2 * mom;
    *jd = jd1 + (k11 + 2.0 * k12 + 2.0 * k13 + k14) / 6.0;
    *jsd = jsd1 + (k21 + 2.0 * k22 + 2.0 * k23 + k24) / 6.0;
}

/* The dynamics block of immersed_boundary_Wu (boundary_condition.f90:
 * 1694-1719) on wing 1's state st (11 fields, restart order): fx is force(1),
 * the raw x force summed over the wings (:1677-1680), mom wing_moment(1). */
void orc_body_step(const lbm_dynamics *d, double scale, double uu, double fx, double mom, double *st)
{
    st[REL_A_PREV] = st[REL_A];
    if (!d->closed_boundary)
        st[REL_A] = fx / (d->solid_den / (scale * scale));
    else
        st[REL_A] = fx / (d->solid_den * d->area / (scale * scale));
    st[REL_V_PREV] = st[REL_V];
    st[REL_S_PREV] = st[REL_S];
    st[REL_THETA_PREV] = st[REL_THETA];
    st[REL_TEMP_THETA_PREV] = st[REL_TEMP_THETA];
    if (!d->rk4) {
        st[REL_V] = st[REL_A] + st[REL_V];
        st[REL_S] = st[REL_S] + st[REL_V_PREV];
        st[REL_POS] = st[REL_POS] + st[REL_V_PREV];
        st[REL_THETA] = st[REL_THETA_PREV] + st[REL_TEMP_THETA_PREV];
        st[REL_TEMP_THETA] = st[REL_TEMP_THETA_PREV] +
                             d->par1 * uu * uu * scale * scale * (st[REL_THETA] + d->alphad) +
                             d->par2 * (scale * scale * scale * scale) * mom;
    } else {
        rk4(st[REL_S_PREV], &st[REL_S], &st[REL_POS], st[REL_V_PREV], &st[REL_V], st[REL_A], st[REL_THETA_PREV],
            &st[REL_THETA], st[REL_TEMP_THETA_PREV], &st[REL_TEMP_THETA], d->par1 * uu * uu * scale * scale,
            d->par2 * (scale * scale * scale * scale), d->alphad, mom);
    }
}

/* ------------------------------------------------------------ test hooks */

/* colission of one node (solution.f90:606-650) with the instance's rates. */
int orc_collide_node(orc_t *o, const double *f, double rt, double ut, double vt,
                     double gfx, double gfy, double *out)
{
    if (!o || !f || !out) return LBM_E_INVALID;
    const double *s = o->s;
    double temp[9], temp1[9], fp[9];
    for (int k = 0; k < 9; k++) temp[k] = f[k];
    matmulvoc(temp1, o->matm, temp);
    temp[0] = 0.0;
    temp[1] = -2.0 * rt + 3.0 * (rt * ut * rt * ut + rt * vt * rt * vt);
    temp[2] = rt - 3.0 * (rt * ut * rt * ut + rt * vt * rt * vt);
    temp[3] = 0.0;
    temp[4] = -rt * ut;
    temp[5] = 0.0;
    temp[6] = -rt * vt;
    temp[7] = (rt * ut * rt * ut - rt * vt * rt * vt);
    temp[8] = (rt * ut * rt * vt);
    fp[0] = 0.0;
    fp[1] = 6.0 * (1.0 - 0.5 * s[1]) * (ut * gfx + vt * gfy);
    fp[2] = -6.0 * (1.0 - 0.5 * s[2]) * (ut * gfx + vt * gfy);
    fp[3] = gfx;
    fp[4] = -(1.0 - 0.5 * s[4]) * gfx;
    fp[5] = gfy;
    fp[6] = -(1.0 - 0.5 * s[6]) * gfy;
    fp[7] = 2.0 * (1.0 - 0.5 * s[7]) * (ut * gfx - vt * gfy);
    fp[8] = (1.0 - 0.5 * s[8]) * (ut * gfy + vt * gfx);
    for (int k = 0; k < 9; k++) temp[k] = temp1[k] + (temp[k] - temp1[k]) * s[k] + fp[k];
    matmulvoc(out, o->minv, temp);
    return 0;
}

/* far-field extrapolation of one boundary node (boundary_condition.f90:443-474). */
int orc_bc_node(orc_t *o, double rb, double ub, double vb, double rn, double un,
                double vn, const double *f0n, double *out)
{
    if (!o || !f0n || !out) return LBM_E_INVALID;
    double t1[9], t2[9], t3[9], tmp[9];
    meq_bc(rb, ub, vb, t1);
    meq_bc(rn, un, vn, t2);
    for (int k = 0; k < 9; k++) tmp[k] = f0n[k];
    matmulvoc(t3, o->matm, tmp);
    for (int k = 0; k < 9; k++) t3[k] = t1[k] + (1.0 - o->s[k]) * (t3[k] - t2[k]);
    matmulvoc(out, o->minv, t3);
    return 0;
}

/* velocity before (u*) and after the IBM correction and density at the
 * adjoint nodes of the last step, in list order. */
int orc_get_ibm_nodes(orc_t *o, int32_t capacity, double *us, double *vs, double *u,
                      double *v, double *rho)
{
    if (!o) return LBM_E_INVALID;
    int n = o->adj_n < capacity ? o->adj_n : capacity;
    for (int i = 0; i < n; i++) {
        size_t id = IDX(o->adj_i[i], o->adj_j[i]);
        if (us) us[i] = o->ibm_us[i];
        if (vs) vs[i] = o->ibm_vs[i];
        if (u) u[i] = o->u[id];
        if (v) v[i] = o->v[id];
        if (rho) rho[i] = o->rho[id];
    }
    return n;
}

/* relaxation rates actually used (s[7], s[8] resolved from tau). */
int orc_get_rates(orc_t *o, double *s)
{
    if (!o || !s) return LBM_E_INVALID;
    for (int k = 0; k < 9; k++) s[k] = o->s[k];
    return 0;
}
