/*
 * fc_oracle.c -- CPU FP64 restatement of the FEICAR3D pressure-Poisson hot
 * path (oracle).  TEST INFRASTRUCTURE ONLY: the checker for tests/, smoke()
 * and bench.py's cpu_baseline leg; never linked into the product library.
 *
 * Sequential, FP64, single flow rank (yb1=zb1=0, yb2=ny+1, zb2=nz+1;
 * yc_start=zc_start=1, yc_end=ny-1, zc_end=nz-1; CORE_PARALLEL.F90:103-185).
 * Each routine cites the reference lines it restates.  Values the reference
 * leaves uninitialised (coarse-level x(0), x(n+1), dxinv(0), the bc corner
 * entries, first-call res ghosts) are defined here as: every array starts at
 * zero, and coarse-level ghost node coordinates are mirrored the way
 * make_grid mirrors the fine ones (CORE_GRID.F90:29-31).
 */
#include "fc_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXLEV 10

typedef struct {
    int nx, ny, nz;               /* node counts at this level            */
    long P, S, N;                 /* strides: j-stride, k-stride, size     */
    double *x, *y, *z, *xc, *yc, *zc;
    double *dxinv, *dyinv, *dzinv, *dxcinv, *dycinv, *dzcinv;
    double *rhs, *phi, *res;
    int8_t *ibk, *igk;
} olev;

struct or_state {
    fc_pson_params prm;
    int ordering;
    int nx, ny, nz;
    long P, S, N;
    /* level-1 grid (CORE_GRID.F90) */
    double *x, *y, *z, *xc, *yc, *zc, *dx, *dy, *dz;
    double *dxinv, *dyinv, *dzinv, *dxc, *dyc, *dzc, *dxcinv, *dycinv, *dzcinv;
    int nlev, mg[3];
    int *mgI, *mgJ, *mgK;
    olev lev[MAXLEV + 1];         /* 1-based like mg_array(n)             */
    int8_t *iblank, *ghost, *dead, *igmark;
    /* GCM lists (0-based copies of the reference's 1-based indices) */
    int nG, *iG, *jG, *kG, *iC, *jC, *kC;
    double *cN, *alpha, *dpdnG;
    int nD, *iD, *jD, *kD, *iDC, *jDC, *kDC;
    double *cDN, *plD, *dpdnD, *ipD;
    /* BC */
    double *pgx1, *pgx2, *pgy1, *pgy2, *pgz1, *pgz2, outflow_area;
    double *bcxp, *bcyp, *bczp;
    double *p_temp;
    /* projection: face interpolation weights (CORE_GRID.F90:184-205) and the
     * outer-face velocities bcxu, bcyv, bczw (CORE_MEMORY_ALLOCATE.F90:66-80) */
    double *fx, *fy, *fz;
    double *bcxu, *bcyv, *bczw;
    /* the tangential outer-face velocities of solve_ad's boundary conditions
     * (bcxv, bcxw, bcyu, bcyw, bczu, bczv), the face values and frequencies
     * of set_outer_velocity_bc (ux1, ux2, vx1, ..., wz2) and the channel
     * profile bounds yOrigin, yout */
    double *bcxv, *bcxw, *bcyu, *bcyw, *bczu, *bczv;
    double vbc_val[18], vbc_freq[18], y_origin, y_out;
    /* GCM_ghostcell_velocity lists: coeffGCMD, coeffGCMDeadD (8 each), the
     * body-intercept velocities (u, v, w per relation), probeLengthNormalizedD */
    double *cD, *cDD, *ubi, *ubiD, probe_norm_d;
    int have_gcm_vel;
    /* advection-diffusion: set_solve_ad coefficients (CORE_SET_SOLVERS.F90:
     * 591-676), half_dt of itsolv_ad (CORE_SOLVE_AD.F90:196-200), omega_ad */
    double *amx_ad, *apx_ad, *amy_ad, *apy_ad, *amz_ad, *apz_ad;
    double half_dt, omega_ad;
    double reinv_ad;              /* nu of set_solve_ad = viscTot (MAIN.F90:926) */
    double *reslap;               /* res_laplacian of the last GCM call   */
    const double *pprev_cur;      /* p of the current solve (dead fallback) */
    /* tdma scratch */
    double *am, *ac, *ap, *rr, *uu, *gam;
};

/* GCM_MEMORY_ALLOCATE.F90:593-623: stencil offsets of the 8-cell box */
static const int incI[8] = {0, 0, 1, 1, 0, 0, 1, 1};
static const int incJ[8] = {0, 1, 1, 0, 0, 1, 1, 0};
static const int incK[8] = {0, 0, 0, 0, 1, 1, 1, 1};

#define IDX(L, i, j, k) ((long)(i) + (L)->P * ((long)(j) + (long)((L)->ny + 2) * (long)(k)))
#define FID(s, i, j, k) ((long)(i) + (s)->P * ((long)(j) + (long)((s)->ny + 2) * (long)(k)))

static void *zalloc(size_t n) { void *p = calloc(n ? n : 1, 1); if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); } return p; }
#define DALLOC(n) ((double *)zalloc(sizeof(double) * (size_t)(n)))
#define IALLOC(n) ((int *)zalloc(sizeof(int) * (size_t)(n)))
#define BALLOC(n) ((int8_t *)zalloc((size_t)(n)))

/* single-precision literals of the reference compared in FP64 */
static const double EPS_GUARD = (double)1.0e-12f;   /* CORE_MG.F90:690,786,881 */
static const double VOL_FRAC = (double)0.01f;       /* CORE_MG.F90:1168,1335   */

/* ------------------------------------------------------------------ */
/* tdma: CORE_TDMA.F90:10-34 (Thomas algorithm, no pivoting)           */
void or_tdma(const double *a, const double *b, const double *c, const double *r,
             double *u, int n1, int n2)
{
    double *gam = (double *)malloc(sizeof(double) * (size_t)(n2 + 2));
    double binv = 1.0 / b[n1];
    u[n1] = r[n1] * binv;
    for (int j = n1 + 1; j <= n2; ++j) {
        gam[j] = c[j - 1] * binv;
        binv = 1.0 / (b[j] - a[j] * gam[j]);
        u[j] = (r[j] - a[j] * u[j - 1]) * binv;
    }
    for (int j = n2 - 1; j >= n1; --j) u[j] = u[j] - gam[j + 1] * u[j + 1];
    free(gam);
}

static void tdma_s(or_state *s, int n1, int n2)
{
    double *a = s->am, *b = s->ac, *c = s->ap, *r = s->rr, *u = s->uu, *gam = s->gam;
    double binv = 1.0 / b[n1];
    u[n1] = r[n1] * binv;
    for (int j = n1 + 1; j <= n2; ++j) {
        gam[j] = c[j - 1] * binv;
        binv = 1.0 / (b[j] - a[j] * gam[j]);
        u[j] = (r[j] - a[j] * u[j - 1]) * binv;
    }
    for (int j = n2 - 1; j >= n1; --j) u[j] = u[j] - gam[j + 1] * u[j + 1];
}

/* ------------------------------------------------------------------ */
/* make_grid / metrics: CORE_GRID.F90:29-31 (ghost nodes given by the
 * caller), 106-150 (centres, spacings) */
static void grid_metrics(int n, const double *x, double *xc, double *dx, double *dxinv,
                         double *dxc, double *dxcinv)
{
    for (int i = 0; i <= n; ++i) {
        xc[i] = 0.5 * (x[i + 1] + x[i]);
        dx[i] = x[i + 1] - x[i];
        dxinv[i] = 1.0 / dx[i];
    }
    xc[n + 1] = x[n + 1];
    for (int i = 1; i <= n; ++i) {
        dxc[i] = xc[i] - xc[i - 1];
        dxcinv[i] = 1.0 / dxc[i];
    }
}

/* MG_initial level count: CORE_MG.F90:69-102 */
static int auto_levels(int n, int mg)
{
    if (mg == 0 || mg > MAXLEV) {
        int ii = n - 1, N;
        for (N = 1; N <= MAXLEV; ++N) {
            ii = ii / 2;
            if (ii >= 4) continue;
            break;
        }
        if (N > MAXLEV) N = MAXLEV;   /* DO loop exit value is max+1; clamp */
        return N;
    }
    return mg;
}

/* coarse node count: CORE_MG.F90:142-175 */
static int coarsen(int n) { return (n % 2 == 0) ? n / 2 + 1 : (n + 1) / 2; }

static void level_alloc(olev *L, int nx, int ny, int nz)
{
    L->nx = nx; L->ny = ny; L->nz = nz;
    L->P = nx + 2; L->S = (long)(nx + 2) * (ny + 2); L->N = L->S * (nz + 2);
    L->x = DALLOC(nx + 2); L->y = DALLOC(ny + 2); L->z = DALLOC(nz + 2);
    L->xc = DALLOC(nx + 2); L->yc = DALLOC(ny + 2); L->zc = DALLOC(nz + 2);
    L->dxinv = DALLOC(nx + 2); L->dyinv = DALLOC(ny + 2); L->dzinv = DALLOC(nz + 2);
    L->dxcinv = DALLOC(nx + 2); L->dycinv = DALLOC(ny + 2); L->dzcinv = DALLOC(nz + 2);
    L->rhs = DALLOC(L->N); L->phi = DALLOC(L->N); L->res = DALLOC(L->N);
    L->ibk = BALLOC(L->N); L->igk = BALLOC(L->N);
}

/* coarse metrics of one axis: CORE_MG.F90:234-330 */
static void coarse_axis(int nc, int nf, const double *xf, const double *xfine1_1, const double *xfine1_n,
                        double *x, double *xc, double *dxinv, double *dxcinv)
{
    for (int i = 1; i <= nc - 1; ++i) {
        int iim = (nc == nf) ? i : 2 * i - 1;
        x[i] = xf[iim];
    }
    x[nc] = xf[nf];
    /* builder-defined ghost nodes (left uninitialised by the reference) */
    x[0] = 2.0 * x[1] - x[2];
    x[nc + 1] = 2.0 * x[nc] - x[nc - 1];
    for (int i = 1; i <= nc - 1; ++i) {
        dxinv[i] = 1.0 / (x[i + 1] - x[i]);
        xc[i] = 0.5 * (x[i] + x[i + 1]);
    }
    dxinv[0] = 1.0 / (x[1] - x[0]);
    dxinv[nc] = 1.0 / (x[nc + 1] - x[nc]);
    xc[0] = 2.0 * (*xfine1_1) - xc[1];
    xc[nc] = 2.0 * (*xfine1_n) - xc[nc - 1];
    xc[nc + 1] = x[nc + 1];
    for (int i = 1; i <= nc; ++i) dxcinv[i] = 1.0 / (xc[i] - xc[i - 1]);
}

or_state *or_create(const fc_grid *g, const fc_pson_params *p, int ordering)
{
    or_state *s = (or_state *)zalloc(sizeof(or_state));
    s->prm = *p;
    s->ordering = ordering;
    int nx = g->nx, ny = g->ny, nz = g->nz;
    s->nx = nx; s->ny = ny; s->nz = nz;
    s->P = nx + 2; s->S = (long)(nx + 2) * (ny + 2); s->N = s->S * (nz + 2);
    s->x = DALLOC(nx + 2); s->y = DALLOC(ny + 2); s->z = DALLOC(nz + 2);
    memcpy(s->x, g->x, sizeof(double) * (nx + 2));
    memcpy(s->y, g->y, sizeof(double) * (ny + 2));
    memcpy(s->z, g->z, sizeof(double) * (nz + 2));
    s->xc = DALLOC(nx + 2); s->yc = DALLOC(ny + 2); s->zc = DALLOC(nz + 2);
    s->dx = DALLOC(nx + 2); s->dy = DALLOC(ny + 2); s->dz = DALLOC(nz + 2);
    s->dxinv = DALLOC(nx + 2); s->dyinv = DALLOC(ny + 2); s->dzinv = DALLOC(nz + 2);
    s->dxc = DALLOC(nx + 2); s->dyc = DALLOC(ny + 2); s->dzc = DALLOC(nz + 2);
    s->dxcinv = DALLOC(nx + 2); s->dycinv = DALLOC(ny + 2); s->dzcinv = DALLOC(nz + 2);
    grid_metrics(nx, s->x, s->xc, s->dx, s->dxinv, s->dxc, s->dxcinv);
    grid_metrics(ny, s->y, s->yc, s->dy, s->dyinv, s->dyc, s->dycinv);
    grid_metrics(nz, s->z, s->zc, s->dz, s->dzinv, s->dzc, s->dzcinv);

    /* MG_initial: CORE_MG.F90:69-105 */
    s->mg[0] = auto_levels(nx, p->mgrids_x);
    s->mg[1] = auto_levels(ny, p->mgrids_y);
    s->mg[2] = auto_levels(nz, p->mgrids_z);
    int ml = s->mg[0];
    if (s->mg[1] > ml) ml = s->mg[1];
    if (s->mg[2] > ml) ml = s->mg[2];
    s->nlev = ml;
    s->mgI = IALLOC(ml + 1); s->mgJ = IALLOC(ml + 1); s->mgK = IALLOC(ml + 1);
    s->mgI[1] = nx; s->mgJ[1] = ny; s->mgK[1] = nz;
    for (int n = 2; n <= ml; ++n) {
        s->mgI[n] = (n <= s->mg[0]) ? coarsen(s->mgI[n - 1]) : s->mgI[n - 1];
        s->mgJ[n] = (n <= s->mg[1]) ? coarsen(s->mgJ[n - 1]) : s->mgJ[n - 1];
        s->mgK[n] = (n <= s->mg[2]) ? coarsen(s->mgK[n - 1]) : s->mgK[n - 1];
    }
    for (int n = 1; n <= ml; ++n) level_alloc(&s->lev[n], s->mgI[n], s->mgJ[n], s->mgK[n]);
    /* level 1 copies: CORE_MG.F90:202-229 */
    olev *L1 = &s->lev[1];
    for (int i = 0; i <= nx; ++i) { L1->dxcinv[i] = s->dxcinv[i]; L1->dxinv[i] = s->dxinv[i]; L1->xc[i] = s->xc[i]; }
    for (int j = 0; j <= ny; ++j) { L1->dycinv[j] = s->dycinv[j]; L1->dyinv[j] = s->dyinv[j]; L1->yc[j] = s->yc[j]; }
    for (int k = 0; k <= nz; ++k) { L1->dzcinv[k] = s->dzcinv[k]; L1->dzinv[k] = s->dzinv[k]; L1->zc[k] = s->zc[k]; }
    for (int i = 0; i <= nx + 1; ++i) L1->x[i] = s->x[i];
    for (int j = 0; j <= ny + 1; ++j) L1->y[j] = s->y[j];
    for (int k = 0; k <= nz + 1; ++k) L1->z[k] = s->z[k];
    for (int n = 2; n <= ml; ++n) {
        olev *L = &s->lev[n], *F = &s->lev[n - 1];
        coarse_axis(L->nx, F->nx, F->x, &s->x[1], &s->x[nx], L->x, L->xc, L->dxinv, L->dxcinv);
        coarse_axis(L->ny, F->ny, F->y, &s->y[1], &s->y[ny], L->y, L->yc, L->dyinv, L->dycinv);
        coarse_axis(L->nz, F->nz, F->z, &s->z[1], &s->z[nz], L->z, L->zc, L->dzinv, L->dzcinv);
    }
    s->iblank = BALLOC(s->N); s->ghost = BALLOC(s->N); s->dead = BALLOC(s->N); s->igmark = BALLOC(s->N);
    s->pgx1 = DALLOC((ny + 2) * (nz + 2)); s->pgx2 = DALLOC((ny + 2) * (nz + 2));
    s->pgy1 = DALLOC((nx + 2) * (nz + 2)); s->pgy2 = DALLOC((nx + 2) * (nz + 2));
    s->pgz1 = DALLOC((nx + 2) * (ny + 2)); s->pgz2 = DALLOC((nx + 2) * (ny + 2));
    s->bcxp = DALLOC(2 * (ny + 2) * (nz + 2));
    s->bcyp = DALLOC((nx + 2) * 2 * (nz + 2));
    s->bczp = DALLOC((nx + 2) * (ny + 2) * 2);
    s->p_temp = DALLOC(s->N);
    /* metrics(): CORE_GRID.F90:184-205, fx(1..nx); 0 at 0 and nx+1 */
    s->fx = DALLOC(nx + 2); s->fy = DALLOC(ny + 2); s->fz = DALLOC(nz + 2);
    for (int i = 1; i <= nx; ++i) s->fx[i] = (s->x[i] - s->xc[i - 1]) / (s->xc[i] - s->xc[i - 1]);
    for (int j = 1; j <= ny; ++j) s->fy[j] = (s->y[j] - s->yc[j - 1]) / (s->yc[j] - s->yc[j - 1]);
    for (int k = 1; k <= nz; ++k) s->fz[k] = (s->z[k] - s->zc[k - 1]) / (s->zc[k] - s->zc[k - 1]);
    s->bcxu = DALLOC(2 * (ny + 2) * (nz + 2));
    s->bcyv = DALLOC((nx + 2) * 2 * (nz + 2));
    s->bczw = DALLOC((nx + 2) * (ny + 2) * 2);
    s->bcxv = DALLOC(2 * (ny + 2) * (nz + 2)); s->bcxw = DALLOC(2 * (ny + 2) * (nz + 2));
    s->bcyu = DALLOC((nx + 2) * 2 * (nz + 2)); s->bcyw = DALLOC((nx + 2) * 2 * (nz + 2));
    s->bczu = DALLOC((nx + 2) * (ny + 2) * 2); s->bczv = DALLOC((nx + 2) * (ny + 2) * 2);
    s->amx_ad = DALLOC(nx + 2); s->apx_ad = DALLOC(nx + 2);
    s->amy_ad = DALLOC(ny + 2); s->apy_ad = DALLOC(ny + 2);
    s->amz_ad = DALLOC(nz + 2); s->apz_ad = DALLOC(nz + 2);
    s->omega_ad = 1.0;
    int m = nx; if (ny > m) m = ny; if (nz > m) m = nz;
    s->am = DALLOC(m + 3); s->ac = DALLOC(m + 3); s->ap = DALLOC(m + 3);
    s->rr = DALLOC(m + 3); s->uu = DALLOC(m + 3); s->gam = DALLOC(m + 3);
    s->outflow_area = 1.0;
    return s;
}

void or_set_params(or_state *s, const fc_pson_params *p, int ordering) { s->prm = *p; s->ordering = ordering; }

int or_levels(const or_state *s, int *nlev, int *nxl, int *nyl, int *nzl)
{
    *nlev = s->nlev;
    for (int n = 1; n <= s->nlev; ++n) {
        if (nxl) nxl[n - 1] = s->mgI[n];
        if (nyl) nyl[n - 1] = s->mgJ[n];
        if (nzl) nzl[n - 1] = s->mgK[n];
    }
    return 0;
}

static void free_lists(or_state *s)
{
    free(s->iG); free(s->jG); free(s->kG); free(s->iC); free(s->jC); free(s->kC);
    free(s->cN); free(s->alpha); free(s->dpdnG);
    free(s->iD); free(s->jD); free(s->kD); free(s->iDC); free(s->jDC); free(s->kDC);
    free(s->cDN); free(s->plD); free(s->dpdnD); free(s->ipD); free(s->reslap);
}

void or_destroy(or_state *s)
{
    if (!s) return;
    for (int n = 1; n <= s->nlev; ++n) {
        olev *L = &s->lev[n];
        free(L->x); free(L->y); free(L->z); free(L->xc); free(L->yc); free(L->zc);
        free(L->dxinv); free(L->dyinv); free(L->dzinv); free(L->dxcinv); free(L->dycinv); free(L->dzcinv);
        free(L->rhs); free(L->phi); free(L->res); free(L->ibk); free(L->igk);
    }
    free(s->mgI); free(s->mgJ); free(s->mgK);
    free(s->x); free(s->y); free(s->z); free(s->xc); free(s->yc); free(s->zc);
    free(s->dx); free(s->dy); free(s->dz); free(s->dxinv); free(s->dyinv); free(s->dzinv);
    free(s->dxc); free(s->dyc); free(s->dzc); free(s->dxcinv); free(s->dycinv); free(s->dzcinv);
    free(s->iblank); free(s->ghost); free(s->dead); free(s->igmark);
    free(s->pgx1); free(s->pgx2); free(s->pgy1); free(s->pgy2); free(s->pgz1); free(s->pgz2);
    free(s->bcxp); free(s->bcyp); free(s->bczp); free(s->p_temp);
    free(s->fx); free(s->fy); free(s->fz); free(s->bcxu); free(s->bcyv); free(s->bczw);
    free(s->bcxv); free(s->bcxw); free(s->bcyu); free(s->bcyw); free(s->bczu); free(s->bczv);
    free(s->cD); free(s->cDD); free(s->ubi); free(s->ubiD);
    free(s->amx_ad); free(s->apx_ad); free(s->amy_ad); free(s->apy_ad); free(s->amz_ad); free(s->apz_ad);
    free(s->am); free(s->ac); free(s->ap); free(s->rr); free(s->uu); free(s->gam);
    free_lists(s);
    free(s);
}

/* ------------------------------------------------------------------ */
/* mg_prepare: CORE_MG.F90:1365-1555 */
static void mg_prepare(or_state *s)
{
    int nx = s->nx, ny = s->ny, nz = s->nz;
    memset(s->igmark, 0, (size_t)s->N);
    for (int k = 0; k <= nz + 1; ++k)
        for (int j = 0; j <= ny + 1; ++j)
            for (int i = 0; i <= nx; ++i) {
                long q = FID(s, i, j, k);
                if (s->iblank[q] == 1 || s->ghost[q] == 1) s->igmark[q] = 1;
            }
    olev *L1 = &s->lev[1];
    for (int k = 0; k <= nz + 1; ++k)
        for (int j = 0; j <= ny + 1; ++j)
            for (int i = 0; i <= nx; ++i) {
                long q = FID(s, i, j, k);
                L1->ibk[q] = s->iblank[q];
                L1->igk[q] = s->igmark[q];
            }
    for (int n = 2; n <= s->nlev; ++n) {
        olev *L = &s->lev[n], *F = &s->lev[n - 1];
        int ic = (L->nx == F->nx) ? 0 : 1, jc = (L->ny == F->ny) ? 0 : 1, kc = (L->nz == F->nz) ? 0 : 1;
        memset(L->igk, 0, (size_t)L->N);
        memset(L->ibk, 0, (size_t)L->N);
        for (int k = 1; k <= L->nz - 1; ++k)
            for (int j = 1; j <= L->ny - 1; ++j)
                for (int i = 1; i <= L->nx - 1; ++i) {
                    int i1 = (1 - ic) * i + ic * (2 * i - 1), i2 = (1 - ic) * i + ic * (2 * i);
                    int j1 = (1 - jc) * j + jc * (2 * j - 1), j2 = (1 - jc) * j + jc * (2 * j);
                    int k1 = (1 - kc) * k + kc * (2 * k - 1), k2 = (1 - kc) * k + kc * (2 * k);
                    double vol0 = 0.0, volibk = 0.0;
                    for (int kk = k1; kk <= k2; ++kk)
                        for (int jj = j1; jj <= j2; ++jj)
                            for (int ii = i1; ii <= i2; ++ii) {
                                long q = IDX(F, ii, jj, kk);
                                if (F->igk[q] == 1) L->igk[IDX(L, i, j, k)] = 1;
                                double dltX = F->x[ii + 1] - F->x[ii];
                                double dltY = F->y[jj + 1] - F->y[jj];
                                double dltZ = F->z[kk + 1] - F->z[kk];
                                double vol_i = dltX * dltY * dltZ;
                                vol0 = vol0 + vol_i;
                                if (F->ibk[q] == 1) volibk = volibk + vol_i;
                            }
                    if (volibk / vol0 >= 0.5) L->ibk[IDX(L, i, j, k)] = 1;
                }
        memset(L->rhs, 0, sizeof(double) * (size_t)L->N);
        memset(L->phi, 0, sizeof(double) * (size_t)L->N);
    }
}

int or_set_masks(or_state *s, const int8_t *iblank, const int8_t *ghost, const int8_t *dead)
{
    memcpy(s->iblank, iblank, (size_t)s->N);
    memcpy(s->ghost, ghost, (size_t)s->N);
    if (dead) memcpy(s->dead, dead, (size_t)s->N); else memset(s->dead, 0, (size_t)s->N);
    mg_prepare(s);
    return 0;
}

int or_set_gcm(or_state *s,
               int nG, const int *iG, const int *jG, const int *kG,
               const int *iC, const int *jC, const int *kC,
               const double *coeffN, const double *alpha, const double *dpdnG,
               int nD, const int *iD, const int *jD, const int *kD,
               const int *iDC, const int *jDC, const int *kDC,
               const double *coeffDN, const double *probeLenD,
               const double *dpdnD, const double *ipD)
{
    free_lists(s);
    s->nG = nG; s->nD = nD;
    s->iG = IALLOC(nG); s->jG = IALLOC(nG); s->kG = IALLOC(nG);
    s->iC = IALLOC(nG); s->jC = IALLOC(nG); s->kC = IALLOC(nG);
    s->cN = DALLOC(8 * (size_t)nG); s->alpha = DALLOC(nG); s->dpdnG = DALLOC(nG);
    s->reslap = DALLOC(nG);
    for (int n = 0; n < nG; ++n) {
        /* the reference's cell index (i) is the array index (i) */
        s->iG[n] = iG[n]; s->jG[n] = jG[n]; s->kG[n] = kG[n];
        s->iC[n] = iC[n]; s->jC[n] = jC[n]; s->kC[n] = kC[n];
        s->alpha[n] = alpha ? alpha[n] : 0.0;
        s->dpdnG[n] = dpdnG ? dpdnG[n] : 0.0;
        for (int r = 0; r < 8; ++r) s->cN[8 * (size_t)n + r] = coeffN[8 * (size_t)n + r];
    }
    s->iD = IALLOC(nD); s->jD = IALLOC(nD); s->kD = IALLOC(nD);
    s->iDC = IALLOC(nD); s->jDC = IALLOC(nD); s->kDC = IALLOC(nD);
    s->cDN = DALLOC(8 * (size_t)nD); s->plD = DALLOC(nD); s->dpdnD = DALLOC(nD); s->ipD = DALLOC(3 * (size_t)nD);
    for (int n = 0; n < nD; ++n) {
        s->iD[n] = iD[n]; s->jD[n] = jD[n]; s->kD[n] = kD[n];
        s->iDC[n] = iDC[n]; s->jDC[n] = jDC[n]; s->kDC[n] = kDC[n];
        s->plD[n] = probeLenD[n];
        s->dpdnD[n] = dpdnD ? dpdnD[n] : 0.0;
        for (int r = 0; r < 8; ++r) s->cDN[8 * (size_t)n + r] = coeffDN[8 * (size_t)n + r];
        for (int r = 0; r < 3; ++r) s->ipD[3 * (size_t)n + r] = ipD[3 * (size_t)n + r];
    }
    return 0;
}

int or_set_bc(or_state *s, const double *pgx1, const double *pgx2, const double *pgy1,
              const double *pgy2, const double *pgz1, const double *pgz2, double outflow_area)
{
    size_t nxz = (size_t)(s->nx + 2) * (s->nz + 2), nyz = (size_t)(s->ny + 2) * (s->nz + 2),
           nxy = (size_t)(s->nx + 2) * (s->ny + 2);
#define CP(dst, src, n) do { if (src) memcpy(dst, src, sizeof(double) * (n)); else memset(dst, 0, sizeof(double) * (n)); } while (0)
    CP(s->pgx1, pgx1, nyz); CP(s->pgx2, pgx2, nyz);
    CP(s->pgy1, pgy1, nxz); CP(s->pgy2, pgy2, nxz);
    CP(s->pgz1, pgz1, nxy); CP(s->pgz2, pgz2, nxy);
#undef CP
    s->outflow_area = outflow_area;
    return 0;
}

int or_get_bc(or_state *s, double *pgx1, double *pgx2, double *pgy1, double *pgy2, double *pgz1, double *pgz2)
{
    size_t nxz = (size_t)(s->nx + 2) * (s->nz + 2), nyz = (size_t)(s->ny + 2) * (s->nz + 2),
           nxy = (size_t)(s->nx + 2) * (s->ny + 2);
    if (pgx1) memcpy(pgx1, s->pgx1, sizeof(double) * nyz);
    if (pgx2) memcpy(pgx2, s->pgx2, sizeof(double) * nyz);
    if (pgy1) memcpy(pgy1, s->pgy1, sizeof(double) * nxz);
    if (pgy2) memcpy(pgy2, s->pgy2, sizeof(double) * nxz);
    if (pgz1) memcpy(pgz1, s->pgz1, sizeof(double) * nxy);
    if (pgz2) memcpy(pgz2, s->pgz2, sizeof(double) * nxy);
    return 0;
}

/* ------------------------------------------------------------------ */
/* set_outer_pressure_bc + set_outer_ghost_pressure:
 * CORE_SET_OUTER_BC.F90:831-1042 and 1044-1186, one rank, MG indexing.
 * bcxp(0:1,0:ny+1,0:nz+1), bcyp(0:nx+1,0:1,0:nz+1), bczp(0:nx+1,0:ny+1,0:1)
 * are level-1 sized as in the reference. */
#define BCX(s, ii, j, k) (s)->bcxp[(ii) + 2 * ((long)(j) + (long)((s)->ny + 2) * (k))]
#define BCY(s, i, jj, k) (s)->bcyp[(i) + (long)((s)->nx + 2) * ((jj) + 2L * (k))]
#define BCZ(s, i, j, kk) (s)->bczp[(i) + (long)((s)->nx + 2) * ((long)(j) + (long)((s)->ny + 2) * (kk))]
#define PGX(a, j, k) (a)[(long)(j) + (long)(s->ny + 2) * (k)]
#define PGY(a, i, k) (a)[(long)(i) + (long)(s->nx + 2) * (k)]
#define PGZ(a, i, j) (a)[(long)(i) + (long)(s->nx + 2) * (j)]

static void bc_fill_arr(or_state *s, int nlev, double *pres, int nxm, int nym, int nzm, long P, long ny2)
{
#define PR(i, j, k) pres[(long)(i) + P * ((long)(j) + ny2 * (long)(k))]
    const fc_pson_params *p = &s->prm;
    double ppx1 = p->ppp[0], ppx2 = p->ppp[1], ppy1 = p->ppp[2], ppy2 = p->ppp[3], ppz1 = p->ppp[4], ppz2 = p->ppp[5];
    int iHomo = (nlev != 1) ? 1 : 0;
    if (iHomo) { ppx1 = ppx2 = ppy1 = ppy2 = ppz1 = ppz2 = 0.0; }
    int zk1 = 1, zk2 = nzm - 1, yj1 = 1, yj2 = nym - 1;
    double h = (double)(1 - iHomo);
    for (int k = zk1; k <= zk2; ++k)
        for (int j = yj1; j <= yj2; ++j) {
            if (p->pbc[0] == FC_PBC_DIRICHLET) BCX(s, 0, j, k) = ppx1;
            else if (p->pbc[0] == FC_PBC_NEUMANN) BCX(s, 0, j, k) = PR(1, j, k) - PGX(s->pgx1, j, k) * s->dxc[1] / 2.0 * h;
            if (p->pbc[1] == FC_PBC_DIRICHLET) BCX(s, 1, j, k) = ppx2;
            else if (p->pbc[1] == FC_PBC_NEUMANN) BCX(s, 1, j, k) = PR(nxm - 1, j, k) + PGX(s->pgx2, j, k) * s->dxc[nxm <= s->nx ? nxm : s->nx] / 2.0 * h;
        }
    for (int k = zk1; k <= zk2; ++k)
        for (int i = 1; i <= nxm - 1; ++i) {
            if (p->pbc[2] == FC_PBC_DIRICHLET) BCY(s, i, 0, k) = ppy1;
            else if (p->pbc[2] == FC_PBC_NEUMANN) BCY(s, i, 0, k) = PR(i, 1, k) - PGY(s->pgy1, i, k) * s->dyc[1] / 2.0 * h;
            if (p->pbc[3] == FC_PBC_DIRICHLET) BCY(s, i, 1, k) = ppy2;
            else if (p->pbc[3] == FC_PBC_NEUMANN) BCY(s, i, 1, k) = PR(i, yj2, k) + PGY(s->pgy2, i, k) * s->dyc[s->ny] / 2.0 * h;
        }
    for (int j = yj1; j <= yj2; ++j)
        for (int i = 1; i <= nxm - 1; ++i) {
            if (p->pbc[4] == FC_PBC_DIRICHLET) BCZ(s, i, j, 0) = ppz1;
            else if (p->pbc[4] == FC_PBC_NEUMANN) BCZ(s, i, j, 0) = PR(i, j, 1) - PGZ(s->pgz1, i, j) * s->dzc[1] / 2.0 * h;
            if (p->pbc[5] == FC_PBC_DIRICHLET) BCZ(s, i, j, 1) = ppz2;
            else if (p->pbc[5] == FC_PBC_NEUMANN) BCZ(s, i, j, 1) = PR(i, j, zk2) + PGZ(s->pgz2, i, j) * s->dzc[s->nz] / 2.0 * h;
        }
    /* periodic x (both x faces) or y (both y faces): the face value is the
     * mean of the two interior cells next to the two faces, so the ghost
     * takes the value of the cell at the other end
     * (CORE_SET_OUTER_BC.F90:1007-1026; the reference's y branch indexes the
     * finest level's ny-1 on every level, here the level's own) */
    if (p->bc_vel[0] == FC_BC_TYPE_PERIODIC && p->bc_vel[1] == FC_BC_TYPE_PERIODIC)
        for (int k = zk1; k <= zk2; ++k)
            for (int j = yj1; j <= yj2; ++j) {
                BCX(s, 0, j, k) = (PR(1, j, k) + PR(nxm - 1, j, k)) / 2.0;
                BCX(s, 1, j, k) = BCX(s, 0, j, k);
            }
    if (p->bc_vel[2] == FC_BC_TYPE_PERIODIC && p->bc_vel[3] == FC_BC_TYPE_PERIODIC)
        for (int k = zk1; k <= zk2; ++k)
            for (int i = 1; i <= nxm - 1; ++i) {
                BCY(s, i, 0, k) = (PR(i, 1, k) + PR(i, yj2, k)) / 2.0;
                BCY(s, i, 1, k) = BCY(s, i, 0, k);
            }
    /* ghost layers: CORE_SET_OUTER_BC.F90:1088-1170 */
    for (int k = zk1; k <= zk2; ++k)
        for (int j = yj1; j <= yj2; ++j) {
            PR(0, j, k) = 2.0 * BCX(s, 0, j, k) - PR(1, j, k);
            PR(nxm, j, k) = 2.0 * BCX(s, 1, j, k) - PR(nxm - 1, j, k);
        }
    for (int k = zk1; k <= zk2; ++k)
        for (int i = 0; i <= nxm; ++i) PR(i, 0, k) = 2.0 * BCY(s, i, 0, k) - PR(i, 1, k);
    for (int k = zk1; k <= zk2; ++k)
        for (int i = 0; i <= nxm; ++i) PR(i, yj2 + 1, k) = 2.0 * BCY(s, i, 1, k) - PR(i, yj2, k);
    for (int j = yj1; j <= yj2; ++j)
        for (int i = 0; i <= nxm; ++i) PR(i, j, 0) = 2.0 * BCZ(s, i, j, 0) - PR(i, j, 1);
    for (int j = yj1; j <= yj2; ++j)
        for (int i = 0; i <= nxm; ++i) PR(i, j, zk2 + 1) = 2.0 * BCZ(s, i, j, 1) - PR(i, j, zk2);
#undef PR
}

void or_bc_fill(or_state *s, int lev)
{
    olev *L = &s->lev[lev];
    bc_fill_arr(s, lev, L->phi, L->nx, L->ny, L->nz, L->P, L->ny + 2);
}

/* ------------------------------------------------------------------ */
/* itsolv_mg: CORE_MG.F90:615-938 (reference ordering) */
static void line_coef(const double *dcinv, const double *dinv, int i, double *m, double *pp, double *c)
{
    *m = dcinv[i] * dinv[i];
    *pp = dcinv[i + 1] * dinv[i];
    *c = -(*m + *pp);
}

static void sweep_x_line(or_state *s, olev *L, int j, int k, double om)
{
    double *var = L->phi, *r = L->rhs; int8_t *ig = L->igk;
    int nxm = L->nx;
    double amz, apz, acz, amy, apy, acy;
    line_coef(L->dzcinv, L->dzinv, k, &amz, &apz, &acz);
    line_coef(L->dycinv, L->dyinv, j, &amy, &apy, &acy);
    for (int i = 1; i <= nxm - 1; ++i) {
        double amx, apx, acx;
        line_coef(L->dxcinv, L->dxinv, i, &amx, &apx, &acx);
        long q = IDX(L, i, j, k);
        double rhs = r[q] - var[q - L->P] * amy - var[q + L->P] * apy - var[q - L->S] * amz - var[q + L->S] * apz;
        double flag = (double)(1 - ig[q]);
        amx = amx * flag; apx = apx * flag;
        acx = (acx + acy + acz) * flag + (1.0 - flag);
        rhs = rhs * flag + (1.0 - flag) * var[q];
        if (fabs(acx) < EPS_GUARD) { amx = 0.0; apx = 0.0; acx = 1.0; rhs = var[q]; }
        s->am[i] = amx; s->ap[i] = apx; s->ac[i] = acx; s->rr[i] = rhs;
    }
    s->am[0] = 0.0; s->ap[0] = 0.0; s->ac[0] = 1.0; s->rr[0] = var[IDX(L, 0, j, k)];
    s->am[nxm] = 0.0; s->ap[nxm] = 0.0; s->ac[nxm] = 1.0; s->rr[nxm] = var[IDX(L, nxm, j, k)];
    tdma_s(s, 0, nxm);
    for (int i = 0; i <= nxm; ++i) { long q = IDX(L, i, j, k); var[q] = var[q] * (1.0 - om) + om * s->uu[i]; }
}

static void sweep_y_line(or_state *s, olev *L, int i, int k, double om)
{
    double *var = L->phi, *r = L->rhs; int8_t *ig = L->igk;
    int nym = L->ny, JL1 = 1, JL2 = nym - 1, j0 = 0, j1 = nym;
    double amz, apz, acz, amx, apx, acx;
    line_coef(L->dzcinv, L->dzinv, k, &amz, &apz, &acz);
    line_coef(L->dxcinv, L->dxinv, i, &amx, &apx, &acx);
    for (int j = JL1; j <= JL2; ++j) {
        double amy, apy, acy;
        line_coef(L->dycinv, L->dyinv, j, &amy, &apy, &acy);
        long q = IDX(L, i, j, k);
        double rhs = r[q] - var[q - 1] * amx - var[q + 1] * apx - var[q - L->S] * amz - var[q + L->S] * apz;
        double flag = (double)(1 - ig[q]);
        amy = amy * flag; apy = apy * flag;
        acy = (acx + acy + acz) * flag + (1.0 - flag);
        rhs = rhs * flag + (1.0 - flag) * var[q];
        if (fabs(acy) < EPS_GUARD) { amy = 0.0; apy = 0.0; acy = 1.0; rhs = var[q]; }
        s->am[j] = amy; s->ap[j] = apy; s->ac[j] = acy; s->rr[j] = rhs;
    }
    s->am[j0] = 0.0; s->ap[j0] = 0.0; s->ac[j0] = 1.0; s->rr[j0] = var[IDX(L, i, JL1 - 1, k)];
    s->am[j1] = 0.0; s->ap[j1] = 0.0; s->ac[j1] = 1.0; s->rr[j1] = var[IDX(L, i, JL2 + 1, k)];
    tdma_s(s, j0, j1);
    for (int j = j0; j <= j1; ++j) { long q = IDX(L, i, j, k); var[q] = var[q] * (1.0 - om) + om * s->uu[j]; }
}

static void sweep_z_line(or_state *s, olev *L, int i, int j, double om)
{
    double *var = L->phi, *r = L->rhs; int8_t *ig = L->igk;
    int nzm = L->nz, KL1 = 1, KL2 = nzm - 1, k0 = 0, k1 = nzm;
    double amy, apy, acy, amx, apx, acx;
    line_coef(L->dycinv, L->dyinv, j, &amy, &apy, &acy);
    line_coef(L->dxcinv, L->dxinv, i, &amx, &apx, &acx);
    for (int k = KL1; k <= KL2; ++k) {
        double amz, apz, acz;
        line_coef(L->dzcinv, L->dzinv, k, &amz, &apz, &acz);
        long q = IDX(L, i, j, k);
        double rhs = r[q] - var[q - 1] * amx - var[q + 1] * apx - var[q - L->P] * amy - var[q + L->P] * apy;
        double flag = (double)(1 - ig[q]);
        amz = amz * flag; apz = apz * flag;
        acz = (acx + acy + acz) * flag + (1.0 - flag);
        rhs = rhs * flag + (1.0 - flag) * var[q];
        if (fabs(acz) < EPS_GUARD) { amz = 0.0; apz = 0.0; acz = 1.0; rhs = var[q]; }
        s->am[k] = amz; s->ap[k] = apz; s->ac[k] = acz; s->rr[k] = rhs;
    }
    s->am[k0] = 0.0; s->ap[k0] = 0.0; s->ac[k0] = 1.0; s->rr[k0] = var[IDX(L, i, j, k0)];
    s->am[k1] = 0.0; s->ap[k1] = 0.0; s->ac[k1] = 1.0; s->rr[k1] = var[IDX(L, i, j, k1)];
    tdma_s(s, k0, k1);
    for (int k = k0; k <= k1; ++k) { long q = IDX(L, i, j, k); var[q] = var[q] * (1.0 - om) + om * s->uu[k]; }
}

/* one directional half sweep of the colored (GPU) ordering */
/* One colour of one direction in the GPU's checkerboard order.  The outer BC
 * is filled before the half sweep (as the reference does before each
 * direction) and again after it: the GPU's sweep rewrites the ghost cells of
 * the cells it updates, so its ghosts are current after every half sweep. */
void or_sweep(or_state *s, int lev, int dir, int color)
{
    olev *L = &s->lev[lev];
    double om = s->prm.omega_pson;
    or_bc_fill(s, lev);
    if (dir == 0) {
        for (int k = 1; k <= L->nz - 1; ++k)
            for (int j = 1; j <= L->ny - 1; ++j)
                if (((j + k) & 1) == color) sweep_x_line(s, L, j, k, om);
    } else if (dir == 1) {
        for (int k = 1; k <= L->nz - 1; ++k)
            for (int i = 1; i <= L->nx - 1; ++i)
                if (((i + k) & 1) == color) sweep_y_line(s, L, i, k, om);
    } else {
        for (int j = 1; j <= L->ny - 1; ++j)
            for (int i = 1; i <= L->nx - 1; ++i)
                if (((i + j) & 1) == color) sweep_z_line(s, L, i, j, om);
    }
    or_bc_fill(s, lev);
}

void or_itsolv_mg(or_state *s, int lev)
{
    olev *L = &s->lev[lev];
    double om = s->prm.omega_pson;
    if (s->ordering == OR_ORDER_COLORED) {
        for (int dir = 0; dir < 3; ++dir) {
            for (int c = 0; c < 2; ++c) or_sweep(s, lev, dir, c);
        }
        return;
    }
    /* x-lines: lexicographic over (k,j) (CORE_MG.F90:650-733) */
    or_bc_fill(s, lev);
    for (int k = 1; k <= L->nz - 1; ++k)
        for (int j = 1; j <= L->ny - 1; ++j) sweep_x_line(s, L, j, k, om);
    /* y-lines: odd i then even i, GS in k (CORE_MG.F90:747-827) */
    for (int color = 1; color <= 2; ++color) {
        or_bc_fill(s, lev);
        for (int k = 1; k <= L->nz - 1; ++k)
            for (int i = color; i <= L->nx - 1; i += 2) sweep_y_line(s, L, i, k, om);
    }
    /* z-lines: lexicographic over (j,i) (CORE_MG.F90:841-925) */
    or_bc_fill(s, lev);
    for (int j = 1; j <= L->ny - 1; ++j)
        for (int i = 1; i <= L->nx - 1; ++i) sweep_z_line(s, L, i, j, om);
}

/* ------------------------------------------------------------------ */
/* residual_mg: CORE_MG.F90:956-1043 (plus the res=rhs copy :1672-1678) */
void or_residual_mg(or_state *s, int lev)
{
    olev *L = &s->lev[lev];
    for (int k = 1; k <= L->nz - 1; ++k)
        for (int j = 1; j <= L->ny - 1; ++j)
            for (int i = 1; i <= L->nx; ++i) L->res[IDX(L, i, j, k)] = L->rhs[IDX(L, i, j, k)];
    double *var = L->phi, *rrr = L->res;
    for (int k = 1; k <= L->nz - 1; ++k)
        for (int j = 1; j <= L->ny - 1; ++j)
            for (int i = 1; i <= L->nx - 1; ++i) {
                double bmz, bpz, bcz, bmy, bpy, bcy, bmx, bpx, bcx;
                line_coef(L->dzcinv, L->dzinv, k, &bmz, &bpz, &bcz);
                line_coef(L->dycinv, L->dyinv, j, &bmy, &bpy, &bcy);
                line_coef(L->dxcinv, L->dxinv, i, &bmx, &bpx, &bcx);
                long q = IDX(L, i, j, k);
                double flag = 1.0 - (double)L->igk[q];
                double bc = (bcx + bcy + bcz);
                double res = rrr[q] - var[q] * bc - var[q - 1] * bmx - var[q + 1] * bpx
                             - var[q - L->P] * bmy - var[q + L->P] * bpy - var[q - L->S] * bmz - var[q + L->S] * bpz;
                rrr[q] = res * flag;
            }
}

/* mg_inject: CORE_MG.F90:1061-1183 (restriction res(N) -> rhs(N+1)) */
void or_mg_inject(or_state *s, int N)
{
    olev *F = &s->lev[N], *C = &s->lev[N + 1];
    const double *rr1 = F->res;
    double *rr2 = C->rhs;
    memset(rr2, 0, sizeof(double) * (size_t)C->N);
    int ic = (F->nx == C->nx) ? 0 : 1, jc = (F->ny == C->ny) ? 0 : 1, kc = (F->nz == C->nz) ? 0 : 1;
    for (int k = 1; k <= C->nz - 1; ++k)
        for (int j = 1; j <= C->ny - 1; ++j)
            for (int i = 1; i <= C->nx - 1; ++i) {
                double xc0 = C->xc[i], yc0 = C->yc[j], zc0 = C->zc[k];
                int i1 = (1 - ic) * i + ic * (2 * i - 1), i2 = (1 - ic) * i + ic * (2 * i);
                int j1 = (1 - jc) * j + jc * (2 * j - 1), j2 = (1 - jc) * j + jc * (2 * j);
                int k1 = (1 - kc) * k + kc * (2 * k - 1), k2 = (1 - kc) * k + kc * (2 * k);
                double aX = (xc0 - F->xc[i1]) * F->dxcinv[i1 + 1];
                double aY = (yc0 - F->yc[j1]) * F->dycinv[j1 + 1];
                double aZ = (zc0 - F->zc[k1]) * F->dzcinv[k1 + 1];
#define R1(a, b, c) rr1[IDX(F, a, b, c)]
                double r_tmp = R1(i1, j1, k1) * (1.0 - aX) * (1.0 - aY) * (1.0 - aZ)
                             + R1(i1 + 1, j1, k1) * aX * (1.0 - aY) * (1.0 - aZ)
                             + R1(i1, j1 + 1, k1) * (1.0 - aX) * aY * (1.0 - aZ)
                             + R1(i1, j1, k1 + 1) * (1.0 - aX) * (1.0 - aY) * aZ
                             + R1(i1 + 1, j1 + 1, k1) * aX * aY * (1.0 - aZ)
                             + R1(i1 + 1, j1, k1 + 1) * aX * (1.0 - aY) * aZ
                             + R1(i1, j1 + 1, k1 + 1) * (1.0 - aX) * aY * aZ
                             + R1(i1 + 1, j1 + 1, k1 + 1) * aX * aY * aZ;
                double vol0 = 0.0, sum0 = 0.0;
                int ibk_inside = 0;
                for (int kk = k1; kk <= k2; ++kk)
                    for (int jj = j1; jj <= j2; ++jj)
                        for (int ii = i1; ii <= i2; ++ii) {
                            if (F->ibk[IDX(F, ii, jj, kk)] == 1) ibk_inside = 1;
                            else {
                                double vol_i = (F->x[ii + 1] - F->x[ii]) * (F->y[jj + 1] - F->y[jj]) * (F->z[kk + 1] - F->z[kk]);
                                vol0 = vol0 + vol_i;
                                sum0 = R1(ii, jj, kk) * vol_i;     /* assignment, as at :1159 */
                            }
                        }
#undef R1
                double vol_c = 1.0 / F->dxinv[i1] / F->dyinv[j1] / F->dzinv[k1];
                double r_avg = (vol0 > VOL_FRAC * vol_c) ? sum0 / vol0 : 0.0;
                rr2[IDX(C, i, j, k)] = r_tmp * (double)(1 - ibk_inside) + r_avg * (double)ibk_inside;
            }
}

/* mg_interp: CORE_MG.F90:1200-1351 (prolongation phi(N) added to phi(N-1),
 * dphi stored in res(N-1)) */
void or_mg_interp(or_state *s, int N)
{
    olev *C = &s->lev[N], *F = &s->lev[N - 1];
    const double *phi1 = C->phi;
    double *phi2 = F->phi, *dphi = F->res;
    memset(dphi, 0, sizeof(double) * (size_t)F->N);
    int ic = (C->nx == F->nx) ? 0 : 1, jc = (C->ny == F->ny) ? 0 : 1, kc = (C->nz == F->nz) ? 0 : 1;
    int nx2 = F->nx, ny2 = F->ny, nz2 = F->nz, JL1 = 1, JL2 = F->ny - 1, KL1 = 1, KL2 = F->nz - 1;
    for (int k = KL1; k <= KL2; ++k)
        for (int j = JL1; j <= JL2; ++j)
            for (int i = 1; i <= nx2 - 1; ++i) {
                if ((i - 1) + (j - 1) == 0) continue;
                if ((i - 1) + (ny2 - 1 - j) == 0) continue;
                if ((nx2 - 1 - i) + (j - 1) == 0) continue;
                if ((nx2 - 1 - i) + (ny2 - 1 - j) == 0) continue;
                if ((i - 1) + (k - KL1) == 0) continue;
                if ((i - 1) + (KL2 - k) == 0) continue;
                if ((nx2 - 1 - i) + (k - KL1) == 0) continue;
                if ((nx2 - 1 - i) + (KL2 - k) == 0) continue;
                if ((j - 1) + (k - KL1) == 0) continue;
                if ((j - 1) + (KL2 - k) == 0) continue;
                if ((ny2 - 1 - j) + (k - KL1) == 0) continue;
                if ((ny2 - 1 - j) + (KL2 - k) == 0) continue;
                if ((k - 1) + (j - JL1) == 0) continue;
                if ((k - 1) + (JL2 - j) == 0) continue;
                if ((nz2 - 1 - k) + (j - JL1) == 0) continue;
                if ((nz2 - 1 - k) + (JL2 - j) == 0) continue;
                if (F->ibk[IDX(F, i, j, k)] == 1) continue;
                double xc0 = F->xc[i], yc0 = F->yc[j], zc0 = F->zc[k];
                int ii = i / 2, jj = j / 2, kk = k / 2;
                int i1 = (1 - ic) * i + ic * ii, i2 = (1 - ic) * i + ic * (ii + 1);
                int j1 = (1 - jc) * j + jc * jj, j2 = (1 - jc) * j + jc * (jj + 1);
                int k1 = (1 - kc) * k + kc * kk, k2 = (1 - kc) * k + kc * (kk + 1);
                double aX = (xc0 - C->xc[i1]) * C->dxcinv[i1 + 1];
                double aY = (yc0 - C->yc[j1]) * C->dycinv[j1 + 1];
                double aZ = (zc0 - C->zc[k1]) * C->dzcinv[k1 + 1];
#define P1(a, b, c) phi1[IDX(C, a, b, c)]
                double r_tmp = P1(i1, j1, k1) * (1.0 - aX) * (1.0 - aY) * (1.0 - aZ)
                             + P1(i1 + 1, j1, k1) * aX * (1.0 - aY) * (1.0 - aZ)
                             + P1(i1, j1 + 1, k1) * (1.0 - aX) * aY * (1.0 - aZ)
                             + P1(i1, j1, k1 + 1) * (1.0 - aX) * (1.0 - aY) * aZ
                             + P1(i1 + 1, j1 + 1, k1) * aX * aY * (1.0 - aZ)
                             + P1(i1 + 1, j1, k1 + 1) * aX * (1.0 - aY) * aZ
                             + P1(i1, j1 + 1, k1 + 1) * (1.0 - aX) * aY * aZ
                             + P1(i1 + 1, j1 + 1, k1 + 1) * aX * aY * aZ;
                double vol0 = 0.0, sum0 = 0.0;
                int ibk_inside = 0;
                for (int k3 = k1; k3 <= k2; ++k3)
                    for (int j3 = j1; j3 <= j2; ++j3)
                        for (int i3 = i1; i3 <= i2; ++i3) {
                            if (C->ibk[IDX(C, i3, j3, k3)] == 1) ibk_inside = 1;
                            else {
                                double vol_i = (C->x[i3 + 1] - C->x[i3]) * (C->y[j3 + 1] - C->y[j3]) * (C->z[k3 + 1] - C->z[k3]);
                                vol0 = vol0 + vol_i;
                                sum0 = P1(i3, j3, k3) * vol_i;     /* assignment, as at :1326 */
                            }
                        }
#undef P1
                double vol_c = 1.0 / C->dxinv[i1] / C->dyinv[j1] / C->dzinv[k1];
                double r_avg = (vol0 > VOL_FRAC * vol_c) ? sum0 / vol0 : 0.0;
                long q = IDX(F, i, j, k);
                dphi[q] = r_tmp * (double)(1 - ibk_inside) + r_avg * (double)ibk_inside;
                phi2[q] = phi2[q] + dphi[q];
            }
}

/* ------------------------------------------------------------------ */
/* calc_residual: CORE_POISSON.F90:460-551 */
double or_calc_residual(or_state *s, const double *var, const double *r)
{
    double res2 = 0.0;
    for (int k = 1; k <= s->nz - 1; ++k)
        for (int j = 1; j <= s->ny - 1; ++j)
            for (int i = 1; i <= s->nx - 1; ++i) {
                double bmx, bpx, bcx, bmy, bpy, bcy, bmz, bpz, bcz;
                line_coef(s->dxcinv, s->dxinv, i, &bmx, &bpx, &bcx);
                line_coef(s->dycinv, s->dyinv, j, &bmy, &bpy, &bcy);
                line_coef(s->dzcinv, s->dzinv, k, &bmz, &bpz, &bcz);
                long q = FID(s, i, j, k);
                double flag = (1.0 - (double)s->iblank[q]) * (1.0 - (double)s->ghost[q]);
                double bc = (bcx + bcy + bcz);
                double res = r[q] - var[q] * bc - var[q - 1] * bmx - var[q + 1] * bpx
                             - var[q - s->P] * bmy - var[q + s->P] * bpy - var[q - s->S] * bmz - var[q + s->S] * bpz;
                res = res * flag;
                res2 = res2 + res * res;
            }
    return sqrt(res2 / (double)((long)s->nx * s->ny * s->nz));
}

/* flag-cell relative residual ||div - A p||_2 / ||div||_2 (the metric used
 * by the benchmarks; same residual as calc_residual) */
double or_rel_residual(or_state *s, const double *var, const double *r)
{
    double res2 = 0.0, b2 = 0.0;
    for (int k = 1; k <= s->nz - 1; ++k)
        for (int j = 1; j <= s->ny - 1; ++j)
            for (int i = 1; i <= s->nx - 1; ++i) {
                double bmx, bpx, bcx, bmy, bpy, bcy, bmz, bpz, bcz;
                line_coef(s->dxcinv, s->dxinv, i, &bmx, &bpx, &bcx);
                line_coef(s->dycinv, s->dyinv, j, &bmy, &bpy, &bcy);
                line_coef(s->dzcinv, s->dzinv, k, &bmz, &bpz, &bcz);
                long q = FID(s, i, j, k);
                double flag = (1.0 - (double)s->iblank[q]) * (1.0 - (double)s->ghost[q]);
                double bc = (bcx + bcy + bcz);
                double res = r[q] - var[q] * bc - var[q - 1] * bmx - var[q + 1] * bpx
                             - var[q - s->P] * bmy - var[q + s->P] * bpy - var[q - s->S] * bmz - var[q + s->S] * bpz;
                res2 += res * res * flag;
                b2 += r[q] * r[q] * flag;
            }
    return b2 > 0.0 ? sqrt(res2 / b2) : sqrt(res2);
}

/* calc_residual2: CORE_POISSON.F90:679-724 */
double or_calc_residual2(or_state *s, const double *var, const double *p_temp)
{
    double res1 = 0.0;
    for (int k = 1; k <= s->nz - 1; ++k)
        for (int j = 1; j <= s->ny - 1; ++j)
            for (int i = 1; i <= s->nx - 1; ++i) {
                long q = FID(s, i, j, k);
                double flag = 1.0 - (double)s->iblank[q];
                double res = fabs(var[q] - p_temp[q]) * flag;
                if (res > res1) res1 = res;
            }
    return res1;
}

/* ------------------------------------------------------------------ */
/* one V-cycle of mg_solve: CORE_MG.F90:1618-1766 */
void or_vcycle(or_state *s)
{
    int ml = s->nlev;
    for (int N = 1; N <= ml; ++N) {
        olev *L = &s->lev[N];
        if (N > 1) memset(L->phi, 0, sizeof(double) * (size_t)L->N);
        or_itsolv_mg(s, N);
        or_residual_mg(s, N);
        if (N != ml) or_mg_inject(s, N);
    }
    for (int N = ml; N >= 2; --N) {
        if (N != ml) or_itsolv_mg(s, N);
        or_mg_interp(s, N);
    }
    /* cycle shape 2 (new, not in the reference): one post-smoothing pass on
     * the finest level; the automatic choice (0) is resolved by the caller */
    if (s->prm.cycle == 2) or_itsolv_mg(s, 1);
}

/* mg_solve: CORE_MG.F90:1572-1848 */
int or_mg_solve(or_state *s, double *var, const double *rr, fc_pson_stats *st)
{
    olev *L1 = &s->lev[1];
    for (int k = 0; k <= s->nz + 1; ++k)
        for (int j = 0; j <= s->ny + 1; ++j)
            for (int i = 0; i <= s->nx; ++i) {
                long q = FID(s, i, j, k);
                L1->phi[q] = var[q];
                L1->rhs[q] = rr[q];
            }
    int iter_mg = 0;
    double resm = 1.0e6;
    while (iter_mg <= s->prm.iterInter && resm > s->prm.restol_pson) {
        or_vcycle(s);
        resm = or_calc_residual(s, L1->phi, L1->rhs);
        if (s->prm.infoconv == 1) printf("MG: residual check at lev 1: %4d  %12.5E\n", iter_mg, resm);
        iter_mg++;
        if (st) st->vcycles_fp64++;
    }
    for (int k = 0; k <= s->nz + 1; ++k)
        for (int j = 0; j <= s->ny + 1; ++j)
            for (int i = 0; i <= s->nx; ++i) {
                long q = FID(s, i, j, k);
                var[q] = L1->phi[q];
            }
    if (st) st->maxres = resm;
    return 0;
}

/* itsolv: CORE_POISSON.F90:205-456 (plain LSOR, interior relaxation, no
 * guard, no BC refresh inside) */
/* one x-, y- or z-line of itsolv (CORE_POISSON.F90:237-450): masked rows
 * (flag = 0) are identity rows holding var, rows 0 and n are the current
 * ghost values, relaxation var + omega (x - var) on the interior rows */
static void itsolv_line(or_state *s, double *var, const double *r, int dir, int i0, int j0, int k0)
{
    double om = s->prm.omega_pson;
    int n = dir == 0 ? s->nx : dir == 1 ? s->ny : s->nz;
    long st = dir == 0 ? 1 : dir == 1 ? s->P : s->S;
    double *am = s->am, *ap = s->ap, *ac = s->ac, *rh = s->rr;
    for (int m = 1; m <= n - 1; ++m) {
        int i = dir == 0 ? m : i0, j = dir == 1 ? m : j0, k = dir == 2 ? m : k0;
        double amx, apx, acx, amy, apy, acy, amz, apz, acz;
        line_coef(s->dxcinv, s->dxinv, i, &amx, &apx, &acx);
        line_coef(s->dycinv, s->dyinv, j, &amy, &apy, &acy);
        line_coef(s->dzcinv, s->dzinv, k, &amz, &apz, &acz);
        long q = FID(s, i, j, k);
        double rhs, aml, apl;
        if (dir == 0) {
            rhs = r[q] - var[q - s->P] * amy - var[q + s->P] * apy - var[q - s->S] * amz - var[q + s->S] * apz;
            aml = amx; apl = apx;
        } else if (dir == 1) {
            rhs = r[q] - var[q - 1] * amx - var[q + 1] * apx - var[q - s->S] * amz - var[q + s->S] * apz;
            aml = amy; apl = apy;
        } else {
            rhs = r[q] - var[q - 1] * amx - var[q + 1] * apx - var[q - s->P] * amy - var[q + s->P] * apy;
            aml = amz; apl = apz;
        }
        double flag = (1.0 - (double)s->iblank[q]) * (1.0 - (double)s->ghost[q]);
        am[m] = aml * flag; ap[m] = apl * flag;
        ac[m] = (acx + acy + acz) * flag + (1.0 - flag);
        rh[m] = rhs * flag + (1.0 - flag) * var[q];
    }
    long q0 = FID(s, dir == 0 ? 0 : i0, dir == 1 ? 0 : j0, dir == 2 ? 0 : k0);
    am[0] = 0; ap[0] = 0; ac[0] = 1; rh[0] = var[q0];
    am[n] = 0; ap[n] = 0; ac[n] = 1; rh[n] = var[q0 + st * n];
    tdma_s(s, 0, n);
    for (int m = 1; m <= n - 1; ++m) { long q = q0 + st * m; var[q] = var[q] + om * (s->uu[m] - var[q]); }
}

/* itsolv: CORE_POISSON.F90:205-456.  ORDER_REFERENCE: x-lines
 * lexicographic over (k, j), y-lines zebra (odd i, then even i) over k,
 * z-lines lexicographic over (j, i).  ORDER_COLORED: the GPU order, per
 * direction the lines of colour 0 then colour 1 (x: (j+k)%2, y: (i+k)%2,
 * z: (i+j)%2).  The ghost layers are read as they are. */
int or_itsolv(or_state *s, double *var, const double *r)
{
    int nx = s->nx, ny = s->ny, nz = s->nz;
    if (s->ordering == OR_ORDER_COLORED) {
        for (int col = 0; col < 2; ++col)
            for (int k = 1; k <= nz - 1; ++k)
                for (int j = 1; j <= ny - 1; ++j)
                    if ((j + k) % 2 == col) itsolv_line(s, var, r, 0, 0, j, k);
        for (int col = 0; col < 2; ++col)
            for (int k = 1; k <= nz - 1; ++k)
                for (int i = 1; i <= nx - 1; ++i)
                    if ((i + k) % 2 == col) itsolv_line(s, var, r, 1, i, 0, k);
        for (int col = 0; col < 2; ++col)
            for (int j = 1; j <= ny - 1; ++j)
                for (int i = 1; i <= nx - 1; ++i)
                    if ((i + j) % 2 == col) itsolv_line(s, var, r, 2, i, j, 0);
        return 0;
    }
    for (int k = 1; k <= nz - 1; ++k)
        for (int j = 1; j <= ny - 1; ++j) itsolv_line(s, var, r, 0, 0, j, k);
    for (int color = 1; color <= 2; ++color) {
        int iinit = (color + 1) % 2 + 1;
        for (int k = 1; k <= nz - 1; ++k)
            for (int i = iinit; i <= nx - 1; i += 2) itsolv_line(s, var, r, 1, i, 0, k);
    }
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) itsolv_line(s, var, r, 2, i, j, 0);
    return 0;
}

/* ------------------------------------------------------------------ */
/* dead-cell value: GCM_GHOSTCELL_UPDATE.F90:427-522 */
static double dead_value(or_state *s, int n, const double *pres, const double *p_prev)
{
    int iG = s->iD[n], jG = s->jD[n], kG = s->kD[n];
    double xIP = s->ipD[3 * (size_t)n], yIP = s->ipD[3 * (size_t)n + 1], zIP = s->ipD[3 * (size_t)n + 2];
    double dpdn = (s->prm.is_PBC_homogeneous == 1) ? 0.0 : s->dpdnD[n];
    double p_inv = 0.0, dist_int = 0.0, pIP = 0.0;
    int sumk = 0;
    for (int r = 0; r < 8; ++r) {
        int ii = s->iDC[n] + incI[r], jj = s->jDC[n] + incJ[r], kk = s->kDC[n] + incK[r];
        if (ii == iG && jj == jG && kk == kG) continue;
        double dx = s->xc[ii] - xIP, dy = s->yc[jj] - yIP, dz = s->zc[kk] - zIP;
        double dist = dx * dx + dy * dy + dz * dz;
        long q = FID(s, ii, jj, kk);
        if (s->iblank[q] == 0 || s->dead[q] == -1) {
            sumk++;
            dist_int = dist_int + (1.0 / dist);
            p_inv = p_inv + (1.0 / dist) * (p_prev ? p_prev[q] : 0.0);
        }
    }
    for (int r = 0; r < 8; ++r) {
        int ii = s->iDC[n] + incI[r], jj = s->jDC[n] + incJ[r], kk = s->kDC[n] + incK[r];
        long q = FID(s, ii, jj, kk);
        if (s->iblank[q] == 1 && s->dead[q] == 0) {
            if (dist_int < 0.000001 || sumk == 1) {
                pIP = 0.0;
            } else {
                pIP = p_inv / dist_int;
                return pIP + dpdn * s->plD[n];
            }
        }
    }
    for (int r = 0; r < 8; ++r) {
        int ii = s->iDC[n] + incI[r], jj = s->jDC[n] + incJ[r], kk = s->kDC[n] + incK[r];
        if (ii == iG && jj == jG && kk == kG) pIP = pIP + s->cDN[8 * (size_t)n + r] * dpdn;
        else pIP = pIP + s->cDN[8 * (size_t)n + r] * pres[FID(s, ii, jj, kk)];
    }
    return pIP + dpdn * s->plD[n];
}

/* ghost-cell value: GCM_GHOSTCELL_UPDATE.F90:563-643 */
static double ghost_value(or_state *s, int n, const double *pres, const double *r, double *reslap)
{
    int iG = s->iG[n], jG = s->jG[n], kG = s->kG[n];
    double dpdn = (s->prm.is_PBC_homogeneous == 1) ? 0.0 : s->dpdnG[n];
    double pG = 0.0;
    for (int row = 0; row < 8; ++row) {
        int ii = s->iC[n] + incI[row], jj = s->jC[n] + incJ[row], kk = s->kC[n] + incK[row];
        if (ii == iG && jj == jG && kk == kG) pG = pG + s->cN[8 * (size_t)n + row] * dpdn;
        else pG = pG + s->cN[8 * (size_t)n + row] * pres[FID(s, ii, jj, kk)];
    }
    if (s->prm.mix_GC_form == 1) {
        long q = FID(s, iG, jG, kG);
        double pc = pres[q];
        double dpdxW = (pc - pres[q - 1]) * s->dxcinv[iG];
        double dpdxE = (pres[q + 1] - pc) * s->dxcinv[iG + 1];
        double dpdyS = (pc - pres[q - s->P]) * s->dycinv[jG];
        double dpdyN = (pres[q + s->P] - pc) * s->dycinv[jG + 1];
        double dpdzB = (pc - pres[q - s->S]) * s->dzcinv[kG];
        double dpdzF = (pres[q + s->S] - pc) * s->dzcinv[kG + 1];
        double flag = s->alpha[n];
        double coeffG = -(s->dxcinv[iG] + s->dxcinv[iG + 1]) * s->dxinv[iG]
                        - (s->dycinv[jG] + s->dycinv[jG + 1]) * s->dyinv[jG]
                        - (s->dzcinv[kG] + s->dzcinv[kG + 1]) * s->dzinv[kG];
        double resP = (dpdxE - dpdxW) * s->dxinv[iG] + (dpdyN - dpdyS) * s->dyinv[jG]
                      + (dpdzF - dpdzB) * s->dzinv[kG] - r[q];
        double pG2 = pc + resP / (-coeffG);
        *reslap = (-coeffG) * (pG2 - pG) * (1.0 - flag);
        return (1.0 - flag) * pG + flag * pG2;
    }
    *reslap = 0.0;
    return pG;
}

/* GCM_ghostcell_pressure: GCM_GHOSTCELL_UPDATE.F90:361-739 */
int or_ghostcell_pressure(or_state *s, double *pres, const double *r, const double *p_prev,
                          double *res_laplacian, fc_pson_stats *st)
{
    int iterGC = 0;
    double resPresMax = 1.0e10;
    double *newD = NULL, *newG = NULL, *newL = NULL;
    if (s->ordering == OR_ORDER_COLORED) {
        newD = DALLOC(s->nD); newG = DALLOC(s->nG); newL = DALLOC(s->nG);
    }
    while (iterGC < s->prm.itermax_gc && resPresMax > s->prm.restol_gcp) {
        resPresMax = 0.0;
        if (s->ordering == OR_ORDER_REFERENCE) {
            for (int n = 0; n < s->nD; ++n) {
                int jG = s->jD[n], kG = s->kD[n];
                if (jG < 1 || jG > s->ny - 1) continue;
                if (kG < 1 || kG > s->nz - 1) continue;
                long q = FID(s, s->iD[n], jG, kG);
                double pPrev = pres[q];
                pres[q] = dead_value(s, n, pres, p_prev);
                double d = fabs(pres[q] - pPrev);
                if (d > resPresMax) resPresMax = d;
            }
            for (int n = 0; n < s->nG; ++n) {
                long q = FID(s, s->iG[n], s->jG[n], s->kG[n]);
                double pPrev = pres[q], rl = 0.0;
                pres[q] = ghost_value(s, n, pres, r, &rl);
                s->reslap[n] = rl;
                if (res_laplacian) res_laplacian[n] = rl;
                double d = fabs(pres[q] - pPrev);
                if (d > resPresMax) resPresMax = d;
            }
        } else {
            for (int n = 0; n < s->nD; ++n) {
                int jG = s->jD[n], kG = s->kD[n];
                if (jG < 1 || jG > s->ny - 1 || kG < 1 || kG > s->nz - 1) continue;
                newD[n] = dead_value(s, n, pres, p_prev);
            }
            for (int n = 0; n < s->nD; ++n) {
                int jG = s->jD[n], kG = s->kD[n];
                if (jG < 1 || jG > s->ny - 1 || kG < 1 || kG > s->nz - 1) continue;
                long q = FID(s, s->iD[n], jG, kG);
                double d = fabs(newD[n] - pres[q]);
                if (d > resPresMax) resPresMax = d;
                pres[q] = newD[n];
            }
            for (int n = 0; n < s->nG; ++n) newG[n] = ghost_value(s, n, pres, r, &newL[n]);
            for (int n = 0; n < s->nG; ++n) {
                long q = FID(s, s->iG[n], s->jG[n], s->kG[n]);
                double d = fabs(newG[n] - pres[q]);
                if (d > resPresMax) resPresMax = d;
                pres[q] = newG[n];
                s->reslap[n] = newL[n];
                if (res_laplacian) res_laplacian[n] = newL[n];
            }
        }
        iterGC++;
        if (st) st->gcm_sweeps++;
    }
    free(newD); free(newG); free(newL);
    if (st) st->gcm_change = resPresMax;
    return (iterGC == s->prm.itermax_gc && resPresMax > s->prm.restol_gcp) ? 1 : 0;
}

/* GCM_enforce_p_compatibility: GCM_GHOSTCELL_UPDATE.F90:865-1003 (one
 * rank; the mix_GC_form ghost term uses res_laplacian of the last GCM call,
 * zero before the first call) */
void or_enforce_p_compatibility(or_state *s, const double *pres)
{
    const fc_pson_params *p = &s->prm;
    for (int f = 0; f < 6; ++f) if (p->pbc[f] == FC_PBC_DIRICHLET) return;
    double flux = 0.0;
    for (int k = 1; k <= s->nz - 1; ++k)
        for (int j = 1; j <= s->ny - 1; ++j)
            for (int i = 1; i <= s->nx - 1; ++i) {
                long q = FID(s, i, j, k);
                double pgxw = (pres[q] - pres[q - 1]) * s->dxcinv[i];
                double pgxe = (pres[q + 1] - pres[q]) * s->dxcinv[i + 1];
                double pgys = (pres[q] - pres[q - s->P]) * s->dycinv[j];
                double pgyn = (pres[q + s->P] - pres[q]) * s->dycinv[j + 1];
                double pgzb = (pres[q] - pres[q - s->S]) * s->dzcinv[k];
                double pgzf = (pres[q + s->S] - pres[q]) * s->dzcinv[k + 1];
                double flag = (1.0 - (double)s->iblank[q]) * (1.0 - (double)s->ghost[q]);
                flux = flux + flag * ((-pgxw + pgxe) * s->dy[j] * s->dz[k] + (-pgys + pgyn) * s->dx[i] * s->dz[k]
                                      + (-pgzb + pgzf) * s->dx[i] * s->dy[j]);
            }
    if (p->mix_GC_form == 1) {
        for (int n = 0; n < s->nG; ++n) {
            int i = s->iG[n], j = s->jG[n], k = s->kG[n];
            long q = FID(s, i, j, k);
            double pgxw = (pres[q] - pres[q - 1]) * s->dxcinv[i];
            double pgxe = (pres[q + 1] - pres[q]) * s->dxcinv[i + 1];
            double pgys = (pres[q] - pres[q - s->P]) * s->dycinv[j];
            double pgyn = (pres[q + s->P] - pres[q]) * s->dycinv[j + 1];
            double pgzb = (pres[q] - pres[q - s->S]) * s->dzcinv[k];
            double pgzf = (pres[q + s->S] - pres[q]) * s->dzcinv[k + 1];
            flux = flux + ((-pgxw + pgxe) * s->dy[j] * s->dz[k] + (-pgys + pgyn) * s->dx[i] * s->dz[k]
                           + (-pgzb + pgzf) * s->dx[i] * s->dy[j]) - s->reslap[n] * s->dx[i] * s->dy[j] * s->dz[k];
        }
    }
    double corr = -flux / s->outflow_area;
    const int *bv = p->bc_vel;
    for (int k = 0; k <= s->nz; ++k)
        for (int j = 0; j <= s->ny; ++j) {
            if (bv[0] == FC_BC_TYPE_ZERO_GRADIENT) PGX(s->pgx1, j, k) -= corr;
            if (bv[1] == FC_BC_TYPE_ZERO_GRADIENT) PGX(s->pgx2, j, k) += corr;
        }
    for (int k = 0; k <= s->nz; ++k)
        for (int i = 0; i <= s->nx; ++i) {
            if (bv[2] == FC_BC_TYPE_ZERO_GRADIENT) PGY(s->pgy1, i, k) -= corr;
            if (bv[3] == FC_BC_TYPE_ZERO_GRADIENT) PGY(s->pgy2, i, k) += corr;
        }
    for (int j = 0; j <= s->ny; ++j)
        for (int i = 0; i <= s->nx; ++i) {
            if (bv[4] == FC_BC_TYPE_ZERO_GRADIENT) PGZ(s->pgz1, i, j) -= corr;
            if (bv[5] == FC_BC_TYPE_ZERO_GRADIENT) PGZ(s->pgz2, i, j) += corr;
        }
}

/* solve_poisson: CORE_POISSON.F90:83-200 */
int or_solve_poisson(or_state *s, double *pPrime, const double *div, const double *p_prev,
                     double *res_laplacian, fc_pson_stats *st)
{
    const fc_pson_params *p = &s->prm;
    int iter = 0;
    double maxres = 1.0e10, maxres2 = 1.0e10;
    int allN = 1;
    for (int f = 0; f < 6; ++f) if (p->pbc[f] != FC_PBC_NEUMANN) allN = 0;
    fc_pson_stats loc; memset(&loc, 0, sizeof(loc));
    /* periodic z stops the reference (CORE_SET_OUTER_BC.F90:1028-1038) */
    if (p->bc_vel[4] == FC_BC_TYPE_PERIODIC && p->bc_vel[5] == FC_BC_TYPE_PERIODIC) return -2;
    s->pprev_cur = p_prev;
    while (iter < p->itermax_pson && maxres2 > p->restol_gcp) {
        if (allN) {
            double pAverage = pPrime[FID(s, 1, 1, 1)];
            for (int k = 0; k <= s->nz; ++k)
                for (int j = 0; j <= s->ny; ++j)
                    for (int i = 0; i <= s->nx; ++i) pPrime[FID(s, i, j, k)] -= pAverage;
        }
        bc_fill_arr(s, 1, pPrime, s->nx, s->ny, s->nz, s->P, s->ny + 2);
        or_enforce_p_compatibility(s, pPrime);
        memcpy(s->p_temp, pPrime, sizeof(double) * (size_t)s->N);
        or_ghostcell_pressure(s, pPrime, div, p_prev, res_laplacian, &loc);
        if (p->it_solver_type == FC_IT_SOLVER_TYPE_LSOR) or_itsolv(s, pPrime, div);
        else or_mg_solve(s, pPrime, div, &loc);
        maxres = or_calc_residual(s, pPrime, div);
        maxres2 = or_calc_residual2(s, pPrime, s->p_temp);
        iter++;
    }
    loc.outer_iters = iter;
    loc.maxres = maxres;
    s->pprev_cur = NULL;
    loc.maxres2 = maxres2;
    loc.rel_resid = or_rel_residual(s, pPrime, div);
    loc.converged = (maxres2 <= p->restol_gcp) ? 1 : 0;
    if (st) *st = loc;
    return loc.converged ? 0 : 1;
}

/* the mass flux through the faces of cell (i,j,k) (CORE_SET_OUTER_BC.F90:
 * 562-566, GCM_GHOSTCELL_UPDATE.F90:771-774) */
static double cell_massflux(const or_state *s, const double *fu, const double *fv, const double *fw, int i, int j,
                            int k)
{
    long q = FID(s, i, j, k);
    return (-fu[q] + fu[q + 1]) * s->dy[j] * s->dz[k] + (-fv[q] + fv[q + s->P]) * s->dx[i] * s->dz[k]
           + (-fw[q] + fw[q + s->S]) * s->dx[i] * s->dy[j];
}

/* correct_ghost_massflux's weight of ghost cell n (GCM_GHOSTCELL_UPDATE.F90:
 * 766-767) */
static double ghost_flux_weight(const or_state *s, int n)
{
    return cos(3.14159265358979323846 / 2.0 * (1.0 - s->alpha[n]));
}

/* enforce_global_mass_consv: CORE_SET_OUTER_BC.F90:515-735.  With a
 * zero-gradient outer face: the net mass flux of the flag cells (plus the
 * weighted ghost-cell fluxes when mix_GC_form = 1) is removed by a uniform
 * velocity correction on the zero-gradient faces, applied to the outer-face
 * velocities bcxu/bcyv/bczw (0..n of the face plane) and to the face
 * velocities face_u/v/w of those faces (interior cells); outflow_area is the
 * fluid area of those faces.  Returns the correction. */
static double mass_consv(or_state *s, double *fu, double *fv, double *fw)
{
    const int *bv = s->prm.bc_vel;
    int any = 0;
    for (int f = 0; f < 6; ++f) any |= (bv[f] == FC_BC_TYPE_ZERO_GRADIENT);
    if (!any) return 0.0;
    int nx = s->nx, ny = s->ny, nz = s->nz;
    double massflux = 0.0;
    for (int k = 1; k <= nz - 1; ++k)
        for (int j = 1; j <= ny - 1; ++j)
            for (int i = 1; i <= nx - 1; ++i) {
                long q = FID(s, i, j, k);
                double flag = (1.0 - (double)s->iblank[q]) * (1.0 - (double)s->ghost[q]);
                massflux = massflux + flag * cell_massflux(s, fu, fv, fw, i, j, k);
            }
    if (s->prm.mix_GC_form == 1)
        for (int n = 0; n < s->nG; ++n)
            massflux = massflux + ghost_flux_weight(s, n) * cell_massflux(s, fu, fv, fw, s->iG[n], s->jG[n], s->kG[n]);
    double aL = 0.0, aR = 0.0, aB = 0.0, aT = 0.0, aK = 0.0, aF = 0.0;
    for (int k = 1; k <= nz - 1; ++k)
        for (int j = 1; j <= ny - 1; ++j) {
            aL = aL + s->dy[j] * s->dz[k] * (1.0 - s->iblank[FID(s, 1, j, k)]);
            aR = aR + s->dy[j] * s->dz[k] * (1.0 - s->iblank[FID(s, nx - 1, j, k)]);
        }
    for (int k = 1; k <= nz - 1; ++k)
        for (int i = 1; i <= nx - 1; ++i) {
            aB = aB + s->dx[i] * s->dz[k] * (1.0 - s->iblank[FID(s, i, 1, k)]);
            aT = aT + s->dx[i] * s->dz[k] * (1.0 - s->iblank[FID(s, i, ny - 1, k)]);
        }
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            aK = aK + s->dx[i] * s->dy[j] * (1.0 - s->iblank[FID(s, i, j, 1)]);
            aF = aF + s->dx[i] * s->dy[j] * (1.0 - s->iblank[FID(s, i, j, nz - 1)]);
        }
    double area = 0.0;
    if (bv[0] == FC_BC_TYPE_ZERO_GRADIENT) area = area + aL;
    if (bv[1] == FC_BC_TYPE_ZERO_GRADIENT) area = area + aR;
    if (bv[2] == FC_BC_TYPE_ZERO_GRADIENT) area = area + aB;
    if (bv[3] == FC_BC_TYPE_ZERO_GRADIENT) area = area + aT;
    if (bv[4] == FC_BC_TYPE_ZERO_GRADIENT) area = area + aK;
    if (bv[5] == FC_BC_TYPE_ZERO_GRADIENT) area = area + aF;
    s->outflow_area = area;
    double corr = -massflux / area;
    int hy = ny + 2, hx = nx + 2;
    for (int k = 0; k <= nz; ++k)
        for (int j = 0; j <= ny; ++j) {
            if (bv[0] == FC_BC_TYPE_ZERO_GRADIENT) s->bcxu[0 + 2 * (j + hy * k)] -= corr;
            if (bv[1] == FC_BC_TYPE_ZERO_GRADIENT) s->bcxu[1 + 2 * (j + hy * k)] += corr;
        }
    for (int k = 0; k <= nz; ++k)
        for (int i = 0; i <= nx; ++i) {
            if (bv[2] == FC_BC_TYPE_ZERO_GRADIENT) s->bcyv[i + hx * (0 + 2 * k)] -= corr;
            if (bv[3] == FC_BC_TYPE_ZERO_GRADIENT) s->bcyv[i + hx * (1 + 2 * k)] += corr;
        }
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i <= nx; ++i) {
            if (bv[4] == FC_BC_TYPE_ZERO_GRADIENT) s->bczw[i + hx * (j + hy * 0)] -= corr;
            if (bv[5] == FC_BC_TYPE_ZERO_GRADIENT) s->bczw[i + hx * (j + hy * 1)] += corr;
        }
    for (int k = 1; k <= nz - 1; ++k)
        for (int j = 1; j <= ny - 1; ++j) {
            if (bv[0] == FC_BC_TYPE_ZERO_GRADIENT) fu[FID(s, 1, j, k)] -= corr;
            if (bv[1] == FC_BC_TYPE_ZERO_GRADIENT) fu[FID(s, nx, j, k)] += corr;
        }
    for (int k = 1; k <= nz - 1; ++k)
        for (int i = 1; i <= nx - 1; ++i) {
            if (bv[2] == FC_BC_TYPE_ZERO_GRADIENT) fv[FID(s, i, 1, k)] -= corr;
            if (bv[3] == FC_BC_TYPE_ZERO_GRADIENT) fv[FID(s, i, ny, k)] += corr;
        }
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            if (bv[4] == FC_BC_TYPE_ZERO_GRADIENT) fw[FID(s, i, j, 1)] -= corr;
            if (bv[5] == FC_BC_TYPE_ZERO_GRADIENT) fw[FID(s, i, j, nz)] += corr;
        }
    return corr;
}

/* rhs_poisson: CORE_POISSON.F90:9-79: enforce_global_mass_consv on the face
 * velocities (which it changes on zero-gradient faces), the divergence of
 * the flag cells, and correct_ghost_massflux (GCM_GHOSTCELL_UPDATE.F90:
 * 742-787) when mix_GC_form = 1 */
int or_rhs_poisson(or_state *s, double *fu, double *fv, double *fw,
                   double dtinv, double *div, double *sumd, double *resDiv)
{
    mass_consv(s, fu, fv, fw);
    memset(div, 0, sizeof(double) * (size_t)s->N);
    double sd = 0.0, rd = 0.0;
    for (int k = 1; k <= s->nz - 1; ++k)
        for (int j = 1; j <= s->ny - 1; ++j)
            for (int i = 1; i <= s->nx - 1; ++i) {
                long q = FID(s, i, j, k);
                double flag = (1.0 - (double)s->iblank[q]) * (1.0 - (double)s->ghost[q]);
                double d = dtinv * ((fu[q + 1] - fu[q]) * s->dxinv[i] + (fv[q + s->P] - fv[q]) * s->dyinv[j]
                                    + (fw[q + s->S] - fw[q]) * s->dzinv[k]);
                d = d * flag;
                div[q] = d;
                sd = sd + d * s->dx[i] * s->dy[j] * s->dz[k];
                double a = fabs(d / dtinv);
                if (a > rd) rd = a;
            }
    sd = sd / dtinv;
    if (s->prm.mix_GC_form == 1) {
        for (int n = 0; n < s->nG; ++n) {
            int i = s->iG[n], j = s->jG[n], k = s->kG[n];
            long q = FID(s, i, j, k);
            double flag = ghost_flux_weight(s, n);
            sd = sd + flag * cell_massflux(s, fu, fv, fw, i, j, k);
            double d = dtinv * ((fu[q + 1] - fu[q]) * s->dxinv[i] + (fv[q + s->P] - fv[q]) * s->dyinv[j]
                                + (fw[q + s->S] - fw[q]) * s->dzinv[k]);
            div[q] = d * flag;
        }
    }
    if (sumd) *sumd = sd;
    if (resDiv) *resDiv = rd;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Velocity projection: the calls after solve_poisson in a time step
 * (CORE_TIME_STEP.F90:211-222), one flow rank. */

int or_set_vel_bc(or_state *s, const double *bcxu, const double *bcyv, const double *bczw)
{
    size_t nu = 2 * (size_t)(s->ny + 2) * (s->nz + 2), nv = (size_t)(s->nx + 2) * 2 * (s->nz + 2),
           nw = (size_t)(s->nx + 2) * (s->ny + 2) * 2;
    if (bcxu) memcpy(s->bcxu, bcxu, sizeof(double) * nu); else memset(s->bcxu, 0, sizeof(double) * nu);
    if (bcyv) memcpy(s->bcyv, bcyv, sizeof(double) * nv); else memset(s->bcyv, 0, sizeof(double) * nv);
    if (bczw) memcpy(s->bczw, bczw, sizeof(double) * nw); else memset(s->bczw, 0, sizeof(double) * nw);
    return 0;
}

int or_get_vel_bc(or_state *s, double *bcxu, double *bcyv, double *bczw)
{
    size_t nu = 2 * (size_t)(s->ny + 2) * (s->nz + 2), nv = (size_t)(s->nx + 2) * 2 * (s->nz + 2),
           nw = (size_t)(s->nx + 2) * (s->ny + 2) * 2;
    if (bcxu) memcpy(bcxu, s->bcxu, sizeof(double) * nu);
    if (bcyv) memcpy(bcyv, s->bcyv, sizeof(double) * nv);
    if (bczw) memcpy(bczw, s->bczw, sizeof(double) * nw);
    return 0;
}

double or_outflow_area(or_state *s) { return s->outflow_area; }

/* pressure gradient at cell (i,j,k) from face-interpolated values,
 * CORE_FRAC_STEP_ROUTINES.F90:165-181 (and GCM_GHOSTCELL_UPDATE.F90:834-844) */
static void face_gradient(const or_state *s, const double *P, int i, int j, int k, double g[3])
{
    long q = FID(s, i, j, k), sy = s->P, sz = s->S;
    double pe = (s->fx[i + 1] * P[q + 1] + (1.0 - s->fx[i + 1]) * P[q]);
    double pw = (s->fx[i] * P[q] + (1.0 - s->fx[i]) * P[q - 1]);
    double pn = (s->fy[j + 1] * P[q + sy] + (1.0 - s->fy[j + 1]) * P[q]);
    double ps = (s->fy[j] * P[q] + (1.0 - s->fy[j]) * P[q - sy]);
    double pf = (s->fz[k + 1] * P[q + sz] + (1.0 - s->fz[k + 1]) * P[q]);
    double pb = (s->fz[k] * P[q] + (1.0 - s->fz[k]) * P[q - sz]);
    g[0] = (pe - pw) * s->dxinv[i];
    g[1] = (pn - ps) * s->dyinv[j];
    g[2] = (pf - pb) * s->dzinv[k];
}

/* correct_vel: CORE_FRAC_STEP_ROUTINES.F90:134-210 (dt_fac = dt, :160) */
int or_correct_vel(or_state *s, const double *pPrime, double dt, double *u, double *v, double *w)
{
    for (int k = 1; k <= s->nz - 1; ++k)
        for (int j = 1; j <= s->ny - 1; ++j)
            for (int i = 1; i <= s->nx - 1; ++i) {
                long q = FID(s, i, j, k);
                double g[3];
                face_gradient(s, pPrime, i, j, k, g);
                double flag = (1.0 - (double)s->iblank[q]);
                u[q] = u[q] - dt * g[0] * flag;
                v[q] = v[q] - dt * g[1] * flag;
                w[q] = w[q] - dt * g[2] * flag;
            }
    return 0;
}

/* correct_ghostcell_velocity: GCM_GHOSTCELL_UPDATE.F90:789-852, in list order */
int or_correct_ghostcell_velocity(or_state *s, const double *pVar, double dt_fac, double *u, double *v, double *w)
{
    for (int n = 0; n < s->nG; ++n) {
        int i = s->iG[n], j = s->jG[n], k = s->kG[n];
        long q = FID(s, i, j, k);
        double flag = (s->prm.mix_GC_form == 1) ? s->alpha[n] : 0.0;
        double g[3];
        face_gradient(s, pVar, i, j, k, g);
        u[q] = u[q] - dt_fac * g[0] * flag;
        v[q] = v[q] - dt_fac * g[1] * flag;
        w[q] = w[q] - dt_fac * g[2] * flag;
    }
    return 0;
}

/* face_vel: CORE_FRAC_STEP_ROUTINES.F90:6-130 with jProc = kProc = 0,
 * nProcY = nProcZ = 1 (j1 = 2, j2 = ny-1, k1 = 2, k2 = nz-1) */
int or_face_vel(or_state *s, const double *u, const double *v, const double *w, double *fu, double *fv, double *fw)
{
    int nx = s->nx, ny = s->ny, nz = s->nz;
    long sy = s->P, sz = s->S, hx = nx + 2, hy = ny + 2;
    for (int k = 1; k <= nz - 1; ++k)
        for (int j = 1; j <= ny - 1; ++j)
            for (int i = 2; i <= nx - 1; ++i) {
                long q = FID(s, i, j, k);
                fu[q] = (s->fx[i] * u[q] + (1.0 - s->fx[i]) * u[q - 1]);
            }
    for (int j = 2; j <= ny - 1; ++j)
        for (int k = 1; k <= nz - 1; ++k)
            for (int i = 1; i <= nx - 1; ++i) {
                long q = FID(s, i, j, k);
                fv[q] = (s->fy[j] * v[q] + (1.0 - s->fy[j]) * v[q - sy]);
            }
    for (int k = 2; k <= nz - 1; ++k)
        for (int j = 1; j <= ny - 1; ++j)
            for (int i = 1; i <= nx - 1; ++i) {
                long q = FID(s, i, j, k);
                fw[q] = (s->fz[k] * w[q] + (1.0 - s->fz[k]) * w[q - sz]);
            }
    for (int k = 1; k <= nz - 1; ++k)
        for (int j = 1; j <= ny - 1; ++j) {
            fu[FID(s, 1, j, k)] = s->bcxu[0 + 2 * (j + hy * k)];
            fu[FID(s, nx, j, k)] = s->bcxu[1 + 2 * (j + hy * k)];
        }
    for (int k = 1; k <= nz - 1; ++k)
        for (int i = 1; i <= nx - 1; ++i) {
            fv[FID(s, i, 1, k)] = s->bcyv[i + hx * (0 + 2 * k)];
            fv[FID(s, i, ny, k)] = s->bcyv[i + hx * (1 + 2 * k)];
        }
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            fw[FID(s, i, j, 1)] = s->bczw[i + hx * (j + hy * 0)];
            fw[FID(s, i, j, nz)] = s->bczw[i + hx * (j + hy * 1)];
        }
    (void)sz;
    return 0;
}

/* update_pressure: CORE_FRAC_STEP_ROUTINES.F90:212-299 with jProc = kProc =
 * 0, nProcY = nProcZ = 1, yc_start = zc_start = 1, yc_end = ny-1,
 * zc_end = nz-1.  The assignments at :279 and :282 name the plane
 * k = zc_start-2 = -1, which is outside pPrime(0:nx+1, 0:ny+1, 0:nz+1) at
 * one rank; only their in-array part (k = 0, also set at :271 and :274) is
 * done. */
int or_update_pressure(or_state *s, double *pPrime, double *p)
{
    int nx = s->nx, ny = s->ny, nz = s->nz;
#define PP(i, j, k) pPrime[FID(s, (i), (j), (k))]
    /* :225-236 */
    for (int k = -1; k <= nz + 1; ++k) {
        if (k == -1) continue;
        for (int j = -1; j <= ny + 1; ++j) {
            if (j == -1) continue;
            for (int i = 0; i <= nx; ++i) {
                long q = FID(s, i, j, k);
                if (s->iblank[q] == 0 || s->dead[q] == -1) continue;
                pPrime[q] = 0.0;
            }
        }
    }
    int j1 = 0, k1 = 0;    /* yc_start-1, zc_start-1 */
    /* :250-268, the four corner-row blocks (all four apply at one rank) */
    for (int i = 0; i <= nx; ++i) {
        for (int k = k1; k <= 0; ++k) { PP(i, 0, k) = 0.0; PP(i, ny, k) = 0.0; }
        for (int k = nz; k <= nz + 1; ++k) { PP(i, 0, k) = 0.0; PP(i, ny, k) = 0.0; }
    }
    /* :270-284 */
    for (int j = j1; j <= ny + 1; ++j) {
        PP(0, j, 0) = 0.0; PP(nx, j, 0) = 0.0;
        for (int k = nz; k <= nz + 1; ++k) { PP(0, j, k) = 0.0; PP(nx, j, k) = 0.0; }
    }
    /* :286-294 */
    for (int k = k1; k <= nz + 1; ++k) {
        PP(0, 0, k) = 0.0; PP(nx, 0, k) = 0.0;
        PP(0, ny, k) = 0.0; PP(nx, ny, k) = 0.0;
    }
#undef PP
    memcpy(p, pPrime, sizeof(double) * (size_t)s->N);    /* :297 */
    return 0;
}

/* ------------------------------------------------------------------ */
/* Advection-diffusion line solves (CORE_SOLVE_AD.F90), one flow rank. */

/* set_solve_ad: CORE_SET_SOLVERS.F90:591-676 (nu = reinv on every face);
 * half_dt = dt/2 under CRANK_NICOLSON2, else 0 (CORE_SOLVE_AD.F90:196-200) */
int or_set_solve_ad(or_state *s, double reinv, double dt, int crank_nicolson, double omega_ad)
{
    memset(s->amx_ad, 0, sizeof(double) * (s->nx + 2)); memset(s->apx_ad, 0, sizeof(double) * (s->nx + 2));
    memset(s->amy_ad, 0, sizeof(double) * (s->ny + 2)); memset(s->apy_ad, 0, sizeof(double) * (s->ny + 2));
    memset(s->amz_ad, 0, sizeof(double) * (s->nz + 2)); memset(s->apz_ad, 0, sizeof(double) * (s->nz + 2));
    for (int i = 1; i <= s->nx - 1; ++i) {
        s->amx_ad[i] = (s->dxcinv[i]) * s->dxinv[i];
        s->apx_ad[i] = (s->dxcinv[i + 1]) * s->dxinv[i];
        s->amx_ad[i] = -(0.50 * dt * reinv) * s->amx_ad[i];
        s->apx_ad[i] = -(0.50 * dt * reinv) * s->apx_ad[i];
    }
    for (int j = 1; j <= s->ny - 1; ++j) {
        s->amy_ad[j] = s->dycinv[j] * s->dyinv[j];
        s->apy_ad[j] = s->dycinv[j + 1] * s->dyinv[j];
        s->amy_ad[j] = -(0.50 * dt * reinv) * s->amy_ad[j];
        s->apy_ad[j] = -(0.50 * dt * reinv) * s->apy_ad[j];
    }
    for (int k = 1; k <= s->nz - 1; ++k) {
        s->amz_ad[k] = s->dzcinv[k] * s->dzinv[k];
        s->apz_ad[k] = s->dzcinv[k + 1] * s->dzinv[k];
        s->amz_ad[k] = -(0.50 * dt * reinv) * s->amz_ad[k];
        s->apz_ad[k] = -(0.50 * dt * reinv) * s->apz_ad[k];
    }
    s->half_dt = crank_nicolson ? 0.5 * dt : 0.0;
    s->omega_ad = omega_ad;
    s->reinv_ad = reinv;
    return 0;
}

void or_get_solve_ad(or_state *s, double *amx, double *apx, double *amy, double *apy, double *amz, double *apz)
{
    memcpy(amx, s->amx_ad, sizeof(double) * (s->nx + 2)); memcpy(apx, s->apx_ad, sizeof(double) * (s->nx + 2));
    memcpy(amy, s->amy_ad, sizeof(double) * (s->ny + 2)); memcpy(apy, s->apy_ad, sizeof(double) * (s->ny + 2));
    memcpy(amz, s->amz_ad, sizeof(double) * (s->nz + 2)); memcpy(apz, s->apz_ad, sizeof(double) * (s->nz + 2));
}

static double sign1(double b) { return (b >= 0.0) ? 1.0 : -1.0; }   /* SIGN(1.0, b); SIGN(1,-0.0) = 1 */

/* the nine coefficients of cell (i,j,k): diffusion plus the Crank-Nicolson
 * advection terms, central for 25 <= i <= 375 and first-order upwind
 * otherwise (the switch tests i on every axis); CORE_SOLVE_AD.F90:212-297
 * (x-sweep; the y- and z-sweeps and calc_residual_ad repeat it).
 * c = {amx, apx, acx, amy, apy, acy, amz, apz, acz} */
static void ad_coef(const or_state *s, const double *fu, const double *fv, const double *fw, int i, int j, int k,
                    double c[9])
{
    long q = FID(s, i, j, k), sy = s->P, sz = s->S;
    double hd = s->half_dt, hyb1, hyb2, tmp1, tmp2;
    int central = (i >= 25) && (i <= 375);
    double amx = s->amx_ad[i], apx = s->apx_ad[i], acx = -(amx + apx);
    double amy = s->amy_ad[j], apy = s->apy_ad[j], acy = -(amy + apy);
    double amz = s->amz_ad[k], apz = s->apz_ad[k], acz = -(amz + apz);
    double ue = fu[q + 1], uw = fu[q], vn = fv[q + sy], vs = fv[q], wf = fw[q + sz], wb = fw[q];

    hyb1 = 0.5 * (1.0 + sign1(uw)); hyb2 = 0.5 * (1.0 - sign1(ue));
    tmp1 = 1.0 - s->fx[i]; tmp2 = s->fx[i + 1];
    if (central) { amx = amx - hd * uw * tmp1 * s->dxinv[i]; apx = apx + hd * ue * tmp2 * s->dxinv[i]; }
    else { amx = amx - hd * hyb1 * uw * s->dxinv[i]; apx = apx + hd * hyb2 * ue * s->dxinv[i]; }
    hyb1 = 0.5 * (1.0 - sign1(uw)); hyb2 = 0.5 * (1.0 + sign1(ue));
    tmp1 = s->fx[i]; tmp2 = (1.0 - s->fx[i + 1]);
    if (central) acx = acx + hd * (ue * tmp2 - uw * tmp1) * s->dxinv[i];
    else acx = acx + hd * (ue * hyb2 - uw * hyb1) * s->dxinv[i];

    hyb1 = 0.5 * (1.0 + sign1(vs)); hyb2 = 0.5 * (1.0 - sign1(vn));
    tmp1 = 1.0 - s->fy[j]; tmp2 = s->fy[j + 1];
    if (central) { amy = amy - hd * vs * tmp1 * s->dyinv[j]; apy = apy + hd * vn * tmp2 * s->dyinv[j]; }
    else { amy = amy - hd * hyb1 * vs * s->dyinv[j]; apy = apy + hd * hyb2 * vn * s->dyinv[j]; }
    hyb1 = 0.5 * (1.0 - sign1(vs)); hyb2 = 0.5 * (1.0 + sign1(vn));
    tmp1 = s->fy[j]; tmp2 = (1.0 - s->fy[j + 1]);
    if (central) acy = acy + hd * (vn * tmp2 - vs * tmp1) * s->dyinv[j];
    else acy = acy + hd * (vn * hyb2 - vs * hyb1) * s->dyinv[j];

    hyb1 = 0.5 * (1.0 + sign1(wb)); hyb2 = 0.5 * (1.0 - sign1(wf));
    tmp1 = 1.0 - s->fz[k]; tmp2 = s->fz[k + 1];
    if (central) { amz = amz - hd * wb * tmp1 * s->dzinv[k]; apz = apz + hd * wf * tmp2 * s->dzinv[k]; }
    else { amz = amz - hd * hyb1 * wb * s->dzinv[k]; apz = apz + hd * hyb2 * wf * s->dzinv[k]; }
    hyb1 = 0.5 * (1.0 - sign1(wb)); hyb2 = 0.5 * (1.0 + sign1(wf));
    tmp1 = s->fz[k]; tmp2 = (1.0 - s->fz[k + 1]);
    if (central) acz = acz + hd * (wf * tmp2 - wb * tmp1) * s->dzinv[k];
    else acz = acz + hd * (wf * hyb2 - wb * hyb1) * s->dzinv[k];

    c[0] = amx; c[1] = apx; c[2] = acx; c[3] = amy; c[4] = apy; c[5] = acy; c[6] = amz; c[7] = apz; c[8] = acz;
}

/* one line of direction dir through (i,j,k) (the line coordinate is
 * ignored): rows 0..n with trivial rows at both ends, off-line neighbours
 * moved to the right-hand side, masked rows made identities, tdma, then
 * var += omega_ad (x - var) on rows 1..n-1.  x: CORE_SOLVE_AD.F90:206-346;
 * y: :400-545; z: :600-772 */
static void ad_line(or_state *s, double *var, const double *r, const double *fu, const double *fv, const double *fw,
                    int dir, int i, int j, int k)
{
    int n = dir == 0 ? s->nx : dir == 1 ? s->ny : s->nz;
    long st = dir == 0 ? 1 : dir == 1 ? s->P : s->S;
    long sa = dir == 0 ? s->P : 1, sb = dir == 2 ? s->P : s->S;   /* the two off-line strides */
    int ca = dir == 0 ? 3 : 0, cb = dir == 2 ? 3 : 6;             /* their coefficient triples */
    int cl = 3 * dir;
    double *a = s->am, *b = s->ac, *c = s->ap, *rh = s->rr;
    for (int m = 1; m <= n - 1; ++m) {
        int ii = dir == 0 ? m : i, jj = dir == 1 ? m : j, kk = dir == 2 ? m : k;
        long q = FID(s, ii, jj, kk);
        double cf[9];
        ad_coef(s, fu, fv, fw, ii, jj, kk, cf);
        double rhs = r[q] - (var[q - sa] * cf[ca] + var[q + sa] * cf[ca + 1] + var[q - sb] * cf[cb] + var[q + sb] * cf[cb + 1]);
        double rib = (1.0 - (double)s->iblank[q]) * (1.0 - (double)abs(s->ghost[q]));
        a[m] = cf[cl] * rib;
        c[m] = cf[cl + 1] * rib;
        b[m] = 1.0 + (cf[2] + cf[5] + cf[8]) * rib;
        rh[m] = rhs * rib + var[q] * (1.0 - rib);
    }
    long q0 = FID(s, dir == 0 ? 0 : i, dir == 1 ? 0 : j, dir == 2 ? 0 : k);
    a[0] = 0.0; c[0] = 0.0; b[0] = 1.0; rh[0] = var[q0];
    a[n] = 0.0; c[n] = 0.0; b[n] = 1.0; rh[n] = var[q0 + st * n];
    tdma_s(s, 0, n);
    for (int m = 1; m <= n - 1; ++m) {
        long q = q0 + st * m;
        var[q] = var[q] + s->omega_ad * (s->uu[m] - var[q]);
    }
}

/* itsolv_ad: x-, then y-, then z-line solves (CORE_SOLVE_AD.F90:131-159).
 * ORDER_REFERENCE: lexicographic line order of the reference (x: k, j;
 * y: k, i; z: j, i).  ORDER_COLORED: the GPU order, per direction the lines
 * of colour 0 then colour 1 (x: (j+k)%2, y: (i+k)%2, z: (i+j)%2). */
int or_itsolv_ad(or_state *s, double *var, const double *r, const double *fu, const double *fv, const double *fw)
{
    int nx = s->nx, ny = s->ny, nz = s->nz;
    int ncol = s->ordering == OR_ORDER_COLORED ? 2 : 1;
    for (int col = 0; col < ncol; ++col)
        for (int k = 1; k <= nz - 1; ++k)
            for (int j = 1; j <= ny - 1; ++j)
                if (ncol == 1 || (j + k) % 2 == col) ad_line(s, var, r, fu, fv, fw, 0, 0, j, k);
    for (int col = 0; col < ncol; ++col)
        for (int k = 1; k <= nz - 1; ++k)
            for (int i = 1; i <= nx - 1; ++i)
                if (ncol == 1 || (i + k) % 2 == col) ad_line(s, var, r, fu, fv, fw, 1, i, 0, k);
    for (int col = 0; col < ncol; ++col)
        for (int j = 1; j <= ny - 1; ++j)
            for (int i = 1; i <= nx - 1; ++i)
                if (ncol == 1 || (i + j) % 2 == col) ad_line(s, var, r, fu, fv, fw, 2, i, j, 0);
    return 0;
}

/* calc_residual_ad: CORE_SOLVE_AD.F90:780-964; max |res| over the cells
 * with iblank = 0 and ghostCellMark = 0, its first location in (k, j, i)
 * loop order (1-based values of the array indices), and the sum of res^2 */
double or_calc_residual_ad(or_state *s, const double *var, const double *r, const double *fu, const double *fv,
                           const double *fw, int *loc, double *res2)
{
    double resm = 0.0, r2 = 0.0;
    long sy = s->P, sz = s->S;
    loc[0] = loc[1] = loc[2] = 0;
    for (int k = 1; k <= s->nz - 1; ++k)
        for (int j = 1; j <= s->ny - 1; ++j)
            for (int i = 1; i <= s->nx - 1; ++i) {
                long q = FID(s, i, j, k);
                double c[9];
                ad_coef(s, fu, fv, fw, i, j, k, c);
                double res = r[q] - var[q] * (1.0 + c[2] + c[5] + c[8])
                             - var[q - 1] * c[0] - var[q + 1] * c[1]
                             - var[q - sy] * c[3] - var[q + sy] * c[4]
                             - var[q - sz] * c[6] - var[q + sz] * c[7];
                res = res * (1.0 - (double)s->iblank[q]) * (1.0 - (double)s->ghost[q]);
                if (fabs(res) > resm) { resm = fabs(res); loc[0] = i; loc[1] = j; loc[2] = k; }
                r2 = r2 + res * res;
            }
    if (res2) *res2 = r2;
    return resm;
}

/* ------------------------------------------------------------------ */
/* rhs_advec_diff (CORE_SET_SOLVERS.F90:10-61): nl = rhs_advec(vel), then
 * rhs_diff(vel, nl), for (uOld, nlu), (vOld, nlv), (wOld, nlw).  Every
 * interior cell (i = 1..nx-1, j = 1..ny-1, k = 1..nz-1) is written; the
 * other entries of nl keep their values. */

/* One axis of rhs_advec (x: CORE_SET_SOLVERS.F90:96-242, y: :244-334,
 * z: :336-434) at cell q, index n on that axis (node count N, stride st,
 * interpolation factors f, inverse spacings dinv, face velocities F):
 * central face values ve/vw, second-order upwind ones (the 2-point
 * extrapolations with the weights of :105-121), first-order upwind ones,
 * the hybrid switch on i (25..375 second order, else first order; the
 * reference tests i on every axis), the inflow special cases at n = 1 and
 * n = N-1, and the blend with alfa = alfa_upwind. */
static double advec_axis(const double *vel, const double *F, long q, long st, int n, int N, const double *f,
                         const double *dinv, int i, double alfa)
{
    int nm = n - 2 > 0 ? n - 2 : 0, np = n + 2 < N ? n + 2 : N;
    double sg = sign1(F[q]), sgp = sign1(F[q + st]);
    double ve = f[n + 1] * vel[q + st] + (1.0 - f[n + 1]) * vel[q];
    double vw = f[n] * vel[q] + (1.0 - f[n]) * vel[q - st];
    double em1 = (dinv[n - 1] + 2.0 * dinv[n]) / (dinv[n - 1] + dinv[n]);
    double em2 = dinv[n] / (dinv[n - 1] + dinv[n]);
    double ep1 = (dinv[np] + 2.0 * dinv[n + 1]) / (dinv[np] + dinv[n + 1]);
    double ep2 = dinv[n + 1] / (dinv[np] + dinv[n + 1]);
    double wm1 = (dinv[nm] + 2.0 * dinv[n - 1]) / (dinv[nm] + dinv[n - 1]);
    double wm2 = dinv[n - 1] / (dinv[nm] + dinv[n - 1]);
    double wp1 = (dinv[n + 1] + 2.0 * dinv[n]) / (dinv[n + 1] + dinv[n]);
    double wp2 = dinv[n] / (dinv[n + 1] + dinv[n]);
    double ve2, vw2, ve1, vw1, veu, vwu;
    if (n == N - 1 && sgp < -0.9) ve2 = ve;
    else ve2 = ((1.0 + sgp) * (em1 * vel[q] - em2 * vel[q - st])
                + (1.0 - sgp) * (ep1 * vel[q + st] - ep2 * vel[q + (long)(np - n) * st])) * 0.5;
    if (n == 1 && sg > 0.9) vw2 = vw;
    else vw2 = ((1.0 + sg) * (wm1 * vel[q - st] - wm2 * vel[q + (long)(nm - n) * st])
                + (1.0 - sg) * (wp1 * vel[q] - wp2 * vel[q + st])) * 0.5;
    if (n == N - 1 && sgp < -0.9) ve1 = ve;
    else ve1 = ((1.0 + sgp) * (vel[q]) + (1.0 - sgp) * (vel[q + st])) * 0.5;
    if (n == 1 && sg > 0.9) vw1 = vw;
    else vw1 = ((1.0 + sg) * (vel[q - st]) + (1.0 - sg) * (vel[q])) * 0.5;
    if (i >= 25 && i <= 375) { veu = ve2; vwu = vw2; }
    else { veu = ve1; vwu = vw1; }
    return (((1.0 - alfa) * ve + alfa * veu) * F[q + st] - ((1.0 - alfa) * vw + alfa * vwu) * F[q]) * dinv[n];
}

/* rhs_diff (CORE_SET_SOLVERS.F90:440-526) at cell q: face viscosities
 * interpolated from viscTot (constant reinv when visc is NULL), the
 * diffusion stencil, nl = vel + dt/2 (-nl + diffvel) */
static double diff_cell(const or_state *s, const double *vel, const double *visc, double nl, double dt, int i,
                        int j, int k)
{
    long q = FID(s, i, j, k), sy = s->P, sz = s->S;
    double c = s->reinv_ad;
#define VT(o) (visc ? visc[q + (o)] : c)
    double nuE = s->fx[i + 1] * VT(1) + (1.0 - s->fx[i + 1]) * VT(0);
    double nuW = s->fx[i] * VT(0) + (1.0 - s->fx[i]) * VT(-1);
    double nuN = s->fy[j + 1] * VT(sy) + (1.0 - s->fy[j + 1]) * VT(0);
    double nuS = s->fy[j] * VT(0) + (1.0 - s->fy[j]) * VT(-sy);
    double nuF = s->fz[k + 1] * VT(sz) + (1.0 - s->fz[k + 1]) * VT(0);
    double nuB = s->fz[k] * VT(0) + (1.0 - s->fz[k]) * VT(-sz);
#undef VT
    double aamx = s->dxcinv[i] * s->dxinv[i], aapx = s->dxcinv[i + 1] * s->dxinv[i];
    aamx = aamx * nuW; aapx = aapx * nuE;
    double aacx = -(aamx + aapx);
    double aamy = s->dycinv[j] * s->dyinv[j], aapy = s->dycinv[j + 1] * s->dyinv[j];
    aamy = aamy * nuS; aapy = aapy * nuN;
    double aacy = -(aamy + aapy);
    double aamz = s->dzcinv[k] * s->dzinv[k], aapz = s->dzcinv[k + 1] * s->dzinv[k];
    aamz = aamz * nuB; aapz = aapz * nuF;
    double aacz = -(aamz + aapz);
    double diffvel = aamx * vel[q - 1] + aacx * vel[q] + aapx * vel[q + 1]
                   + aamy * vel[q - sy] + aacy * vel[q] + aapy * vel[q + sy]
                   + aamz * vel[q - sz] + aacz * vel[q] + aapz * vel[q + sz];
    return vel[q] + 0.5 * dt * (-nl + diffvel);
}

int or_rhs_advec_diff(or_state *s, const double *uo, const double *vo, const double *wo, const double *fu,
                      const double *fv, const double *fw, const double *visc, double alfa_upwind, double dt,
                      double *nlu, double *nlv, double *nlw)
{
    const double *vel[3] = {uo, vo, wo};
    double *nl[3] = {nlu, nlv, nlw};
    int nx = s->nx, ny = s->ny, nz = s->nz;
    long sy = s->P, sz = s->S;
    for (int c = 0; c < 3; ++c) {          /* rhs_advec(uOld, nlu), (vOld, nlv), (wOld, nlw) */
        for (int k = 1; k <= nz - 1; ++k)
            for (int j = 1; j <= ny - 1; ++j)
                for (int i = 1; i <= nx - 1; ++i)
                    nl[c][FID(s, i, j, k)] = advec_axis(vel[c], fu, FID(s, i, j, k), 1, i, nx, s->fx, s->dxinv, i,
                                                        alfa_upwind);
        for (int k = 1; k <= nz - 1; ++k)
            for (int j = 1; j <= ny - 1; ++j)
                for (int i = 1; i <= nx - 1; ++i) {
                    long q = FID(s, i, j, k);
                    nl[c][q] = nl[c][q] + advec_axis(vel[c], fv, q, sy, j, ny, s->fy, s->dyinv, i, alfa_upwind);
                }
        for (int k = 1; k <= nz - 1; ++k)
            for (int j = 1; j <= ny - 1; ++j)
                for (int i = 1; i <= nx - 1; ++i) {
                    long q = FID(s, i, j, k);
                    nl[c][q] = nl[c][q] + advec_axis(vel[c], fw, q, sz, k, nz, s->fz, s->dzinv, i, alfa_upwind);
                }
    }
    for (int c = 0; c < 3; ++c)            /* rhs_diff(uOld, nlu), ... */
        for (int k = 1; k <= nz - 1; ++k)
            for (int j = 1; j <= ny - 1; ++j)
                for (int i = 1; i <= nx - 1; ++i) {
                    long q = FID(s, i, j, k);
                    nl[c][q] = diff_cell(s, vel[c], visc, nl[c][q], dt, i, j, k);
                }
    return 0;
}

/* ------------------------------------------------------------------ */
void *or_level_ptr(or_state *s, int lev, int which)
{
    if (lev < 1 || lev > s->nlev) return NULL;
    olev *L = &s->lev[lev];
    switch (which) {
    case 0: return L->phi;
    case 1: return L->rhs;
    case 2: return L->res;
    case 3: return L->ibk;
    case 4: return L->igk;
    default: return NULL;
    }
}

void or_level_metric(or_state *s, int lev, int which, double **ptr)
{
    olev *L = &s->lev[lev];
    double *arr[12] = {L->x, L->y, L->z, L->xc, L->yc, L->zc, L->dxinv, L->dyinv, L->dzinv, L->dxcinv, L->dycinv, L->dzcinv};
    *ptr = (which >= 0 && which < 12) ? arr[which] : NULL;
}

/* ------------------------------------------------------------------ */
/* solve_ad's loop (CORE_SOLVE_AD.F90:11-127) and the routines it calls
 * besides itsolv_ad: set_outer_velocity_bc, set_outer_ghost_velocity,
 * GCM_ghostcell_velocity.  One flow rank: the buffer ranges yb1..yb2 and
 * zb1..zb2 are 0..ny+1 and 0..nz+1. */

int or_set_outer_velocity(or_state *s, const double *vals, const double *freq, double y_origin, double y_out)
{
    for (int t = 0; t < 18; ++t) {
        s->vbc_val[t] = vals ? vals[t] : 0.0;
        s->vbc_freq[t] = freq ? freq[t] : 0.0;
    }
    s->y_origin = y_origin;
    s->y_out = y_out;
    return 0;
}

/* the nine outer-face arrays: axis a (x, y, z) and component c (u, v, w) */
static double *vbc_arr(or_state *s, int a, int c)
{
    double *t[3][3] = {{s->bcxu, s->bcxv, s->bcxw}, {s->bcyu, s->bcyv, s->bcyw}, {s->bczu, s->bczv, s->bczw}};
    return t[a][c];
}

/* the value one face type sets (CORE_SET_OUTER_BC.F90:37-506) for component
 * c at face f = 2 axis + side, from the interior value vi, the value vp at
 * the other end (periodic) and the node coordinate yj of the row (x faces,
 * user-specified channel profile) */
static double vbc_value(const or_state *s, int f, int c, double vi, double vp, double yj, double time)
{
    const int a = f / 2, side = f % 2;
    const double val = s->vbc_val[6 * a + 2 * c + side], fr = s->vbc_freq[6 * a + 2 * c + side];
    const double pi = 4.0 * atan(1.0);
    switch (s->prm.bc_vel[f]) {
    case FC_BC_TYPE_DIRICHLET: return val;
    case FC_BC_TYPE_ZERO_GRADIENT: return vi;
    case 3: /* pulsatile inflow; the x1 face's u ramps linearly in time */
        if (f == 0 && c == 0) return val * time / 30.0;
        return val * sin(2.0 * pi * fr * time);
    case 4: /* symmetry: the normal component vanishes */
        return (c == a) ? 0.0 : vi;
    case FC_BC_TYPE_PERIODIC: return 0.5 * (vi + vp);
    case 6: /* user specified, x faces: channel profile in y for u */
        if (c == 0) {
            const double yo = s->y_origin, yt = s->y_out;
            return val * 6.0 * (yj - yo) * (yt - yj) / ((yt - yo) * (yt - yo));
        }
        return val;
    }
    return 0.0;
}

/* set_outer_velocity_bc: CORE_SET_OUTER_BC.F90:18-509.  Returns -2 for a
 * face type this restatement does not carry: shear (7, the reference stops)
 * and user specified (6) on y or z faces (the reference writes the x arrays
 * there). */
int or_set_outer_velocity_bc(or_state *s, const double *u, const double *v, const double *w, double time)
{
    const int nx = s->nx, ny = s->ny, nz = s->nz, hx = nx + 2, hy = ny + 2;
    const double *U[3] = {u, v, w};
    for (int f = 0; f < 6; ++f) {
        const int t = s->prm.bc_vel[f];
        if (t == 7 || (t == 6 && f >= 2) || t < 1 || t > 7) return -2;
    }
    /* x faces: k, j over the whole planes */
    for (int side = 0; side < 2; ++side) {
        const int f = side, i = side ? nx - 1 : 1, ip = side ? 1 : nx - 1;
        for (int k = 0; k <= nz + 1; ++k)
            for (int j = 0; j <= ny + 1; ++j)
                for (int c = 0; c < 3; ++c)
                    vbc_arr(s, 0, c)[side + 2 * (j + (long)hy * k)] =
                        vbc_value(s, f, c, U[c][FID(s, i, j, k)], U[c][FID(s, ip, j, k)], s->y[j], time);
    }
    /* y faces: k, i = 0..nx */
    for (int side = 0; side < 2; ++side) {
        const int f = 2 + side, j = side ? ny - 1 : 1, jp = side ? 1 : ny - 1;
        for (int k = 0; k <= nz + 1; ++k)
            for (int i = 0; i <= nx; ++i)
                for (int c = 0; c < 3; ++c)
                    vbc_arr(s, 1, c)[i + hx * (side + 2 * (long)k)] =
                        vbc_value(s, f, c, U[c][FID(s, i, j, k)], U[c][FID(s, i, jp, k)], 0.0, time);
    }
    /* z faces: i = 0..nx, j over the whole plane */
    for (int side = 0; side < 2; ++side) {
        const int f = 4 + side, k = side ? nz - 1 : 1, kp = side ? 1 : nz - 1;
        for (int i = 0; i <= nx; ++i)
            for (int j = 0; j <= ny + 1; ++j)
                for (int c = 0; c < 3; ++c)
                    vbc_arr(s, 2, c)[i + hx * (j + (long)hy * side)] =
                        vbc_value(s, f, c, U[c][FID(s, i, j, k)], U[c][FID(s, i, j, kp)], 0.0, time);
    }
    return 0;
}

/* set_outer_ghost_velocity: CORE_SET_OUTER_BC.F90:739-823: ghost = 2 bc -
 * interior, x planes first (whole planes), then y (i = 0..nx), then z
 * (i = 0..nx) */
int or_set_outer_ghost_velocity(or_state *s, double *u, double *v, double *w)
{
    const int nx = s->nx, ny = s->ny, nz = s->nz, hx = nx + 2, hy = ny + 2;
    double *U[3] = {u, v, w};
    for (int k = 0; k <= nz + 1; ++k)
        for (int j = 0; j <= ny + 1; ++j)
            for (int c = 0; c < 3; ++c) {
                const double *b = vbc_arr(s, 0, c);
                U[c][FID(s, 0, j, k)] = 2.0 * b[0 + 2 * (j + (long)hy * k)] - U[c][FID(s, 1, j, k)];
                U[c][FID(s, nx, j, k)] = 2.0 * b[1 + 2 * (j + (long)hy * k)] - U[c][FID(s, nx - 1, j, k)];
            }
    for (int k = 0; k <= nz + 1; ++k)
        for (int i = 0; i <= nx; ++i)
            for (int c = 0; c < 3; ++c) {
                const double *b = vbc_arr(s, 1, c);
                U[c][FID(s, i, 0, k)] = 2.0 * b[i + hx * (0 + 2 * (long)k)] - U[c][FID(s, i, 1, k)];
                U[c][FID(s, i, ny, k)] = 2.0 * b[i + hx * (1 + 2 * (long)k)] - U[c][FID(s, i, ny - 1, k)];
            }
    for (int i = 0; i <= nx; ++i)
        for (int j = 0; j <= ny + 1; ++j)
            for (int c = 0; c < 3; ++c) {
                const double *b = vbc_arr(s, 2, c);
                U[c][FID(s, i, j, 0)] = 2.0 * b[i + hx * (j + (long)hy * 0)] - U[c][FID(s, i, j, 1)];
                U[c][FID(s, i, j, nz)] = 2.0 * b[i + hx * (j + (long)hy * 1)] - U[c][FID(s, i, j, nz - 1)];
            }
    return 0;
}

int or_get_outer_velocity_bc(or_state *s, double *const *out9)
{
    size_t n[3] = {2 * (size_t)(s->ny + 2) * (s->nz + 2), (size_t)(s->nx + 2) * 2 * (s->nz + 2),
                   (size_t)(s->nx + 2) * (s->ny + 2) * 2};
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c)
            if (out9[3 * a + c]) memcpy(out9[3 * a + c], vbc_arr(s, a, c), sizeof(double) * n[a]);
    return 0;
}

int or_set_gcm_vel(or_state *s, const double *coeffD, const double *coeffDeadD, const double *ubi,
                   const double *ubiDead, double probe_norm_d)
{
    free(s->cD); free(s->cDD); free(s->ubi); free(s->ubiD);
    s->cD = DALLOC(8 * (size_t)s->nG); s->cDD = DALLOC(8 * (size_t)s->nD);
    s->ubi = DALLOC(3 * (size_t)s->nG); s->ubiD = DALLOC(3 * (size_t)s->nD);
    if (coeffD) memcpy(s->cD, coeffD, sizeof(double) * 8 * (size_t)s->nG);
    if (coeffDeadD) memcpy(s->cDD, coeffDeadD, sizeof(double) * 8 * (size_t)s->nD);
    if (ubi) memcpy(s->ubi, ubi, sizeof(double) * 3 * (size_t)s->nG);
    if (ubiDead) memcpy(s->ubiD, ubiDead, sizeof(double) * 3 * (size_t)s->nD);
    s->probe_norm_d = probe_norm_d;
    s->have_gcm_vel = 1;
    return 0;
}

/* GCM_ghostcell_velocity(iOpt, u, v, w): GCM_GHOSTCELL_UPDATE.F90:1-359, one
 * rank: Gauss-Seidel sweeps in list order (dead cells, then, with iOpt = 2,
 * ghost cells) until the largest |du|+|dv|+|dw| of a sweep is at most
 * restol_gcv.  Dead cells next to a pure solid cell take the inverse
 * distance average of the fluid and dead-candidate neighbours (unless there
 * is at most one of them), others the body-intercept interpolation.  With
 * mix_GC_form = 1 the ghost value blends in the advection-diffusion row of
 * the cell (ghostNScoeff: the coefficients of itsolv_ad's z-sweep,
 * CORE_SOLVE_AD.F90:715-724, and nlu/nlv/nlw, CORE_SET_SOLVERS.F90:46-58).
 * Returns 0, or 1 when itermax_gc sweeps did not converge (the reference
 * stops there); *sweeps and *resmax report the sweeps and the last change. */
int or_ghostcell_velocity(or_state *s, int iOpt, double *u, double *v, double *w, const double *nlu,
                          const double *nlv, const double *nlw, const double *fu, const double *fv,
                          const double *fw, double restol_gcv, int *sweeps, double *resmax)
{
    if (!s->have_gcm_vel) return -1;
    double *U[3] = {u, v, w};
    const double *NL[3] = {nlu, nlv, nlw};
    const double pl = s->probe_norm_d;
    const double biw = pl / (pl - 1.0), ipw = 1.0 / (pl - 1.0);
    int iterGC = 0;
    double resVelMax = 1.0e10;
    while (iterGC < s->prm.itermax_gc && resVelMax > restol_gcv) {
        resVelMax = 0.0;
        for (int n = 0; n < s->nD; ++n) {
            int i = s->iD[n], j = s->jD[n], k = s->kD[n];
            if (j < 1 || j > s->ny - 1 || k < 1 || k > s->nz - 1) continue;
            double xIP = s->ipD[3 * (size_t)n], yIP = s->ipD[3 * (size_t)n + 1], zIP = s->ipD[3 * (size_t)n + 2];
            double prev[3], ip[3] = {0.0, 0.0, 0.0}, inv[3] = {0.0, 0.0, 0.0}, dist_int = 0.0;
            long q0 = FID(s, i, j, k);
            for (int c = 0; c < 3; ++c) prev[c] = U[c][q0];
            int sumk = 0;
            for (int r = 0; r < 8; ++r) {
                int ii = s->iDC[n] + incI[r], jj = s->jDC[n] + incJ[r], kk = s->kDC[n] + incK[r];
                if (ii == i && jj == j && kk == k) continue;
                double dx = s->xc[ii] - xIP, dy = s->yc[jj] - yIP, dz = s->zc[kk] - zIP;
                double dist = dx * dx + dy * dy + dz * dz;
                long q = FID(s, ii, jj, kk);
                if (s->iblank[q] == 0 || s->dead[q] == -1) {
                    sumk++;
                    dist_int = dist_int + (1.0 / dist);
                    for (int c = 0; c < 3; ++c) inv[c] = inv[c] + (1.0 / dist) * U[c][q];
                }
            }
            int idw = 0;
            for (int r = 0; r < 8 && !idw; ++r) {
                int ii = s->iDC[n] + incI[r], jj = s->jDC[n] + incJ[r], kk = s->kDC[n] + incK[r];
                long q = FID(s, ii, jj, kk);
                if (s->iblank[q] == 1 && s->dead[q] == 0) {
                    if (dist_int < 0.000001 || sumk == 1) {
                        for (int c = 0; c < 3; ++c) ip[c] = 0.0;
                    } else {
                        for (int c = 0; c < 3; ++c) ip[c] = inv[c] / dist_int;
                        idw = 1;
                    }
                }
            }
            if (!idw)
                for (int r = 0; r < 8; ++r) {
                    int ii = s->iDC[n] + incI[r], jj = s->jDC[n] + incJ[r], kk = s->kDC[n] + incK[r];
                    double cf = s->cDD[8 * (size_t)n + r];
                    if (ii != i || jj != j || kk != k)
                        for (int c = 0; c < 3; ++c) ip[c] = ip[c] + cf * U[c][FID(s, ii, jj, kk)];
                    else
                        for (int c = 0; c < 3; ++c) ip[c] = ip[c] + cf * s->ubiD[3 * (size_t)n + c];
                }
            for (int c = 0; c < 3; ++c) U[c][q0] = s->ubiD[3 * (size_t)n + c] * biw - ip[c] * ipw;
            double rv = fabs(U[0][q0] - prev[0]) + fabs(U[1][q0] - prev[1]) + fabs(U[2][q0] - prev[2]);
            if (rv > resVelMax) resVelMax = rv;
        }
        if (iOpt != 1)
            for (int n = 0; n < s->nG; ++n) {
                int i = s->iG[n], j = s->jG[n], k = s->kG[n];
                long q0 = FID(s, i, j, k);
                double prev[3], g[3] = {0.0, 0.0, 0.0}, tmp[3];
                for (int c = 0; c < 3; ++c) prev[c] = U[c][q0];
                for (int r = 0; r < 8; ++r) {
                    int ii = s->iC[n] + incI[r], jj = s->jC[n] + incJ[r], kk = s->kC[n] + incK[r];
                    double cf = s->cD[8 * (size_t)n + r];
                    if (ii != i || jj != j || kk != k)
                        for (int c = 0; c < 3; ++c) g[c] = g[c] + cf * U[c][FID(s, ii, jj, kk)];
                    else
                        for (int c = 0; c < 3; ++c) g[c] = g[c] + cf * s->ubi[3 * (size_t)n + c];
                }
                double flag = s->alpha[n];
                if (s->prm.mix_GC_form == 1) {
                    double cf[9];
                    ad_coef(s, fu, fv, fw, i, j, k, cf);
                    double bcx = 1.0 + (cf[2] + cf[5] + cf[8]);
                    long sy = s->P, sz = s->S;
                    for (int c = 0; c < 3; ++c)
                        tmp[c] = (NL[c][q0] - cf[0] * U[c][q0 - 1] - cf[1] * U[c][q0 + 1] - cf[3] * U[c][q0 - sy]
                                  - cf[4] * U[c][q0 + sy] - cf[6] * U[c][q0 - sz] - cf[7] * U[c][q0 + sz]) / bcx;
                } else {
                    for (int c = 0; c < 3; ++c) tmp[c] = g[c];
                }
                for (int c = 0; c < 3; ++c) U[c][q0] = (1.0 - flag) * g[c] + flag * tmp[c];
                double rv = fabs(U[0][q0] - prev[0]) + fabs(U[1][q0] - prev[1]) + fabs(U[2][q0] - prev[2]);
                if (rv > resVelMax) resVelMax = rv;
            }
        iterGC++;
    }
    if (sweeps) *sweeps = iterGC;
    if (resmax) *resmax = resVelMax;
    return (resVelMax > restol_gcv) ? 1 : 0;
}

/* solve_ad: CORE_SOLVE_AD.F90:11-127, one rank.  Each pass: itsolv_ad on u,
 * v, w with nlu, nlv, nlw; GCM_ghostcell_velocity(2); face_vel;
 * enforce_global_mass_consv; set_outer_velocity_bc; set_outer_ghost_velocity;
 * calc_residual_ad of the three components; until the largest residual is
 * at most restol_ad or itermax_ad passes ran.  Returns 0 when converged, 1
 * at the cap, negative for errors; iters, res and loc report the passes,
 * the final largest residual and its location. */
int or_solve_ad(or_state *s, double *u, double *v, double *w, const double *nlu, const double *nlv,
                const double *nlw, double *fu, double *fv, double *fw, double time, int itermax_ad,
                double restol_ad, double restol_gcv, int *iters, double *res, int *loc)
{
    int iter = 0, rc;
    double res2 = 1.0e10;
    int lc[3] = {0, 0, 0};
    while (iter < itermax_ad && res2 > restol_ad) {
        or_itsolv_ad(s, u, nlu, fu, fv, fw);
        or_itsolv_ad(s, v, nlv, fu, fv, fw);
        or_itsolv_ad(s, w, nlw, fu, fv, fw);
        if ((rc = or_ghostcell_velocity(s, 2, u, v, w, nlu, nlv, nlw, fu, fv, fw, restol_gcv, NULL, NULL)) != 0)
            return rc > 0 ? -5 : rc;
        or_face_vel(s, u, v, w, fu, fv, fw);
        mass_consv(s, fu, fv, fw);
        if ((rc = or_set_outer_velocity_bc(s, u, v, w, time)) != 0) return rc;
        or_set_outer_ghost_velocity(s, u, v, w);
        int lu[3], lv[3], lw[3];
        double ru = fabs(or_calc_residual_ad(s, u, nlu, fu, fv, fw, lu, NULL));
        double rv = fabs(or_calc_residual_ad(s, v, nlv, fu, fv, fw, lv, NULL));
        double rw = fabs(or_calc_residual_ad(s, w, nlw, fu, fv, fw, lw, NULL));
        double r = ru;
        memcpy(lc, lu, sizeof(lc));
        if (rv > r) { r = rv; memcpy(lc, lv, sizeof(lc)); }
        if (rw > r) { r = rw; memcpy(lc, lw, sizeof(lc)); }
        res2 = r;
        iter++;
    }
    if (iters) *iters = iter;
    if (res) *res = res2;
    if (loc) memcpy(loc, lc, sizeof(lc));
    return (res2 > restol_ad) ? 1 : 0;
}
