/*
 * fc_geom_oracle.c -- CPU FP64 restatement of the immersed-boundary
 * geometry refresh (TEST INFRASTRUCTURE ONLY, see fc_geom_oracle.h).
 * Each routine follows the reference statement by statement, in its loop
 * orders and with its operation order, so that index results are exact and
 * floating-point results round as the reference's do.
 */
#include "fc_geom_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define UNDECIDED 100                     /* BOUNDARY_IBLANK_FAST.F90:14 */

typedef struct {
    int npt, nel;
    double *x, *y, *z;
    int *tri;                             /* 3*nel, 1-based marker numbers */
    double *enx, *eny, *enz, *area, *ecx, *ecy, *ecz;
    double *vnx, *vny, *vnz;
    double *mu, *mv, *mw, *muo, *mvo, *mwo;            /* marker velocities */
    double *fx, *fy, *fz, *sx, *sy, *sz, *ma, *tx, *ty, *tz;   /* marker forces */
} og_body;

struct og_state {
    int nx, ny, nz;
    long n;
    fc_geom_params p;
    double *x, *y, *z, *xc, *yc, *zc, *dx, *dy, *dz, *dxc, *dyc, *dzc;
    og_body b[FC_GEOM_MAX_BODIES];
    int8_t *iblank, *ghost, *dead, *bodynum, *undec, *tmp, *ps1, *ps2;
    int *gidx;
    double *dcip;                         /* 3 * n */
    int nG, nD;
    int *gi, *gj, *gk, *gci, *gcj, *gck, *gel;
    double *gbx, *gby, *gbz, *gnx, *gny, *gnz, *gpl, *gipx, *gipy, *gipz, *gcd, *gcn, *galpha;
    int *di, *dj, *dk, *dci, *dcj, *dck, *del;
    double *dbx, *dby, *dbz, *dnx, *dny, *dnz, *dpl, *dcd, *dcn;
    double *bgu, *bgv, *bgw, *bgd, *bdu, *bdv, *bdw, *bdd;   /* body-intercept values */
    int unselected;
};

#define ID(s, i, j, k) ((long)(i) + (long)((s)->nx + 2) * ((long)(j) + (long)((s)->ny + 2) * (long)(k)))

/* CORE_GRID.F90:106-150 */
static void axis(int n, const double *x, double *xc, double *dx, double *dxc)
{
    for (int i = 0; i <= n; ++i) {
        xc[i] = 0.5 * (x[i + 1] + x[i]);
        dx[i] = x[i + 1] - x[i];
    }
    xc[n + 1] = x[n + 1];
    dx[n + 1] = 0.0;
    dxc[0] = 0.0;
    for (int i = 1; i <= n; ++i) dxc[i] = xc[i] - xc[i - 1];
    dxc[n + 1] = 0.0;
}

og_state *og_create(const fc_grid *g, const fc_geom_params *p)
{
    if (!g || !p || p->nBody < 1 || p->nBody > FC_GEOM_MAX_BODIES) return NULL;
    og_state *s = calloc(1, sizeof(*s));
    s->nx = g->nx; s->ny = g->ny; s->nz = g->nz;
    s->p = *p;
    s->n = (long)(g->nx + 2) * (g->ny + 2) * (g->nz + 2);
    int m = (g->nx > g->ny ? g->nx : g->ny);
    m = (m > g->nz ? m : g->nz) + 2;
    double **ax[12] = {&s->x, &s->y, &s->z, &s->xc, &s->yc, &s->zc, &s->dx, &s->dy, &s->dz, &s->dxc, &s->dyc, &s->dzc};
    for (int a = 0; a < 12; ++a) *ax[a] = calloc(m, sizeof(double));
    memcpy(s->x, g->x, (g->nx + 2) * sizeof(double));
    memcpy(s->y, g->y, (g->ny + 2) * sizeof(double));
    memcpy(s->z, g->z, (g->nz + 2) * sizeof(double));
    axis(g->nx, s->x, s->xc, s->dx, s->dxc);
    axis(g->ny, s->y, s->yc, s->dy, s->dyc);
    axis(g->nz, s->z, s->zc, s->dz, s->dzc);
    int8_t **mk[8] = {&s->iblank, &s->ghost, &s->dead, &s->bodynum, &s->undec, &s->tmp, &s->ps1, &s->ps2};
    for (int a = 0; a < 8; ++a) *mk[a] = calloc(s->n, 1);
    s->gidx = calloc(s->n, sizeof(int));
    s->dcip = calloc(3 * s->n, sizeof(double));
    return s;
}

static void free_lists(og_state *s)
{
    int **il[14] = {&s->gi, &s->gj, &s->gk, &s->gci, &s->gcj, &s->gck, &s->gel,
                    &s->di, &s->dj, &s->dk, &s->dci, &s->dcj, &s->dck, &s->del};
    for (int a = 0; a < 14; ++a) { free(*il[a]); *il[a] = NULL; }
    double **dl[30] = {&s->gbx, &s->gby, &s->gbz, &s->gnx, &s->gny, &s->gnz, &s->gpl, &s->gipx, &s->gipy, &s->gipz,
                       &s->gcd, &s->gcn, &s->galpha, &s->dbx, &s->dby, &s->dbz, &s->dnx, &s->dny, &s->dnz, &s->dpl,
                       &s->dcd, &s->dcn, &s->bgu, &s->bgv, &s->bgw, &s->bgd, &s->bdu, &s->bdv, &s->bdw, &s->bdd};
    for (int a = 0; a < 30; ++a) { free(*dl[a]); *dl[a] = NULL; }
}

static void free_body(og_body *b)
{
    double **d[29] = {&b->x, &b->y, &b->z, &b->enx, &b->eny, &b->enz, &b->area, &b->ecx, &b->ecy, &b->ecz,
                      &b->vnx, &b->vny, &b->vnz, &b->mu, &b->mv, &b->mw, &b->muo, &b->mvo, &b->mwo,
                      &b->fx, &b->fy, &b->fz, &b->sx, &b->sy, &b->sz, &b->ma, &b->tx, &b->ty, &b->tz};
    for (int a = 0; a < 29; ++a) { free(*d[a]); *d[a] = NULL; }
    free(b->tri);
    b->tri = NULL;
}

void og_destroy(og_state *s)
{
    if (!s) return;
    free_lists(s);
    for (int b = 0; b < FC_GEOM_MAX_BODIES; ++b) free_body(&s->b[b]);
    double **ax[12] = {&s->x, &s->y, &s->z, &s->xc, &s->yc, &s->zc, &s->dx, &s->dy, &s->dz, &s->dxc, &s->dyc, &s->dzc};
    for (int a = 0; a < 12; ++a) free(*ax[a]);
    int8_t **mk[8] = {&s->iblank, &s->ghost, &s->dead, &s->bodynum, &s->undec, &s->tmp, &s->ps1, &s->ps2};
    for (int a = 0; a < 8; ++a) free(*mk[a]);
    free(s->gidx);
    free(s->dcip);
    free(s);
}

int og_set_body(og_state *s, int ibody, int npts, int nelem, const double *x, const double *y, const double *z,
                const int *tri)
{
    if (!s || ibody < 1 || ibody > s->p.nBody || npts < 3 || nelem < 1) return -1;
    for (long e = 0; e < 3L * nelem; ++e)
        if (tri[e] < 1 || tri[e] > npts) return -1;
    og_body *b = &s->b[ibody - 1];
    free_body(b);
    b->npt = npts;
    b->nel = nelem;
    double **d[3] = {&b->x, &b->y, &b->z};
    for (int a = 0; a < 3; ++a) *d[a] = malloc(npts * sizeof(double));
    double **e[7] = {&b->enx, &b->eny, &b->enz, &b->area, &b->ecx, &b->ecy, &b->ecz};
    for (int a = 0; a < 7; ++a) *e[a] = calloc(nelem, sizeof(double));
    double **v[19] = {&b->vnx, &b->vny, &b->vnz, &b->mu, &b->mv, &b->mw, &b->muo, &b->mvo, &b->mwo,
                      &b->fx, &b->fy, &b->fz, &b->sx, &b->sy, &b->sz, &b->ma, &b->tx, &b->ty, &b->tz};
    for (int a = 0; a < 19; ++a) *v[a] = calloc(npts, sizeof(double));
    b->tri = malloc(3 * (size_t)nelem * sizeof(int));
    memcpy(b->tri, tri, 3 * (size_t)nelem * sizeof(int));
    return og_move_body(s, ibody, x, y, z);
}

int og_move_body(og_state *s, int ibody, const double *x, const double *y, const double *z)
{
    if (!s || ibody < 1 || ibody > s->p.nBody || !s->b[ibody - 1].x) return -1;
    og_body *b = &s->b[ibody - 1];
    memcpy(b->x, x, b->npt * sizeof(double));
    memcpy(b->y, y, b->npt * sizeof(double));
    memcpy(b->z, z, b->npt * sizeof(double));
    return 0;
}

static inline double sq(double a) { return a * a; }

/* calculate_arclength_norm_ds, BOUNDARY_SET_ARCLENGTH_NORM.F90:38-186 (the
 * ntime > 0 part: element normals, areas, centroids, vertex normals) */
static void surface_metrics(og_body *b)
{
    for (int m = 0; m < b->npt; ++m) { b->vnx[m] = 0.0; b->vny[m] = 0.0; b->vnz[m] = 0.0; }
    for (int m = 0; m < b->nel; ++m) {
        const int n1 = b->tri[3 * m] - 1, n2 = b->tri[3 * m + 1] - 1, n3 = b->tri[3 * m + 2] - 1;
        const double v1x = b->x[n2] - b->x[n1], v1y = b->y[n2] - b->y[n1], v1z = b->z[n2] - b->z[n1];
        const double v2x = b->x[n3] - b->x[n2], v2y = b->y[n3] - b->y[n2], v2z = b->z[n3] - b->z[n2];
        double nx = v1y * v2z - v1z * v2y;
        double ny = v1z * v2x - v1x * v2z;
        double nz = v1x * v2y - v1y * v2x;
        const double mag = sqrt(sq(nx) + sq(ny) + sq(nz));
        b->area[m] = 0.5 * mag;
        nx = nx / mag; ny = ny / mag; nz = nz / mag;
        b->enx[m] = nx; b->eny[m] = ny; b->enz[m] = nz;
        b->vnx[n1] += nx; b->vnx[n2] += nx; b->vnx[n3] += nx;
        b->vny[n1] += ny; b->vny[n2] += ny; b->vny[n3] += ny;
        b->vnz[n1] += nz; b->vnz[n2] += nz; b->vnz[n3] += nz;
        b->ecx[m] = (b->x[n1] + b->x[n2] + b->x[n3]) / 3.0;
        b->ecy[m] = (b->y[n1] + b->y[n2] + b->y[n3]) / 3.0;
        b->ecz[m] = (b->z[n1] + b->z[n2] + b->z[n3]) / 3.0;
    }
    for (int m = 0; m < b->npt; ++m) {
        const double mag = sqrt(sq(b->vnx[m]) + sq(b->vny[m]) + sq(b->vnz[m]));
        b->vnx[m] = b->vnx[m] / mag;
        b->vny[m] = b->vny[m] / mag;
        b->vnz[m] = b->vnz[m] / mag;
    }
}

/* check_BIInsideTriangle, BOUNDARY_IBLANK_FAST.F90:1220-1313 */
static void inside_triangle(const double *xv, const double *yv, const double *zv, double xb, double yb, double zb,
                            double *area123, double *areaDiff)
{
    const double s12 = sqrt(sq(xv[1] - xv[0]) + sq(yv[1] - yv[0]) + sq(zv[1] - zv[0]));
    const double s23 = sqrt(sq(xv[2] - xv[1]) + sq(yv[2] - yv[1]) + sq(zv[2] - zv[1]));
    const double s31 = sqrt(sq(xv[0] - xv[2]) + sq(yv[0] - yv[2]) + sq(zv[0] - zv[2]));
    const double s14 = sqrt(sq(xb - xv[0]) + sq(yb - yv[0]) + sq(zb - zv[0]));
    const double s24 = sqrt(sq(xb - xv[1]) + sq(yb - yv[1]) + sq(zb - zv[1]));
    const double s34 = sqrt(sq(xb - xv[2]) + sq(yb - yv[2]) + sq(zb - zv[2]));
    const double p123 = 0.5 * (s12 + s23 + s31);
    const double p124 = 0.5 * (s12 + s24 + s14);
    const double p234 = 0.5 * (s23 + s24 + s34);
    const double p314 = 0.5 * (s31 + s34 + s14);
    const double a123 = sqrt(p123 * fabs(p123 - s12) * fabs(p123 - s23) * fabs(p123 - s31));
    const double a124 = sqrt(p124 * fabs(p124 - s12) * fabs(p124 - s24) * fabs(p124 - s14));
    const double a234 = sqrt(p234 * fabs(p234 - s23) * fabs(p234 - s24) * fabs(p234 - s34));
    const double a314 = sqrt(p314 * fabs(p314 - s31) * fabs(p314 - s34) * fabs(p314 - s14));
    *area123 = a123;
    *areaDiff = a124 + a234 + a314 - a123;
}

/* calc_crossProduct, BOUNDARY_IBLANK_FAST.F90:1539-1555 */
static void cross(const double *r, const double *s, double *c)
{
    c[0] = r[1] * s[2] - s[1] * r[2];
    c[1] = r[2] * s[0] - s[2] * r[0];
    c[2] = r[0] * s[1] - s[0] * r[1];
}

/* calc_BIOutsideTriangle, BOUNDARY_IBLANK_FAST.F90:1316-1536.  Returns 0 when
 * no side is selected (the reference then reads an unset isideSelect); side 1
 * is used in that case. */
static int outside_triangle(const double *xv, const double *yv, const double *zv, double xc, double yc, double zc,
                            const double *xn, const double *yn, const double *zn, double *xb, double *yb, double *zb,
                            double *dist, double *bnx, double *bny, double *bnz)
{
    const double v4[3] = {*xb, *yb, *zb};
    int sel = 0;
    for (int is = 1; is <= 3; ++is) {
        const int ms = (is == 3) ? 1 : is + 1;
        const int ks = (is == 1) ? 3 : (is == 2) ? 1 : 2;
        const double v1[3] = {xv[is - 1], yv[is - 1], zv[is - 1]};
        const double v2[3] = {xv[ms - 1], yv[ms - 1], zv[ms - 1]};
        const double v3[3] = {xv[ks - 1], yv[ks - 1], zv[ks - 1]};
        const double a[3] = {v3[0] - v1[0], v3[1] - v1[1], v3[2] - v1[2]};
        const double bb[3] = {v2[0] - v1[0], v2[1] - v1[1], v2[2] - v1[2]};
        const double c[3] = {v4[0] - v1[0], v4[1] - v1[1], v4[2] - v1[2]};
        double axb[3], cxb[3];
        cross(a, bb, axb);
        cross(c, bb, cxb);
        const double dot = cxb[0] * axb[0] + cxb[1] * axb[1] + cxb[2] * axb[2];
        if (dot <= 0.0) { sel = is; break; }
    }
    const int found = sel != 0;
    if (!found) sel = 1;
    const int n1 = sel - 1, n2 = (sel == 3) ? 0 : sel;
    double ex = xv[n2] - xv[n1], ey = yv[n2] - yv[n1], ez = zv[n2] - zv[n1];
    const double mag = sqrt(sq(ex) + sq(ey) + sq(ez));
    ex = ex / mag; ey = ey / mag; ez = ez / mag;
    const double pl = (*xb - xv[n1]) * ex + (*yb - yv[n1]) * ey + (*zb - zv[n1]) * ez;
    const double px = xv[n1] + pl * ex, py = yv[n1] + pl * ey, pz = zv[n1] + pl * ez;
    const double d1 = sqrt(sq(xv[n1] - px) + sq(yv[n1] - py) + sq(zv[n1] - pz)) / mag;
    const double d2 = sqrt(sq(xv[n2] - px) + sq(yv[n2] - py) + sq(zv[n2] - pz)) / mag;
    if (d1 <= 1.0 && d2 <= 1.0) {
        *xb = px; *yb = py; *zb = pz;
        *bnx = 0.5 * (xn[n1] + xn[n2]);
        *bny = 0.5 * (yn[n1] + yn[n2]);
        *bnz = 0.5 * (zn[n1] + zn[n2]);
    } else if (d2 > 1.0) {
        *xb = xv[n1]; *yb = yv[n1]; *zb = zv[n1];
        *bnx = xn[n1]; *bny = yn[n1]; *bnz = zn[n1];
    } else {
        *xb = xv[n2]; *yb = yv[n2]; *zb = zv[n2];
        *bnx = xn[n2]; *bny = yn[n2]; *bnz = zn[n2];
    }
    *dist = sqrt(sq(xc - *xb) + sq(yc - *yb) + sq(zc - *zb));
    return found;
}

/* calc_bodyIntercept_Unstruc, BOUNDARY_IBLANK_FAST.F90:857-1216, with
 * nCheck = 1: the closest marker (first of equal distances, MINLOC), then
 * every element that has it as a vertex in element order; the last element
 * of equal distance wins. */
static void intercept(og_state *s, const og_body *b, double xC, double yC, double zC, double *bi, int *elem,
                      double *dist, double *dotNorm)
{
    int cv = 0;
    double dmin = 0.0;
    for (int m = 0; m < b->npt; ++m) {
        const double d = sq(b->x[m] - xC) + sq(b->y[m] - yC) + sq(b->z[m] - zC);
        if (m == 0 || d < dmin) { dmin = d; cv = m; }
    }
    const int cv1 = cv + 1;
    double dbest = 1.0e16, xB = 0.0, yB = 0.0, zB = 0.0, nB[3] = {0.0, 0.0, 0.0};
    int ce = 0;
    for (int m = 0; m < b->nel; ++m) {
        const int *t = b->tri + 3 * m;
        if (t[0] != cv1 && t[1] != cv1 && t[2] != cv1) continue;
        double xv[3], yv[3], zv[3], xn[3], yn[3], zn[3];
        for (int q = 0; q < 3; ++q) {
            const int v = t[q] - 1;
            xv[q] = b->x[v]; yv[q] = b->y[v]; zv[q] = b->z[v];
            xn[q] = b->vnx[v]; yn[q] = b->vny[v]; zn[q] = b->vnz[v];
        }
        const double ex = b->enx[m], ey = b->eny[m], ez = b->enz[m];
        const double pc = -ex * xv[0] - ey * yv[0] - ez * zv[0];
        const double dp = -(ex * xC + ey * yC + ez * zC + pc);
        double xt = xC + ex * dp, yt = yC + ey * dp, zt = zC + ez * dp;
        double a123, adiff;
        inside_triangle(xv, yv, zv, xt, yt, zt, &a123, &adiff);
        double de = sqrt(sq(xC - xt) + sq(yC - yt) + sq(zC - zt));
        double tx, ty, tz;
        if (fabs(adiff) < 1.0e-6 * a123) {
            tx = ex; ty = ey; tz = ez;
        } else {
            if (!outside_triangle(xv, yv, zv, xC, yC, zC, xn, yn, zn, &xt, &yt, &zt, &de, &tx, &ty, &tz))
                s->unselected++;
        }
        if (de <= dbest) {
            dbest = de;
            ce = m + 1;
            xB = xt; yB = yt; zB = zt;
            nB[0] = tx; nB[1] = ty; nB[2] = tz;
        }
    }
    bi[0] = xB; bi[1] = yB; bi[2] = zB;
    *elem = ce;
    *dist = sqrt(sq(xB - xC) + sq(yB - yC) + sq(zB - zC));
    *dotNorm = (xC - xB) * nB[0] + (yC - yB) * nB[1] + (zC - zB) * nB[2];
}

int og_body_intercept(const og_state *s, int ibody, int i, int j, int k, double *bi, int *elem, double *dist,
                      double *dotNorm)
{
    if (!s || ibody < 1 || ibody > s->p.nBody) return -1;
    intercept((og_state *)s, &s->b[ibody - 1], s->xc[i], s->yc[j], s->zc[k], bi, elem, dist, dotNorm);
    return 0;
}

/* the cell range of one bound along an axis: the last i in [lo, hi-1] with
 * (xc(i)-v)*(xc(i+1)-v) <= 0 (BOUNDARY_IBLANK_FAST.F90:228-290) */
static void bracket(const double *c, int lo, int hi, double vmin, double vmax, int init, int *cmin, int *cmax)
{
    int a = init, b = init;
    for (int i = lo; i <= hi - 1; ++i) {
        if ((c[i] - vmin) * (c[i + 1] - vmin) <= 0.0) a = i;
        if ((c[i] - vmax) * (c[i + 1] - vmax) <= 0.0) b = i + 1;
    }
    *cmin = a;
    *cmax = b;
}

/* set_iblank_body_fast, BOUNDARY_IBLANK_FAST.F90:2-854, one rank, 3-D */
static void set_iblank(og_state *s, fc_geom_stats *st)
{
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    const double tk = s->p.membrane_tkns;
    memset(s->bodynum, 0, s->n);
    memset(s->tmp, 0, s->n);
    memset(s->undec, 0, s->n);
    for (int k = 0; k <= nz + 1; ++k)
        for (int j = 0; j <= ny + 1; ++j)
            for (int i = 0; i <= nx; ++i) s->undec[ID(s, i, j, k)] = UNDECIDED;
    for (int ib = s->p.nBody; ib >= 1; --ib) {
        const og_body *b = &s->b[ib - 1];
        const int mem = s->p.unstruc_surface_type[ib - 1] == FC_SURFACE_MEMBRANE;
        const double th = tk * (mem ? 3.0 : 0.0);
        for (int m = 0; m < b->nel; ++m) {
            double xv[3], yv[3], zv[3];
            for (int q = 0; q < 3; ++q) {
                const int v = b->tri[3 * m + q] - 1;
                xv[q] = b->x[v]; yv[q] = b->y[v]; zv[q] = b->z[v];
            }
            if ((xv[0] <= s->x[1] && xv[1] <= s->x[1] && xv[2] <= s->x[1]) ||
                (xv[0] >= s->x[nx] && xv[1] >= s->x[nx] && xv[2] >= s->x[nx]) ||
                (yv[0] + th <= s->y[1] && yv[1] + th <= s->y[1] && yv[2] + th <= s->y[1]) ||
                (yv[0] - th >= s->y[ny] && yv[1] - th >= s->y[ny] && yv[2] - th >= s->y[ny]) ||
                (zv[0] + th <= s->z[1] && zv[1] + th <= s->z[1] && zv[2] + th <= s->z[1]) ||
                (zv[0] - th >= s->z[nz] && zv[1] - th >= s->z[nz] && zv[2] - th >= s->z[nz]))
                continue;
            double x0 = fmin(fmin(xv[0], xv[1]), xv[2]), x1 = fmax(fmax(xv[0], xv[1]), xv[2]);
            double y0 = fmin(fmin(yv[0], yv[1]), yv[2]), y1 = fmax(fmax(yv[0], yv[1]), yv[2]);
            double z0 = fmin(fmin(zv[0], zv[1]), zv[2]), z1 = fmax(fmax(zv[0], zv[1]), zv[2]);
            if (mem) { x0 -= th; x1 += th; y0 -= th; y1 += th; z0 -= th; z1 += th; }
            int i0, i1, j0, j1, k0, k1;
            bracket(s->xc, 0, nx, x0, x1, 0, &i0, &i1);
            if (i0 == 0) i0 = 1;
            if (i1 == 0 || i1 == nx) i1 = nx - 1;
            bracket(s->yc, 0, ny, y0, y1, 0, &j0, &j1);
            if (j0 == 0) j0 = 1;
            if (j1 == 0 || j1 == ny) j1 = ny - 1;
            bracket(s->zc, 0, nz, z0, z1, 0, &k0, &k1);
            if (k0 == 0) k0 = 1;
            if (k1 == 0 || k1 == nz) k1 = nz - 1;
            for (int k = k0; k <= k1; ++k)
                for (int j = j0; j <= j1; ++j)
                    for (int i = i0; i <= i1; ++i) s->undec[ID(s, i, j, k)] = -UNDECIDED;
        }
    }
    int nund = 0;
    for (int k = 1; k <= nz - 1; ++k)
        for (int j = 1; j <= ny - 1; ++j)
            for (int i = 1; i <= nx - 1; ++i) {
                const long q = ID(s, i, j, k);
                if (s->undec[q] != -UNDECIDED) continue;
                ++nund;
                int hit = 0;
                for (int ib = s->p.nBody; ib >= 1 && !hit; --ib) {
                    double bi[3], dist, dn;
                    int el;
                    intercept(s, &s->b[ib - 1], s->xc[i], s->yc[j], s->zc[k], bi, &el, &dist, &dn);
                    if (s->p.unstruc_surface_type[ib - 1] == FC_SURFACE_MEMBRANE) {
                        const double d = sqrt(sq(s->xc[i] - bi[0]) + sq(s->yc[j] - bi[1]) + sq(s->zc[k] - bi[2]));
                        if (d < tk / 2.0) hit = ib;
                    } else if (dn >= 0.0) {
                        hit = ib;
                    }
                }
                s->tmp[q] = hit ? 1 : 0;
                if (hit) s->bodynum[q] = (int8_t)hit;
            }
    if (st) st->n_undecided = nund;
    for (int k = 1; k <= nz - 1; ++k)
        for (int j = 1; j <= ny - 1; ++j) {
            for (int ii = 1; ii <= nx - 1; ++ii) {
                if (s->undec[ID(s, ii, j, k)] == UNDECIDED) continue;
                s->tmp[ID(s, 0, j, k)] = s->tmp[ID(s, ii, j, k)];
                s->bodynum[ID(s, 0, j, k)] = s->bodynum[ID(s, ii, j, k)];
                break;
            }
            for (int i = 1; i <= nx - 1; ++i) {
                const long q = ID(s, i, j, k);
                if (s->undec[q] != UNDECIDED) continue;
                s->tmp[q] = s->tmp[q - 1];
                if (s->tmp[q] == 1) s->bodynum[q] = s->bodynum[q - 1];
            }
            s->tmp[ID(s, nx, j, k)] = s->tmp[ID(s, nx - 1, j, k)];
            s->bodynum[ID(s, nx, j, k)] = s->bodynum[ID(s, nx - 1, j, k)];
        }
    for (int k = 1; k <= nz - 1; ++k)
        for (int j = 1; j <= ny - 1; ++j)
            for (int i = 0; i <= nx; ++i) s->iblank[ID(s, i, j, k)] = s->tmp[ID(s, i, j, k)];
}

/* One in-place smoothing pass over the interior in the order of its loops
 * (outer, middle, inner axes), counting the iblank face neighbours listed in
 * nb[] (offsets in that order); thr of them make the cell solid with the
 * bodyNum of the last such neighbour with bodyNum > 0.  smooth_iblank:
 * GCM_SET_BOUNDARY.F90:1618-1697; the 2-D passes: :120-197. */
static int smooth_pass(og_state *s, const int (*nb)[3], int nn, int thr, int order)
{
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    int changed = 0, ibody = 0;
    int lo[3] = {1, 1, 1}, hi[3] = {nx - 1, ny - 1, nz - 1};
    /* order 0: k, j, i (i inner); 1: i, j, k (k inner); 2: j, i, k (k inner) */
    const int ax[3][3] = {{2, 1, 0}, {0, 1, 2}, {1, 0, 2}};
    const int *a = ax[order];
    int c[3];
    for (c[a[0]] = lo[a[0]]; c[a[0]] <= hi[a[0]]; ++c[a[0]])
        for (c[a[1]] = lo[a[1]]; c[a[1]] <= hi[a[1]]; ++c[a[1]])
            for (c[a[2]] = lo[a[2]]; c[a[2]] <= hi[a[2]]; ++c[a[2]]) {
                const long q = ID(s, c[0], c[1], c[2]);
                if (s->iblank[q] == 1) continue;
                int sum = 0;
                for (int t = 0; t < nn; ++t) {
                    const long r = ID(s, c[0] + nb[t][0], c[1] + nb[t][1], c[2] + nb[t][2]);
                    if (s->iblank[r] == 1) {
                        ++sum;
                        if (s->bodynum[r] > 0) ibody = s->bodynum[r];
                    }
                }
                if (sum >= thr) {
                    s->iblank[q] = 1;
                    s->bodynum[q] = (int8_t)ibody;
                    ++changed;
                }
            }
    return changed;
}

/* GCM_Calc_VanMatrixDN, GCM_SET_BOUNDARY.F90:607-797.  The inverses are
 * formed by LU factorisation with partial pivoting (first largest pivot,
 * row interchanges) and column solves against the identity; LAPACK's
 * DGETRF/DGETRI compute the same inverse with other roundings. */
static void invert8(double A[8][8], double X[8][8])
{
    int piv[8];
    for (int c = 0; c < 8; ++c) {
        int p = c;
        double best = fabs(A[c][c]);
        for (int r = c + 1; r < 8; ++r)
            if (fabs(A[r][c]) > best) { best = fabs(A[r][c]); p = r; }
        piv[c] = p;
        if (p != c)
            for (int q = 0; q < 8; ++q) { double t = A[c][q]; A[c][q] = A[p][q]; A[p][q] = t; }
        const double inv = 1.0 / A[c][c];
        for (int r = c + 1; r < 8; ++r) {
            const double l = A[r][c] * inv;
            A[r][c] = l;
            for (int q = c + 1; q < 8; ++q) A[r][q] = A[r][q] - l * A[c][q];
        }
    }
    for (int col = 0; col < 8; ++col) {
        double e[8];
        for (int r = 0; r < 8; ++r) e[r] = (r == col) ? 1.0 : 0.0;
        for (int c = 0; c < 8; ++c)
            if (piv[c] != c) { double t = e[c]; e[c] = e[piv[c]]; e[piv[c]] = t; }
        for (int r = 1; r < 8; ++r)
            for (int q = 0; q < r; ++q) e[r] = e[r] - A[r][q] * e[q];
        for (int r = 7; r >= 0; --r) {
            for (int q = r + 1; q < 8; ++q) e[r] = e[r] - A[r][q] * e[q];
            e[r] = e[r] / A[r][r];
        }
        for (int r = 0; r < 8; ++r) X[r][col] = e[r];
    }
}

static const int INC_I[8] = {0, 0, 1, 1, 0, 0, 1, 1};
static const int INC_J[8] = {0, 1, 1, 0, 0, 1, 1, 0};
static const int INC_K[8] = {0, 0, 0, 0, 1, 1, 1, 1};

static void van_matrix(const og_state *s, int iG, int jG, int kG, int iC, int jC, int kC, const double *ev,
                       const double *bi, const double *bn, double *cD, double *cN)
{
    double VD[8][8], VN[8][8], ID_[8][8], IN_[8][8];
    for (int r = 0; r < 8; ++r) {
        const int i = iC + INC_I[r], j = jC + INC_J[r], k = kC + INC_K[r];
        double x1 = s->xc[i], x2 = s->yc[j], x3 = s->zc[k];
        VD[r][0] = x1 * x2 * x3; VD[r][1] = x1 * x2; VD[r][2] = x1 * x3; VD[r][3] = x2 * x3;
        VD[r][4] = x1; VD[r][5] = x2; VD[r][6] = x3; VD[r][7] = 1.0;
        for (int q = 0; q < 8; ++q) VN[r][q] = VD[r][q];
        if (i == iG && j == jG && k == kG) {
            x1 = bi[0]; x2 = bi[1]; x3 = bi[2];
            const double n1 = bn[0], n2 = bn[1], n3 = bn[2];
            VD[r][0] = x1 * x2 * x3; VD[r][1] = x1 * x2; VD[r][2] = x1 * x3; VD[r][3] = x2 * x3;
            VD[r][4] = x1; VD[r][5] = x2; VD[r][6] = x3; VD[r][7] = 1.0;
            VN[r][0] = n1 * x2 * x3 + n2 * x1 * x3 + n3 * x1 * x2;
            VN[r][1] = n1 * x2 + n2 * x1;
            VN[r][2] = n1 * x3 + n3 * x1;
            VN[r][3] = n2 * x3 + n3 * x2;
            VN[r][4] = n1; VN[r][5] = n2; VN[r][6] = n3; VN[r][7] = 0.0;
        }
    }
    invert8(VD, ID_);
    invert8(VN, IN_);
    const double xI = ev[0], yI = ev[1], zI = ev[2];
    for (int r = 0; r < 8; ++r) {
        cD[r] = ID_[0][r] * xI * yI * zI + ID_[1][r] * xI * yI + ID_[2][r] * xI * zI + ID_[3][r] * yI * zI +
                ID_[4][r] * xI + ID_[5][r] * yI + ID_[6][r] * zI + ID_[7][r];
        cN[r] = IN_[0][r] * xI * yI * zI + IN_[1][r] * xI * yI + IN_[2][r] * xI * zI + IN_[3][r] * yI * zI +
                IN_[4][r] * xI + IN_[5][r] * yI + IN_[6][r] * zI + IN_[7][r];
    }
}

int og_van_matrix(const og_state *s, int iG, int jG, int kG, int iC, int jC, int kC, const double *eval,
                  const double *bi, const double *bin, double *cD, double *cN)
{
    if (!s) return -1;
    van_matrix(s, iG, jG, kG, iC, jC, kC, eval, bi, bin, cD, cN);
    return 0;
}

/* NINT: nearest integer, halves away from zero */
static int nint(double v) { return (int)lround(v); }

/* host cell of an image point along one axis: the last i in [lo, hi] with
 * c(i) <= v < c(i+1) (strict = 1, ghost cells, GCM_SET_BOUNDARY.F90:
 * 418-463) or c(i) <= v <= c(i+1) (dead cells, :1323-1359); indices below 0
 * are outside the reference's arrays and never match */
static int host(const double *c, int lo, int hi, double v, int strict)
{
    int r = -1;
    for (int i = lo; i <= hi; ++i) {
        if (i < 0) continue;
        if (c[i] <= v && (strict ? c[i + 1] > v : c[i + 1] >= v)) r = i;
    }
    return r;
}

static void ranges(const og_state *s, int iG, int jG, int kG, int *iR, int *jR, int *kR)
{
    const double md = fmax(fmax(s->dx[iG], s->dy[jG]), s->dz[kG]);
    *iR = 2 * nint(md / s->dx[iG] + 0.5);
    *jR = 2 * nint(md / s->dy[jG] + 0.5);
    *kR = 2 * nint(md / s->dz[kG] + 0.5);
}

static int imax(int a, int b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }

int og_refresh(og_state *s, fc_geom_stats *st)
{
    if (!s) return -1;
    for (int ib = 0; ib < s->p.nBody; ++ib)
        if (!s->b[ib].x) return -1;
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    fc_geom_stats local;
    if (!st) st = &local;
    memset(st, 0, sizeof(*st));
    s->unselected = 0;
    for (int ib = s->p.nBody; ib >= 1; --ib) surface_metrics(&s->b[ib - 1]);
    set_iblank(s, st);
    /* smooth_iblank (3-D): 6 face neighbours in its kk, jj, ii loop order */
    static const int nb3[6][3] = {{0, 0, -1}, {0, -1, 0}, {-1, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    static const int nbA[4][3] = {{0, 0, -1}, {0, -1, 0}, {0, 1, 0}, {0, 0, 1}};
    static const int nbB[4][3] = {{0, 0, -1}, {-1, 0, 0}, {1, 0, 0}, {0, 0, 1}};
    int nsm = 0;
    for (int l = 0; l < 5; ++l) nsm += smooth_pass(s, nb3, 6, 5, 0);
    for (int l = 0; l < 5; ++l) nsm += smooth_pass(s, nbA, 4, 3, 1);
    for (int l = 0; l < 5; ++l) nsm += smooth_pass(s, nbB, 4, 3, 2);
    st->n_smoothed = nsm;
    /* GCM_define_ghostcells */
    memset(s->ghost, 0, s->n);
    for (int k = 0; k <= nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 1; i <= nx - 1; ++i) {
                const long q = ID(s, i, j, k);
                if (s->iblank[q] != 1) continue;
                for (int t = 0; t < 6; ++t) {
                    const int i2 = i + nb3[t][0], j2 = j + nb3[t][1], k2 = k + nb3[t][2];
                    if (i2 < 1 || i2 > nx - 1 || j2 < 1 || j2 > ny - 1 || k2 < 1 || k2 > nz - 1) continue;
                    const long r = ID(s, i2, j2, k2);
                    if (s->iblank[r] == 0 && s->ghost[r] == 0) {
                        s->ghost[r] = 1;
                        s->bodynum[r] = s->bodynum[q];
                    }
                }
            }
    free_lists(s);
    int nG = 0;
    for (long q = 0; q < s->n; ++q) nG += s->ghost[q] == 1;
    s->nG = nG;
    int **gl[7] = {&s->gi, &s->gj, &s->gk, &s->gci, &s->gcj, &s->gck, &s->gel};
    for (int a = 0; a < 7; ++a) *gl[a] = calloc(nG + 1, sizeof(int));
    double **gd[10] = {&s->gbx, &s->gby, &s->gbz, &s->gnx, &s->gny, &s->gnz, &s->gpl, &s->gipx, &s->gipy, &s->gipz};
    for (int a = 0; a < 10; ++a) *gd[a] = calloc(nG + 1, sizeof(double));
    s->gcd = calloc(8 * (size_t)nG + 1, sizeof(double));
    s->gcn = calloc(8 * (size_t)nG + 1, sizeof(double));
    s->galpha = calloc(nG + 1, sizeof(double));
    int n = 0;
    for (int k = 1; k <= nz - 1; ++k)
        for (int j = 1; j <= ny - 1; ++j)
            for (int i = 1; i <= nx - 1; ++i)
                if (s->ghost[ID(s, i, j, k)] == 1) {
                    s->gi[n] = i; s->gj[n] = j; s->gk[n] = k;
                    s->gidx[ID(s, i, j, k)] = n + 1;
                    ++n;
                }
    int rc = 0;
    const double tk = s->p.membrane_tkns, pr = s->p.probeLengthNormalized, prD = s->p.probeLengthNormalizedD;
    for (n = 0; n < nG; ++n) {
        const int iG = s->gi[n], jG = s->gj[n], kG = s->gk[n];
        const int ib = s->bodynum[ID(s, iG, jG, kG)];
        const double xG = s->xc[iG], yG = s->yc[jG], zG = s->zc[kG];
        double bi[3], dist, dn;
        int el = 0;
        if (ib < 1 || ib > s->p.nBody) { rc = -5; continue; }
        intercept(s, &s->b[ib - 1], xG, yG, zG, bi, &el, &dist, &dn);
        if (s->p.body_dim[ib - 1] == 2) bi[2] = zG;
        s->gel[n] = el;
        s->gbx[n] = bi[0]; s->gby[n] = bi[1]; s->gbz[n] = bi[2];
        double ds = sqrt(sq(xG - bi[0]) + sq(yG - bi[1]) + sq(zG - bi[2]));
        const double sx = (xG - bi[0]) / ds, sy = (yG - bi[1]) / ds, sz = (zG - bi[2]) / ds;
        if (s->p.unstruc_surface_type[ib - 1] == FC_SURFACE_MEMBRANE) {
            bi[0] = bi[0] + tk / 2.0 * sx;
            bi[1] = bi[1] + tk / 2.0 * sy;
            bi[2] = bi[2] + tk / 2.0 * sz;
            ds = sqrt(sq(xG - bi[0]) + sq(yG - bi[1]) + sq(zG - bi[2]));
        }
        const double bn[3] = {-sx, -sy, -sz};
        s->gnx[n] = bn[0]; s->gny[n] = bn[1]; s->gnz[n] = bn[2];
        s->gpl[n] = ds * pr;
        const double xI = bi[0] - (bi[0] - xG) * pr, yI = bi[1] - (bi[1] - yG) * pr, zI = bi[2] - (bi[2] - zG) * pr;
        s->gipx[n] = xI; s->gipy[n] = yI; s->gipz[n] = zI;
        const long q = ID(s, iG, jG, kG);
        s->dcip[q] = xI; s->dcip[q + s->n] = yI; s->dcip[q + 2 * s->n] = zI;
        int iR, jR, kR;
        ranges(s, iG, jG, kG, &iR, &jR, &kR);
        const int iC = host(s->xc, imax(iG - iR, 0), imin(iG + iR, nx), xI, 1);
        const int jC = host(s->yc, imax(jG - jR, -1), imin(jG + jR, ny), yI, 1);
        const int kC = host(s->zc, imax(kG - kR, -1), imin(kG + kR, nz), zI, 1);
        s->gci[n] = iC; s->gcj[n] = jC; s->gck[n] = kC;
        if (iC < 0 || jC < 0 || kC < 0) { rc = -5; continue; }
        const double ev[3] = {xG, yG, zG};
        van_matrix(s, iG, jG, kG, iC, jC, kC, ev, bi, bn, s->gcd + 8 * n, s->gcn + 8 * n);
    }
    /* GCM_set_deadcells */
    memset(s->dead, 0, s->n);
    memset(s->ps1, 0, s->n);
    memset(s->ps2, 0, s->n);
    for (int k = 1; k <= nz - 1; ++k)
        for (int j = 1; j <= ny - 1; ++j)
            for (int i = 1; i <= nx - 1; ++i) {
                if (s->ghost[ID(s, i, j, k)] != 1) continue;
                static const int o6[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
                for (int t = 0; t < 6; ++t) {
                    const long r = ID(s, i + o6[t][0], j + o6[t][1], k + o6[t][2]);
                    if (s->iblank[r] == 1) s->dead[r] = -1;
                }
            }
    int nD = 0;
    for (int k = 0; k <= nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 1; i <= nx - 1; ++i) nD += abs(s->dead[ID(s, i, j, k)]) == 1;
    s->nD = nD;
    int **dl[7] = {&s->di, &s->dj, &s->dk, &s->dci, &s->dcj, &s->dck, &s->del};
    for (int a = 0; a < 7; ++a) *dl[a] = calloc(nD + 1, sizeof(int));
    double **dd[7] = {&s->dbx, &s->dby, &s->dbz, &s->dnx, &s->dny, &s->dnz, &s->dpl};
    for (int a = 0; a < 7; ++a) *dd[a] = calloc(nD + 1, sizeof(double));
    s->dcd = calloc(8 * (size_t)nD + 1, sizeof(double));
    s->dcn = calloc(8 * (size_t)nD + 1, sizeof(double));
    n = 0;
    for (int k = 0; k <= nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 1; i <= nx - 1; ++i)
                if (abs(s->dead[ID(s, i, j, k)]) == 1) { s->di[n] = i; s->dj[n] = j; s->dk[n] = k; ++n; }
    for (n = 0; n < nD; ++n) {
        const int iG = s->di[n], jG = s->dj[n], kG = s->dk[n];
        const long q = ID(s, iG, jG, kG);
        const int ib = s->bodynum[q];
        if (ib < 1 || ib > s->p.nBody) { rc = -5; continue; }
        const double xD = s->xc[iG], yD = s->yc[jG], zD = s->zc[kG];
        double bi[3], dist, dn;
        int el = 0;
        intercept(s, &s->b[ib - 1], xD, yD, zD, bi, &el, &dist, &dn);
        if (s->p.body_dim[ib - 1] == 2) bi[2] = zD;
        s->del[n] = el;
        s->dbx[n] = bi[0]; s->dby[n] = bi[1]; s->dbz[n] = bi[2];
        double ds = sqrt(sq(xD - bi[0]) + sq(yD - bi[1]) + sq(zD - bi[2]));
        double sx = (bi[0] - xD) / ds, sy = (bi[1] - yD) / ds, sz = (bi[2] - zD) / ds;
        const double bn[3] = {-sx, -sy, -sz};
        s->dnx[n] = bn[0]; s->dny[n] = bn[1]; s->dnz[n] = bn[2];
        if (s->p.unstruc_surface_type[ib - 1] == FC_SURFACE_MEMBRANE) {
            sx = -sx; sy = -sy; sz = -sz;
            if (ds >= tk / 2.0) {
                const double f = ds + (double)0.1f * (ds - tk / 2.0);
                bi[0] = bi[0] + f * sx; bi[1] = bi[1] + f * sy; bi[2] = bi[2] + f * sz;
            } else {
                bi[0] = bi[0] + tk / 2.0 * sx; bi[1] = bi[1] + tk / 2.0 * sy; bi[2] = bi[2] + tk / 2.0 * sz;
            }
            double zI = zD + (bi[2] - zD) * prD;
            /* The y loop (:1183-1193) tests yIP before this cell sets it: the
               value the previous dead cell left, which the clamp below keeps
               inside [yc(0), yc(ny+1)], so the loop runs for the first dead
               cell of the list only, where yIP is unset.  The compiled
               reference runs its body there until yIP is inside the range;
               so does this restatement. */
            if (n == 0) {
                double yI;
                int it = 0;
                do {
                    const double d1 = sqrt(sq(xD - bi[0]) + sq(yD - bi[1]) + sq(zD - bi[2]));
                    const double t = -0.5 * d1;
                    bi[0] = bi[0] + t * sx; bi[1] = bi[1] + t * sy; bi[2] = bi[2] + t * sz;
                    yI = yD + (bi[1] - yD) * prD;
                } while ((yI < s->yc[0] || yI > s->yc[ny + 1]) && ++it < 2200);
            }
            for (int it = 0; (zI < s->zc[0] || zI > s->zc[nz + 1]) && it < 2200; ++it) {
                const double d1 = sqrt(sq(xD - bi[0]) + sq(yD - bi[1]) + sq(zD - bi[2]));
                const double t = -0.5 * d1;
                bi[0] = bi[0] + t * sx; bi[1] = bi[1] + t * sy; bi[2] = bi[2] + t * sz;
                zI = zD + (bi[2] - zD) * prD;
            }
            ds = sqrt(sq(xD - bi[0]) + sq(yD - bi[1]) + sq(zD - bi[2]));
        }
        /* alphaGhost (:1155-1213); the buffer-slice tests never hold at one rank */
        const long nbq[6] = {ID(s, iG + 1, jG, kG), ID(s, iG - 1, jG, kG), ID(s, iG, jG + 1, kG),
                             ID(s, iG, jG - 1, kG), ID(s, iG, jG, kG + 1), ID(s, iG, jG, kG - 1)};
        const double dc[6] = {s->dxc[iG + 1], s->dxc[iG], s->dyc[jG + 1], s->dyc[jG], s->dzc[kG + 1], s->dzc[kG]};
        for (int t = 0; t < 6; ++t) {
            if (abs(s->ghost[nbq[t]]) != 1) continue;
            const double tt = ds / dc[t];
            const int g = s->gidx[nbq[t]];
            s->galpha[g - 1] = s->galpha[g - 1] + tt * tt;
        }
        double xI = xD + (bi[0] - xD) * prD, yI = yD + (bi[1] - yD) * prD, zI = zD + (bi[2] - zD) * prD;
        s->dpl[n] = ds * prD;
        if (yI < s->yc[0] || yI > s->yc[ny + 1]) {
            yI = (fabs(yI - s->yc[0]) < fabs(yI - s->yc[ny + 1])) ? s->yc[0] : s->yc[ny + 1];
            s->dpl[n] = sqrt(sq(xD - xI) + sq(yD - yI) + sq(zD - zI));
        }
        if (zI < s->zc[0] || zI > s->zc[nz + 1]) {
            zI = (fabs(zI - s->zc[0]) < fabs(zI - s->zc[nz + 1])) ? s->zc[0] : s->zc[nz + 1];
            s->dpl[n] = sqrt(sq(xD - xI) + sq(yD - yI) + sq(zD - zI));
        }
        s->dcip[q] = xI; s->dcip[q + s->n] = yI; s->dcip[q + 2 * s->n] = zI;
        int iR, jR, kR;
        ranges(s, iG, jG, kG, &iR, &jR, &kR);
        const int iC = host(s->xc, imax(iG - iR, 0), imin(iG + iR, nx), xI, 0);
        const int jC = host(s->yc, imax(jG - jR, -1), imin(jG + jR, ny), yI, 0);
        const int kC = host(s->zc, imax(kG - kR, -1), imin(kG + kR, nz), zI, 0);
        s->dci[n] = iC; s->dcj[n] = jC; s->dck[n] = kC;
        if (iC < 0 || jC < 0 || kC < 0) { rc = -5; continue; }
        /* ps1 / ps2: dead cells whose image-point stencil holds a solid cell
           that is not dead, and their ghost neighbours (:1362-1437) */
        for (int r = 0; r < 8; ++r) {
            const long c = ID(s, iC + INC_I[r], jC + INC_J[r], kC + INC_K[r]);
            if (s->iblank[c] == 1 && s->dead[c] != -1) s->ps1[q] = 10;
        }
        if (s->ps1[q] == 10) {
            if (iG - 1 >= 0 && iG - 1 <= nx && s->ghost[ID(s, iG - 1, jG, kG)] == 1) s->ps2[ID(s, iG - 1, jG, kG)] = 10;
            if (iG + 1 >= 0 && iG + 1 <= nx && s->ghost[ID(s, iG + 1, jG, kG)] == 1) s->ps2[ID(s, iG + 1, jG, kG)] = 10;
            if (jG - 1 >= 0 && jG - 1 <= ny && s->ghost[ID(s, iG, jG - 1, kG)] == 1) s->ps2[ID(s, iG, jG - 1, kG)] = 10;
            if (jG + 1 >= 0 && jG + 1 <= ny && s->ghost[ID(s, iG, jG + 1, kG)] == 1) s->ps2[ID(s, iG, jG + 1, kG)] = 10;
            if (kG - 1 >= 1 && kG - 1 <= nz - 1 && s->ghost[ID(s, iG, jG, kG - 1)] == 1) s->ps2[ID(s, iG, jG, kG - 1)] = 10;
            if (kG + 1 >= 1 && kG + 1 <= nz - 1 && s->ghost[ID(s, iG, jG, kG + 1)] == 1) s->ps2[ID(s, iG, jG, kG + 1)] = 10;
        }
        const double ev[3] = {xI, yI, zI};
        van_matrix(s, iG, jG, kG, iC, jC, kC, ev, bi, bn, s->dcd + 8 * n, s->dcn + 8 * n);
    }
    for (n = 0; n < nG; ++n) s->galpha[n] = 1.0 - sqrt(s->galpha[n]);
    st->nGhost = nG;
    st->nDead = nD;
    st->n_unselected_side = s->unselected;
    return rc;
}

int og_get_counts(const og_state *s, int *nGhost, int *nDead)
{
    if (!s) return -1;
    if (nGhost) *nGhost = s->nG;
    if (nDead) *nDead = s->nD;
    return 0;
}

int og_get_masks(const og_state *s, int8_t *iblank, int8_t *ghost, int8_t *dead, int8_t *bodyNum, int8_t *ps1,
                 int8_t *ps2)
{
    if (!s) return -1;
    if (iblank) memcpy(iblank, s->iblank, s->n);
    if (ghost) memcpy(ghost, s->ghost, s->n);
    if (dead) memcpy(dead, s->dead, s->n);
    if (bodyNum) memcpy(bodyNum, s->bodynum, s->n);
    if (ps1) memcpy(ps1, s->ps1, s->n);
    if (ps2) memcpy(ps2, s->ps2, s->n);
    return 0;
}

#define CPI(dst, src, cnt) do { if (dst) memcpy((dst), (src), (size_t)(cnt) * sizeof(int)); } while (0)
#define CPD(dst, src, cnt) do { if (dst) memcpy((dst), (src), (size_t)(cnt) * sizeof(double)); } while (0)

int og_get_ghosts(const og_state *s, fc_geom_ghosts *g)
{
    if (!s || !g) return -1;
    const int n = s->nG;
    CPI(g->iGhost, s->gi, n); CPI(g->jGhost, s->gj, n); CPI(g->kGhost, s->gk, n);
    CPI(g->iCellIndex, s->gci, n); CPI(g->jCellIndex, s->gcj, n); CPI(g->kCellIndex, s->gck, n);
    CPI(g->closestElementGC, s->gel, n);
    CPD(g->xBodyIntercept, s->gbx, n); CPD(g->yBodyIntercept, s->gby, n); CPD(g->zBodyIntercept, s->gbz, n);
    CPD(g->xBodyInterceptNorm, s->gnx, n); CPD(g->yBodyInterceptNorm, s->gny, n);
    CPD(g->zBodyInterceptNorm, s->gnz, n); CPD(g->probeLength, s->gpl, n);
    CPD(g->xImagePoint, s->gipx, n); CPD(g->yImagePoint, s->gipy, n); CPD(g->zImagePoint, s->gipz, n);
    CPD(g->coeffGCMD, s->gcd, 8 * n); CPD(g->coeffGCMN, s->gcn, 8 * n); CPD(g->alphaGhost, s->galpha, n);
    return 0;
}

int og_get_dead(const og_state *s, fc_geom_dead *d)
{
    if (!s || !d) return -1;
    const int n = s->nD;
    CPI(d->iDead, s->di, n); CPI(d->jDead, s->dj, n); CPI(d->kDead, s->dk, n);
    CPI(d->iDeadCellIndex, s->dci, n); CPI(d->jDeadCellIndex, s->dcj, n); CPI(d->kDeadCellIndex, s->dck, n);
    CPI(d->closestElementDead, s->del, n);
    CPD(d->xBodyInterceptDead, s->dbx, n); CPD(d->yBodyInterceptDead, s->dby, n);
    CPD(d->zBodyInterceptDead, s->dbz, n); CPD(d->xNormInterceptDead, s->dnx, n);
    CPD(d->yNormInterceptDead, s->dny, n); CPD(d->zNormInterceptDead, s->dnz, n);
    CPD(d->probeLengthDead, s->dpl, n); CPD(d->coeffGCMDeadD, s->dcd, 8 * n); CPD(d->coeffGCMDeadN, s->dcn, 8 * n);
    return 0;
}

int og_get_surface(const og_state *s, int ibody, fc_geom_surface *f)
{
    if (!s || !f || ibody < 1 || ibody > s->p.nBody || !s->b[ibody - 1].x) return -1;
    const og_body *b = &s->b[ibody - 1];
    CPD(f->triElemNormX, b->enx, b->nel); CPD(f->triElemNormY, b->eny, b->nel); CPD(f->triElemNormZ, b->enz, b->nel);
    CPD(f->triElemArea, b->area, b->nel); CPD(f->triElemCentX, b->ecx, b->nel); CPD(f->triElemCentY, b->ecy, b->nel);
    CPD(f->triElemCentZ, b->ecz, b->nel); CPD(f->xNormBodyMarker, b->vnx, b->npt);
    CPD(f->yNormBodyMarker, b->vny, b->npt); CPD(f->zNormBodyMarker, b->vnz, b->npt);
    return 0;
}

int og_get_dcip(const og_state *s, double *xyzDCIP)
{
    if (!s || !xyzDCIP) return -1;
    memcpy(xyzDCIP, s->dcip, 3 * (size_t)s->n * sizeof(double));
    return 0;
}

/* ------------------------------------------------------------------------ */
int og_set_marker_velocity(og_state *s, int ibody, const double *u, const double *v, const double *w,
                           const double *uOld, const double *vOld, const double *wOld)
{
    if (!s || ibody < 1 || ibody > s->p.nBody || !s->b[ibody - 1].x) return -1;
    og_body *b = &s->b[ibody - 1];
    const size_t n = (size_t)b->npt * sizeof(double);
    memcpy(b->mu, u, n); memcpy(b->mv, v, n); memcpy(b->mw, w, n);
    memcpy(b->muo, uOld, n); memcpy(b->mvo, vOld, n); memcpy(b->mwo, wOld, n);
    return 0;
}

/* LU with partial pivoting (first largest pivot) and column solves: the
 * inverse of an n x n matrix, n <= 8 (the reference calls DGETRF/DGETRI) */
static void invert_n(int n, double A[8][8], double X[8][8])
{
    int piv[8];
    for (int c = 0; c < n; ++c) {
        int p = c;
        double best = fabs(A[c][c]);
        for (int r = c + 1; r < n; ++r)
            if (fabs(A[r][c]) > best) { best = fabs(A[r][c]); p = r; }
        piv[c] = p;
        if (p != c)
            for (int q = 0; q < n; ++q) { double t = A[c][q]; A[c][q] = A[p][q]; A[p][q] = t; }
        const double inv = 1.0 / A[c][c];
        for (int r = c + 1; r < n; ++r) {
            const double l = A[r][c] * inv;
            A[r][c] = l;
            for (int q = c + 1; q < n; ++q) A[r][q] = A[r][q] - l * A[c][q];
        }
    }
    for (int col = 0; col < n; ++col) {
        double e[8];
        for (int r = 0; r < n; ++r) e[r] = (r == col) ? 1.0 : 0.0;
        for (int c = 0; c < n; ++c)
            if (piv[c] != c) { double t = e[c]; e[c] = e[piv[c]]; e[piv[c]] = t; }
        for (int r = 1; r < n; ++r)
            for (int q = 0; q < r; ++q) e[r] = e[r] - A[r][q] * e[q];
        for (int r = n - 1; r >= 0; --r) {
            for (int q = r + 1; q < n; ++q) e[r] = e[r] - A[r][q] * e[q];
            e[r] = e[r] / A[r][r];
        }
        for (int r = 0; r < n; ++r) X[r][col] = e[r];
    }
}

/* GCM_calc_BIVelocity_Unstruc, GCM_BODYINTERCEPT_VALUES.F90:198-388: the
 * body velocity, linear over the closest element (its three markers and
 * zero normal derivative), at the intercept; current and previous step */
static void bi_velocity(const og_body *b, int elem, double xB, double yB, double zB, double *vel, double *old)
{
    double A[8][8], X[8][8];
    const int m = elem - 1;
    for (int q = 0; q < 3; ++q) {
        const int v = b->tri[3 * m + q] - 1;
        A[q][0] = b->x[v]; A[q][1] = b->y[v]; A[q][2] = b->z[v]; A[q][3] = 1.0;
    }
    A[3][0] = b->enx[m]; A[3][1] = b->eny[m]; A[3][2] = b->enz[m]; A[3][3] = 0.0;
    invert_n(4, A, X);
    const double *src[6] = {b->mu, b->mv, b->mw, b->muo, b->mvo, b->mwo};
    for (int c = 0; c < 6; ++c) {
        double rhs[4];
        for (int q = 0; q < 3; ++q) rhs[q] = src[c][b->tri[3 * m + q] - 1];
        rhs[3] = 0.0;
        double cX = 0.0, cY = 0.0, cZ = 0.0, cC = 0.0;
        for (int i = 0; i < 4; ++i) {
            cX = cX + X[0][i] * rhs[i];
            cY = cY + X[1][i] * rhs[i];
            cZ = cZ + X[2][i] * rhs[i];
            cC = cC + X[3][i] * rhs[i];
        }
        const double r = cX * xB + cY * yB + cZ * zB + cC;
        if (c < 3) vel[c] = r; else old[c - 3] = r;
    }
}

/* GCM_SetBodyInterceptValues, GCM_BODYINTERCEPT_VALUES.F90:22-178 */
int og_body_intercept_values(og_state *s, double dt, int ntime)
{
    if (!s) return -1;
    double **a[8] = {&s->bgu, &s->bgv, &s->bgw, &s->bgd, &s->bdu, &s->bdv, &s->bdw, &s->bdd};
    for (int q = 0; q < 8; ++q) { free(*a[q]); *a[q] = calloc((q < 4 ? s->nG : s->nD) + 1, sizeof(double)); }
    for (int kind = 0; kind < 2; ++kind) {
        const int nn = kind ? s->nD : s->nG;
        for (int n = 0; n < nn; ++n) {
            const int iG = kind ? s->di[n] : s->gi[n], jG = kind ? s->dj[n] : s->gj[n], kG = kind ? s->dk[n] : s->gk[n];
            const int ib = s->bodynum[ID(s, iG, jG, kG)];
            if (ib < 1 || ib > s->p.nBody) return -5;
            const int el = kind ? s->del[n] : s->gel[n];
            const double xB = kind ? s->dbx[n] : s->gbx[n], yB = kind ? s->dby[n] : s->gby[n],
                         zB = kind ? s->dbz[n] : s->gbz[n];
            double vel[3], old[3];
            bi_velocity(&s->b[ib - 1], el, xB, yB, zB, vel, old);
            double *U = kind ? s->bdu : s->bgu, *V = kind ? s->bdv : s->bgv, *W = kind ? s->bdw : s->bgw;
            double *D = kind ? s->bdd : s->bgd;
            U[n] = vel[0]; V[n] = vel[1]; W[n] = vel[2];
            if (ntime > 1) {
                const double nx = kind ? s->dnx[n] : s->gnx[n], ny = kind ? s->dny[n] : s->gny[n],
                             nz = kind ? s->dnz[n] : s->gnz[n];
                const double d = (vel[0] - old[0]) / dt * nx + (vel[1] - old[1]) / dt * ny + (vel[2] - old[2]) / dt * nz;
                D[n] = -d;
            } else {
                D[n] = 0.0;
            }
        }
    }
    return 0;
}

int og_get_bi_values(const og_state *s, double *uG, double *vG, double *wG, double *dpdnG, double *uD, double *vD,
                     double *wD, double *dpdnD)
{
    if (!s || !s->bgu) return -1;
    CPD(uG, s->bgu, s->nG); CPD(vG, s->bgv, s->nG); CPD(wG, s->bgw, s->nG); CPD(dpdnG, s->bgd, s->nG);
    CPD(uD, s->bdu, s->nD); CPD(vD, s->bdv, s->nD); CPD(wD, s->bdw, s->nD); CPD(dpdnD, s->bdd, s->nD);
    return 0;
}

/* marker_interpolation, UTIL_DRAG_LIFT.F90:375-686, sign = -1 */
static void marker_interpolation(const og_state *s, int ib, double xE0, double yE0, double zE0, double xN, double yN,
                                 double zN, double sign, const double *p, const double *u, const double *v,
                                 const double *w, double *dUdn, double *pEc)
{
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    int ii = -1, jj = -1, kk = -1;
    *pEc = 0.0;
    dUdn[0] = dUdn[1] = dUdn[2] = 0.0;
    const int j1 = 1, j2 = ny - 1, k1 = 1, k2 = nz - 1;
    double xE = xE0, yE = yE0, zE = zE0;
    if (s->p.unstruc_surface_type[ib - 1] == FC_SURFACE_MEMBRANE) {
        const double tk = s->p.membrane_tkns;
        xE = xE0 + sign * xN * tk / 2.0;
        yE = yE0 + sign * yN * tk / 2.0;
        zE = zE0 + sign * zN * tk / 2.0;
    }
    double dmin = 1.0e8;
    int nG = 0, iG = 0, jG = 0, kG = 0;
    for (int n = 0; n < s->nD; ++n) {
        const int i = s->di[n], j = s->dj[n], k = s->dk[n];
        const double d = sq(s->xc[i] - xE) + sq(s->yc[j] - yE) + sq(s->zc[k] - zE);
        if (d <= dmin) { dmin = d; nG = n; iG = i; jG = j; kG = k; }
    }
    const int i0 = imax(1, iG - 20), i1 = imin(nx - 1, iG + 20);
    const int jA = imax(j1, jG - 20), jB = imin(j2, jG + 20);
    const int kA = imax(k1, kG - 20), kB = imin(k2, kG + 20);
    for (int i = i0; i <= i1; ++i) if ((s->xc[i] - xE) * (s->xc[i + 1] - xE) <= 0.0) ii = i;
    for (int j = jA; j <= jB; ++j) if ((s->yc[j] - yE) * (s->yc[j + 1] - yE) <= 0.0) jj = j;
    for (int k = kA; k <= kB; ++k) if ((s->zc[k] - zE) * (s->zc[k + 1] - zE) <= 0.0) kk = k;
    if (ii == -1 || jj == -1 || kk == -1) return;
    double pinv = 0.0, dint = 0.0;
    for (int k = 0; k <= 1; ++k)
        for (int j = 0; j <= 1; ++j)
            for (int i = 0; i <= 1; ++i) {
                const long q = ID(s, ii + i, jj + j, kk + k);
                const double d = sq(s->xc[ii + i] - xE) + sq(s->yc[jj + j] - yE) + sq(s->zc[kk + k] - zE);
                if (s->iblank[q] == 0 || s->dead[q] == -1) {
                    dint = dint + (1.0 / d);
                    pinv = pinv + (1.0 / d) * p[q];
                }
            }
    int idw = 0;
    for (int k = 0; k <= 1 && !idw; ++k)
        for (int j = 0; j <= 1 && !idw; ++j)
            for (int i = 0; i <= 1 && !idw; ++i) {
                const long q = ID(s, ii + i, jj + j, kk + k);
                if (s->iblank[q] == 1 && s->dead[q] == 0 && !(dint < 0.000001)) idw = 1;
            }
    if (idw) {
        *pEc = pinv / dint;
    } else {
        const double aX = (xE - s->xc[ii]) * (1.0 / s->dxc[ii + 1]);
        const double aY = (yE - s->yc[jj]) * (1.0 / s->dyc[jj + 1]);
        const double aZ = (zE - s->zc[kk]) * (1.0 / s->dzc[kk + 1]);
#define P3(a, b, c) p[ID(s, ii + (a), jj + (b), kk + (c))]
        *pEc = P3(0, 0, 0) * (1.0 - aX) * (1.0 - aY) * (1.0 - aZ) + P3(1, 0, 0) * aX * (1.0 - aY) * (1.0 - aZ) +
               P3(0, 1, 0) * (1.0 - aX) * aY * (1.0 - aZ) + P3(0, 0, 1) * (1.0 - aX) * (1.0 - aY) * aZ +
               P3(1, 1, 0) * aX * aY * (1.0 - aZ) + P3(1, 0, 1) * aX * (1.0 - aY) * aZ +
               P3(0, 1, 1) * (1.0 - aX) * aY * aZ + P3(1, 1, 1) * aX * aY * aZ;
#undef P3
    }
    double uI = 0.0, vI = 0.0, wI = 0.0;
    for (int r = 0; r < 8; ++r) {
        const int i = s->dci[nG] + INC_I[r], j = s->dcj[nG] + INC_J[r], k = s->dck[nG] + INC_K[r];
        const double c = s->dcd[8 * nG + r];
        if (i != iG || j != jG || k != kG) {
            const long q = ID(s, i, j, k);
            uI = uI + c * u[q]; vI = vI + c * v[q]; wI = wI + c * w[q];
        } else {
            uI = uI + c * s->bdu[nG]; vI = vI + c * s->bdv[nG]; wI = wI + c * s->bdw[nG];
        }
    }
    const long qG = ID(s, iG, jG, kG);
    dUdn[0] = (uI - u[qG]) / s->dpl[nG] * (-sign);
    dUdn[1] = (vI - v[qG]) / s->dpl[nG] * (-sign);
    dUdn[2] = (wI - w[qG]) / s->dpl[nG] * (-sign);
}

/* compute_marker_stress, UTIL_DRAG_LIFT.F90:29-342 (one rank; the FSI
 * blending of :318-339 is not applied) */
int og_marker_stress(og_state *s, const double *p, const double *u, const double *v, const double *w, double reinv,
                     double *totals)
{
    if (!s || !s->bdu || s->nD < 1) return -1;
    const int nx = s->nx;
    for (int ib = 1; ib <= s->p.nBody; ++ib) {
        og_body *b = &s->b[ib - 1];
        double c[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};   /* cxt cyt czt cmxt cmyt cmzt cxs cys czs */
        double *acc[7] = {b->ma, b->sx, b->sy, b->sz, b->fx, b->fy, b->fz};
        for (int a = 0; a < 7; ++a) memset(acc[a], 0, b->npt * sizeof(double));
        for (int m = 0; m < b->nel; ++m) {
            const double xE = b->ecx[m], yE = b->ecy[m], zE = b->ecz[m];
            const double xN = b->enx[m], yN = b->eny[m], zN = b->enz[m];
            const double ar = b->area[m];
            double dU[3] = {0.0, 0.0, 0.0}, pE = 0.0;
            if (zE <= s->z[s->nz] && zE > s->z[1] && yE <= s->y[s->ny] && yE > s->y[1] && xE <= s->xc[nx - 1] &&
                xE > s->xc[1])
                marker_interpolation(s, ib, xE, yE, zE, xN, yN, zN, -1.0, p, u, v, w, dU, &pE);
            dU[0] = dU[0] - 0.0; dU[1] = dU[1] - 0.0; dU[2] = dU[2] - 0.0;
            pE = pE - 0.0;
            const double fn = dU[0] * xN + dU[1] * yN + dU[2] * zN;
            dU[0] = dU[0] - fn * xN;
            dU[1] = dU[1] - fn * yN;
            dU[2] = dU[2] - fn * zN;
            const double xpf = pE * xN * ar, ypf = pE * yN * ar, zpf = pE * zN * ar;
            const double xsf = dU[0] * reinv * ar, ysf = dU[1] * reinv * ar, zsf = dU[2] * reinv * ar;
            c[0] = c[0] + xpf + xsf;
            c[1] = c[1] + ypf + ysf;
            c[2] = c[2] + zpf + zsf;
            c[3] = c[3] + yE * (zsf + zpf) - zE * (ysf + ypf);
            c[4] = c[4] + zE * (xsf + xpf) - xE * (zsf + zpf);
            c[5] = c[5] + xE * (ysf + ypf) - yE * (xsf + xpf);
            c[6] = c[6] + xsf;
            c[7] = c[7] + ysf;
            c[8] = c[8] + zsf;
            for (int q = 0; q < 3; ++q) {
                const int n = b->tri[3 * m + q] - 1;
                b->ma[n] = b->ma[n] + ar / 3.0;
                b->sx[n] = b->sx[n] + xsf / 3.0;
                b->sy[n] = b->sy[n] + ysf / 3.0;
                b->sz[n] = b->sz[n] + zsf / 3.0;
                b->fx[n] = b->fx[n] + (xpf + xsf) / 3.0;
                b->fy[n] = b->fy[n] + (ypf + ysf) / 3.0;
                b->fz[n] = b->fz[n] + (zpf + zsf) / 3.0;
            }
        }
        for (int n = 0; n < b->npt; ++n) {
            b->tx[n] = b->fx[n] / b->ma[n];
            b->ty[n] = b->fy[n] / b->ma[n];
            b->tz[n] = b->fz[n] / b->ma[n];
        }
        if (totals) memcpy(totals + 9 * (ib - 1), c, sizeof(c));
    }
    return 0;
}

int og_get_marker_forces(const og_state *s, int ibody, fc_geom_forces *f)
{
    if (!s || !f || ibody < 1 || ibody > s->p.nBody || !s->b[ibody - 1].x) return -1;
    const og_body *b = &s->b[ibody - 1];
    const int n = b->npt;
    CPD(f->xMarkerForce, b->fx, n); CPD(f->yMarkerForce, b->fy, n); CPD(f->zMarkerForce, b->fz, n);
    CPD(f->xMarkerShear, b->sx, n); CPD(f->yMarkerShear, b->sy, n); CPD(f->zMarkerShear, b->sz, n);
    CPD(f->MarkerArea, b->ma, n); CPD(f->xMarkerStress, b->tx, n); CPD(f->yMarkerStress, b->ty, n);
    CPD(f->zMarkerStress, b->tz, n);
    return 0;
}
