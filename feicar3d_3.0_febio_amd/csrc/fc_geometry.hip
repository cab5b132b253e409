// fc_geometry.hip -- the immersed-boundary geometry refresh of one time step
// of a moving triangulated body, on the device (include/fc_geometry.h):
// surface normals (calculate_arclength_norm_ds), iblank from the surface
// (set_iblank_body_fast with calc_bodyIntercept_Unstruc), the smoothing
// passes, ghost and dead cells with their image points and GCM
// interpolation coefficients (GCM_set_internal_boundary).
//
// This file is compiled with -ffp-contract=off: every product and sum is
// rounded on its own, in the reference's order, so the floating-point
// results match the CPU restatement of these routines bit for bit.
//
// Parallel forms of the reference's sequential loops:
//   * the closest marker of a cell: the reference scans every marker
//     (MINLOC, the first of equal distances); here the markers are binned
//     on a uniform grid each time they move, and a cell scans the rings of
//     bins around its own until no unvisited bin can hold a marker as close
//     as the best one found, keeping the lowest marker number among equal
//     distances, which is the same marker;
//   * the elements around that marker come from a marker -> element
//     incidence list in element order, so ties between elements resolve as
//     in the reference's element loop;
//   * the vertex normals gather their elements' normals in element order,
//     which is the reference's order of accumulation;
//   * the in-place smoothing sweeps (Gauss-Seidel over the grid) are
//     reproduced exactly by rounds of a Jacobi iteration in which a cell reads
//     its lower-index face neighbours from the previous round and its
//     higher-index ones from the sweep's start: a face neighbour comes
//     earlier in any of the sweeps' loop orders exactly when it has the lower
//     index, so the rounds settle on the sweep's result after as many rounds
//     as the longest chain of cells that turn solid one after another;
//   * the ghost and dead lists are compacted in array order (k, j, i) from
//     per-row counts and a scan of the row counts;
//   * alphaGhost gathers the contributions of a ghost cell's dead
//     neighbours in dead-list order (increasing array index).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fc_geometry.h"
#include "fc_handle.h"

// element rows padded to a multiple of 16 doubles (the force-totals kernel
// streams them in batches of 16)
#define FC_ROWPAD(n) ((((n) + 15) / 16) * 16)

namespace {

constexpr int UNDEC = 100;          // BOUNDARY_IBLANK_FAST.F90:14

struct GGrid {
    int nx, ny, nz;
    long n;
    const double *x, *y, *z, *xc, *yc, *zc, *dx, *dy, *dz, *dxc, *dyc, *dzc;
};

struct GBody {
    int npt, nel, kind, dim;
    const double *x, *y, *z;        // markers
    const int *tri;                 // 3*nel, 1-based markers
    const int *voff, *vel;          // marker -> incident elements (0-based), element order
    double *enx, *eny, *enz, *area, *ecx, *ecy, *ecz, *vnx, *vny, *vnz;
    // marker bins: origin, bin size, bins per axis; bin b holds markers
    // bmark[bstart[b] .. bstart[b+1]), whose coordinates bxyz holds in that
    // order (x, y, z of slot t at 3t .. 3t+2)
    double bo[3], bh, bih;
    int bd[3];
    const int *bstart, *bmark;
    const double *bxyz;
    const double *mvel;             // u, v, w, uOld, vOld, wOld of the markers (npt each)
};

struct GBodies {
    GBody b[FC_GEOM_MAX_BODIES];
    int nb;
    double tkns, probe, probeD;
};

struct GMasks {
    int8_t *iblank, *ghost, *dead, *bodynum, *undec, *tmp, *ps1, *ps2;
};

// device counters: [0] undecided cells, [1] changed flag, [2] smoothed cells,
// [3] error flags, [4] unselected sides, [5] scan total, [6..11] the box of
// the interior solid cells (-min i, max i, -min j, max j, -min k, max k)
enum { C_UNDEC = 0, C_FLAG, C_FLIPS, C_ERR, C_UNSEL, C_TOTAL, C_BOX, C_N = C_BOX + 6 };
enum { E_GHOST_HOST = 1, E_DEAD_HOST = 2, E_DEAD_BODY = 4, E_GHOST_BODY = 8 };

__device__ __forceinline__ long gid(const GGrid &g, int i, int j, int k)
{
    return (long)i + (long)(g.nx + 2) * ((long)j + (long)(g.ny + 2) * (long)k);
}
__device__ __forceinline__ double sq(double a) { return a * a; }

// ---------------------------------------------------------------------------
// calculate_arclength_norm_ds (BOUNDARY_SET_ARCLENGTH_NORM.F90:38-186):
// element normals, areas and centroids, then vertex normals
__global__ void k_surf_elem(GBody b)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= b.nel) return;
    const int n1 = b.tri[3 * m] - 1, n2 = b.tri[3 * m + 1] - 1, n3 = b.tri[3 * m + 2] - 1;
    const double v1x = b.x[n2] - b.x[n1], v1y = b.y[n2] - b.y[n1], v1z = b.z[n2] - b.z[n1];
    const double v2x = b.x[n3] - b.x[n2], v2y = b.y[n3] - b.y[n2], v2z = b.z[n3] - b.z[n2];
    const double nx = v1y * v2z - v1z * v2y;
    const double ny = v1z * v2x - v1x * v2z;
    const double nz = v1x * v2y - v1y * v2x;
    const double mag = sqrt(sq(nx) + sq(ny) + sq(nz));
    b.area[m] = 0.5 * mag;
    b.enx[m] = nx / mag;
    b.eny[m] = ny / mag;
    b.enz[m] = nz / mag;
    b.ecx[m] = (b.x[n1] + b.x[n2] + b.x[n3]) / 3.0;
    b.ecy[m] = (b.y[n1] + b.y[n2] + b.y[n3]) / 3.0;
    b.ecz[m] = (b.z[n1] + b.z[n2] + b.z[n3]) / 3.0;
}

__global__ void k_surf_vert(GBody b)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= b.npt) return;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int e = b.voff[v]; e < b.voff[v + 1]; ++e) {
        const int m = b.vel[e];
        sx += b.enx[m];
        sy += b.eny[m];
        sz += b.enz[m];
    }
    const double mag = sqrt(sq(sx) + sq(sy) + sq(sz));
    b.vnx[v] = sx / mag;
    b.vny[v] = sy / mag;
    b.vnz[v] = sz / mag;
}

// ---------------------------------------------------------------------------
// calc_bodyIntercept_Unstruc (BOUNDARY_IBLANK_FAST.F90:857-1216), nCheck = 1.
// The closest marker (the first of equal squared distances, as the
// reference's MINLOC): rings of bins at Chebyshev distance r = 0, 1, ...
// around the cell's bin; every marker of a bin outside the rings visited so
// far lies at least r * bh from the cell, so the search stops once the best
// squared distance is below (r * bh)^2, reduced by a relative 1e-12 that
// covers the rounding of the bin assignment.  Ties keep the lower marker
// number, whatever order the bins hold their markers in.
__device__ __forceinline__ void scan_bin(const GBody &b, int bin, double xC, double yC, double zC, double &best,
                                         int &cv, int &cnt)
{
    cnt += b.bstart[bin + 1] - b.bstart[bin];
    for (int t = b.bstart[bin]; t < b.bstart[bin + 1]; ++t) {
        const int m = b.bmark[t];
        const double d = sq(b.bxyz[3 * t] - xC) + sq(b.bxyz[3 * t + 1] - yC) + sq(b.bxyz[3 * t + 2] - zC);
        if (cv < 0 || d < best || (d == best && m < cv)) {
            best = d;
            cv = m;
        }
    }
}

__device__ int closest_marker(const GBody &b, double xC, double yC, double zC, int &cnt)
{
    int q[3];
    const double P[3] = {xC, yC, zC};
    for (int a = 0; a < 3; ++a) {
        const double f = floor((P[a] - b.bo[a]) * b.bih);
        q[a] = (int)fmax(fmin(f, 1.0e8), -1.0e8);
    }
    // rings closer than the bin grid are empty
    int r0 = 0;
    for (int a = 0; a < 3; ++a) r0 = max(r0, max(-q[a], q[a] - (b.bd[a] - 1)));
    double best = 0.0;
    int cv = -1;
    for (int r = r0;; ++r) {
        for (int dk = -r; dk <= r; ++dk) {
            const int k = q[2] + dk;
            if (k < 0 || k >= b.bd[2]) continue;
            for (int dj = -r; dj <= r; ++dj) {
                const int j = q[1] + dj;
                if (j < 0 || j >= b.bd[1]) continue;
                const bool face = (dk == -r || dk == r || dj == -r || dj == r);
                const int istep = face ? 1 : 2 * r;
                for (int di = -r; di <= r; di += (istep > 0 ? istep : 1)) {
                    const int i = q[0] + di;
                    if (i < 0 || i >= b.bd[0]) continue;
                    scan_bin(b, i + b.bd[0] * (j + b.bd[1] * k), xC, yC, zC, best, cv, cnt);
                }
            }
        }
        const double lim = (double)r * b.bh * (1.0 - 1.0e-12);
        if (cv >= 0 && best < lim * lim) break;
        // every bin visited
        if (q[0] - r <= 0 && q[0] + r >= b.bd[0] - 1 && q[1] - r <= 0 && q[1] + r >= b.bd[1] - 1 && q[2] - r <= 0 &&
            q[2] + r >= b.bd[2] - 1)
            break;
    }
    return cv;
}

// check_BIInsideTriangle (BOUNDARY_IBLANK_FAST.F90:1220-1313)
__device__ void inside_triangle(const double *xv, const double *yv, const double *zv, double xb, double yb,
                                double zb, double &area123, double &areaDiff)
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
    area123 = a123;
    areaDiff = a124 + a234 + a314 - a123;
}

// calc_BIOutsideTriangle (BOUNDARY_IBLANK_FAST.F90:1316-1536) with
// calc_crossProduct (:1539-1555).  Returns false when no side is selected
// (the reference then reads an unset isideSelect); side 1 is used there.
__device__ bool outside_triangle(const double *xv, const double *yv, const double *zv, double xc, double yc,
                                 double zc, const double *xn, const double *yn, const double *zn, double &xb,
                                 double &yb, double &zb, double &dist, double &bnx, double &bny, double &bnz)
{
    const double v4x = xb, v4y = yb, v4z = zb;
    int sel = 0;
    for (int is = 1; is <= 3; ++is) {
        const int a1 = is - 1, a2 = (is == 3) ? 0 : is, a3 = (is == 1) ? 2 : (is == 2) ? 0 : 1;
        const double ax = xv[a3] - xv[a1], ay = yv[a3] - yv[a1], az = zv[a3] - zv[a1];
        const double bx = xv[a2] - xv[a1], by = yv[a2] - yv[a1], bz = zv[a2] - zv[a1];
        const double cx = v4x - xv[a1], cy = v4y - yv[a1], cz = v4z - zv[a1];
        const double p0 = ay * bz - by * az, p1 = az * bx - bz * ax, p2 = ax * by - bx * ay;
        const double q0 = cy * bz - by * cz, q1 = cz * bx - bz * cx, q2 = cx * by - bx * cy;
        const double dot = q0 * p0 + q1 * p1 + q2 * p2;
        if (dot <= 0.0) {
            sel = is;
            break;
        }
    }
    const bool found = sel != 0;
    if (!found) sel = 1;
    const int n1 = sel - 1, n2 = (sel == 3) ? 0 : sel;
    double ex = xv[n2] - xv[n1], ey = yv[n2] - yv[n1], ez = zv[n2] - zv[n1];
    const double mag = sqrt(sq(ex) + sq(ey) + sq(ez));
    ex = ex / mag;
    ey = ey / mag;
    ez = ez / mag;
    const double pl = (xb - xv[n1]) * ex + (yb - yv[n1]) * ey + (zb - zv[n1]) * ez;
    const double px = xv[n1] + pl * ex, py = yv[n1] + pl * ey, pz = zv[n1] + pl * ez;
    const double d1 = sqrt(sq(xv[n1] - px) + sq(yv[n1] - py) + sq(zv[n1] - pz)) / mag;
    const double d2 = sqrt(sq(xv[n2] - px) + sq(yv[n2] - py) + sq(zv[n2] - pz)) / mag;
    if (d1 <= 1.0 && d2 <= 1.0) {
        xb = px; yb = py; zb = pz;
        bnx = 0.5 * (xn[n1] + xn[n2]);
        bny = 0.5 * (yn[n1] + yn[n2]);
        bnz = 0.5 * (zn[n1] + zn[n2]);
    } else if (d2 > 1.0) {
        xb = xv[n1]; yb = yv[n1]; zb = zv[n1];
        bnx = xn[n1]; bny = yn[n1]; bnz = zn[n1];
    } else {
        xb = xv[n2]; yb = yv[n2]; zb = zv[n2];
        bnx = xn[n2]; bny = yn[n2]; bnz = zn[n2];
    }
    dist = sqrt(sq(xc - xb) + sq(yc - yb) + sq(zc - zb));
    return found;
}

// the body intercept from the closest marker cv: the elements around it in
// element order, the last of equal distance wins
struct Intercept {
    double x, y, z, dot;
    int elem;
};

__device__ Intercept intercept_from(const GBody &b, int cv, double xC, double yC, double zC, int *unsel)
{
    double dbest = 1.0e16, xB = 0.0, yB = 0.0, zB = 0.0, nbx = 0.0, nby = 0.0, nbz = 0.0;
    int ce = 0;
    for (int e = b.voff[cv]; e < b.voff[cv + 1]; ++e) {
        const int m = b.vel[e];
        if (e > b.voff[cv] && b.vel[e - 1] == m) continue;      // an element listed twice counts once
        double xv[3], yv[3], zv[3], xn[3], yn[3], zn[3];
        for (int q = 0; q < 3; ++q) {
            const int v = b.tri[3 * m + q] - 1;
            xv[q] = b.x[v]; yv[q] = b.y[v]; zv[q] = b.z[v];
            xn[q] = b.vnx[v]; yn[q] = b.vny[v]; zn[q] = b.vnz[v];
        }
        const double ex = b.enx[m], ey = b.eny[m], ez = b.enz[m];
        const double pc = -ex * xv[0] - ey * yv[0] - ez * zv[0];
        const double dp = -(ex * xC + ey * yC + ez * zC + pc);
        double xt = xC + ex * dp, yt = yC + ey * dp, zt = zC + ez * dp;
        double a123, adiff;
        inside_triangle(xv, yv, zv, xt, yt, zt, a123, adiff);
        double de = sqrt(sq(xC - xt) + sq(yC - yt) + sq(zC - zt));
        double tx, ty, tz;
        if (fabs(adiff) < 1.0e-6 * a123) {
            tx = ex; ty = ey; tz = ez;
        } else if (!outside_triangle(xv, yv, zv, xC, yC, zC, xn, yn, zn, xt, yt, zt, de, tx, ty, tz)) {
            atomicAdd(unsel, 1);
        }
        if (de <= dbest) {
            dbest = de;
            ce = m + 1;
            xB = xt; yB = yt; zB = zt;
            nbx = tx; nby = ty; nbz = tz;
        }
    }
    Intercept r;
    r.x = xB; r.y = yB; r.z = zB;
    r.elem = ce;
    r.dot = (xC - xB) * nbx + (yC - yB) * nby + (zC - zB) * nbz;
    return r;
}

// ---------------------------------------------------------------------------
// set_iblank_body_fast (BOUNDARY_IBLANK_FAST.F90:2-854), one rank, 3-D
// Full-grid kernels: a grid-stride loop over the array index q (below 2^31),
// with 32-bit index arithmetic.
__device__ __forceinline__ void cell3(const GGrid &g, unsigned q, int &i, int &j, int &k)
{
    const unsigned sx = (unsigned)g.nx + 2u, sy = (unsigned)g.ny + 2u;
    const unsigned r = q / sx;
    i = (int)(q - r * sx);
    k = (int)(r / sy);
    j = (int)(r - (unsigned)k * sy);
}

__global__ void k_iblank_init(GGrid g, GMasks m)
{
    const unsigned sx = (unsigned)g.nx + 2u;
    for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < (unsigned)g.n; q += gridDim.x * blockDim.x) {
        const unsigned i = q % sx;
        m.undec[q] = (int8_t)((i <= (unsigned)g.nx) ? UNDEC : 0);
        m.tmp[q] = 0;
        m.bodynum[q] = 0;
    }
}

// lane-aggregated append: one atomic per wavefront
__device__ __forceinline__ void wave_append(bool want, int value, int *list, int *counter)
{
    const unsigned long long mask = __ballot(want);
    if (!mask) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(mask));
    base = __shfl(base, leader);
    if (want) list[base + __popcll(mask & ((1ull << lane) - 1ull))] = value;
}

// the last i in [0, n-1] with (c(i)-v)(c(i+1)-v) <= 0 (+1 for the upper
// bound), or init when there is none: lanes scan strided i, the wavefront
// takes the largest
__device__ __forceinline__ int wave_bracket(const double *c, int n, double v, int plus, int init, int lane)
{
    int a = -1;
    for (int i = lane; i <= n - 1; i += 64)
        if ((c[i] - v) * (c[i + 1] - v) <= 0.0) a = i + plus;
    for (int off = 32; off > 0; off >>= 1) a = max(a, __shfl_xor(a, off));
    return a < 0 ? init : a;
}

// the cells in the bounding box of each element (widened by three
// thicknesses for a membrane) become undecided (:94-305); one wavefront per
// element
__global__ void __launch_bounds__(256) k_bbox_mark(GGrid g, GBody b, double tkns, GMasks m)
{
    const int e = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (e >= b.nel) return;
    const bool mem = b.kind == FC_SURFACE_MEMBRANE;
    const double th = tkns * (mem ? 3.0 : 0.0);
    double xv[3], yv[3], zv[3];
    for (int q = 0; q < 3; ++q) {
        const int v = b.tri[3 * e + q] - 1;
        xv[q] = b.x[v]; yv[q] = b.y[v]; zv[q] = b.z[v];
    }
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if ((xv[0] <= g.x[1] && xv[1] <= g.x[1] && xv[2] <= g.x[1]) ||
        (xv[0] >= g.x[nx] && xv[1] >= g.x[nx] && xv[2] >= g.x[nx]) ||
        (yv[0] + th <= g.y[1] && yv[1] + th <= g.y[1] && yv[2] + th <= g.y[1]) ||
        (yv[0] - th >= g.y[ny] && yv[1] - th >= g.y[ny] && yv[2] - th >= g.y[ny]) ||
        (zv[0] + th <= g.z[1] && zv[1] + th <= g.z[1] && zv[2] + th <= g.z[1]) ||
        (zv[0] - th >= g.z[nz] && zv[1] - th >= g.z[nz] && zv[2] - th >= g.z[nz]))
        return;
    double x0 = fmin(fmin(xv[0], xv[1]), xv[2]), x1 = fmax(fmax(xv[0], xv[1]), xv[2]);
    double y0 = fmin(fmin(yv[0], yv[1]), yv[2]), y1 = fmax(fmax(yv[0], yv[1]), yv[2]);
    double z0 = fmin(fmin(zv[0], zv[1]), zv[2]), z1 = fmax(fmax(zv[0], zv[1]), zv[2]);
    if (mem) { x0 -= th; x1 += th; y0 -= th; y1 += th; z0 -= th; z1 += th; }
    int i0 = wave_bracket(g.xc, nx, x0, 0, 0, lane), i1 = wave_bracket(g.xc, nx, x1, 1, 0, lane);
    if (i0 == 0) i0 = 1;
    if (i1 == 0 || i1 == nx) i1 = nx - 1;
    int j0 = wave_bracket(g.yc, ny, y0, 0, 0, lane), j1 = wave_bracket(g.yc, ny, y1, 1, 0, lane);
    if (j0 == 0) j0 = 1;
    if (j1 == 0 || j1 == ny) j1 = ny - 1;
    int k0 = wave_bracket(g.zc, nz, z0, 0, 0, lane), k1 = wave_bracket(g.zc, nz, z1, 1, 0, lane);
    if (k0 == 0) k0 = 1;
    if (k1 == 0 || k1 == nz) k1 = nz - 1;
    const int ni = i1 - i0 + 1, nj = j1 - j0 + 1, nk = k1 - k0 + 1;
    if (ni <= 0 || nj <= 0 || nk <= 0) return;
    const long tot = (long)ni * nj * nk;
    for (long t = lane; t < tot; t += 64) {
        const int i = i0 + (int)(t % ni);
        const long r = t / ni;
        m.undec[gid(g, i, j0 + (int)(r % nj), k0 + (int)(r / nj))] = (int8_t)(-UNDEC);
    }
}

// interior cells marked undecided, listed (in any order; the element boxes
// lie inside the interior)
__global__ void k_collect_undecided(GGrid g, GMasks m, int *list, int *ctr)
{
    const unsigned stride = gridDim.x * blockDim.x;
    for (unsigned base = blockIdx.x * blockDim.x; base < (unsigned)g.n; base += stride) {
        const unsigned q = base + threadIdx.x;
        const bool want = q < (unsigned)g.n && m.undec[q] == -UNDEC;
        wave_append(want, (int)q, list, ctr + C_UNDEC);
    }
}


// solid or fluid by the closest body intercept: dotNorm >= 0 inside a solid
// body, distance below half the thickness for a membrane (:307-368); bodies
// from the last to the first, the first hit wins
// markers scanned by the closest-marker searches (the work of the search),
// summed per wavefront
__device__ __forceinline__ void add_pairs(int c, unsigned long long *pairs)
{
    unsigned long long v = (unsigned long long)c;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(pairs, v);
}

__global__ void __launch_bounds__(256) k_classify(GGrid g, GBodies B, GMasks m, const int *list, int cnt, int *ctr,
                                                  unsigned long long *pairs)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int scanned = 0;
    if (t < cnt) {
        const unsigned q = (unsigned)list[t];
        int i = 0, j = 0, k = 0;
        cell3(g, q, i, j, k);
        const double xP = g.xc[i], yP = g.yc[j], zP = g.zc[k];
        int hit = 0;
        for (int ib = B.nb; ib >= 1 && !hit; --ib) {
            const GBody &b = B.b[ib - 1];
            const int cv = closest_marker(b, xP, yP, zP, scanned);
            const Intercept r = intercept_from(b, cv, xP, yP, zP, ctr + C_UNSEL);
            if (b.kind == FC_SURFACE_MEMBRANE) {
                const double d = sqrt(sq(xP - r.x) + sq(yP - r.y) + sq(zP - r.z));
                if (d < B.tkns / 2.0) hit = ib;
            } else if (r.dot >= 0.0) {
                hit = ib;
            }
        }
        m.tmp[q] = (int8_t)(hit ? 1 : 0);
        if (hit) m.bodynum[q] = (int8_t)hit;
    }
    add_pairs(scanned, pairs);
}

// the cells no element box reached copy their left neighbour along i,
// starting from the first decided cell of the row (:372-411)
__global__ void __launch_bounds__(256) k_rowfill(GGrid g, GMasks m)
{
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int nj = g.ny - 1;
    if (r >= nj * (g.nz - 1)) return;
    const int j = 1 + r % nj, k = 1 + r / nj;
    const long q0 = gid(g, 0, j, k);
    const int last = g.nx - 1;
    // cell 0 takes the first decided cell of the row (stays 0 if none)
    int8_t cur = 0, curb = m.bodynum[q0];
    for (int i0 = 1; i0 <= last; i0 += 64) {
        const int i = i0 + lane;
        const unsigned long long bal = __ballot(i <= last && m.undec[q0 + i] != UNDEC);
        if (bal) {
            const int f = i0 + __ffsll((long long)bal) - 1;
            cur = m.tmp[q0 + f];
            curb = m.bodynum[q0 + f];
            break;
        }
    }
    if (lane == 0) {
        m.tmp[q0] = cur;
        m.bodynum[q0] = curb;
        m.iblank[q0] = cur;
    }
    // each undecided cell copies the last decided cell before it (the carry
    // of the previous chunk when there is none in its chunk), and its bodyNum
    // when that value is solid; the chunk's last value carries on
    for (int i0 = 1; i0 <= last; i0 += 64) {
        const int i = i0 + lane;
        const bool in = i <= last;
        const long q = q0 + i;
        const bool dec = in && m.undec[q] != UNDEC;
        int8_t t = dec ? m.tmp[q] : (int8_t)0, b = in ? m.bodynum[q] : (int8_t)0;
        int src = dec ? lane : -1;
        for (int off = 1; off < 64; off <<= 1) {
            const int o = __shfl_up(src, off);
            if (lane >= off) src = max(src, o);
        }
        const int st = __shfl((int)t, src < 0 ? 0 : src), sb = __shfl((int)b, src < 0 ? 0 : src);
        const int8_t vt = (src < 0) ? cur : (int8_t)st;
        const int8_t vb = (src < 0) ? curb : (int8_t)sb;
        if (in && !dec) {
            t = vt;
            if (vt == 1) b = vb;
            m.tmp[q] = t;
            if (vt == 1) m.bodynum[q] = b;
        }
        if (in) m.iblank[q] = t;
        // carry: the last cell of the chunk (its own value and bodyNum)
        const int lastlane = min(63, last - i0);
        cur = (int8_t)__shfl((int)t, lastlane);
        curb = (int8_t)__shfl((int)b, lastlane);
    }
    if (lane == 0) {
        m.tmp[q0 + g.nx] = cur;
        m.bodynum[q0 + g.nx] = curb;
        m.iblank[q0 + g.nx] = cur;
    }
}

// ---------------------------------------------------------------------------
// smoothing passes: smooth_iblank (GCM_SET_BOUNDARY.F90:1618-1697, 6 face
// neighbours, 5 solid make a cell solid) and the two 2-D passes of
// GCM_set_internal_boundary (:120-197, the 4 neighbours in the j-k and the
// i-k planes, 3 make it solid).  State per cell: iblank | bodyNum << 8.
// A cell outside the box of the solid cells has at most one solid face
// neighbour, so no pass turns it solid: the passes run on that box, and read
// the box widened by one cell.
struct Box {
    int i0, i1, j0, j1, k0, k1;
    long count() const { return (long)(i1 - i0 + 1) * (j1 - j0 + 1) * (k1 - k0 + 1); }
};

__device__ __forceinline__ bool box_cell(const Box &b, long t, int &i, int &j, int &k)
{
    const int ni = b.i1 - b.i0 + 1, nj = b.j1 - b.j0 + 1;
    if (t >= (long)ni * nj * (b.k1 - b.k0 + 1)) return false;
    i = b.i0 + (int)(t % ni);
    const long r = t / ni;
    j = b.j0 + (int)(r % nj);
    k = b.k0 + (int)(r / nj);
    return true;
}

__global__ void __launch_bounds__(256) k_solid_box(GGrid g, GMasks m, int *ctr)
{
    constexpr int NONE = -(1 << 30);
    int v[6] = {NONE, NONE, NONE, NONE, NONE, NONE};
    for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < (unsigned)g.n; q += gridDim.x * blockDim.x) {
        if (m.iblank[q] != 1) continue;
        int i, j, k;
        cell3(g, q, i, j, k);
        if (i < 1 || i > g.nx - 1 || j < 1 || j > g.ny - 1 || k < 1 || k > g.nz - 1) continue;
        v[0] = max(v[0], -i); v[1] = max(v[1], i);
        v[2] = max(v[2], -j); v[3] = max(v[3], j);
        v[4] = max(v[4], -k); v[5] = max(v[5], k);
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int a = 0; a < 6; ++a) v[a] = max(v[a], __shfl_xor(v[a], off));
    __shared__ int part[4][6];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int a = 0; a < 6; ++a) part[w][a] = v[a];
    __syncthreads();
    if (threadIdx.x < 6) {
        int r = part[0][threadIdx.x];
        for (int t = 1; t < 4; ++t) r = max(r, part[t][threadIdx.x]);
        if (r != NONE) atomicMax(ctr + C_BOX + threadIdx.x, r);
    }
}

__global__ void k_smooth_pack(GGrid g, Box b, GMasks m, int16_t *A, int16_t *P, int16_t *N)
{
    int i, j, k;
    if (!box_cell(b, (long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    const long q = gid(g, i, j, k);
    const int16_t v = (int16_t)((m.iblank[q] & 0xff) | ((int)(uint8_t)m.bodynum[q] << 8));
    A[q] = v;
    P[q] = v;
    N[q] = v;
}

// neighbour offsets in the order of each pass's inner loops
struct SmoothPass {
    int nn, thr;
    int di[6], dj[6], dk[6];
};

__global__ void k_smooth_round(GGrid g, Box b, SmoothPass sp, const int16_t *A, const int16_t *P, int16_t *N,
                               int *ctr)
{
    int i, j, k;
    if (!box_cell(b, (long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    const long q = gid(g, i, j, k);
    const int16_t a = A[q];
    if ((a & 0xff) == 1) return;
    int sum = 0, ibody = 0;
    for (int t = 0; t < sp.nn; ++t) {
        const long r = gid(g, i + sp.di[t], j + sp.dj[t], k + sp.dk[t]);
        const bool lower = sp.di[t] + sp.dj[t] + sp.dk[t] < 0;
        const int16_t v = lower ? P[r] : A[r];
        if ((v & 0xff) == 1) {
            ++sum;
            const int bn = (int8_t)(v >> 8);
            if (bn > 0) ibody = bn;
        }
    }
    const int16_t nv = (sum >= sp.thr) ? (int16_t)(1 | (ibody << 8)) : a;
    if (nv != P[q]) ctr[C_FLAG] = 1;
    N[q] = nv;
}

__global__ void k_smooth_unpack(GGrid g, Box b, GMasks m, const int16_t *A, const int16_t *P, int *ctr)
{
    int i, j, k;
    if (!box_cell(b, (long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    const long q = gid(g, i, j, k);
    if (P[q] == A[q]) return;
    m.iblank[q] = 1;
    m.bodynum[q] = (int8_t)(P[q] >> 8);
    atomicAdd(ctr + C_FLIPS, 1);
}

// ---------------------------------------------------------------------------
// GCM_define_ghostcells (GCM_SET_BOUNDARY.F90:800-876): interior fluid cells
// with an iblank face neighbour (of i in 1..nx-1); bodyNum from the first
// such neighbour in array order, the one the reference's loop visits first
__global__ void k_ghost_mark(GGrid g, GMasks m)
{
    const unsigned q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (unsigned)g.n) return;
    int i, j, k;
    cell3(g, q, i, j, k);
    int8_t gh = 0;
    if (i >= 1 && i <= g.nx - 1 && j >= 1 && j <= g.ny - 1 && k >= 1 && k <= g.nz - 1 && m.iblank[q] == 0) {
        const int oi[6] = {0, 0, -1, 1, 0, 0}, oj[6] = {0, -1, 0, 0, 1, 0}, ok[6] = {-1, 0, 0, 0, 0, 1};
        for (int t = 0; t < 6; ++t) {
            const int i2 = i + oi[t];
            if (i2 < 1 || i2 > g.nx - 1) continue;
            const long r = gid(g, i2, j + oj[t], k + ok[t]);
            if (m.iblank[r] == 1) {
                gh = 1;
                m.bodynum[q] = m.bodynum[r];
                break;
            }
        }
    }
    m.ghost[q] = gh;
}

// dead cells (GCM_set_deadcells :905-921): iblank face neighbours of
// interior ghost cells; ps1 and ps2 cleared
__global__ void k_dead_mark(GGrid g, GMasks m)
{
    const unsigned q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (unsigned)g.n) return;
    int i, j, k;
    cell3(g, q, i, j, k);
    int8_t dv = 0;
    if (m.iblank[q] == 1) {
        const int oi[6] = {1, -1, 0, 0, 0, 0}, oj[6] = {0, 0, 1, -1, 0, 0}, ok[6] = {0, 0, 0, 0, 1, -1};
        for (int t = 0; t < 6; ++t) {
            const int i2 = i + oi[t], j2 = j + oj[t], k2 = k + ok[t];
            if (i2 < 1 || i2 > g.nx - 1 || j2 < 1 || j2 > g.ny - 1 || k2 < 1 || k2 > g.nz - 1) continue;
            if (m.ghost[gid(g, i2, j2, k2)] == 1) dv = -1;
        }
    }
    m.dead[q] = dv;
    m.ps1[q] = 0;
    m.ps2[q] = 0;
}

// ---------------------------------------------------------------------------
// list compaction in array order: rows (j, k) in [j0, j1] x [k0, k1], cells i
// in 1..nx-1 with |mask| == 1
struct Rows {
    int j0, j1, k0, k1;
    int count() const { return (j1 - j0 + 1) * (k1 - k0 + 1); }
};

// one wavefront per row: the count of the row's listed cells
__global__ void __launch_bounds__(256) k_row_count(GGrid g, Rows R, const int8_t *mask, int *cnt)
{
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int nj = R.j1 - R.j0 + 1;
    if (r >= nj * (R.k1 - R.k0 + 1)) return;
    const long q0 = gid(g, 0, R.j0 + r % nj, R.k0 + r / nj);
    int c = 0;
    for (int i = 1 + lane; i <= g.nx - 1; i += 64) c += (mask[q0 + i] == 1 || mask[q0 + i] == -1);
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if (lane == 0) cnt[r] = c;
}

// exclusive scan of the row counts: each block of 256 threads scans 1024
// rows (offsets within the block's rows, block totals) ...
__global__ void __launch_bounds__(256) k_scan_local(const int *cnt, int *loc, int *bsum, int nr)
{
    __shared__ int part[256];
    const int base = blockIdx.x * 1024 + threadIdx.x * 4;
    int v[4], s = 0;
    for (int a = 0; a < 4; ++a) {
        v[a] = (base + a < nr) ? cnt[base + a] : 0;
        s += v[a];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int st = 1; st < 256; st <<= 1) {
        const int t = (threadIdx.x >= st) ? part[threadIdx.x - st] : 0;
        __syncthreads();
        part[threadIdx.x] += t;
        __syncthreads();
    }
    int o = part[threadIdx.x] - s;
    for (int a = 0; a < 4; ++a) {
        if (base + a < nr) loc[base + a] = o;
        o += v[a];
    }
    if (threadIdx.x == 255) bsum[blockIdx.x] = part[255];
}

// ... then one block scans the block totals (total in ctr)
__global__ void __launch_bounds__(1024) k_scan_top(int *bsum, int nb, int *ctr)
{
    __shared__ int part[1024];
    const int per = (nb + 1023) / 1024;
    const int lo = threadIdx.x * per, hi = min(nb, lo + per);
    int s = 0;
    for (int b = lo; b < hi; ++b) s += bsum[b];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int st = 1; st < 1024; st <<= 1) {
        const int t = (threadIdx.x >= st) ? part[threadIdx.x - st] : 0;
        __syncthreads();
        part[threadIdx.x] += t;
        __syncthreads();
    }
    int o = part[threadIdx.x] - s;
    for (int b = lo; b < hi; ++b) {
        const int v = bsum[b];
        bsum[b] = o;
        o += v;
    }
    if (threadIdx.x == 1023) ctr[C_TOTAL] = part[1023];
}

// one wavefront per row writes the row's cells at its offset: 1-based i, j,
// k and the array-index map (1-based list position)
__global__ void __launch_bounds__(256) k_row_write(GGrid g, Rows R, const int8_t *mask, const int *loc,
                                                   const int *boff, int *li, int *lj, int *lk, int *map)
{
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int nj = R.j1 - R.j0 + 1;
    if (r >= nj * (R.k1 - R.k0 + 1)) return;
    const int j = R.j0 + r % nj, k = R.k0 + r / nj;
    const long q0 = gid(g, 0, j, k);
    int n = loc[r] + boff[r >> 10];
    for (int i0 = 1; i0 <= g.nx - 1; i0 += 64) {
        const int i = i0 + lane;
        const bool want = i <= g.nx - 1 && (mask[q0 + i] == 1 || mask[q0 + i] == -1);
        const unsigned long long bal = __ballot(want);
        if (want) {
            const int p = n + __popcll(bal & ((1ull << lane) - 1ull));
            li[p] = i;
            lj[p] = j;
            lk[p] = k;
            map[q0 + i] = p + 1;
        }
        n += __popcll(bal);
    }
}

// ---------------------------------------------------------------------------
// GCM_Calc_VanMatrixDN (GCM_SET_BOUNDARY.F90:607-797).  The inverse of each
// 8x8 matrix is formed column by column from its LU factors (partial
// pivoting, first largest pivot), and each column is folded into the
// coefficient at the evaluation point as soon as it is formed.
__device__ __constant__ int c_incI[8] = {0, 0, 1, 1, 0, 0, 1, 1};
__device__ __constant__ int c_incJ[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__device__ __constant__ int c_incK[8] = {0, 0, 0, 0, 1, 1, 1, 1};

__device__ void lu_coeff(double (&A)[8][8], double xI, double yI, double zI, double *out)
{
    int piv[8];
    for (int c = 0; c < 8; ++c) {
        int p = c;
        double best = fabs(A[c][c]);
        for (int r = c + 1; r < 8; ++r)
            if (fabs(A[r][c]) > best) { best = fabs(A[r][c]); p = r; }
        piv[c] = p;
        if (p != c)
            for (int q = 0; q < 8; ++q) { const double t = A[c][q]; A[c][q] = A[p][q]; A[p][q] = t; }
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
            if (piv[c] != c) { const double t = e[c]; e[c] = e[piv[c]]; e[piv[c]] = t; }
        for (int r = 1; r < 8; ++r)
            for (int q = 0; q < r; ++q) e[r] = e[r] - A[r][q] * e[q];
        for (int r = 7; r >= 0; --r) {
            for (int q = r + 1; q < 8; ++q) e[r] = e[r] - A[r][q] * e[q];
            e[r] = e[r] / A[r][r];
        }
        out[col] = e[0] * xI * yI * zI + e[1] * xI * yI + e[2] * xI * zI + e[3] * yI * zI + e[4] * xI + e[5] * yI +
                   e[6] * zI + e[7];
    }
}

__device__ void van_matrix(const GGrid &g, int iG, int jG, int kG, int iC, int jC, int kC, double xI, double yI,
                           double zI, const double *bi, const double *bn, double *cD, double *cN)
{
    for (int form = 0; form < 2; ++form) {
        double A[8][8];
        for (int r = 0; r < 8; ++r) {
            const int i = iC + c_incI[r], j = jC + c_incJ[r], k = kC + c_incK[r];
            double x1 = g.xc[i], x2 = g.yc[j], x3 = g.zc[k];
            const bool self = (i == iG && j == jG && k == kG);
            if (self) { x1 = bi[0]; x2 = bi[1]; x3 = bi[2]; }
            if (self && form == 1) {
                const double n1 = bn[0], n2 = bn[1], n3 = bn[2];
                A[r][0] = n1 * x2 * x3 + n2 * x1 * x3 + n3 * x1 * x2;
                A[r][1] = n1 * x2 + n2 * x1;
                A[r][2] = n1 * x3 + n3 * x1;
                A[r][3] = n2 * x3 + n3 * x2;
                A[r][4] = n1; A[r][5] = n2; A[r][6] = n3; A[r][7] = 0.0;
            } else {
                A[r][0] = x1 * x2 * x3; A[r][1] = x1 * x2; A[r][2] = x1 * x3; A[r][3] = x2 * x3;
                A[r][4] = x1; A[r][5] = x2; A[r][6] = x3; A[r][7] = 1.0;
            }
        }
        lu_coeff(A, xI, yI, zI, form == 0 ? cD : cN);
    }
}

__device__ __forceinline__ int nint(double v) { return (int)lround(v); }

// host cell along one axis: the last i in [lo, hi] with c(i) <= v < c(i+1)
// (strict, ghost cells) or c(i) <= v <= c(i+1) (dead cells); negative
// indices never match
__device__ __forceinline__ int host_cell(const double *c, int lo, int hi, double v, bool strict)
{
    int r = -1;
    for (int i = max(lo, 0); i <= hi; ++i)
        if (c[i] <= v && (strict ? c[i + 1] > v : c[i + 1] >= v)) r = i;
    return r;
}

struct GhostOut {
    int *gi, *gj, *gk, *ci, *cj, *ck, *el;
    double *bx, *by, *bz, *nx, *ny, *nz, *pl, *ix, *iy, *iz, *cd, *cn, *alpha;
};
struct DeadOut {
    int *di, *dj, *dk, *ci, *cj, *ck, *el;
    double *bx, *by, *bz, *nx, *ny, *nz, *pl, *cd, *cn, *ds;
};

// the ghost-cell loop of GCM_set_internal_boundary (:236-600)
__global__ void __launch_bounds__(256) k_ghost_rel(GGrid g, GBodies B, GMasks m, GhostOut o, int nG, double *dcip,
                                                   unsigned long long *pairs,
                                                   int *ctr)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nG) return;
    const int iG = o.gi[n], jG = o.gj[n], kG = o.gk[n];
    const long q = gid(g, iG, jG, kG);
    const int ib = m.bodynum[q];
    const double xG = g.xc[iG], yG = g.yc[jG], zG = g.zc[kG];
    if (ib < 1 || ib > B.nb) {
        atomicOr(ctr + C_ERR, E_GHOST_BODY);
        return;
    }
    const GBody &b = B.b[ib - 1];
    int scanned = 0;
    const int cv = closest_marker(b, xG, yG, zG, scanned);
    atomicAdd(pairs, (unsigned long long)scanned);
    const Intercept r = intercept_from(b, cv, xG, yG, zG, ctr + C_UNSEL);
    double bi[3] = {r.x, r.y, r.z};
    if (b.dim == 2) bi[2] = zG;
    o.el[n] = r.elem;
    o.bx[n] = bi[0]; o.by[n] = bi[1]; o.bz[n] = bi[2];
    double ds = sqrt(sq(xG - bi[0]) + sq(yG - bi[1]) + sq(zG - bi[2]));
    const double sx = (xG - bi[0]) / ds, sy = (yG - bi[1]) / ds, sz = (zG - bi[2]) / ds;
    if (b.kind == FC_SURFACE_MEMBRANE) {
        bi[0] = bi[0] + B.tkns / 2.0 * sx;
        bi[1] = bi[1] + B.tkns / 2.0 * sy;
        bi[2] = bi[2] + B.tkns / 2.0 * sz;
        ds = sqrt(sq(xG - bi[0]) + sq(yG - bi[1]) + sq(zG - bi[2]));
    }
    const double bn[3] = {-sx, -sy, -sz};
    o.nx[n] = bn[0]; o.ny[n] = bn[1]; o.nz[n] = bn[2];
    o.pl[n] = ds * B.probe;
    const double xI = bi[0] - (bi[0] - xG) * B.probe, yI = bi[1] - (bi[1] - yG) * B.probe,
                 zI = bi[2] - (bi[2] - zG) * B.probe;
    o.ix[n] = xI; o.iy[n] = yI; o.iz[n] = zI;
    dcip[q] = xI;
    dcip[q + g.n] = yI;
    dcip[q + 2 * g.n] = zI;
    const double md = fmax(fmax(g.dx[iG], g.dy[jG]), g.dz[kG]);
    const int iR = 2 * nint(md / g.dx[iG] + 0.5), jR = 2 * nint(md / g.dy[jG] + 0.5), kR = 2 * nint(md / g.dz[kG] + 0.5);
    const int iC = host_cell(g.xc, max(iG - iR, 0), min(iG + iR, g.nx), xI, true);
    const int jC = host_cell(g.yc, max(jG - jR, -1), min(jG + jR, g.ny), yI, true);
    const int kC = host_cell(g.zc, max(kG - kR, -1), min(kG + kR, g.nz), zI, true);
    o.ci[n] = iC; o.cj[n] = jC; o.ck[n] = kC;
    if (iC < 0 || jC < 0 || kC < 0) {
        atomicOr(ctr + C_ERR, E_GHOST_HOST);
        return;
    }
    van_matrix(g, iG, jG, kG, iC, jC, kC, xG, yG, zG, bi, bn, o.cd + 8 * n, o.cn + 8 * n);
}

// the dead-cell loop of GCM_set_deadcells (:979-1481)
__global__ void __launch_bounds__(256) k_dead_rel(GGrid g, GBodies B, GMasks m, DeadOut o, int nD, double *dcip,
                                                   unsigned long long *pairs,
                                                  int *ctr)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nD) return;
    const int iG = o.di[n], jG = o.dj[n], kG = o.dk[n];
    const long q = gid(g, iG, jG, kG);
    const int ib = m.bodynum[q];
    const double xD = g.xc[iG], yD = g.yc[jG], zD = g.zc[kG];
    if (ib < 1 || ib > B.nb) {
        atomicOr(ctr + C_ERR, E_DEAD_BODY);
        o.ds[n] = 0.0;
        return;
    }
    const GBody &b = B.b[ib - 1];
    int scanned = 0;
    const int cv = closest_marker(b, xD, yD, zD, scanned);
    atomicAdd(pairs, (unsigned long long)scanned);
    const Intercept r = intercept_from(b, cv, xD, yD, zD, ctr + C_UNSEL);
    double bi[3] = {r.x, r.y, r.z};
    if (b.dim == 2) bi[2] = zD;
    o.el[n] = r.elem;
    o.bx[n] = bi[0]; o.by[n] = bi[1]; o.bz[n] = bi[2];
    double ds = sqrt(sq(xD - bi[0]) + sq(yD - bi[1]) + sq(zD - bi[2]));
    double sx = (bi[0] - xD) / ds, sy = (bi[1] - yD) / ds, sz = (bi[2] - zD) / ds;
    const double bn[3] = {-sx, -sy, -sz};
    o.nx[n] = bn[0]; o.ny[n] = bn[1]; o.nz[n] = bn[2];
    const int ny = g.ny, nz = g.nz;
    if (b.kind == FC_SURFACE_MEMBRANE) {
        const double tk = B.tkns;
        sx = -sx; sy = -sy; sz = -sz;
        if (ds >= tk / 2.0) {
            const double f = ds + (double)0.1f * (ds - tk / 2.0);
            bi[0] = bi[0] + f * sx; bi[1] = bi[1] + f * sy; bi[2] = bi[2] + f * sz;
        } else {
            bi[0] = bi[0] + tk / 2.0 * sx; bi[1] = bi[1] + tk / 2.0 * sy; bi[2] = bi[2] + tk / 2.0 * sz;
        }
        double zI = zD + (bi[2] - zD) * B.probeD;
        // the y loop runs for the first dead cell of the list only: later
        // cells test the image point the previous cell left inside the range
        // (each step halves the distance to the cell centre, which lies inside
        // the range, so the loops end; the step caps only bound them)
        if (n == 0) {
            double yI;
            int it = 0;
            do {
                const double d1 = sqrt(sq(xD - bi[0]) + sq(yD - bi[1]) + sq(zD - bi[2]));
                const double t = -0.5 * d1;
                bi[0] = bi[0] + t * sx; bi[1] = bi[1] + t * sy; bi[2] = bi[2] + t * sz;
                yI = yD + (bi[1] - yD) * B.probeD;
            } while ((yI < g.yc[0] || yI > g.yc[ny + 1]) && ++it < 2200);
        }
        for (int it = 0; (zI < g.zc[0] || zI > g.zc[nz + 1]) && it < 2200; ++it) {
            const double d1 = sqrt(sq(xD - bi[0]) + sq(yD - bi[1]) + sq(zD - bi[2]));
            const double t = -0.5 * d1;
            bi[0] = bi[0] + t * sx; bi[1] = bi[1] + t * sy; bi[2] = bi[2] + t * sz;
            zI = zD + (bi[2] - zD) * B.probeD;
        }
        ds = sqrt(sq(xD - bi[0]) + sq(yD - bi[1]) + sq(zD - bi[2]));
    }
    o.ds[n] = ds;
    const double xI = xD + (bi[0] - xD) * B.probeD;
    double yI = yD + (bi[1] - yD) * B.probeD, zI = zD + (bi[2] - zD) * B.probeD;
    double pl = ds * B.probeD;
    if (yI < g.yc[0] || yI > g.yc[ny + 1]) {
        yI = (fabs(yI - g.yc[0]) < fabs(yI - g.yc[ny + 1])) ? g.yc[0] : g.yc[ny + 1];
        pl = sqrt(sq(xD - xI) + sq(yD - yI) + sq(zD - zI));
    }
    if (zI < g.zc[0] || zI > g.zc[nz + 1]) {
        zI = (fabs(zI - g.zc[0]) < fabs(zI - g.zc[nz + 1])) ? g.zc[0] : g.zc[nz + 1];
        pl = sqrt(sq(xD - xI) + sq(yD - yI) + sq(zD - zI));
    }
    o.pl[n] = pl;
    dcip[q] = xI;
    dcip[q + g.n] = yI;
    dcip[q + 2 * g.n] = zI;
    const double md = fmax(fmax(g.dx[iG], g.dy[jG]), g.dz[kG]);
    const int iR = 2 * nint(md / g.dx[iG] + 0.5), jR = 2 * nint(md / g.dy[jG] + 0.5), kR = 2 * nint(md / g.dz[kG] + 0.5);
    const int iC = host_cell(g.xc, max(iG - iR, 0), min(iG + iR, g.nx), xI, false);
    const int jC = host_cell(g.yc, max(jG - jR, -1), min(jG + jR, ny), yI, false);
    const int kC = host_cell(g.zc, max(kG - kR, -1), min(kG + kR, nz), zI, false);
    o.ci[n] = iC; o.cj[n] = jC; o.ck[n] = kC;
    if (iC < 0 || jC < 0 || kC < 0) {
        atomicOr(ctr + C_ERR, E_DEAD_HOST);
        return;
    }
    // ps1 / ps2 (:1362-1437)
    bool pure = false;
    for (int rr = 0; rr < 8; ++rr) {
        const long c = gid(g, iC + c_incI[rr], jC + c_incJ[rr], kC + c_incK[rr]);
        if (m.iblank[c] == 1 && m.dead[c] != -1) pure = true;
    }
    if (pure) {
        m.ps1[q] = 10;
        const int nx = g.nx;
        if (iG - 1 >= 0 && iG - 1 <= nx && m.ghost[q - 1] == 1) m.ps2[q - 1] = 10;
        if (iG + 1 >= 0 && iG + 1 <= nx && m.ghost[q + 1] == 1) m.ps2[q + 1] = 10;
        const long sy = nx + 2, sz = (long)(nx + 2) * (ny + 2);
        if (jG - 1 >= 0 && jG - 1 <= ny && m.ghost[q - sy] == 1) m.ps2[q - sy] = 10;
        if (jG + 1 >= 0 && jG + 1 <= ny && m.ghost[q + sy] == 1) m.ps2[q + sy] = 10;
        if (kG - 1 >= 1 && kG - 1 <= nz - 1 && m.ghost[q - sz] == 1) m.ps2[q - sz] = 10;
        if (kG + 1 >= 1 && kG + 1 <= nz - 1 && m.ghost[q + sz] == 1) m.ps2[q + sz] = 10;
    }
    van_matrix(g, iG, jG, kG, iC, jC, kC, xI, yI, zI, bi, bn, o.cd + 8 * n, o.cn + 8 * n);
}

// alphaGhost = 1 - sqrt(sum of (ds/dc)^2 over the dead face neighbours), the
// contributions added in dead-list order (:1155-1213, 1483-1489)
__global__ void k_alpha(GGrid g, GhostOut o, int nG, const int *dmap, const double *dds)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nG) return;
    const int i = o.gi[n], j = o.gj[n], k = o.gk[n];
    const long q = gid(g, i, j, k);
    const long sy = g.nx + 2, sz = (long)(g.nx + 2) * (g.ny + 2);
    const long nb[6] = {q - sz, q - sy, q - 1, q + 1, q + sy, q + sz};
    const double dc[6] = {g.dzc[k], g.dyc[j], g.dxc[i], g.dxc[i + 1], g.dyc[j + 1], g.dzc[k + 1]};
    double a = 0.0;
    for (int t = 0; t < 6; ++t) {
        const int d = dmap[nb[t]];
        if (d <= 0) continue;
        const double tt = dds[d - 1] / dc[t];
        a = a + tt * tt;
    }
    o.alpha[n] = 1.0 - sqrt(a);
}

// clears the dead-list map where the previous refresh left entries
__global__ void k_clear_map(const int *li, const int *lj, const int *lk, int cnt, GGrid g, int *map)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= cnt) return;
    map[gid(g, li[n], lj[n], lk[n])] = 0;
}


// ---------------------------------------------------------------------------
// the image point of every dead cell, in list order (x, y, z per cell): what
// the pressure update reads of xyzDCIP
__global__ void k_dead_ip(GGrid g, const int *di, const int *dj, const int *dk, int nD, const double *dcip,
                          double *out)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nD) return;
    const long q = gid(g, di[n], dj[n], dk[n]);
    out[3 * (long)n] = dcip[q];
    out[3 * (long)n + 1] = dcip[q + g.n];
    out[3 * (long)n + 2] = dcip[q + 2 * g.n];
}

// GCM_SetBodyInterceptValues (GCM_BODYINTERCEPT_VALUES.F90:22-178) with
// GCM_calc_BIVelocity_Unstruc (:198-388): one thread per ghost or dead cell;
// the 4x4 matrix of the closest element's markers and normal is inverted as
// the oracle does (LU, partial pivoting, column solves) and its rows fold the
// marker velocities into the linear fit evaluated at the intercept.
__device__ void invert4(double (&A)[4][4], double (&X)[4][4])
{
    int piv[4];
    for (int c = 0; c < 4; ++c) {
        int p = c;
        double best = fabs(A[c][c]);
        for (int r = c + 1; r < 4; ++r)
            if (fabs(A[r][c]) > best) { best = fabs(A[r][c]); p = r; }
        piv[c] = p;
        if (p != c)
            for (int q = 0; q < 4; ++q) { const double t = A[c][q]; A[c][q] = A[p][q]; A[p][q] = t; }
        const double inv = 1.0 / A[c][c];
        for (int r = c + 1; r < 4; ++r) {
            const double l = A[r][c] * inv;
            A[r][c] = l;
            for (int q = c + 1; q < 4; ++q) A[r][q] = A[r][q] - l * A[c][q];
        }
    }
    for (int col = 0; col < 4; ++col) {
        double e[4];
        for (int r = 0; r < 4; ++r) e[r] = (r == col) ? 1.0 : 0.0;
        for (int c = 0; c < 4; ++c)
            if (piv[c] != c) { const double t = e[c]; e[c] = e[piv[c]]; e[piv[c]] = t; }
        for (int r = 1; r < 4; ++r)
            for (int q = 0; q < r; ++q) e[r] = e[r] - A[r][q] * e[q];
        for (int r = 3; r >= 0; --r) {
            for (int q = r + 1; q < 4; ++q) e[r] = e[r] - A[r][q] * e[q];
            e[r] = e[r] / A[r][r];
        }
        for (int r = 0; r < 4; ++r) X[r][col] = e[r];
    }
}

struct BiOut {
    double *u, *v, *w, *dpdn;
};

__global__ void k_bi_values(GGrid g, GBodies B, GMasks m, const int *li, const int *lj, const int *lk,
                            const int *el, const double *bx, const double *by, const double *bz, const double *nx,
                            const double *ny, const double *nz, int cnt, double dt, int ntime, BiOut o, int *ctr)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= cnt) return;
    const int ib = m.bodynum[gid(g, li[n], lj[n], lk[n])];
    if (ib < 1 || ib > B.nb || el[n] < 1) {
        atomicOr(ctr + C_ERR, E_DEAD_BODY);
        return;
    }
    const GBody &b = B.b[ib - 1];
    const int e = el[n] - 1;
    double A[4][4], X[4][4];
    int node[3];
    for (int q = 0; q < 3; ++q) {
        node[q] = b.tri[3 * e + q] - 1;
        A[q][0] = b.x[node[q]]; A[q][1] = b.y[node[q]]; A[q][2] = b.z[node[q]]; A[q][3] = 1.0;
    }
    A[3][0] = b.enx[e]; A[3][1] = b.eny[e]; A[3][2] = b.enz[e]; A[3][3] = 0.0;
    invert4(A, X);
    double r[6];
    for (int c = 0; c < 6; ++c) {
        const double *src = b.mvel + (size_t)c * b.npt;
        const double rhs[4] = {src[node[0]], src[node[1]], src[node[2]], 0.0};
        double cX = 0.0, cY = 0.0, cZ = 0.0, cC = 0.0;
        for (int i = 0; i < 4; ++i) {
            cX = cX + X[0][i] * rhs[i];
            cY = cY + X[1][i] * rhs[i];
            cZ = cZ + X[2][i] * rhs[i];
            cC = cC + X[3][i] * rhs[i];
        }
        r[c] = cX * bx[n] + cY * by[n] + cZ * bz[n] + cC;
    }
    o.u[n] = r[0];
    o.v[n] = r[1];
    o.w[n] = r[2];
    if (ntime > 1) {
        const double d = (r[0] - r[3]) / dt * nx[n] + (r[1] - r[4]) / dt * ny[n] + (r[2] - r[5]) / dt * nz[n];
        o.dpdn[n] = -d;
    } else {
        o.dpdn[n] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// compute_marker_stress (UTIL_DRAG_LIFT.F90:29-342) with marker_interpolation
// (:375-686), sign -1.  One thread per surface element: the pressure at the
// centroid (inverse-distance or trilinear from the host cell) and the
// velocity gradient along the normal from the nearest dead cell's image
// point.  The nearest dead cell: the reference scans the dead list (the last
// of equal distances wins); here rings of grid cells around the centroid's
// cell are scanned through the dead-list map until the distance to every
// unvisited cell centre exceeds the best one, keeping the higher list index
// among equal distances, which is the same cell.
struct Flow {
    const double *p, *u, *v, *w;
};
// per element, rows of FC_ROWPAD(nel) doubles: 0..5 the pressure and viscous
// force xp yp zp xs ys zs, 6..11 the moment products of the totals with the
// subtracted one negated: yE(zs+zp), -zE(ys+yp), zE(xs+xp), -xE(zs+zp),
// xE(ys+yp), -yE(xs+xp); 12 zeros.  The padding stays zero.
struct ElemForce {
    double *xp, *yp, *zp, *xs, *ys, *zs;
    double *mom;                              // rows 6..11
};

__device__ __forceinline__ int axis_cell(const double *c, int n, double v)
{
    // the cell i in [0, n] with c(i) <= v < c(i+1) (clamped): the ring centre
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (c[mid] <= v) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ int nearest_dead(const GGrid &g, const int *dmap, double xE, double yE, double zE)
{
    const int c0[3] = {axis_cell(g.xc, g.nx + 1, xE), axis_cell(g.yc, g.ny + 1, yE), axis_cell(g.zc, g.nz + 1, zE)};
    const int hi[3] = {g.nx + 1, g.ny + 1, g.nz + 1};
    const double *cc[3] = {g.xc, g.yc, g.zc};
    const double P[3] = {xE, yE, zE};
    double best = 0.0;
    int bn = 0;
    for (int r = 0;; ++r) {
        for (int dk = -r; dk <= r; ++dk) {
            const int k = c0[2] + dk;
            if (k < 0 || k > hi[2]) continue;
            for (int dj = -r; dj <= r; ++dj) {
                const int j = c0[1] + dj;
                if (j < 0 || j > hi[1]) continue;
                const bool face = (dk == -r || dk == r || dj == -r || dj == r);
                const int istep = face ? 1 : 2 * r;
                for (int di = -r; di <= r; di += (istep > 0 ? istep : 1)) {
                    const int i = c0[0] + di;
                    if (i < 0 || i > hi[0]) continue;
                    const int d1 = dmap[gid(g, i, j, k)];
                    if (d1 <= 0) continue;
                    const double d = sq(g.xc[i] - xE) + sq(g.yc[j] - yE) + sq(g.zc[k] - zE);
                    if (bn == 0 || d < best || (d == best && d1 > bn)) {
                        best = d;
                        bn = d1;
                    }
                }
            }
        }
        // every cell outside the box of rings 0..r differs along one axis by
        // more than r: it lies at least the gap to the next centre away
        double lim = 1.0e300;
        bool more = false;
        for (int a = 0; a < 3; ++a) {
            if (c0[a] - r - 1 >= 0) { lim = fmin(lim, P[a] - cc[a][c0[a] - r - 1]); more = true; }
            if (c0[a] + r + 1 <= hi[a]) { lim = fmin(lim, cc[a][c0[a] + r + 1] - P[a]); more = true; }
        }
        if (!more) break;
        if (bn > 0 && lim > 0.0 && best < lim * lim * (1.0 - 1.0e-12)) break;
    }
    return bn - 1;
}

__global__ void k_marker_stress(GGrid g, GBody b, int ibody, double tkns, GMasks m, Flow f, const int *dmap,
                                DeadOut d, BiOut bd, double reinv, ElemForce out)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= b.nel) return;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const double xE0 = b.ecx[e], yE0 = b.ecy[e], zE0 = b.ecz[e];
    const double xN = b.enx[e], yN = b.eny[e], zN = b.enz[e];
    const double ar = b.area[e];
    double dU[3] = {0.0, 0.0, 0.0}, pE = 0.0;
    if (zE0 <= g.z[nz] && zE0 > g.z[1] && yE0 <= g.y[ny] && yE0 > g.y[1] && xE0 <= g.xc[nx - 1] && xE0 > g.xc[1]) {
        const double sign = -1.0;
        double xE = xE0, yE = yE0, zE = zE0;
        if (b.kind == FC_SURFACE_MEMBRANE) {
            xE = xE0 + sign * xN * tkns / 2.0;
            yE = yE0 + sign * yN * tkns / 2.0;
            zE = zE0 + sign * zN * tkns / 2.0;
        }
        const int nG = nearest_dead(g, dmap, xE, yE, zE);
        if (nG >= 0) {
            const int iG = d.di[nG], jG = d.dj[nG], kG = d.dk[nG];
            int ii = -1, jj = -1, kk = -1;
            for (int i = max(1, iG - 20); i <= min(nx - 1, iG + 20); ++i)
                if ((g.xc[i] - xE) * (g.xc[i + 1] - xE) <= 0.0) ii = i;
            for (int j = max(1, jG - 20); j <= min(ny - 1, jG + 20); ++j)
                if ((g.yc[j] - yE) * (g.yc[j + 1] - yE) <= 0.0) jj = j;
            for (int k = max(1, kG - 20); k <= min(nz - 1, kG + 20); ++k)
                if ((g.zc[k] - zE) * (g.zc[k + 1] - zE) <= 0.0) kk = k;
            if (ii != -1 && jj != -1 && kk != -1) {
                double pinv = 0.0, dint = 0.0;
                bool pure = false;
                for (int k = 0; k <= 1; ++k)
                    for (int j = 0; j <= 1; ++j)
                        for (int i = 0; i <= 1; ++i) {
                            const long q = gid(g, ii + i, jj + j, kk + k);
                            const double dd = sq(g.xc[ii + i] - xE) + sq(g.yc[jj + j] - yE) + sq(g.zc[kk + k] - zE);
                            if (m.iblank[q] == 0 || m.dead[q] == -1) {
                                dint = dint + (1.0 / dd);
                                pinv = pinv + (1.0 / dd) * f.p[q];
                            }
                            if (m.iblank[q] == 1 && m.dead[q] == 0) pure = true;
                        }
                if (pure && !(dint < 0.000001)) {
                    pE = pinv / dint;
                } else {
                    const double aX = (xE - g.xc[ii]) * (1.0 / g.dxc[ii + 1]);
                    const double aY = (yE - g.yc[jj]) * (1.0 / g.dyc[jj + 1]);
                    const double aZ = (zE - g.zc[kk]) * (1.0 / g.dzc[kk + 1]);
                    const double *p = f.p;
                    pE = p[gid(g, ii, jj, kk)] * (1.0 - aX) * (1.0 - aY) * (1.0 - aZ) +
                         p[gid(g, ii + 1, jj, kk)] * aX * (1.0 - aY) * (1.0 - aZ) +
                         p[gid(g, ii, jj + 1, kk)] * (1.0 - aX) * aY * (1.0 - aZ) +
                         p[gid(g, ii, jj, kk + 1)] * (1.0 - aX) * (1.0 - aY) * aZ +
                         p[gid(g, ii + 1, jj + 1, kk)] * aX * aY * (1.0 - aZ) +
                         p[gid(g, ii + 1, jj, kk + 1)] * aX * (1.0 - aY) * aZ +
                         p[gid(g, ii, jj + 1, kk + 1)] * (1.0 - aX) * aY * aZ +
                         p[gid(g, ii + 1, jj + 1, kk + 1)] * aX * aY * aZ;
                }
                double uI = 0.0, vI = 0.0, wI = 0.0;
                for (int r = 0; r < 8; ++r) {
                    const int i = d.ci[nG] + c_incI[r], j = d.cj[nG] + c_incJ[r], k = d.ck[nG] + c_incK[r];
                    const double c = d.cd[8 * nG + r];
                    if (i != iG || j != jG || k != kG) {
                        const long q = gid(g, i, j, k);
                        uI = uI + c * f.u[q];
                        vI = vI + c * f.v[q];
                        wI = wI + c * f.w[q];
                    } else {
                        uI = uI + c * bd.u[nG];
                        vI = vI + c * bd.v[nG];
                        wI = wI + c * bd.w[nG];
                    }
                }
                const long qG = gid(g, iG, jG, kG);
                dU[0] = (uI - f.u[qG]) / d.pl[nG] * (-sign);
                dU[1] = (vI - f.v[qG]) / d.pl[nG] * (-sign);
                dU[2] = (wI - f.w[qG]) / d.pl[nG] * (-sign);
            }
        }
    }
    dU[0] = dU[0] - 0.0;
    dU[1] = dU[1] - 0.0;
    dU[2] = dU[2] - 0.0;
    pE = pE - 0.0;
    const double fn = dU[0] * xN + dU[1] * yN + dU[2] * zN;
    dU[0] = dU[0] - fn * xN;
    dU[1] = dU[1] - fn * yN;
    dU[2] = dU[2] - fn * zN;
    out.xp[e] = pE * xN * ar;
    out.yp[e] = pE * yN * ar;
    out.zp[e] = pE * zN * ar;
    out.xs[e] = dU[0] * reinv * ar;
    out.ys[e] = dU[1] * reinv * ar;
    out.zs[e] = dU[2] * reinv * ar;
    const double xpf = out.xp[e], ypf = out.yp[e], zpf = out.zp[e];
    const double xsf = out.xs[e], ysf = out.ys[e], zsf = out.zs[e];
    const double xC = b.ecx[e], yC = b.ecy[e], zC = b.ecz[e];
    const size_t r = (size_t)FC_ROWPAD(b.nel);
    out.mom[e] = yC * (zsf + zpf);
    out.mom[r + e] = -(zC * (ysf + ypf));
    out.mom[2 * r + e] = zC * (xsf + xpf);
    out.mom[3 * r + e] = -(xC * (zsf + zpf));
    out.mom[4 * r + e] = xC * (ysf + ypf);
    out.mom[5 * r + e] = -(yC * (xsf + xpf));
}

// the body totals, summed in element order as the reference does: lanes
// 0..8 of one wave each run one of the nine sequential chains
// c = (c + A) + B over the element rows, where B is the second term negated
// for the moment chains (x - b == x + (-b) exactly) and the zero row for the
// single-term chains (c is never -0.0, so adding +0.0 changes nothing; the
// zero padding past nel changes nothing either).  Batches of 16 elements
// are loaded a batch ahead so that the chain waits on its adds only.
__device__ __forceinline__ void load16(const double *A, const double *B, int e, double (&a)[16], double (&b)[16])
{
#pragma unroll
    for (int j = 0; j < 16; j += 2) {
        const double2 x = *reinterpret_cast<const double2 *>(A + e + j);
        const double2 y = *reinterpret_cast<const double2 *>(B + e + j);
        a[j] = x.x; a[j + 1] = x.y;
        b[j] = y.x; b[j + 1] = y.y;
    }
}

__device__ __forceinline__ double chain16(double c, const double (&a)[16], const double (&b)[16])
{
#pragma unroll
    for (int j = 0; j < 16; ++j) c = (c + a[j]) + b[j];
    return c;
}

__global__ void __launch_bounds__(64) k_force_totals(int nel, const double *rows, double *totals)
{
    const int l = threadIdx.x;
    if (l >= 9) return;
    const int arow[9] = {0, 1, 2, 6, 8, 10, 3, 4, 5};
    const int brow[9] = {3, 4, 5, 7, 9, 11, 12, 12, 12};
    const size_t r = (size_t)FC_ROWPAD(nel);
    const double *A = rows + arow[l] * r, *B = rows + brow[l] * r;
    const int n = (int)r;
    double c = 0.0;
    double a0[16], b0[16], a1[16], b1[16];
    load16(A, B, 0, a0, b0);
    for (int e = 0; e < n; e += 32) {
        if (e + 16 < n) load16(A, B, e + 16, a1, b1);
        c = chain16(c, a0, b0);
        if (e + 16 >= n) break;
        if (e + 32 < n) load16(A, B, e + 32, a0, b0);
        c = chain16(c, a1, b1);
    }
    totals[l] = c;
}

// per marker: a third of each incident element's area, shear and force, in
// element order; the stress
struct MarkerForces {
    double *fx, *fy, *fz, *sx, *sy, *sz, *ma, *tx, *ty, *tz;
};

__global__ void k_marker_forces(GBody b, ElemForce fe, MarkerForces o)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= b.npt) return;
    double ma = 0.0, sx = 0.0, sy = 0.0, sz = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
    for (int t = b.voff[v]; t < b.voff[v + 1]; ++t) {
        const int e = b.vel[t];
        ma = ma + b.area[e] / 3.0;
        sx = sx + fe.xs[e] / 3.0;
        sy = sy + fe.ys[e] / 3.0;
        sz = sz + fe.zs[e] / 3.0;
        fx = fx + (fe.xp[e] + fe.xs[e]) / 3.0;
        fy = fy + (fe.yp[e] + fe.ys[e]) / 3.0;
        fz = fz + (fe.zp[e] + fe.zs[e]) / 3.0;
    }
    o.ma[v] = ma; o.sx[v] = sx; o.sy[v] = sy; o.sz[v] = sz; o.fx[v] = fx; o.fy[v] = fy; o.fz[v] = fz;
    o.tx[v] = fx / ma; o.ty[v] = fy / ma; o.tz[v] = fz / ma;
}

}  // namespace

// ---------------------------------------------------------------------------
// host side

struct fc_geom {
    int dev = 0;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr, ev[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t es[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};   // BI values, marker stress
    bool es_bi = false, es_stress = false;
    int nx = 0, ny = 0, nz = 0;
    long n = 0;
    fc_geom_params p{};
    std::vector<void *> owned;
    double *grid = nullptr;               // 12 axis arrays of length m
    int m = 0;
    struct Body {
        int npt = 0, nel = 0;
        double *buf = nullptr;            // x, y, z, vnx, vny, vnz (npt each); 7 element arrays
        int *tri = nullptr, *voff = nullptr, *vel = nullptr;
        std::vector<int> tri_h;           // host copy of the connectivity
        int *bstart = nullptr, *bmark = nullptr;   // marker bins (bstart: capb + 1)
        double *bxyz = nullptr;                    // binned marker coordinates
        double *mvel = nullptr;                    // u, v, w, uOld, vOld, wOld (npt each)
        double *ef = nullptr;                      // per-element forces (6 nel)
        double *mf = nullptr;                      // per-marker forces (10 npt)
        size_t capb = 0;
        double bo[3] = {0, 0, 0}, bh = 1.0;
        int bd[3] = {1, 1, 1};
    } body[FC_GEOM_MAX_BODIES];
    int8_t *m8 = nullptr;                 // 8 full-grid int8 arrays
    int *gidx = nullptr, *didx = nullptr;
    double *dcip = nullptr;
    int16_t *sA = nullptr, *sP = nullptr, *sN = nullptr;
    int *qlist = nullptr, *rowcnt = nullptr, *rowoff = nullptr, *bsum = nullptr, *ctr = nullptr;
    int ctr_init[C_N] = {0};              // host copy the counters start from
    unsigned long long *pairs = nullptr;  // markers scanned by the searches
    double *flow = nullptr;               // p, u, v, w uploaded for the marker stresses
    double *totals = nullptr;             // 9 per body
    bool bi_ready = false;
    int capG = 0, capD = 0;
    int *gint = nullptr;                  // 7 * capG
    double *gdbl = nullptr;               // (13 + 16 + 4) * capG
    int *dint = nullptr;                  // 7 * capD
    double *ddbl = nullptr;               // (8 + 16 + 4) * capD
    int nG = 0, nD = 0;
    bool fresh = false;
};

namespace {

int herr(hipError_t e, const char *what)
{
    return fc::set_error(-3, std::string("HIP error: ") + hipGetErrorString(e) + " at " + what);
}
#define GCHK(x)                                  \
    do {                                         \
        hipError_t e_ = (x);                     \
        if (e_ != hipSuccess) return herr(e_, #x); \
    } while (0)

template <typename T> int dalloc(fc_geom *h, T **p, size_t count)
{
    void *q = nullptr;
    GCHK(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
    h->owned.push_back(q);
    *p = (T *)q;
    return 0;
}

void dfree(fc_geom *h, void *p)
{
    if (!p) return;
    auto it = std::find(h->owned.begin(), h->owned.end(), p);
    if (it != h->owned.end()) {
        (void)hipFree(p);
        h->owned.erase(it);
    }
}

GGrid ggrid(const fc_geom *h)
{
    GGrid g;
    g.nx = h->nx; g.ny = h->ny; g.nz = h->nz; g.n = h->n;
    const double *a = h->grid;
    const int m = h->m;
    g.x = a; g.y = a + m; g.z = a + 2 * m; g.xc = a + 3 * m; g.yc = a + 4 * m; g.zc = a + 5 * m;
    g.dx = a + 6 * m; g.dy = a + 7 * m; g.dz = a + 8 * m; g.dxc = a + 9 * m; g.dyc = a + 10 * m; g.dzc = a + 11 * m;
    return g;
}

GBody gbody(const fc_geom *h, int b)
{
    const fc_geom::Body &B = h->body[b];
    GBody r;
    r.npt = B.npt; r.nel = B.nel;
    r.kind = h->p.unstruc_surface_type[b];
    r.dim = h->p.body_dim[b];
    double *a = B.buf;
    r.x = a; r.y = a + B.npt; r.z = a + 2 * B.npt;
    r.vnx = a + 3 * B.npt; r.vny = a + 4 * B.npt; r.vnz = a + 5 * B.npt;
    double *e = a + 6 * B.npt;
    r.enx = e; r.eny = e + B.nel; r.enz = e + 2 * B.nel; r.area = e + 3 * B.nel;
    r.ecx = e + 4 * B.nel; r.ecy = e + 5 * B.nel; r.ecz = e + 6 * B.nel;
    r.tri = B.tri; r.voff = B.voff; r.vel = B.vel;
    for (int a = 0; a < 3; ++a) { r.bo[a] = B.bo[a]; r.bd[a] = B.bd[a]; }
    r.bh = B.bh;
    r.bih = 1.0 / B.bh;
    r.bstart = B.bstart;
    r.bmark = B.bmark;
    r.bxyz = B.bxyz;
    r.mvel = B.mvel;
    return r;
}

// Marker bins of a body at its current positions (host arrays): cubic bins
// of 2.5 mean marker spacings (sqrt of the surface area per marker), at most
// 512 per axis and 2^22 in all, over the markers' bounding box; a counting
// sort in marker order.  The device search only needs every marker to lie
// in (or within rounding of) its bin.
int rebin(fc_geom *h, fc_geom::Body &B, const double *x, const double *y, const double *z)
{
    double area = 0.0;
    for (int m = 0; m < B.nel; ++m) {
        const int a = B.tri_h[3 * m] - 1, b = B.tri_h[3 * m + 1] - 1, c = B.tri_h[3 * m + 2] - 1;
        const double ux = x[b] - x[a], uy = y[b] - y[a], uz = z[b] - z[a];
        const double vx = x[c] - x[a], vy = y[c] - y[a], vz = z[c] - z[a];
        const double cx = uy * vz - uz * vy, cy = uz * vx - ux * vz, cz = ux * vy - uy * vx;
        area += 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
    }
    double lo[3] = {x[0], y[0], z[0]}, hi[3] = {x[0], y[0], z[0]};
    for (int m = 1; m < B.npt; ++m) {
        const double p[3] = {x[m], y[m], z[m]};
        for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], p[a]); hi[a] = std::max(hi[a], p[a]); }
    }
    double ext = std::max(std::max(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
    if (!(ext > 0.0)) ext = 1.0;
    double bh = 2.5 * std::sqrt(area / B.npt);
    if (!(bh > ext * 1e-6)) bh = ext / 64.0;
    bh = std::max(bh, ext / 512.0);
    for (;;) {
        size_t tot = 1;
        for (int a = 0; a < 3; ++a) {
            B.bd[a] = (int)std::floor((hi[a] - lo[a]) / bh) + 1;
            tot *= (size_t)B.bd[a];
        }
        if (tot <= (size_t(1) << 22)) break;
        bh *= 1.25;
    }
    B.bh = bh;
    for (int a = 0; a < 3; ++a) B.bo[a] = lo[a];
    const double ih = 1.0 / bh;
    const size_t nb = (size_t)B.bd[0] * B.bd[1] * B.bd[2];
    std::vector<int> bin(B.npt), start(nb + 1, 0), mark(B.npt);
    for (int m = 0; m < B.npt; ++m) {
        const double p[3] = {x[m], y[m], z[m]};
        int q[3];
        for (int a = 0; a < 3; ++a)
            q[a] = std::min(std::max((int)std::floor((p[a] - B.bo[a]) * ih), 0), B.bd[a] - 1);
        bin[m] = q[0] + B.bd[0] * (q[1] + B.bd[1] * q[2]);
        start[bin[m] + 1]++;
    }
    for (size_t b = 0; b < nb; ++b) start[b + 1] += start[b];
    std::vector<int> fill(start.begin(), start.end() - 1);
    for (int m = 0; m < B.npt; ++m) mark[fill[bin[m]]++] = m;
    if (nb > B.capb) {
        dfree(h, B.bstart);
        B.bstart = nullptr;
        B.capb = std::max(nb, 2 * B.capb);
        if (int rc = dalloc(h, &B.bstart, B.capb + 1)) return rc;
    }
    if (!B.bmark)
        if (int rc = dalloc(h, &B.bmark, B.npt)) return rc;
    if (!B.bxyz)
        if (int rc = dalloc(h, &B.bxyz, 3 * (size_t)B.npt)) return rc;
    std::vector<double> xyz(3 * (size_t)B.npt);
    for (int t = 0; t < B.npt; ++t) {
        xyz[3 * t] = x[mark[t]];
        xyz[3 * t + 1] = y[mark[t]];
        xyz[3 * t + 2] = z[mark[t]];
    }
    GCHK(hipMemcpyAsync(B.bstart, start.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice, h->st));
    GCHK(hipMemcpyAsync(B.bmark, mark.data(), sizeof(int) * B.npt, hipMemcpyHostToDevice, h->st));
    GCHK(hipMemcpyAsync(B.bxyz, xyz.data(), sizeof(double) * xyz.size(), hipMemcpyHostToDevice, h->st));
    GCHK(hipStreamSynchronize(h->st));
    return 0;
}

GMasks gmasks(const fc_geom *h)
{
    GMasks m;
    int8_t *a = h->m8;
    const long n = h->n;
    m.iblank = a; m.ghost = a + n; m.dead = a + 2 * n; m.bodynum = a + 3 * n;
    m.undec = a + 4 * n; m.tmp = a + 5 * n; m.ps1 = a + 6 * n; m.ps2 = a + 7 * n;
    return m;
}

GhostOut gout(const fc_geom *h)
{
    GhostOut o;
    const int c = h->capG;
    int *a = h->gint;
    o.gi = a; o.gj = a + c; o.gk = a + 2 * c; o.ci = a + 3 * c; o.cj = a + 4 * c; o.ck = a + 5 * c; o.el = a + 6 * c;
    double *d = h->gdbl;
    o.bx = d; o.by = d + c; o.bz = d + 2 * c; o.nx = d + 3 * c; o.ny = d + 4 * c; o.nz = d + 5 * c;
    o.pl = d + 6 * c; o.ix = d + 7 * c; o.iy = d + 8 * c; o.iz = d + 9 * c; o.alpha = d + 10 * c;
    o.cd = d + 13 * c; o.cn = d + 21 * c;
    return o;
}

DeadOut dout(const fc_geom *h)
{
    DeadOut o;
    const int c = h->capD;
    int *a = h->dint;
    o.di = a; o.dj = a + c; o.dk = a + 2 * c; o.ci = a + 3 * c; o.cj = a + 4 * c; o.ck = a + 5 * c; o.el = a + 6 * c;
    double *d = h->ddbl;
    o.bx = d; o.by = d + c; o.bz = d + 2 * c; o.nx = d + 3 * c; o.ny = d + 4 * c; o.nz = d + 5 * c;
    o.pl = d + 6 * c; o.ds = d + 7 * c; o.cd = d + 8 * c; o.cn = d + 16 * c;
    return o;
}

inline unsigned nblk(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

int read_ctr(fc_geom *h, int *hc)
{
    GCHK(hipMemcpyAsync(hc, h->ctr, sizeof(int) * C_N, hipMemcpyDeviceToHost, h->st));
    GCHK(hipStreamSynchronize(h->st));
    return 0;
}

int grow_lists(fc_geom *h, int nG, int nD)
{
    if (nG > h->capG) {
        dfree(h, h->gint);
        dfree(h, h->gdbl);
        h->capG = std::max(nG, 2 * h->capG);
        if (int rc = dalloc(h, &h->gint, 7 * (size_t)h->capG)) return rc;
        if (int rc = dalloc(h, &h->gdbl, 33 * (size_t)h->capG)) return rc;
    }
    if (nD > h->capD) {
        dfree(h, h->dint);
        dfree(h, h->ddbl);
        h->capD = std::max(nD, 2 * h->capD);
        if (int rc = dalloc(h, &h->dint, 7 * (size_t)h->capD)) return rc;
        if (int rc = dalloc(h, &h->ddbl, 28 * (size_t)h->capD)) return rc;
    }
    return 0;
}

// compacts the cells of the rows R with |mask| == 1 into three lists (and
// the map) after growing them to the count: count, scan, read the total,
// grow, write
int compact(fc_geom *h, const Rows &R, const int8_t *mask, bool ghost, int *map, int *total)
{
    const GGrid g = ggrid(h);
    const int nr = R.count();
    const int nb = (nr + 1023) / 1024;
    hipLaunchKernelGGL(k_row_count, dim3(nblk(nr, 4)), dim3(256), 0, h->st, g, R, mask, h->rowcnt);
    hipLaunchKernelGGL(k_scan_local, dim3(nb), dim3(256), 0, h->st, h->rowcnt, h->rowoff, h->bsum, nr);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, h->st, h->bsum, nb, h->ctr);
    int hc[C_N];
    if (int rc = read_ctr(h, hc)) return rc;
    *total = hc[C_TOTAL];
    if (int rc = ghost ? grow_lists(h, *total, 0) : grow_lists(h, 0, *total)) return rc;
    int *L[3];
    if (ghost) {
        const GhostOut o = gout(h);
        L[0] = o.gi; L[1] = o.gj; L[2] = o.gk;
    } else {
        const DeadOut o = dout(h);
        L[0] = o.di; L[1] = o.dj; L[2] = o.dk;
    }
    hipLaunchKernelGGL(k_row_write, dim3(nblk(nr, 4)), dim3(256), 0, h->st, g, R, mask, h->rowoff, h->bsum, L[0],
                       L[1], L[2], map);
    GCHK(hipGetLastError());
    return 0;
}


}  // namespace

extern "C" int fc_geom_create(const fc_grid *gr, const fc_geom_params *p, fc_geom **out)
{
    if (!gr || !p || !out || !gr->x || !gr->y || !gr->z) return fc::set_error(-1, "fc_geom_create: null argument");
    *out = nullptr;
    if (gr->nx < 3 || gr->ny < 3 || gr->nz < 3) return fc::set_error(-1, "fc_geom_create: at least 3 nodes per axis");
    if (p->nBody < 1 || p->nBody > FC_GEOM_MAX_BODIES)
        return fc::set_error(-1, "fc_geom_create: nBody must be 1.." + std::to_string(FC_GEOM_MAX_BODIES));
    for (int b = 0; b < p->nBody; ++b) {
        if (p->unstruc_surface_type[b] != FC_SURFACE_SOLID_BODY && p->unstruc_surface_type[b] != FC_SURFACE_MEMBRANE)
            return fc::set_error(-1, "fc_geom_create: unstruc_surface_type must be SOLID_BODY (1) or MEMBRANE (2)");
        if (p->body_dim[b] != 2 && p->body_dim[b] != 3) return fc::set_error(-1, "fc_geom_create: body_dim must be 2 or 3");
    }
    const long n = (long)(gr->nx + 2) * (gr->ny + 2) * (gr->nz + 2);
    if (n > (1L << 31) - 1) return fc::set_error(-1, "fc_geom_create: grid above 2^31 cells");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fc::set_error(-5, "no HIP device available");
    fc_geom *h = new fc_geom();
    struct Guard {
        fc_geom *h;
        ~Guard() { if (h) fc_geom_destroy(h); }
    } guard{h};
    GCHK(hipGetDevice(&h->dev));
    GCHK(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking));
    GCHK(hipEventCreate(&h->e0));
    GCHK(hipEventCreate(&h->e1));
    for (int q = 0; q < 5; ++q) GCHK(hipEventCreate(&h->es[q]));
    for (auto &e : h->ev) GCHK(hipEventCreate(&e));
    h->nx = gr->nx; h->ny = gr->ny; h->nz = gr->nz; h->n = n;
    h->p = *p;
    h->m = std::max(std::max(gr->nx, gr->ny), gr->nz) + 2;
    // CORE_GRID.F90:106-150 on the host, one buffer of 12 axis arrays
    std::vector<double> ax(12 * (size_t)h->m, 0.0);
    const double *src[3] = {gr->x, gr->y, gr->z};
    const int nn[3] = {gr->nx, gr->ny, gr->nz};
    for (int a = 0; a < 3; ++a) {
        double *x = &ax[a * h->m], *xc = &ax[(3 + a) * h->m], *dx = &ax[(6 + a) * h->m], *dxc = &ax[(9 + a) * h->m];
        const int nA = nn[a];
        for (int i = 0; i <= nA + 1; ++i) x[i] = src[a][i];
        for (int i = 0; i <= nA; ++i) {
            xc[i] = 0.5 * (x[i + 1] + x[i]);
            dx[i] = x[i + 1] - x[i];
        }
        xc[nA + 1] = x[nA + 1];
        for (int i = 1; i <= nA; ++i) dxc[i] = xc[i] - xc[i - 1];
    }
    if (int rc = dalloc(h, &h->grid, ax.size())) return rc;
    GCHK(hipMemcpyAsync(h->grid, ax.data(), ax.size() * sizeof(double), hipMemcpyHostToDevice, h->st));
    if (int rc = dalloc(h, &h->m8, 8 * (size_t)n)) return rc;
    GCHK(hipMemsetAsync(h->m8, 0, 8 * (size_t)n, h->st));
    if (int rc = dalloc(h, &h->gidx, n)) return rc;
    if (int rc = dalloc(h, &h->didx, n)) return rc;
    GCHK(hipMemsetAsync(h->gidx, 0, sizeof(int) * (size_t)n, h->st));
    GCHK(hipMemsetAsync(h->didx, 0, sizeof(int) * (size_t)n, h->st));
    if (int rc = dalloc(h, &h->dcip, 3 * (size_t)n)) return rc;
    GCHK(hipMemsetAsync(h->dcip, 0, sizeof(double) * 3 * (size_t)n, h->st));
    if (int rc = dalloc(h, &h->sA, 3 * (size_t)n)) return rc;
    h->sP = h->sA + n;
    h->sN = h->sA + 2 * n;
    if (int rc = dalloc(h, &h->qlist, n)) return rc;
    const int rows = (gr->ny + 1) * (gr->nz + 1);
    if (int rc = dalloc(h, &h->rowcnt, rows)) return rc;
    if (int rc = dalloc(h, &h->bsum, (rows + 1023) / 1024)) return rc;
    if (int rc = dalloc(h, &h->rowoff, rows)) return rc;
    if (int rc = dalloc(h, &h->ctr, C_N)) return rc;
    if (int rc = dalloc(h, &h->pairs, 1)) return rc;
    if (int rc = dalloc(h, &h->totals, 9 * FC_GEOM_MAX_BODIES)) return rc;
    GCHK(hipStreamSynchronize(h->st));
    guard.h = nullptr;
    *out = h;
    return 0;
}

extern "C" int fc_geom_destroy(fc_geom *h)
{
    if (!h) return 0;
    if (h->st) (void)hipStreamSynchronize(h->st);
    for (void *p : h->owned) (void)hipFree(p);
    h->owned.clear();
    if (h->e0) (void)hipEventDestroy(h->e0);
    if (h->e1) (void)hipEventDestroy(h->e1);
    for (int q = 0; q < 5; ++q)
        if (h->es[q]) (void)hipEventDestroy(h->es[q]);
    for (auto &e : h->ev)
        if (e) (void)hipEventDestroy(e);
    if (h->st) (void)hipStreamDestroy(h->st);
    delete h;
    return 0;
}

extern "C" int fc_geom_set_body(fc_geom *h, int ibody, int nPts, int nElem, const double *x, const double *y,
                                const double *z, const int *tri)
{
    if (!h || !x || !y || !z || !tri) return fc::set_error(-1, "fc_geom_set_body: null argument");
    if (ibody < 1 || ibody > h->p.nBody) return fc::set_error(-1, "fc_geom_set_body: ibody out of range");
    if (nPts < 3 || nElem < 1) return fc::set_error(-1, "fc_geom_set_body: at least 3 markers and 1 element");
    for (long e = 0; e < 3L * nElem; ++e)
        if (tri[e] < 1 || tri[e] > nPts) return fc::set_error(-1, "fc_geom_set_body: triElemNeig entry out of 1..nPts");
    fc_geom::Body &B = h->body[ibody - 1];
    dfree(h, B.buf); dfree(h, B.tri); dfree(h, B.voff); dfree(h, B.vel); dfree(h, B.bstart); dfree(h, B.bmark); dfree(h, B.bxyz);
    dfree(h, B.mvel); dfree(h, B.ef); dfree(h, B.mf);
    B = fc_geom::Body();
    B.npt = nPts;
    B.nel = nElem;
    B.tri_h.assign(tri, tri + 3 * (size_t)nElem);
    // marker -> element incidence in element order (counting sort)
    std::vector<int> off(nPts + 1, 0), lst(3 * (size_t)nElem);
    for (long e = 0; e < 3L * nElem; ++e) off[tri[e]]++;
    for (int v = 0; v < nPts; ++v) off[v + 1] += off[v];
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int m = 0; m < nElem; ++m)
        for (int q = 0; q < 3; ++q) lst[fill[tri[3 * m + q] - 1]++] = m;
    if (int rc = dalloc(h, &B.buf, 6 * (size_t)nPts + 7 * (size_t)nElem)) return rc;
    if (int rc = dalloc(h, &B.tri, 3 * (size_t)nElem)) return rc;
    if (int rc = dalloc(h, &B.voff, nPts + 1)) return rc;
    if (int rc = dalloc(h, &B.vel, 3 * (size_t)nElem)) return rc;
    if (int rc = dalloc(h, &B.mvel, 6 * (size_t)nPts)) return rc;
    if (int rc = dalloc(h, &B.ef, 13 * (size_t)FC_ROWPAD(nElem))) return rc;
    GCHK(hipMemsetAsync(B.ef, 0, sizeof(double) * 13 * (size_t)FC_ROWPAD(nElem), h->st));
    if (int rc = dalloc(h, &B.mf, 10 * (size_t)nPts)) return rc;
    GCHK(hipMemsetAsync(B.mvel, 0, sizeof(double) * 6 * (size_t)nPts, h->st));
    GCHK(hipMemcpyAsync(B.tri, tri, sizeof(int) * 3 * (size_t)nElem, hipMemcpyHostToDevice, h->st));
    GCHK(hipMemcpyAsync(B.voff, off.data(), sizeof(int) * (nPts + 1), hipMemcpyHostToDevice, h->st));
    GCHK(hipMemcpyAsync(B.vel, lst.data(), sizeof(int) * 3 * (size_t)nElem, hipMemcpyHostToDevice, h->st));
    GCHK(hipStreamSynchronize(h->st));
    return fc_geom_move_body(h, ibody, x, y, z);
}

extern "C" int fc_geom_move_body(fc_geom *h, int ibody, const double *x, const double *y, const double *z)
{
    if (!h || !x || !y || !z) return fc::set_error(-1, "fc_geom_move_body: null argument");
    if (ibody < 1 || ibody > h->p.nBody || !h->body[ibody - 1].buf)
        return fc::set_error(-1, "fc_geom_move_body: body not set");
    fc_geom::Body &B = h->body[ibody - 1];
    const size_t b = sizeof(double) * (size_t)B.npt;
    GCHK(hipMemcpyAsync(B.buf, x, b, hipMemcpyHostToDevice, h->st));
    GCHK(hipMemcpyAsync(B.buf + B.npt, y, b, hipMemcpyHostToDevice, h->st));
    GCHK(hipMemcpyAsync(B.buf + 2 * (size_t)B.npt, z, b, hipMemcpyHostToDevice, h->st));
    h->bi_ready = false;
    return rebin(h, B, x, y, z);
}

extern "C" int fc_geom_refresh(fc_geom *h, fc_geom_stats *st)
{
    if (!h) return fc::set_error(-1, "fc_geom_refresh: null handle");
    for (int b = 0; b < h->p.nBody; ++b)
        if (!h->body[b].buf) return fc::set_error(-1, "fc_geom_refresh: body " + std::to_string(b + 1) + " not set");
    GCHK(hipSetDevice(h->dev));
    const GGrid g = ggrid(h);
    const GMasks m = gmasks(h);
    GBodies B;
    std::memset(&B, 0, sizeof(B));
    B.nb = h->p.nBody;
    B.tkns = h->p.membrane_tkns;
    B.probe = h->p.probeLengthNormalized;
    B.probeD = h->p.probeLengthNormalizedD;
    for (int b = 0; b < B.nb; ++b) B.b[b] = gbody(h, b);
    const long n = h->n;
    fc_geom_stats S;
    std::memset(&S, 0, sizeof(S));
    for (int a = 0; a < C_N; ++a) h->ctr_init[a] = (a >= C_BOX) ? -(1 << 30) : 0;
    GCHK(hipMemcpyAsync(h->ctr, h->ctr_init, sizeof(h->ctr_init), hipMemcpyHostToDevice, h->st));
    GCHK(hipMemsetAsync(h->pairs, 0, sizeof(unsigned long long), h->st));
    GCHK(hipEventRecord(h->e0, h->st));
    // surface metrics, bodies from the last to the first
    for (int b = B.nb - 1; b >= 0; --b) {
        hipLaunchKernelGGL(k_surf_elem, dim3(nblk(B.b[b].nel, 256)), dim3(256), 0, h->st, B.b[b]);
        hipLaunchKernelGGL(k_surf_vert, dim3(nblk(B.b[b].npt, 256)), dim3(256), 0, h->st, B.b[b]);
    }
    // set_iblank_body_fast
    const dim3 gstride(std::min<unsigned>(nblk(n, 256), 8192)), gall(nblk(n, 256));
    hipLaunchKernelGGL(k_iblank_init, gstride, dim3(256), 0, h->st, g, m);
    for (int b = B.nb - 1; b >= 0; --b)
        hipLaunchKernelGGL(k_bbox_mark, dim3(nblk(B.b[b].nel, 4)), dim3(256), 0, h->st, g, B.b[b], B.tkns, m);
    hipLaunchKernelGGL(k_collect_undecided, gstride, dim3(256), 0, h->st, g, m, h->qlist, h->ctr);
    int hc[C_N];
    if (int rc = read_ctr(h, hc)) return rc;
    S.n_undecided = hc[C_UNDEC];
    GCHK(hipEventRecord(h->ev[0], h->st));
    if (S.n_undecided > 0)
        hipLaunchKernelGGL(k_classify, dim3(nblk(S.n_undecided, 256)), dim3(256), 0, h->st, g, B, m, h->qlist,
                           S.n_undecided, h->ctr, h->pairs);
    GCHK(hipEventRecord(h->ev[1], h->st));
    hipLaunchKernelGGL(k_rowfill, dim3(nblk((long)(h->ny - 1) * (h->nz - 1), 4)), dim3(256), 0, h->st, g, m);
    GCHK(hipGetLastError());
    // the smoothing passes (5 x 3-D, 5 x j-k planes, 5 x i-k planes)
    SmoothPass passes[3];
    {
        const int o3[6][3] = {{0, 0, -1}, {0, -1, 0}, {-1, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        const int oA[4][3] = {{0, 0, -1}, {0, -1, 0}, {0, 1, 0}, {0, 0, 1}};
        const int oB[4][3] = {{0, 0, -1}, {-1, 0, 0}, {1, 0, 0}, {0, 0, 1}};
        passes[0].nn = 6; passes[0].thr = 5;
        for (int t = 0; t < 6; ++t) { passes[0].di[t] = o3[t][0]; passes[0].dj[t] = o3[t][1]; passes[0].dk[t] = o3[t][2]; }
        passes[1].nn = 4; passes[1].thr = 3;
        passes[2].nn = 4; passes[2].thr = 3;
        for (int t = 0; t < 4; ++t) {
            passes[1].di[t] = oA[t][0]; passes[1].dj[t] = oA[t][1]; passes[1].dk[t] = oA[t][2];
            passes[2].di[t] = oB[t][0]; passes[2].dj[t] = oB[t][1]; passes[2].dk[t] = oB[t][2];
        }
        for (int t = 4; t < 6; ++t)
            for (int a = 1; a < 3; ++a) { passes[a].di[t] = 0; passes[a].dj[t] = 0; passes[a].dk[t] = 0; }
    }
    int rounds = 0;
    hipLaunchKernelGGL(k_solid_box, gstride, dim3(256), 0, h->st, g, m, h->ctr);
    if (int rc = read_ctr(h, hc)) return rc;
    const int *bx = hc + C_BOX;
    if (bx[1] > 0) {       // there are interior solid cells
        const Box inner{-bx[0], bx[1], -bx[2], bx[3], -bx[4], bx[5]};
        const Box wide{inner.i0 - 1, inner.i1 + 1, inner.j0 - 1, inner.j1 + 1, inner.k0 - 1, inner.k1 + 1};
        for (int ps = 0; ps < 15; ++ps) {
            const SmoothPass &sp = passes[ps / 5];
            hipLaunchKernelGGL(k_smooth_pack, dim3(nblk(wide.count(), 256)), dim3(256), 0, h->st, g, wide, m, h->sA,
                               h->sP, h->sN);
            int16_t *P = h->sP, *N = h->sN;
            for (int r = 0;; ++r) {
                if (r > 4096) return fc::set_error(-3, "fc_geom_refresh: smoothing rounds did not settle");
                GCHK(hipMemsetAsync(h->ctr + C_FLAG, 0, sizeof(int), h->st));
                hipLaunchKernelGGL(k_smooth_round, dim3(nblk(inner.count(), 256)), dim3(256), 0, h->st, g, inner, sp,
                                   h->sA, P, N, h->ctr);
                if (int rc = read_ctr(h, hc)) return rc;
                ++rounds;
                std::swap(P, N);
                if (!hc[C_FLAG]) break;
            }
            hipLaunchKernelGGL(k_smooth_unpack, dim3(nblk(inner.count(), 256)), dim3(256), 0, h->st, g, inner, m,
                               h->sA, P, h->ctr);
        }
    }
    S.smooth_rounds = rounds;
    // ghost cells
    hipLaunchKernelGGL(k_ghost_mark, gall, dim3(256), 0, h->st, g, m);
    GCHK(hipGetLastError());
    const Rows RG{1, h->ny - 1, 1, h->nz - 1};
    int nG = 0;
    if (int rc = compact(h, RG, m.ghost, true, h->gidx, &nG)) return rc;
    h->nG = nG;
    // dead cells: marks, list (rows j 0..ny, k 0..nz), relations
    hipLaunchKernelGGL(k_dead_mark, gall, dim3(256), 0, h->st, g, m);
    if (h->nD > 0) {
        DeadOut od = dout(h);
        hipLaunchKernelGGL(k_clear_map, dim3(nblk(h->nD, 256)), dim3(256), 0, h->st, od.di, od.dj, od.dk, h->nD, g,
                           h->didx);
    }
    const Rows RD{0, h->ny, 0, h->nz};
    int nD = 0;
    if (int rc = compact(h, RD, m.dead, false, h->didx, &nD)) return rc;
    h->nD = nD;
    GCHK(hipEventRecord(h->ev[2], h->st));
    if (nG > 0)
        hipLaunchKernelGGL(k_ghost_rel, dim3(nblk(nG, 256)), dim3(256), 0, h->st, g, B, m, gout(h), nG, h->dcip,
                           h->pairs, h->ctr);
    if (nD > 0)
        hipLaunchKernelGGL(k_dead_rel, dim3(nblk(nD, 256)), dim3(256), 0, h->st, g, B, m, dout(h), nD, h->dcip,
                           h->pairs, h->ctr);
    GCHK(hipEventRecord(h->ev[3], h->st));
    if (nG > 0) {
        DeadOut od = dout(h);
        hipLaunchKernelGGL(k_alpha, dim3(nblk(nG, 256)), dim3(256), 0, h->st, g, gout(h), nG, h->didx, od.ds);
    }
    GCHK(hipGetLastError());
    GCHK(hipEventRecord(h->e1, h->st));
    if (int rc = read_ctr(h, hc)) return rc;
    float ms = 0.f;
    GCHK(hipEventElapsedTime(&ms, h->e0, h->e1));
    float ms1 = 0.f, ms2 = 0.f;
    GCHK(hipEventElapsedTime(&ms1, h->ev[0], h->ev[1]));
    GCHK(hipEventElapsedTime(&ms2, h->ev[2], h->ev[3]));
    S.scan_ms = (double)ms1 + (double)ms2;
    unsigned long long np = 0;
    GCHK(hipMemcpy(&np, h->pairs, sizeof(np), hipMemcpyDeviceToHost));
    S.scan_pairs = (double)np;
    S.nGhost = nG;
    S.nDead = nD;
    S.n_smoothed = hc[C_FLIPS];
    S.n_unselected_side = hc[C_UNSEL];
    S.time_ms = ms;
    h->fresh = true;
    h->bi_ready = false;
    if (st) *st = S;
    if (hc[C_ERR] & (E_GHOST_BODY | E_DEAD_BODY))
        return fc::set_error(-5, "fc_geom_refresh: a ghost or dead cell has no body number");
    if (hc[C_ERR] & (E_GHOST_HOST | E_DEAD_HOST))
        return fc::set_error(-5, "fc_geom_refresh: failed to find the hosting box for an image point");
    return 0;
}

extern "C" int fc_geom_get_counts(fc_geom *h, int *nGhost, int *nDead)
{
    if (!h) return fc::set_error(-1, "fc_geom_get_counts: null handle");
    if (nGhost) *nGhost = h->nG;
    if (nDead) *nDead = h->nD;
    return 0;
}

extern "C" int fc_geom_get_masks(fc_geom *h, int8_t *iblank, int8_t *ghost, int8_t *dead, int8_t *bodyNum,
                                 int8_t *ps1, int8_t *ps2)
{
    if (!h) return fc::set_error(-1, "fc_geom_get_masks: null handle");
    const GMasks m = gmasks(h);
    int8_t *dst[6] = {iblank, ghost, dead, bodyNum, ps1, ps2};
    const int8_t *src[6] = {m.iblank, m.ghost, m.dead, m.bodynum, m.ps1, m.ps2};
    for (int a = 0; a < 6; ++a)
        if (dst[a]) GCHK(hipMemcpyAsync(dst[a], src[a], (size_t)h->n, hipMemcpyDeviceToHost, h->st));
    GCHK(hipStreamSynchronize(h->st));
    return 0;
}

namespace {
int copy_out(fc_geom *h, void *dst, const void *src, size_t bytes)
{
    if (!dst || bytes == 0) return 0;
    GCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->st));
    return 0;
}
}  // namespace

extern "C" int fc_geom_get_ghosts(fc_geom *h, const fc_geom_ghosts *o)
{
    if (!h || !o) return fc::set_error(-1, "fc_geom_get_ghosts: null argument");
    if (h->nG == 0) return 0;
    const GhostOut g = gout(h);
    const size_t ni = sizeof(int) * h->nG, nd = sizeof(double) * h->nG;
    int rc = 0;
    rc |= copy_out(h, o->iGhost, g.gi, ni); rc |= copy_out(h, o->jGhost, g.gj, ni); rc |= copy_out(h, o->kGhost, g.gk, ni);
    rc |= copy_out(h, o->iCellIndex, g.ci, ni); rc |= copy_out(h, o->jCellIndex, g.cj, ni);
    rc |= copy_out(h, o->kCellIndex, g.ck, ni); rc |= copy_out(h, o->closestElementGC, g.el, ni);
    rc |= copy_out(h, o->xBodyIntercept, g.bx, nd); rc |= copy_out(h, o->yBodyIntercept, g.by, nd);
    rc |= copy_out(h, o->zBodyIntercept, g.bz, nd); rc |= copy_out(h, o->xBodyInterceptNorm, g.nx, nd);
    rc |= copy_out(h, o->yBodyInterceptNorm, g.ny, nd); rc |= copy_out(h, o->zBodyInterceptNorm, g.nz, nd);
    rc |= copy_out(h, o->probeLength, g.pl, nd); rc |= copy_out(h, o->xImagePoint, g.ix, nd);
    rc |= copy_out(h, o->yImagePoint, g.iy, nd); rc |= copy_out(h, o->zImagePoint, g.iz, nd);
    rc |= copy_out(h, o->coeffGCMD, g.cd, 8 * nd); rc |= copy_out(h, o->coeffGCMN, g.cn, 8 * nd);
    rc |= copy_out(h, o->alphaGhost, g.alpha, nd);
    if (rc) return -3;
    GCHK(hipStreamSynchronize(h->st));
    return 0;
}

extern "C" int fc_geom_get_dead(fc_geom *h, const fc_geom_dead *o)
{
    if (!h || !o) return fc::set_error(-1, "fc_geom_get_dead: null argument");
    if (h->nD == 0) return 0;
    const DeadOut d = dout(h);
    const size_t ni = sizeof(int) * h->nD, nd = sizeof(double) * h->nD;
    int rc = 0;
    rc |= copy_out(h, o->iDead, d.di, ni); rc |= copy_out(h, o->jDead, d.dj, ni); rc |= copy_out(h, o->kDead, d.dk, ni);
    rc |= copy_out(h, o->iDeadCellIndex, d.ci, ni); rc |= copy_out(h, o->jDeadCellIndex, d.cj, ni);
    rc |= copy_out(h, o->kDeadCellIndex, d.ck, ni); rc |= copy_out(h, o->closestElementDead, d.el, ni);
    rc |= copy_out(h, o->xBodyInterceptDead, d.bx, nd); rc |= copy_out(h, o->yBodyInterceptDead, d.by, nd);
    rc |= copy_out(h, o->zBodyInterceptDead, d.bz, nd); rc |= copy_out(h, o->xNormInterceptDead, d.nx, nd);
    rc |= copy_out(h, o->yNormInterceptDead, d.ny, nd); rc |= copy_out(h, o->zNormInterceptDead, d.nz, nd);
    rc |= copy_out(h, o->probeLengthDead, d.pl, nd); rc |= copy_out(h, o->coeffGCMDeadD, d.cd, 8 * nd);
    rc |= copy_out(h, o->coeffGCMDeadN, d.cn, 8 * nd);
    if (rc) return -3;
    GCHK(hipStreamSynchronize(h->st));
    return 0;
}

extern "C" int fc_geom_get_surface(fc_geom *h, int ibody, const fc_geom_surface *f)
{
    if (!h || !f) return fc::set_error(-1, "fc_geom_get_surface: null argument");
    if (ibody < 1 || ibody > h->p.nBody || !h->body[ibody - 1].buf)
        return fc::set_error(-1, "fc_geom_get_surface: body not set");
    const GBody b = gbody(h, ibody - 1);
    const size_t ne = sizeof(double) * b.nel, nv = sizeof(double) * b.npt;
    int rc = 0;
    rc |= copy_out(h, f->triElemNormX, b.enx, ne); rc |= copy_out(h, f->triElemNormY, b.eny, ne);
    rc |= copy_out(h, f->triElemNormZ, b.enz, ne); rc |= copy_out(h, f->triElemArea, b.area, ne);
    rc |= copy_out(h, f->triElemCentX, b.ecx, ne); rc |= copy_out(h, f->triElemCentY, b.ecy, ne);
    rc |= copy_out(h, f->triElemCentZ, b.ecz, ne); rc |= copy_out(h, f->xNormBodyMarker, b.vnx, nv);
    rc |= copy_out(h, f->yNormBodyMarker, b.vny, nv); rc |= copy_out(h, f->zNormBodyMarker, b.vnz, nv);
    if (rc) return -3;
    GCHK(hipStreamSynchronize(h->st));
    return 0;
}

extern "C" int fc_geom_get_dcip(fc_geom *h, double *xyzDCIP)
{
    if (!h || !xyzDCIP) return fc::set_error(-1, "fc_geom_get_dcip: null argument");
    GCHK(hipMemcpyAsync(xyzDCIP, h->dcip, sizeof(double) * 3 * (size_t)h->n, hipMemcpyDeviceToHost, h->st));
    GCHK(hipStreamSynchronize(h->st));
    return 0;
}

extern "C" int fc_geom_get_dead_ip(fc_geom *h, double *xyz)
{
    if (!h || !xyz) return fc::set_error(-1, "fc_geom_get_dead_ip: null argument");
    if (!h->fresh) return fc::set_error(-1, "fc_geom_get_dead_ip: call fc_geom_refresh first");
    if (h->nD < 1) return 0;
    GCHK(hipSetDevice(h->dev));
    double *buf = nullptr;
    if (int rc = dalloc(h, &buf, 3 * (size_t)h->nD)) return rc;
    const DeadOut o = dout(h);
    hipLaunchKernelGGL(k_dead_ip, dim3(nblk(h->nD, 256)), dim3(256), 0, h->st, ggrid(h), o.di, o.dj, o.dk, h->nD,
                       h->dcip, buf);
    int rc = 0;
    if (hipGetLastError() != hipSuccess) rc = fc::set_error(-3, "fc_geom_get_dead_ip: launch failed");
    if (!rc) rc = copy_out(h, xyz, buf, sizeof(double) * 3 * (size_t)h->nD);
    (void)hipStreamSynchronize(h->st);
    dfree(h, buf);
    return rc ? (rc < 0 ? rc : -3) : 0;
}

namespace {
BiOut gbi(const fc_geom *h)
{
    const size_t c = h->capG;
    double *d = h->gdbl + 29 * c;
    return BiOut{d, d + c, d + 2 * c, d + 3 * c};
}
BiOut dbi(const fc_geom *h)
{
    const size_t c = h->capD;
    double *d = h->ddbl + 24 * c;
    return BiOut{d, d + c, d + 2 * c, d + 3 * c};
}
}  // namespace

extern "C" int fc_geom_set_marker_velocity(fc_geom *h, int ibody, const double *u, const double *v, const double *w,
                                           const double *uOld, const double *vOld, const double *wOld)
{
    if (!h || !u || !v || !w || !uOld || !vOld || !wOld) return fc::set_error(-1, "fc_geom_set_marker_velocity: null argument");
    if (ibody < 1 || ibody > h->p.nBody || !h->body[ibody - 1].buf)
        return fc::set_error(-1, "fc_geom_set_marker_velocity: body not set");
    fc_geom::Body &B = h->body[ibody - 1];
    const double *src[6] = {u, v, w, uOld, vOld, wOld};
    for (int c = 0; c < 6; ++c)
        GCHK(hipMemcpyAsync(B.mvel + (size_t)c * B.npt, src[c], sizeof(double) * B.npt, hipMemcpyHostToDevice, h->st));
    GCHK(hipStreamSynchronize(h->st));
    h->bi_ready = false;
    return 0;
}

extern "C" int fc_geom_body_intercept_values(fc_geom *h, double dt, int ntime)
{
    if (!h) return fc::set_error(-1, "fc_geom_body_intercept_values: null handle");
    if (!h->fresh) return fc::set_error(-1, "fc_geom_body_intercept_values: call fc_geom_refresh first");
    if (ntime > 1 && !(dt != 0.0)) return fc::set_error(-1, "fc_geom_body_intercept_values: dt must be nonzero");
    GCHK(hipSetDevice(h->dev));
    const GGrid g = ggrid(h);
    const GMasks m = gmasks(h);
    GBodies B;
    std::memset(&B, 0, sizeof(B));
    B.nb = h->p.nBody;
    for (int b = 0; b < B.nb; ++b) B.b[b] = gbody(h, b);
    int hc[C_N];
    for (int a = 0; a < C_N; ++a) h->ctr_init[a] = 0;
    GCHK(hipMemcpyAsync(h->ctr, h->ctr_init, sizeof(h->ctr_init), hipMemcpyHostToDevice, h->st));
    GCHK(hipEventRecord(h->es[0], h->st));
    if (h->nG > 0) {
        const GhostOut o = gout(h);
        hipLaunchKernelGGL(k_bi_values, dim3(nblk(h->nG, 128)), dim3(128), 0, h->st, g, B, m, o.gi, o.gj, o.gk, o.el,
                           o.bx, o.by, o.bz, o.nx, o.ny, o.nz, h->nG, dt, ntime, gbi(h), h->ctr);
    }
    if (h->nD > 0) {
        const DeadOut o = dout(h);
        hipLaunchKernelGGL(k_bi_values, dim3(nblk(h->nD, 128)), dim3(128), 0, h->st, g, B, m, o.di, o.dj, o.dk, o.el,
                           o.bx, o.by, o.bz, o.nx, o.ny, o.nz, h->nD, dt, ntime, dbi(h), h->ctr);
    }
    GCHK(hipGetLastError());
    GCHK(hipEventRecord(h->es[1], h->st));
    h->es_bi = true;
    if (int rc = read_ctr(h, hc)) return rc;
    if (hc[C_ERR]) return fc::set_error(-5, "fc_geom_body_intercept_values: a relation has no body or element");
    h->bi_ready = true;
    return 0;
}

extern "C" int fc_geom_get_bi_values(fc_geom *h, double *uG, double *vG, double *wG, double *dpdnG, double *uD,
                                     double *vD, double *wD, double *dpdnD)
{
    if (!h) return fc::set_error(-1, "fc_geom_get_bi_values: null handle");
    if (!h->bi_ready) return fc::set_error(-1, "fc_geom_get_bi_values: call fc_geom_body_intercept_values first");
    const BiOut a = gbi(h), b = dbi(h);
    const size_t ng = sizeof(double) * h->nG, nd = sizeof(double) * h->nD;
    int rc = 0;
    rc |= copy_out(h, uG, a.u, ng); rc |= copy_out(h, vG, a.v, ng); rc |= copy_out(h, wG, a.w, ng);
    rc |= copy_out(h, dpdnG, a.dpdn, ng);
    rc |= copy_out(h, uD, b.u, nd); rc |= copy_out(h, vD, b.v, nd); rc |= copy_out(h, wD, b.w, nd);
    rc |= copy_out(h, dpdnD, b.dpdn, nd);
    if (rc) return -3;
    GCHK(hipStreamSynchronize(h->st));
    return 0;
}

namespace {
int marker_stress(fc_geom *h, const double *p, const double *u, const double *v, const double *w, double reinv,
                  double *totals)
{
    if (!h->bi_ready) return fc::set_error(-1, "fc_geom_marker_stress: call fc_geom_body_intercept_values first");
    if (h->nD < 1) return fc::set_error(-1, "fc_geom_marker_stress: no dead cells");
    const GGrid g = ggrid(h);
    const GMasks m = gmasks(h);
    const Flow f{p, u, v, w};
    ElemForce ef[FC_GEOM_MAX_BODIES];
    for (int b = 0; b < h->p.nBody; ++b) {
        fc_geom::Body &B = h->body[b];
        const size_t ne = FC_ROWPAD(B.nel);
        ef[b] = ElemForce{B.ef, B.ef + ne, B.ef + 2 * ne, B.ef + 3 * ne, B.ef + 4 * ne, B.ef + 5 * ne, B.ef + 6 * ne};
    }
    GCHK(hipEventRecord(h->es[2], h->st));
    for (int b = 0; b < h->p.nBody; ++b)
        hipLaunchKernelGGL(k_marker_stress, dim3(nblk(h->body[b].nel, 128)), dim3(128), 0, h->st, g, gbody(h, b),
                           b + 1, h->p.membrane_tkns, m, f, h->didx, dout(h), dbi(h), reinv, ef[b]);
    GCHK(hipEventRecord(h->es[3], h->st));
    for (int b = 0; b < h->p.nBody; ++b) {
        const GBody gb = gbody(h, b);
        fc_geom::Body &B = h->body[b];
        const size_t np = B.npt;
        const MarkerForces mf{B.mf, B.mf + np, B.mf + 2 * np, B.mf + 3 * np, B.mf + 4 * np, B.mf + 5 * np,
                              B.mf + 6 * np, B.mf + 7 * np, B.mf + 8 * np, B.mf + 9 * np};
        hipLaunchKernelGGL(k_force_totals, dim3(1), dim3(64), 0, h->st, B.nel, B.ef, h->totals + 9 * b);
        hipLaunchKernelGGL(k_marker_forces, dim3(nblk(B.npt, 256)), dim3(256), 0, h->st, gb, ef[b], mf);
    }
    GCHK(hipEventRecord(h->es[4], h->st));
    h->es_stress = true;
    GCHK(hipGetLastError());
    if (totals) {
        GCHK(hipMemcpyAsync(totals, h->totals, sizeof(double) * 9 * h->p.nBody, hipMemcpyDeviceToHost, h->st));
    }
    GCHK(hipStreamSynchronize(h->st));
    return 0;
}
}  // namespace

extern "C" int fc_geom_marker_stress(fc_geom *h, const double *p, const double *u, const double *v, const double *w,
                                     double reinv, double *totals)
{
    if (!h || !p || !u || !v || !w) return fc::set_error(-1, "fc_geom_marker_stress: null argument");
    GCHK(hipSetDevice(h->dev));
    if (!h->flow)
        if (int rc = dalloc(h, &h->flow, 4 * (size_t)h->n)) return rc;
    const double *src[4] = {p, u, v, w};
    for (int c = 0; c < 4; ++c)
        GCHK(hipMemcpyAsync(h->flow + (size_t)c * h->n, src[c], sizeof(double) * h->n, hipMemcpyHostToDevice, h->st));
    return marker_stress(h, h->flow, h->flow + h->n, h->flow + 2 * h->n, h->flow + 3 * h->n, reinv, totals);
}

extern "C" int fc_geom_marker_stress_device(fc_geom *h, const double *p, const double *u, const double *v,
                                            const double *w, double reinv, double *totals)
{
    if (!h || !p || !u || !v || !w) return fc::set_error(-1, "fc_geom_marker_stress_device: null argument");
    GCHK(hipSetDevice(h->dev));
    return marker_stress(h, p, u, v, w, reinv, totals);
}

extern "C" int fc_geom_marker_stress_resident(fc_geom *h, fc_pson *flow, double reinv, double *totals)
{
    if (!h || !flow) return fc::set_error(-1, "fc_geom_marker_stress_resident: null argument");
    int nlev = 0, nxl[64], nyl[64], nzl[64];
    if (int rc = fc_pson_levels(flow, &nlev, nxl, nyl, nzl)) return rc;
    if (nlev < 1 || nxl[0] != h->nx || nyl[0] != h->ny || nzl[0] != h->nz)
        return fc::set_error(-1, "fc_geom_marker_stress_resident: the flow handle's grid is not this grid");
    GCHK(hipSetDevice(h->dev));
    if (!h->flow)
        if (int rc = dalloc(h, &h->flow, 4 * (size_t)h->n)) return rc;
    // p (update_pressure's), u, v, w as dense copies on this handle
    const int which[4] = {6, 0, 1, 2};
    for (int q = 0; q < 4; ++q)
        if (int rc = fc_pson_resident_copy(flow, which[q], h->flow + (size_t)q * h->n)) return rc;
    return marker_stress(h, h->flow, h->flow + h->n, h->flow + 2 * h->n, h->flow + 3 * h->n, reinv, totals);
}

extern "C" int fc_geom_get_marker_forces(fc_geom *h, int ibody, const fc_geom_forces *o)
{
    if (!h || !o) return fc::set_error(-1, "fc_geom_get_marker_forces: null argument");
    if (ibody < 1 || ibody > h->p.nBody || !h->body[ibody - 1].buf)
        return fc::set_error(-1, "fc_geom_get_marker_forces: body not set");
    const fc_geom::Body &B = h->body[ibody - 1];
    const size_t np = B.npt, b = sizeof(double) * np;
    double *dst[10] = {o->xMarkerForce, o->yMarkerForce, o->zMarkerForce, o->xMarkerShear, o->yMarkerShear,
                       o->zMarkerShear, o->MarkerArea, o->xMarkerStress, o->yMarkerStress, o->zMarkerStress};
    const int order[10] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9};
    // device layout: fx fy fz sx sy sz ma tx ty tz
    int rc = 0;
    for (int a = 0; a < 10; ++a) rc |= copy_out(h, dst[a], B.mf + order[a] * np, b);
    if (rc) return -3;
    GCHK(hipStreamSynchronize(h->st));
    return 0;
}

extern "C" int fc_geom_get_stress_timing(fc_geom *h, double *bi_ms, double *interp_ms, double *sum_ms)
{
    if (!h) return fc::set_error(-1, "fc_geom_get_stress_timing: null handle");
    if (!h->es_bi || !h->es_stress) return fc::set_error(-1, "fc_geom_get_stress_timing: nothing timed yet");
    float a = 0.f, b = 0.f, c = 0.f;
    GCHK(hipEventSynchronize(h->es[4]));
    GCHK(hipEventElapsedTime(&a, h->es[0], h->es[1]));
    GCHK(hipEventElapsedTime(&b, h->es[2], h->es[3]));
    GCHK(hipEventElapsedTime(&c, h->es[3], h->es[4]));
    if (bi_ms) *bi_ms = a;
    if (interp_ms) *interp_ms = b;
    if (sum_ms) *sum_ms = c;
    return 0;
}
