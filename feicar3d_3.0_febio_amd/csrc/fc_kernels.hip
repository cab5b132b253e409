// fc_kernels.hip -- gfx950 kernels of the pressure-Poisson hot path.
//
// All kernels are bandwidth-bound stencil / line-solve / gather work; MFMA
// is not used (the operator is a 7-point sparse stencil).  Each kernel cites
// the reference routine it implements.  T is double (FP64 path) or float
// (FP32 V-cycles of the mixed-precision path); masks are int8.
#include "fc_kernels.h"

#include <algorithm>

namespace fc {


__constant__ int c_incI[8] = {0, 0, 1, 1, 0, 0, 1, 1};   // GCM_MEMORY_ALLOCATE.F90:593-623
__constant__ int c_incJ[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__constant__ int c_incK[8] = {0, 0, 0, 0, 1, 1, 1, 1};

template <typename T> __device__ __forceinline__ T rcp(T x);
template <> __device__ __forceinline__ double rcp<double>(double x) { return 1.0 / x; }
template <> __device__ __forceinline__ float rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ __forceinline__ long idx(const LevelDims &d, int i, int j, int k) {
    return (long)i + d.sy * (long)j + d.sz * (long)k;
}

// Cell mask byte read by the line sweeps and the residuals.  Bit 0: the
// cell is held (an identity row: masked by igk, or a wall cell); bits 1..6:
// the -x, +x, -y, +y, -z, +z face of a free cell is a zero-flux wall (the
// correction V-cycles that treat bodies as walls).  The reference's igk
// masks are the special case with bit 0 only.
constexpr unsigned MB_HELD = 1u, MB_WALLS = 0x7eu;
constexpr unsigned MB_XM = 2u, MB_XP = 4u, MB_YM = 8u, MB_YP = 16u, MB_ZM = 32u, MB_ZP = 64u;

// Block coordinates and grid extents of one grid's share of a launch: the
// hardware's for a single-grid launch, or the slice of a batched launch that
// belongs to one request (serving runs the same phase of many grids in one
// launch).  Kernel bodies take it instead of reading blockIdx / gridDim.
struct VBlock {
    int x, y, z;
    int nx, ny, nz;
};
__device__ __forceinline__ VBlock hw_block()
{
    return VBlock{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, (int)gridDim.x, (int)gridDim.y, (int)gridDim.z};
}

// Row-strided kernels: block b owns the contiguous rows [b*R, (b+1)*R) of the
// (j,k) row list, R = ceil(nrow / number of blocks).  Blocks b and b+8 run on
// the same XCD (round-robin dispatch), so the k-1/k+1 neighbour rows of a
// block are read by a block sharing its L2.
__device__ __forceinline__ long row_begin(long nrow, const VBlock &vb)
{
    long R = (nrow + vb.nx - 1) / vb.nx;
    return (long)vb.x * R;
}
__device__ __forceinline__ long row_end(long nrow, const VBlock &vb)
{
    long R = (nrow + vb.nx - 1) / vb.nx;
    long e = ((long)vb.x + 1) * R;
    return e < nrow ? e : nrow;
}

// ---------------------------------------------------------------------------
// In-wavefront tridiagonal solve of one line by LPL lanes, E rows per lane
// (rows p*E .. p*E+E-1 of lane p held in registers).  Each lane reduces its
// interior rows, which turns its first and last rows into relations between
// the interface unknowns Y_p = x(first row) and X_p = x(last row); the Y's
// are eliminated with the neighbour lane's first-row relation (one shuffle),
// leaving a tridiagonal system in X_0..X_{LPL-1} (one row per lane) that is
// solved by parallel cyclic reduction across the lanes; Y_p and the interior
// rows follow by substitution.  No pivoting, as in the reference's tdma:
// the line systems of itsolv_mg are diagonally dominant (CORE_MG.F90:686,
// 783,878); the Crank-Nicolson advection-diffusion rows of itsolv_ad are not
// once the cell Courant number exceeds about 2, and there the partition
// solve's agreement with the sequential elimination is what the parity tests
// at Courant numbers up to 10 establish, not a stability guarantee.
// Parallel cyclic reduction of one row per lane, rows kept normalised to a
// unit diagonal: a X_{p-st} + X_p + c X_{p+st} = d, so a step exchanges three
// values with each neighbour and takes one reciprocal.
template <typename T>
__device__ __forceinline__ T pcr_solve(T RA, T RB, T RC, T RD, int p, int LPL)
{
    const T rB0 = rcp(RB);
    T a = RA * rB0, c = RC * rB0, dv = RD * rB0;
    for (int st = 1; st < LPL; st <<= 1) {
        T am = __shfl_up(a, st, LPL), cm = __shfl_up(c, st, LPL), dm = __shfl_up(dv, st, LPL);
        T ap = __shfl_down(a, st, LPL), cp = __shfl_down(c, st, LPL), dp = __shfl_down(dv, st, LPL);
        if (p < st) { am = 0; cm = 0; dm = 0; }
        if (p + st >= LPL) { ap = 0; cp = 0; dp = 0; }
        const T r = rcp((T)1 - a * cm - c * ap);
        const T na = -a * am * r, nc = -c * cp * r;
        dv = (dv - a * dm - c * dp) * r;
        a = na;
        c = nc;
    }
    return dv;
}

template <typename T, int E>
__device__ __forceinline__ void line_solve(const T (&oa)[E], const T (&ob)[E], const T (&oc)[E], const T (&od)[E],
                                           T (&x)[E], int p, int LPL)
{
    if constexpr (E == 1) {
        x[0] = pcr_solve<T>(p == 0 ? (T)0 : oa[0], ob[0], p == LPL - 1 ? (T)0 : oc[0], od[0], p, LPL);
    } else {
        // downward through rows 1..E-1 with x(first) kept as the left unknown
        T f[E], bb[E], dd[E], rb[E];
        f[1] = oa[1]; bb[1] = ob[1]; dd[1] = od[1]; rb[1] = rcp(bb[1]);
        for (int i = 2; i < E; ++i) {
            T m = oa[i] * rb[i - 1];
            f[i] = -m * f[i - 1];
            bb[i] = ob[i] - m * oc[i - 1];
            dd[i] = od[i] - m * dd[i - 1];
            rb[i] = rcp(bb[i]);
        }
        // last row: A2 Y_p + B2 X_p + C2 Y_{p+1} = D2
        const T A2 = f[E - 1], B2 = bb[E - 1], C2 = oc[E - 1], D2 = dd[E - 1];
        // upward through rows E-2..1 with x(last) kept as the right unknown;
        // first row: a_s X_{p-1} + B1 Y_p + C1 X_p = D1
        T B1, C1, D1;
        if constexpr (E > 2) {
            T bu = ob[E - 2], g = oc[E - 2], du = od[E - 2];
            for (int i = E - 3; i >= 1; --i) {
                T m = oc[i] * rcp(bu);
                bu = ob[i] - m * oa[i + 1];
                g = -m * g;
                du = od[i] - m * du;
            }
            T m = oc[0] * rcp(bu);
            B1 = ob[0] - m * oa[1];
            C1 = -m * g;
            D1 = od[0] - m * du;
        } else {
            B1 = ob[0]; C1 = oc[0]; D1 = od[0];
        }
        const T a_s = (p == 0) ? (T)0 : oa[0];
        const T rB1 = rcp(B1);
        T na = __shfl_down(a_s * rB1, 1, LPL), nc = __shfl_down(C1 * rB1, 1, LPL), nd = __shfl_down(D1 * rB1, 1, LPL);
        if (p == LPL - 1) { na = 0; nc = 0; nd = 0; }
        const T C2e = (p == LPL - 1) ? (T)0 : C2;
        T RA = -A2 * a_s * rB1;
        T RB = B2 - A2 * C1 * rB1 - C2e * na;
        T RC = -C2e * nc;
        T RD = D2 - A2 * D1 * rB1 - C2e * nd;
        T X = pcr_solve<T>(RA, RB, RC, RD, p, LPL);
        T Xm = __shfl_up(X, 1, LPL);
        if (p == 0) Xm = 0;
        T Y = (D1 - a_s * Xm - C1 * X) * rB1;
        x[E - 1] = X;
        x[0] = Y;
        for (int i = E - 2; i >= 1; --i) x[i] = (dd[i] - f[i] * Y - oc[i] * x[i + 1]) * rb[i];
    }
}

// ---------------------------------------------------------------------------
// set_outer_pressure_bc + set_outer_ghost_pressure, one rank
// (CORE_SET_OUTER_BC.F90:831-1042, 1044-1186).  One thread per face cell;
// the ghost-layer corner entries the reference never sets (bcyp/bczp at
// i=0 and i=nx_mg) are taken as zero, which only affects edge ghosts that
// no interior stencil reads.  Both x (y) faces periodic: the face value is
// the mean of the interior cells next to the two faces (CORE_SET_OUTER_BC.
// F90:1007-1026), so a ghost takes the value of the cell at the other end.
template <typename T>
__device__ __forceinline__ void bc_fill_at(const LevelDims &d, T *p, const BcInfo &bc, long t)
{
    const int nxm = d.nx, nym = d.ny, nzm = d.nz;
    const long nX = (long)(nym - 1) * (nzm - 1);
    const long nY = (long)(nxm + 1) * (nzm - 1);
    const long nZ = (long)(nxm + 1) * (nym - 1);
    const double h = bc.homo ? 0.0 : 1.0;
    const int fny2 = bc.fny + 2, fnx2 = bc.fnx + 2;
    auto bcx0 = [&](int j, int k) -> double {
        if (bc.pbc[0] == 1) return bc.homo ? 0.0 : bc.ppp[0];
        return (double)p[idx(d, 1, j, k)] - (bc.homo ? 0.0 : bc.pgx1[j + (long)fny2 * k]) * bc.dxc1 / 2.0 * h;
    };
    auto bcx1 = [&](int j, int k) -> double {
        if (bc.pbc[1] == 1) return bc.homo ? 0.0 : bc.ppp[1];
        return (double)p[idx(d, nxm - 1, j, k)] + (bc.homo ? 0.0 : bc.pgx2[j + (long)fny2 * k]) * bc.dxcn / 2.0 * h;
    };
    auto xghost = [&](int i, int j, int k) -> double {   // value set for (i,j,k), i in {0, nxm}
        if (bc.per & 1) {
            const double a = (double)p[idx(d, 1, j, k)], b = (double)p[idx(d, nxm - 1, j, k)];
            const double m = (a + b) / 2.0;
            return (i == 0) ? 2.0 * m - a : 2.0 * m - b;
        }
        if (i == 0) return 2.0 * bcx0(j, k) - (double)p[idx(d, 1, j, k)];
        return 2.0 * bcx1(j, k) - (double)p[idx(d, nxm - 1, j, k)];
    };
    auto val = [&](int i, int j, int k) -> double {       // interior value or x-ghost
        if (i == 0 || i == nxm) return xghost(i, j, k);
        return (double)p[idx(d, i, j, k)];
    };
    if (t < nX) {
        int j = 1 + (int)(t % (nym - 1)), k = 1 + (int)(t / (nym - 1));
        double g0 = xghost(0, j, k), g1 = xghost(nxm, j, k);
        p[idx(d, 0, j, k)] = (T)g0;
        p[idx(d, nxm, j, k)] = (T)g1;
        return;
    }
    t -= nX;
    if (t < nY) {
        int i = (int)(t % (nxm + 1)), k = 1 + (int)(t / (nxm + 1));
        bool inner = (i >= 1 && i <= nxm - 1);
        double b0 = 0.0, b1 = 0.0;
        double v1 = val(i, 1, k), vn = val(i, nym - 1, k);
        if (inner) {
            if (bc.pbc[2] == 1) b0 = bc.homo ? 0.0 : bc.ppp[2];
            else b0 = v1 - (bc.homo ? 0.0 : bc.pgy1[i + (long)fnx2 * k]) * bc.dyc1 / 2.0 * h;
            if (bc.pbc[3] == 1) b1 = bc.homo ? 0.0 : bc.ppp[3];
            else b1 = vn + (bc.homo ? 0.0 : bc.pgy2[i + (long)fnx2 * k]) * bc.dycn / 2.0 * h;
            if (bc.per & 2) { b0 = (v1 + vn) / 2.0; b1 = b0; }
        }
        p[idx(d, i, 0, k)] = (T)(2.0 * b0 - v1);
        p[idx(d, i, nym, k)] = (T)(2.0 * b1 - vn);
        return;
    }
    t -= nY;
    if (t < nZ) {
        int i = (int)(t % (nxm + 1)), j = 1 + (int)(t / (nxm + 1));
        bool inner = (i >= 1 && i <= nxm - 1);
        double b0 = 0.0, b1 = 0.0;
        double v1 = val(i, j, 1), vn = val(i, j, nzm - 1);
        if (inner) {
            if (bc.pbc[4] == 1) b0 = bc.homo ? 0.0 : bc.ppp[4];
            else b0 = v1 - (bc.homo ? 0.0 : bc.pgz1[i + (long)fnx2 * j]) * bc.dzc1 / 2.0 * h;
            if (bc.pbc[5] == 1) b1 = bc.homo ? 0.0 : bc.ppp[5];
            else b1 = vn + (bc.homo ? 0.0 : bc.pgz2[i + (long)fnx2 * j]) * bc.dzcn / 2.0 * h;
        }
        p[idx(d, i, j, 0)] = (T)(2.0 * b0 - v1);
        p[idx(d, i, j, nzm)] = (T)(2.0 * b1 - vn);
    }
}

template <typename T>
__device__ __forceinline__ void dk_bc_fill(const LevelDims &d, T *p, const BcInfo &bc, const VBlock &vb)
{
    bc_fill_at<T>(d, p, bc, (long)vb.x * blockDim.x + threadIdx.x);
}
template <typename T>
__global__ void k_bc_fill(LevelDims d, T *p, BcInfo bc)
{
    dk_bc_fill<T>(d, p, bc, hw_block());
}

// number of face cells bc_fill_at covers (t ranges over 0..n-1)
__host__ __device__ inline long bc_cells(const LevelDims &d)
{
    return (long)(d.ny - 1) * (d.nz - 1) + (long)(d.nx + 1) * (d.nz - 1) + (long)(d.nx + 1) * (d.ny - 1);
}

// The face-ghost values bc_fill_at sets, as functions of the one interior
// value v each depends on (x faces: v = p(1,j,k) or p(nx-1,j,k); y faces at
// interior i: v = p(i,1,k) or p(i,ny-1,k); z faces at interior i: v =
// p(i,j,1) or p(i,j,nz-1)).  side 0 = low face, 1 = high face.
__device__ __forceinline__ double ghost_x(const BcInfo &bc, int side, int j, int k, double v)
{
    const double h = bc.homo ? 0.0 : 1.0;
    const long o = j + (long)(bc.fny + 2) * k;
    double b;
    if (side == 0) b = (bc.pbc[0] == 1) ? (bc.homo ? 0.0 : bc.ppp[0]) : v - (bc.homo ? 0.0 : bc.pgx1[o]) * bc.dxc1 / 2.0 * h;
    else b = (bc.pbc[1] == 1) ? (bc.homo ? 0.0 : bc.ppp[1]) : v + (bc.homo ? 0.0 : bc.pgx2[o]) * bc.dxcn / 2.0 * h;
    return 2.0 * b - v;
}
__device__ __forceinline__ double ghost_y(const BcInfo &bc, int side, int i, int k, double v)
{
    const double h = bc.homo ? 0.0 : 1.0;
    const long o = i + (long)(bc.fnx + 2) * k;
    double b;
    if (side == 0) b = (bc.pbc[2] == 1) ? (bc.homo ? 0.0 : bc.ppp[2]) : v - (bc.homo ? 0.0 : bc.pgy1[o]) * bc.dyc1 / 2.0 * h;
    else b = (bc.pbc[3] == 1) ? (bc.homo ? 0.0 : bc.ppp[3]) : v + (bc.homo ? 0.0 : bc.pgy2[o]) * bc.dycn / 2.0 * h;
    return 2.0 * b - v;
}
__device__ __forceinline__ double ghost_z(const BcInfo &bc, int side, int i, int j, double v)
{
    const double h = bc.homo ? 0.0 : 1.0;
    const long o = i + (long)(bc.fnx + 2) * j;
    double b;
    if (side == 0) b = (bc.pbc[4] == 1) ? (bc.homo ? 0.0 : bc.ppp[4]) : v - (bc.homo ? 0.0 : bc.pgz1[o]) * bc.dzc1 / 2.0 * h;
    else b = (bc.pbc[5] == 1) ? (bc.homo ? 0.0 : bc.ppp[5]) : v + (bc.homo ? 0.0 : bc.pgz2[o]) * bc.dzcn / 2.0 * h;
    return 2.0 * b - v;
}

// After interior cell (i,j,k) took the value v: rewrite every ghost cell
// bc_fill_at derives from it (faces, and the edge cells the y- and z-face
// passes derive from x-face ghosts), so the ghosts stay what a BC fill now
// would give.  Each such ghost depends on this one cell only, except on
// periodic faces, whose ghosts depend on the cells at both ends: those are
// left alone here and refreshed by a boundary fill after the half sweep.
template <typename T>
__device__ __forceinline__ void bc_post(const LevelDims &d, T *p, const BcInfo &bc, int i, int j, int k, T v)
{
    const int nxm = d.nx, nym = d.ny, nzm = d.nz;
    const double vd = (double)v;
    double gx = 0.0;
    const bool xs = ((i == 1) || (i == nxm - 1)) && !(bc.per & 1);
    if (bc.per & 1) {
        if (j == 1 && !(bc.per & 2)) p[idx(d, i, 0, k)] = (T)ghost_y(bc, 0, i, k, vd);
        if (j == nym - 1 && !(bc.per & 2)) p[idx(d, i, nym, k)] = (T)ghost_y(bc, 1, i, k, vd);
        if (k == 1) p[idx(d, i, j, 0)] = (T)ghost_z(bc, 0, i, j, vd);
        if (k == nzm - 1) p[idx(d, i, j, nzm)] = (T)ghost_z(bc, 1, i, j, vd);
        return;
    }
    if (i == 1) { gx = ghost_x(bc, 0, j, k, vd); p[idx(d, 0, j, k)] = (T)gx; }
    else if (i == nxm - 1) { gx = ghost_x(bc, 1, j, k, vd); p[idx(d, nxm, j, k)] = (T)gx; }
    const int ie = (i == 1) ? 0 : nxm;                 // x-face edge index
    if (j == 1 && !(bc.per & 2)) {
        p[idx(d, i, 0, k)] = (T)ghost_y(bc, 0, i, k, vd);
        if (xs) p[idx(d, ie, 0, k)] = (T)(-gx);
    }
    if (j == nym - 1 && !(bc.per & 2)) {
        p[idx(d, i, nym, k)] = (T)ghost_y(bc, 1, i, k, vd);
        if (xs) p[idx(d, ie, nym, k)] = (T)(-gx);
    }
    if (k == 1) {
        p[idx(d, i, j, 0)] = (T)ghost_z(bc, 0, i, j, vd);
        if (xs) p[idx(d, ie, j, 0)] = (T)(-gx);
    }
    if (k == nzm - 1) {
        p[idx(d, i, j, nzm)] = (T)ghost_z(bc, 1, i, j, vd);
        if (xs) p[idx(d, ie, j, nzm)] = (T)(-gx);
    }
}

// What the wall faces of mask byte b add to a 7-point residual: a wall face
// f removes c_f (phi_f - phi_c) from A phi, so the residual gains it back.
template <typename T>
__device__ __forceinline__ T wall_terms(unsigned b, T pc, T xm, T xp, T ym, T yp, T zm, T zp, T cxm, T cxp, T cym,
                                        T cyp, T czm, T czp)
{
    T a = (T)0;
    if (b & MB_XM) a += cxm * (xm - pc);
    if (b & MB_XP) a += cxp * (xp - pc);
    if (b & MB_YM) a += cym * (ym - pc);
    if (b & MB_YP) a += cyp * (yp - pc);
    if (b & MB_ZM) a += czm * (zm - pc);
    if (b & MB_ZP) a += czp * (zp - pc);
    return a;
}

// ---------------------------------------------------------------------------
// residual_mg with the res=rhs copy (CORE_MG.F90:956-1043, 1672-1678):
// res = (rhs - A phi)*(1-igk) on interior cells, res(nx_mg,j,k) = rhs.
template <typename T>
__device__ __forceinline__ void residual_at(const LevelDims &d, const Metrics<T> &m, const T *__restrict__ phi,
                                            const T *__restrict__ rhs, T *__restrict__ res,
                                            const int8_t *__restrict__ igk, int i, int j, int k)
{
    long q = idx(d, i, j, k);
    if (i == d.nx) { res[q] = rhs[q]; return; }
    T bmx = m.amx[i], bpx = m.apx[i], bmy = m.amy[j], bpy = m.apy[j], bmz = m.amz[k], bpz = m.apz[k];
    T bc = (m.acx[i] + m.acy[j]) + m.acz[k];
    T r = rhs[q] - phi[q] * bc - phi[q - 1] * bmx - phi[q + 1] * bpx - phi[q - d.sy] * bmy - phi[q + d.sy] * bpy
          - phi[q - d.sz] * bmz - phi[q + d.sz] * bpz;
    const unsigned b = (unsigned char)igk[q];
    if (b & MB_WALLS) r += wall_terms<T>(b, phi[q], phi[q - 1], phi[q + 1], phi[q - d.sy], phi[q + d.sy],
                                         phi[q - d.sz], phi[q + d.sz], bmx, bpx, bmy, bpy, bmz, bpz);
    res[q] = (b & MB_HELD) ? (T)0 : r;
}

// i = 1..nx (the i = nx column takes the res = rhs copy)
template <typename T>
__global__ void __launch_bounds__(256) k_residual(LevelDims d, Metrics<T> m, const T *__restrict__ phi,
                                                  const T *__restrict__ rhs, T *__restrict__ res,
                                                  const int8_t *__restrict__ igk)
{
    int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i > d.nx) return;
    residual_at<T>(d, m, phi, rhs, res, igk, i, 1 + blockIdx.y, 1 + blockIdx.z);
}

// ---------------------------------------------------------------------------
// calc_residual (CORE_POISSON.F90:460-551): per-block partial sums of
// (res*flag)^2 and of (rhs*flag)^2; the final pass adds them in block order
// (deterministic).  flag = 1 - igk equals (1-iblank)(1-ghostCellMark) on the
// finest level.
template <typename T>
__global__ void __launch_bounds__(256) k_resnorm(LevelDims d, Metrics<T> m, const T *__restrict__ phi,
                                                 const T *__restrict__ rhs, const int8_t *__restrict__ igk,
                                                 double *__restrict__ partial)
{
    __shared__ double s1[256], s2[256];
    int j = 1 + blockIdx.x;
    int k = 1 + blockIdx.y;
    double a1 = 0.0, a2 = 0.0;
    T bmy = m.amy[j], bpy = m.apy[j], bmz = m.amz[k], bpz = m.apz[k];
    for (int i = 1 + threadIdx.x; i <= d.nx - 1; i += blockDim.x) {
        long q = idx(d, i, j, k);
        if (igk[q]) continue;
        T bc = (m.acx[i] + m.acy[j]) + m.acz[k];
        T r = rhs[q] - phi[q] * bc - phi[q - 1] * m.amx[i] - phi[q + 1] * m.apx[i] - phi[q - d.sy] * bmy
              - phi[q + d.sy] * bpy - phi[q - d.sz] * bmz - phi[q + d.sz] * bpz;
        a1 += (double)r * (double)r;
        a2 += (double)rhs[q] * (double)rhs[q];
    }
    s1[threadIdx.x] = a1;
    s2[threadIdx.x] = a2;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) { s1[threadIdx.x] += s1[threadIdx.x + s]; s2[threadIdx.x] += s2[threadIdx.x + s]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        long b = blockIdx.x + (long)gridDim.x * blockIdx.y;
        partial[2 * b] = s1[0];
        partial[2 * b + 1] = s2[0];
    }
}

// Deterministic final reduction of n (sum or max) partials into out[slot].
__global__ void k_reduce(const double *__restrict__ partial, long n, int stride, int comp, int op,
                         double *__restrict__ out)
{
    __shared__ double s[1024];
    double a = 0.0;
    for (long b = threadIdx.x; b < n; b += blockDim.x) {
        double v = partial[b * stride + comp];
        a = op ? fmax(a, v) : a + v;
    }
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] = op ? fmax(s[threadIdx.x], s[threadIdx.x + w]) : s[threadIdx.x] + s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

// calc_residual2 (CORE_POISSON.F90:679-724): max |p - p_temp|*(1-iblank).
__global__ void __launch_bounds__(256) k_maxdiff(LevelDims d, const double *__restrict__ p,
                                                 const double *__restrict__ pt, const int8_t *__restrict__ iblank,
                                                 double *__restrict__ partial)
{
    __shared__ double s[256];
    int j = 1 + blockIdx.x, k = 1 + blockIdx.y;
    double a = 0.0;
    for (int i = 1 + threadIdx.x; i <= d.nx - 1; i += blockDim.x) {
        long q = idx(d, i, j, k);
        double v = iblank[q] ? 0.0 : fabs(p[q] - pt[q]);
        a = fmax(a, v);
    }
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x + (long)gridDim.x * blockIdx.y] = s[0];
}

// Mean pinning of solve_poisson (CORE_POISSON.F90:129-150): subtract the
// value at (1,1,1), read once into *corner, from i=0..nx, j=0..ny, k=0..nz.
__global__ void k_read_corner(LevelDims d, const double *p, double *corner) { *corner = p[idx(d, 1, 1, 1)]; }
__global__ void k_pin(LevelDims d, double *p, const double *corner)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y, k = blockIdx.z;
    if (i > d.nx) return;
    p[idx(d, i, j, k)] -= *corner;
}

// ---------------------------------------------------------------------------
// mg_inject (CORE_MG.F90:1061-1183): res(N) -> rhs(N+1).  Coarse rhs ghosts
// stay zero (the reference zeroes rr2 first).
template <typename T>
__device__ __forceinline__ void restrict_at(const LevelDims &f, const LevelDims &c, const Metrics<T> &mf,
                                            const Metrics<T> &mc, const T *__restrict__ rr1, T *__restrict__ rr2,
                                            const int8_t *__restrict__ ibkf, const int8_t *__restrict__ rflag, int i,
                                            int j, int k)
{
    const int ic = c.ic, jc = c.jc, kc = c.kc;
    int i1 = (1 - ic) * i + ic * (2 * i - 1), i2 = (1 - ic) * i + ic * (2 * i);
    int j1 = (1 - jc) * j + jc * (2 * j - 1), j2 = (1 - jc) * j + jc * (2 * j);
    int k1 = (1 - kc) * k + kc * (2 * k - 1), k2 = (1 - kc) * k + kc * (2 * k);
    T aX = (mc.xc[i] - mf.xc[i1]) * mf.dxcinv[i1 + 1];
    T aY = (mc.yc[j] - mf.yc[j1]) * mf.dycinv[j1 + 1];
    T aZ = (mc.zc[k] - mf.zc[k1]) * mf.dzcinv[k1 + 1];
    const T one = (T)1;
    long q = idx(f, i1, j1, k1);
    T r_tmp = rr1[q] * (one - aX) * (one - aY) * (one - aZ) + rr1[q + 1] * aX * (one - aY) * (one - aZ)
              + rr1[q + f.sy] * (one - aX) * aY * (one - aZ) + rr1[q + f.sz] * (one - aX) * (one - aY) * aZ
              + rr1[q + 1 + f.sy] * aX * aY * (one - aZ) + rr1[q + 1 + f.sz] * aX * (one - aY) * aZ
              + rr1[q + f.sy + f.sz] * (one - aX) * aY * aZ + rr1[q + 1 + f.sy + f.sz] * aX * aY * aZ;
    T out = r_tmp;
    // volume branch only where a child is iblank (rare; rflag precomputed)
    if (rflag && rflag[idx(c, i, j, k)]) {
        T vol0 = 0, sum0 = 0;
        for (int kk = k1; kk <= k2; ++kk)
            for (int jj = j1; jj <= j2; ++jj)
                for (int ii = i1; ii <= i2; ++ii) {
                    long qq = idx(f, ii, jj, kk);
                    if (ibkf[qq] != 1) {
                        T vol_i = (mf.x[ii + 1] - mf.x[ii]) * (mf.y[jj + 1] - mf.y[jj]) * (mf.z[kk + 1] - mf.z[kk]);
                        vol0 = vol0 + vol_i;
                        sum0 = rr1[qq] * vol_i;          // assignment, as CORE_MG.F90:1159
                    }
                }
        T vol_c = one / mf.dxinv[i1] / mf.dyinv[j1] / mf.dzinv[k1];
        out = (vol0 > (T)0.01f * vol_c) ? sum0 / vol0 : (T)0;
    }
    rr2[idx(c, i, j, k)] = out;
}

// The value residual_mg leaves in res(i,j,k) for a restriction child: the
// masked residual on interior cells, the res = rhs copy on the i = nx column
// (CORE_MG.F90:1672-1678), and zero on the j = ny / k = nz ghost rows, which
// residual_mg never writes.
template <typename T>
__device__ __forceinline__ T res_value(const LevelDims &d, const Metrics<T> &m, const T *__restrict__ phi,
                                       const T *__restrict__ rhs, const int8_t *__restrict__ igk, int i, int j, int k)
{
    if (j > d.ny - 1 || k > d.nz - 1) return (T)0;
    const long q = idx(d, i, j, k);
    if (i == d.nx) return rhs[q];
    T bmx = m.amx[i], bpx = m.apx[i], bmy = m.amy[j], bpy = m.apy[j], bmz = m.amz[k], bpz = m.apz[k];
    T bc = (m.acx[i] + m.acy[j]) + m.acz[k];
    T r = rhs[q] - phi[q] * bc - phi[q - 1] * bmx - phi[q + 1] * bpx - phi[q - d.sy] * bmy - phi[q + d.sy] * bpy
          - phi[q - d.sz] * bmz - phi[q + d.sz] * bpz;
    const unsigned b = (unsigned char)igk[q];
    if (b & MB_WALLS) r += wall_terms<T>(b, phi[q], phi[q - 1], phi[q + 1], phi[q - d.sy], phi[q + d.sy],
                                         phi[q - d.sz], phi[q + d.sz], bmx, bpx, bmy, bpy, bmz, bpz);
    return (b & MB_HELD) ? (T)0 : r;
}

// Four and two consecutive elements with one vector load (the row pitch is a
// multiple of 16 elements, so a 4-aligned index is 16-byte aligned in FP32
// and 32-byte aligned in FP64; an even index is 8 / 16-byte aligned).
template <typename T> struct V4 { T v[4]; };
template <typename T> struct Pair;
template <> struct Pair<float> { using type = float2; };
template <> struct Pair<double> { using type = double2; };

// raw buffer access (gfx9 resource word 3 = 0x00020000); out-of-range loads
// return zero, offsets are bytes
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *p, long bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)bytes, 0x00020000);
}
template <typename T> struct Buf;
template <> struct Buf<float> {
    static constexpr int shift = 2;
    __device__ static float ld(__amdgpu_buffer_rsrc_t r, int v, int so)
    {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, v, so, 0));
    }
    __device__ static void st(__amdgpu_buffer_rsrc_t r, float x, int v, int so)
    {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, v, so, 0);
    }
};
template <> struct Buf<double> {
    static constexpr int shift = 3;
    __device__ static double ld(__amdgpu_buffer_rsrc_t r, int v, int so)
    {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, v, so, 0));
    }
    __device__ static void st(__amdgpu_buffer_rsrc_t r, double x, int v, int so)
    {
        typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), r, v, so, 0);
    }
};

// four consecutive elements from byte offset v with whole 16-byte buffer
// loads: the compiler cannot split them per element and sink a part into a
// per-element branch
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
template <typename T> __device__ __forceinline__ V4<T> bld4(__amdgpu_buffer_rsrc_t r, int v);
template <> __device__ __forceinline__ V4<float> bld4<float>(__amdgpu_buffer_rsrc_t r, int v)
{
    const float4 a = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, v, 0, 0));
    return V4<float>{{a.x, a.y, a.z, a.w}};
}
template <> __device__ __forceinline__ V4<double> bld4<double>(__amdgpu_buffer_rsrc_t r, int v)
{
    const double2 a = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, v, 0, 0));
    const double2 b = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, v + 16, 0, 0));
    return V4<double>{{a.x, a.y, b.x, b.y}};
}
// the metric arrays hold 4 entries from any multiple of 4 up to n+1, and at
// least n+6 entries in all
template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t met_rsrc(const T *a, int n) { return buf_rsrc(a, (long)(n + 6) * sizeof(T)); }
template <typename T> __device__ __forceinline__ V4<T> ldv4(const T *p);
template <> __device__ __forceinline__ V4<float> ldv4<float>(const float *p)
{
    const float4 a = *(const float4 *)p;
    return V4<float>{{a.x, a.y, a.z, a.w}};
}
template <> __device__ __forceinline__ V4<double> ldv4<double>(const double *p)
{
    const double2 a = *(const double2 *)p, b = *(const double2 *)(p + 2);
    return V4<double>{{a.x, a.y, b.x, b.y}};
}

// res_value of the four cells i0..i0+3 of row (j,k), i0 a multiple of 4 and
// i0 <= nx; cells beyond nx are left 0 (never used).
template <typename T>
__device__ __forceinline__ void res_chunk(const LevelDims &d, const Metrics<T> &m, const T *__restrict__ phi,
                                          const T *__restrict__ rhs, const int8_t *__restrict__ igk, int i0, int j,
                                          int k, T (&r)[4])
{
#pragma unroll
    for (int e = 0; e < 4; ++e) r[e] = (T)0;
    if (j > d.ny - 1 || k > d.nz - 1) return;
    const int sh = Buf<T>::shift;
    const long bytes = d.n << sh;
    const __amdgpu_buffer_rsrc_t rp = buf_rsrc(phi, bytes), rr_ = buf_rsrc(rhs, bytes), rg = buf_rsrc(igk, d.n);
    const int q0 = (int)idx(d, i0, j, k), v = q0 << sh, sy = (int)d.sy << sh, sz = (int)d.sz << sh;
    const V4<T> pc = bld4<T>(rp, v), pjm = bld4<T>(rp, v - sy), pjp = bld4<T>(rp, v + sy), pkm = bld4<T>(rp, v - sz),
                pkp = bld4<T>(rp, v + sz), rv = bld4<T>(rr_, v);
    const T pl = Buf<T>::ld(rp, v - (1 << sh), 0), pr = Buf<T>::ld(rp, v + (4 << sh), 0);
    const unsigned g = __builtin_amdgcn_raw_buffer_load_b32(rg, q0, 0, 0);
    const T bmy = m.amy[j], bpy = m.apy[j], bmz = m.amz[k], bpz = m.apz[k];
    const T acy = m.acy[j], acz = m.acz[k];
    const int vm = i0 << sh;
    const V4<T> amx = bld4<T>(met_rsrc(m.amx, d.nx), vm), apx = bld4<T>(met_rsrc(m.apx, d.nx), vm),
                acx = bld4<T>(met_rsrc(m.acx, d.nx), vm);
    // branch-free (no load is sunk into a per-element branch)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int i = i0 + e;
        const T left = (e == 0) ? pl : pc.v[e - 1], right = (e == 3) ? pr : pc.v[e + 1];
        const T bc = (acx.v[e] + acy) + acz;
        const T rr = rv.v[e] - pc.v[e] * bc - left * amx.v[e] - right * apx.v[e] - pjm.v[e] * bmy -
                     pjp.v[e] * bpy - pkm.v[e] * bmz - pkp.v[e] * bpz;
        const unsigned b = (g >> (8 * e)) & 0xffu;
        r[e] = (i == d.nx) ? rv.v[e] : (i >= 1 && i < d.nx && !(b & MB_HELD)) ? rr : (T)0;
    }
    // wall faces (rare: only cells next to a body in the correction cycles)
    if (g & 0x7e7e7e7eu) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const unsigned b = (g >> (8 * e)) & 0xffu;
            const int i = i0 + e;
            if ((b & MB_WALLS) && !(b & MB_HELD) && i >= 1 && i < d.nx) {
                const T left = (e == 0) ? pl : pc.v[e - 1], right = (e == 3) ? pr : pc.v[e + 1];
                r[e] += wall_terms<T>(b, pc.v[e], left, right, pjm.v[e], pjp.v[e], pkm.v[e], pkp.v[e], amx.v[e],
                                      apx.v[e], bmy, bpy, bmz, bpz);
            }
        }
    }
}

// residual_mg on level f fused with mg_inject into level c (CORE_MG.F90:
// 956-1043, 1061-1183) and the zeroing of the coarse phi (CORE_MG.F90:1539-
// 1540 per V-cycle).  A block takes 64 x 4 coarse cells of one coarse plane.
// Its threads first compute the residuals of all the fine children (the
// values res would hold, res_value) four consecutive cells at a time into
// LDS, then each thread restricts the eight children (i1..i1+1, j1..j1+1,
// k1..k1+1) of its coarse cell.  The fine tile is at most 2 planes x 8 rows
// x 33 chunks.
constexpr int RR_ROWS = 8, RR_PLANES = 2, RR_W = 33 * 4;
template <typename T>
__device__ __forceinline__ void dk_resrestrict(const LevelDims &f, const LevelDims &c, const Metrics<T> &mf,
                                               const Metrics<T> &mc, const T *__restrict__ phi_f,
                                               const T *__restrict__ rhs_f, const int8_t *__restrict__ igk_f,
                                               const int8_t *__restrict__ ibkf, const int8_t *__restrict__ rflag,
                                               T *__restrict__ rhs_c, T *__restrict__ phi_c, const VBlock &vb)
{
    __shared__ T rs[RR_PLANES * RR_ROWS * RR_W];
    const int ic = c.ic, jc = c.jc, kc = c.kc;
    const int ci0 = 1 + vb.x * 64, cj0 = 1 + vb.y * 4, k = 1 + vb.z;
    const int ci1 = min(ci0 + 63, c.nx - 1), cj1 = min(cj0 + 3, c.ny - 1);
    // fine tile: the children i1 .. i1+1 of the block's first .. last coarse cell
    const int fi0 = ic ? 2 * ci0 - 1 : ci0, fi1 = (ic ? 2 * ci1 - 1 : ci1) + 1;
    const int fj0 = jc ? 2 * cj0 - 1 : cj0, fj1 = (jc ? 2 * cj1 - 1 : cj1) + 1;
    const int fk0 = kc ? 2 * k - 1 : k;
    const int cb = fi0 >> 2, ncx = (fi1 >> 2) - cb + 1, nry = fj1 - fj0 + 1;
    const int nch = ncx * nry * RR_PLANES;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    for (int t = tid; t < nch; t += blockDim.x * blockDim.y) {
        const int cx = t % ncx, rr = t / ncx;
        const int jy = rr % nry, kz = rr / nry;
        T r4[4];
        res_chunk<T>(f, mf, phi_f, rhs_f, igk_f, 4 * (cb + cx), fj0 + jy, fk0 + kz, r4);
        T *o = rs + (kz * RR_ROWS + jy) * RR_W + 4 * cx;
#pragma unroll
        for (int e = 0; e < 4; ++e) o[e] = r4[e];
    }
    __syncthreads();
    const int i = ci0 + threadIdx.x, j = cj0 + threadIdx.y;
    if (i > c.nx - 1 || j > c.ny - 1) return;
    const int i1 = (1 - ic) * i + ic * (2 * i - 1), i2 = (1 - ic) * i + ic * (2 * i);
    const int j1 = (1 - jc) * j + jc * (2 * j - 1), j2 = (1 - jc) * j + jc * (2 * j);
    const int k1 = (1 - kc) * k + kc * (2 * k - 1), k2 = (1 - kc) * k + kc * (2 * k);
    T rv[2][2][2];
    const T *base = rs + (j1 - fj0) * RR_W + (i1 - 4 * cb);
#pragma unroll
    for (int dk = 0; dk < 2; ++dk)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int di = 0; di < 2; ++di) rv[dk][dj][di] = base[(dk * RR_ROWS + dj) * RR_W + di];
    T aX = (mc.xc[i] - mf.xc[i1]) * mf.dxcinv[i1 + 1];
    T aY = (mc.yc[j] - mf.yc[j1]) * mf.dycinv[j1 + 1];
    T aZ = (mc.zc[k] - mf.zc[k1]) * mf.dzcinv[k1 + 1];
    const T one = (T)1;
    T out = rv[0][0][0] * (one - aX) * (one - aY) * (one - aZ) + rv[0][0][1] * aX * (one - aY) * (one - aZ)
            + rv[0][1][0] * (one - aX) * aY * (one - aZ) + rv[1][0][0] * (one - aX) * (one - aY) * aZ
            + rv[0][1][1] * aX * aY * (one - aZ) + rv[1][0][1] * aX * (one - aY) * aZ
            + rv[1][1][0] * (one - aX) * aY * aZ + rv[1][1][1] * aX * aY * aZ;
    if (rflag && rflag[idx(c, i, j, k)]) {
        T vol0 = 0, sum0 = 0;
        // children k1..k2, j1..j2, i1..i2 in the reference's loop order;
        // constant indices keep rv in registers
#pragma unroll
        for (int dk = 0; dk < 2; ++dk)
#pragma unroll
            for (int dj = 0; dj < 2; ++dj)
#pragma unroll
                for (int di = 0; di < 2; ++di) {
                    const int ii = i1 + di, jj = j1 + dj, kk = k1 + dk;
                    if (ii > i2 || jj > j2 || kk > k2) continue;
                    if (ibkf[idx(f, ii, jj, kk)] != 1) {
                        T vol_i = (mf.x[ii + 1] - mf.x[ii]) * (mf.y[jj + 1] - mf.y[jj]) * (mf.z[kk + 1] - mf.z[kk]);
                        vol0 = vol0 + vol_i;
                        sum0 = rv[dk][dj][di] * vol_i;   // assignment, as CORE_MG.F90:1159
                    }
                }
        T vol_c = one / mf.dxinv[i1] / mf.dyinv[j1] / mf.dzinv[k1];
        out = (vol0 > (T)0.01f * vol_c) ? sum0 / vol0 : (T)0;
    }
    const long qc = idx(c, i, j, k);
    rhs_c[qc] = out;
    phi_c[qc] = (T)0;
    // With the coarse interior zero, the homogeneous outer-BC fill
    // (set_outer_pressure_bc, CORE_SET_OUTER_BC.F90:831-1186) of the coarse
    // level sets +0 on exactly these face ghosts (bc_fill_at: x faces at
    // interior j,k; y faces at i = 0..nx, interior k; z faces at i = 0..nx,
    // interior j).  Writing them here makes that fill unnecessary before
    // the level's first smoothing.
    const bool xl = (i == 1), xh = (i == c.nx - 1);
    if (xl) phi_c[idx(c, 0, j, k)] = (T)0;
    if (xh) phi_c[idx(c, c.nx, j, k)] = (T)0;
    if (j == 1 || j == c.ny - 1) {
        const int jg = (j == 1) ? 0 : c.ny;
        phi_c[idx(c, i, jg, k)] = (T)0;
        if (xl) phi_c[idx(c, 0, jg, k)] = (T)0;
        if (xh) phi_c[idx(c, c.nx, jg, k)] = (T)0;
        if (c.ny - 1 == 1) {     // a single interior row: both y faces
            phi_c[idx(c, i, c.ny, k)] = (T)0;
            if (xl) phi_c[idx(c, 0, c.ny, k)] = (T)0;
            if (xh) phi_c[idx(c, c.nx, c.ny, k)] = (T)0;
        }
    }
    if (k == 1 || k == c.nz - 1) {
        const int kg = (k == 1) ? 0 : c.nz;
        phi_c[idx(c, i, j, kg)] = (T)0;
        if (xl) phi_c[idx(c, 0, j, kg)] = (T)0;
        if (xh) phi_c[idx(c, c.nx, j, kg)] = (T)0;
        if (c.nz - 1 == 1) {     // a single interior plane: both z faces
            phi_c[idx(c, i, j, c.nz)] = (T)0;
            if (xl) phi_c[idx(c, 0, j, c.nz)] = (T)0;
            if (xh) phi_c[idx(c, c.nx, j, c.nz)] = (T)0;
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_resrestrict(LevelDims f, LevelDims c, Metrics<T> mf, Metrics<T> mc,
                                                     const T *__restrict__ phi_f, const T *__restrict__ rhs_f,
                                                     const int8_t *__restrict__ igk_f,
                                                     const int8_t *__restrict__ ibkf,
                                                     const int8_t *__restrict__ rflag, T *__restrict__ rhs_c,
                                                     T *__restrict__ phi_c)
{
    dk_resrestrict<T>(f, c, mf, mc, phi_f, rhs_f, igk_f, ibkf, rflag, rhs_c, phi_c, hw_block());
}

template <typename T>
__global__ void __launch_bounds__(256) k_restrict(LevelDims f, LevelDims c, Metrics<T> mf, Metrics<T> mc,
                                                  const T *__restrict__ rr1, T *__restrict__ rr2,
                                                  const int8_t *__restrict__ ibkf, const int8_t *__restrict__ rflag)
{
    int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    if (i > c.nx - 1 || j > c.ny - 1) return;
    restrict_at<T>(f, c, mf, mc, rr1, rr2, ibkf, rflag, i, j, 1 + blockIdx.z);
}

// mg_interp (CORE_MG.F90:1200-1351): phi(N) added into phi(N-1); skips
// fine cells on two domain faces (:1254-1273) and fine ibk cells (:1275).
// Thread (gi,gj,gk) takes the fine cells whose
// interpolation block starts at coarse cell (gi,gj,gk) -- fine i = 2gi,
// 2gi+1 where x is coarsened, i = gi where it is not, likewise in j and k.
// They share the block's eight coarse values, its pflag and (where pflag is
// set) its volume-weighted value, so each is loaded or computed once; the
// fine cells of a group move with paired loads and stores in x.
// walls != 0 (correction cycles with bodies as walls; ibkf / ibkc are then
// the wall masks, bit 0 = held cell, pflag marks blocks with a held coarse
// cell): where the block has held coarse cells, each fine cell takes the
// trilinear weights of the free ones, renormalised to sum to one (no
// correction if none is free), instead of the reference's volume branch.
template <typename T, int FX>
__device__ __forceinline__ void prolong_group(const LevelDims &f, const LevelDims &c, const Metrics<T> &mf,
                                              const Metrics<T> &mc, const T *__restrict__ phi1, T *__restrict__ phi2,
                                              const int8_t *__restrict__ ibkf, const int8_t *__restrict__ ibkc,
                                              const int8_t *__restrict__ pflag, int gi, int gj, int gk, int walls)
{
    const int jc = c.jc, kc = c.kc;
    const long q = idx(c, gi, gj, gk);
    const T c000 = phi1[q], c001 = phi1[q + 1], c010 = phi1[q + c.sy], c100 = phi1[q + c.sz],
            c011 = phi1[q + 1 + c.sy], c101 = phi1[q + 1 + c.sz], c110 = phi1[q + c.sy + c.sz],
            c111 = phi1[q + 1 + c.sy + c.sz];
    const T one = (T)1;
    // the volume-weighted value of the coarse block (CORE_MG.F90:1305-1335)
    const bool part = pflag[q] != 0;
    const bool vol = part && !walls;
    // free flags of the eight coarse cells (walls: renormalised weights)
    T f000 = (T)1, f001 = (T)1, f010 = (T)1, f100 = (T)1, f011 = (T)1, f101 = (T)1, f110 = (T)1, f111 = (T)1;
    if (part && walls) {
        auto fr = [&](long o) { return (ibkc[q + o] & 1) ? (T)0 : (T)1; };
        f000 = fr(0); f001 = fr(1); f010 = fr(c.sy); f100 = fr(c.sz);
        f011 = fr(1 + c.sy); f101 = fr(1 + c.sz); f110 = fr(c.sy + c.sz); f111 = fr(1 + c.sy + c.sz);
    }
    T dvol = (T)0;
    if (vol) {
        const int i2 = gi + c.ic, j2 = gj + jc, k2 = gk + kc;
        T vol0 = 0, sum0 = 0;
        for (int kk = gk; kk <= k2; ++kk)
            for (int jj = gj; jj <= j2; ++jj)
                for (int ii = gi; ii <= i2; ++ii) {
                    long qq = idx(c, ii, jj, kk);
                    if (ibkc[qq] != 1) {
                        T vol_i = (mc.x[ii + 1] - mc.x[ii]) * (mc.y[jj + 1] - mc.y[jj]) * (mc.z[kk + 1] - mc.z[kk]);
                        vol0 = vol0 + vol_i;
                        sum0 = phi1[qq] * vol_i;         // assignment, as CORE_MG.F90:1326
                    }
                }
        T vol_c = one / mc.dxinv[gi] / mc.dyinv[gj] / mc.dzinv[gk];
        dvol = (vol0 > (T)0.01f * vol_c) ? sum0 / vol0 : (T)0;
    }
    const int i0 = (FX == 2) ? 2 * gi : gi;
    T aX[FX];
    bool inx[FX];
#pragma unroll
    for (int e = 0; e < FX; ++e) {
        const int i = i0 + e;
        inx[e] = (i >= 1) && (i <= f.nx - 1);
        aX[e] = inx[e] ? (mf.xc[i] - mc.xc[gi]) * mc.dxcinv[gi + 1] : (T)0;
    }
    for (int a = 0; a <= kc; ++a) {
        const int k = kc ? 2 * gk + a : gk;
        if (k < 1 || k > f.nz - 1) continue;
        const T aZ = (mf.zc[k] - mc.zc[gk]) * mc.dzcinv[gk + 1];
        const int fk = (k == 1) || (k == f.nz - 1);
        for (int b = 0; b <= jc; ++b) {
            const int j = jc ? 2 * gj + b : gj;
            if (j < 1 || j > f.ny - 1) continue;
            const T aY = (mf.yc[j] - mc.yc[gj]) * mc.dycinv[gj + 1];
            const int fj = (j == 1) || (j == f.ny - 1);
            const long qf = idx(f, i0, j, k);
            T old[FX];
            int8_t ib[FX];
            if (FX == 2) {
                const typename Pair<T>::type o2 = *(const typename Pair<T>::type *)(phi2 + qf);
                old[0] = o2.x;
                old[FX - 1] = o2.y;
                const char2 b2 = *(const char2 *)(ibkf + qf);
                ib[0] = (int8_t)b2.x;
                ib[FX - 1] = (int8_t)b2.y;
            } else {
                old[0] = phi2[qf];
                ib[0] = ibkf[qf];
            }
            bool w[FX];
            T nv[FX];
#pragma unroll
            for (int e = 0; e < FX; ++e) {
                const int i = i0 + e;
                const int fi = (i == 1) || (i == f.nx - 1);
                w[e] = inx[e] && (fi + fj + fk < 2) && ib[e] != 1;
                T r_tmp = c000 * (one - aX[e]) * (one - aY) * (one - aZ) + c001 * aX[e] * (one - aY) * (one - aZ)
                          + c010 * (one - aX[e]) * aY * (one - aZ) + c100 * (one - aX[e]) * (one - aY) * aZ
                          + c011 * aX[e] * aY * (one - aZ) + c101 * aX[e] * (one - aY) * aZ
                          + c110 * (one - aX[e]) * aY * aZ + c111 * aX[e] * aY * aZ;
                if (part && walls) {
                    const T w000 = f000 * (one - aX[e]) * (one - aY) * (one - aZ),
                            w001 = f001 * aX[e] * (one - aY) * (one - aZ), w010 = f010 * (one - aX[e]) * aY * (one - aZ),
                            w100 = f100 * (one - aX[e]) * (one - aY) * aZ, w011 = f011 * aX[e] * aY * (one - aZ),
                            w101 = f101 * aX[e] * (one - aY) * aZ, w110 = f110 * (one - aX[e]) * aY * aZ,
                            w111 = f111 * aX[e] * aY * aZ;
                    const T ws = ((w000 + w001) + (w010 + w100)) + ((w011 + w101) + (w110 + w111));
                    const T num = ((c000 * w000 + c001 * w001) + (c010 * w010 + c100 * w100))
                                  + ((c011 * w011 + c101 * w101) + (c110 * w110 + c111 * w111));
                    r_tmp = (ws > (T)0) ? num / ws : (T)0;
                }
                nv[e] = old[e] + (vol ? dvol : r_tmp);
            }
            if (FX == 2 && w[0] && w[FX - 1]) {
                typename Pair<T>::type o2;
                o2.x = nv[0];
                o2.y = nv[FX - 1];
                *(typename Pair<T>::type *)(phi2 + qf) = o2;
            } else {
#pragma unroll
                for (int e = 0; e < FX; ++e)
                    if (w[e]) phi2[qf + e] = nv[e];
            }
        }
    }
}

// group ranges of a prolongation from c into f along one axis: first group
// and count
__host__ __device__ inline int prolong_g0(int coarsened) { return coarsened ? 0 : 1; }
__host__ __device__ inline int prolong_ng(int nf, int coarsened) { return coarsened ? (nf - 1) / 2 + 1 : nf - 1; }

template <typename T>
__device__ __forceinline__ void dk_prolong(const LevelDims &f, const LevelDims &c, const Metrics<T> &mf,
                                           const Metrics<T> &mc, const T *__restrict__ phi1, T *__restrict__ phi2,
                                           const int8_t *__restrict__ ibkf, const int8_t *__restrict__ ibkc,
                                           const int8_t *__restrict__ pflag, int walls, const VBlock &vb)
{
    const int gx = vb.x * blockDim.x + threadIdx.x, gy = vb.y * blockDim.y + threadIdx.y;
    if (gx >= prolong_ng(f.nx, c.ic) || gy >= prolong_ng(f.ny, c.jc)) return;
    const int gi = prolong_g0(c.ic) + gx, gj = prolong_g0(c.jc) + gy, gk = prolong_g0(c.kc) + vb.z;
    if (c.ic) prolong_group<T, 2>(f, c, mf, mc, phi1, phi2, ibkf, ibkc, pflag, gi, gj, gk, walls);
    else prolong_group<T, 1>(f, c, mf, mc, phi1, phi2, ibkf, ibkc, pflag, gi, gj, gk, walls);
}
template <typename T>
__global__ void __launch_bounds__(256) k_prolong(LevelDims f, LevelDims c, Metrics<T> mf, Metrics<T> mc,
                                                 const T *__restrict__ phi1, T *__restrict__ phi2,
                                                 const int8_t *__restrict__ ibkf, const int8_t *__restrict__ ibkc,
                                                 const int8_t *__restrict__ pflag, int walls)
{
    dk_prolong<T>(f, c, mf, mc, phi1, phi2, ibkf, ibkc, pflag, walls, hw_block());
}

// Block -> work mapping.  Blocks are dealt round-robin to the 8 XCDs
// (block b runs on XCD b % 8); the remap gives each XCD a contiguous run of
// tiles so that neighbouring lines (which read each other's values) share
// an L2.
__device__ __forceinline__ int xcd_remap(int b, int nb)
{
    int per = (nb + 7) >> 3;
    int q = (b & 7) * per + (b >> 3);
    return q;
}

// ---------------------------------------------------------------------------
// Line-SOR half sweep: one kernel per direction (DIR 0 = x, 1 = y, 2 = z
// lines) over the interior rows 1..R of every line of one colour.  Rows 0
// and R+1 are the reference's trivial ghost rows (value = current ghost
// value, left unchanged); their coupling to rows 1 and R is moved to the
// right-hand side.
//   gather  : one 32-bit cell offset per element serves all eight loads
//             (raw buffer loads, neighbour offsets in SGPRs); the off-line
//             neighbours are folded into the row's right-hand side and the
//             pair (b, d) goes to LDS, b = 0 marking an identity row (masked
//             or guarded, CORE_MG.F90:686,783,878)
//   solve   : lane p of a line holds rows p*E .. p*E+E-1; the line
//             coefficients of those rows stay in registers for all the
//             block's lines; partition solve with PCR across the LPL lanes
//   scatter : relaxed (1-omega)*old + omega*x (the old value is reloaded
//             only when omega != 1)
// LDS pair index of interior row rr (0-based) of line s: s*Rp + rr + rr/16.
// (LineSweep, the launch shape of one half sweep, is in fc_device.h.)
__device__ __forceinline__ int lds_row(int s, int Rp, int rr) { return s * Rp + rr + (rr >> 4); }


// line descriptors shared by the phases of a block
struct LineSlots {
    int i[64], j[64], k[64];
    int live[64];
    double g0[64], g1[64];    // the line's ghost rows 0 and R+1 (loaded at the start)
};

// the line coefficients (am, ap) of one lane's rows, the same on every line
// of the block; loaded before the gather so that the load overlaps it
template <typename T, int E> struct LineCoef {
    T am[E], ap[E];
};
template <typename T, int E, int DIR>
__device__ __forceinline__ LineCoef<T, E> ls_coef(const Metrics<T> &m, const LineSweep &g)
{
    const int p = (threadIdx.x & 63) & (g.LPL - 1);
    const T *am = (DIR == 0) ? m.amx : (DIR == 1) ? m.amy : m.amz;
    const T *ap = (DIR == 0) ? m.apx : (DIR == 1) ? m.apy : m.apz;
    LineCoef<T, E> c;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int rr = p * E + e;
        const int r = (rr < g.R) ? rr + 1 : 1;
        c.am[e] = am[r];
        c.ap[e] = ap[r];
    }
    return c;
}

// solve the NL lines held in LDS; wave w takes lines w*lpw.. pass by pass
template <typename T, int E, int DIR>
__device__ __forceinline__ void ls_solve(typename Pair<T>::type *sp, const LineSweep &g, const LineSlots &L,
                                         const LineCoef<T, E> &cf)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    const int LPL = g.LPL, R = g.R, Rp = g.Rp;
    const int lpw = 64 >> g.lpl_shift;
    const int p = lane & (LPL - 1);
    const T(&amr)[E] = cf.am;
    const T(&apr)[E] = cf.ap;
    const int pl = (R - 1) / E, el = (R - 1) - pl * E;     // lane and slot of row R
    for (int base = wave * lpw; base < g.NL; base += nwave * lpw) {
        const int s = base + (lane >> g.lpl_shift);
        const int ss = (s < g.NL) ? s : 0;
        const bool live = (s < g.NL) && L.live[ss];
        T oa[E], ob[E], oc[E], od[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int rr = p * E + e;
            const bool v = live && rr < R;
            typename Pair<T>::type bd = sp[lds_row(ss, Rp, v ? rr : 0)];
            const bool reg = v && bd.x != (T)0;
            oa[e] = reg ? amr[e] : (T)0;
            ob[e] = reg ? bd.x : (T)1;
            oc[e] = reg ? apr[e] : (T)0;
            od[e] = v ? bd.y : (T)0;
        }
        // ghost rows 0 and R+1 are known values: move their coupling right
        if (live) {
            if (p == 0) { od[0] -= oa[0] * (T)L.g0[ss]; oa[0] = (T)0; }
            if (p == pl) {
                const T gh = (T)L.g1[ss];
#pragma unroll
                for (int e = 0; e < E; ++e)
                    if (e == el) { od[e] -= oc[e] * gh; oc[e] = (T)0; }
            }
        }
        T x[E];
        line_solve<T, E>(oa, ob, oc, od, x, p, LPL);
        if (live) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int rr = p * E + e;
                if (rr < R) sp[lds_row(s, Rp, rr)].y = x[e];
            }
        }
    }
}

// the along-line coefficients and the neighbour stride of a sweep's rows (for
// walls across the line)
template <typename T> struct AlongLine {
    const T *am, *ap;
    int step;
};

// one element of the gather: cell offset vq = q - o1 - o2 (elements); row r
// along the line.  A wall face drops that face's coupling: an off-line
// neighbour leaves the right-hand side and its coefficient leaves the
// diagonal; an along-line neighbour behind a wall is a held row that keeps
// its value in the line solve, so its coupling is cancelled on the
// right-hand side with that value.
template <typename T, int DIR>
__device__ __forceinline__ typename Pair<T>::type sweep_row(__amdgpu_buffer_rsrc_t rphi, __amdgpu_buffer_rsrc_t rrhs,
                                                            __amdgpu_buffer_rsrc_t rigk, int vq, int so_q, int so_m1,
                                                            int so_p1, int so_m2, int so_p2, int sb_q, T c_m1,
                                                            T c_p1, T c_m2, T c_p2, T diag, bool live,
                                                            const T *__restrict__ phi, const AlongLine<T> &al, int r)
{
    const int v = vq << Buf<T>::shift;
    const T old = Buf<T>::ld(rphi, v, so_q);
    const T rv = Buf<T>::ld(rrhs, v, so_q);
    const T n0 = Buf<T>::ld(rphi, v, so_m1), n1 = Buf<T>::ld(rphi, v, so_p1);
    const T n2 = Buf<T>::ld(rphi, v, so_m2), n3 = Buf<T>::ld(rphi, v, so_p2);
    const unsigned gk = __builtin_amdgcn_raw_buffer_load_b8(rigk, vq, sb_q, 0);
    const T EPS = (T)1.0e-12f;
    T rh = rv - n0 * c_m1 - n1 * c_p1 - n2 * c_m2 - n3 * c_p2;
    T dg = diag;
    if (gk & MB_WALLS) {
        // off-line neighbours: m1/p1 are y (x-lines) or x (y-, z-lines); m2/p2
        // are z (x-, y-lines) or y (z-lines)
        const unsigned bm1 = (DIR == 0) ? MB_YM : MB_XM, bp1 = (DIR == 0) ? MB_YP : MB_XP;
        const unsigned bm2 = (DIR == 2) ? MB_YM : MB_ZM, bp2 = (DIR == 2) ? MB_YP : MB_ZP;
        const unsigned bml = (DIR == 0) ? MB_XM : (DIR == 1) ? MB_YM : MB_ZM, bpl = bml << 1;
        if (gk & bm1) { rh += n0 * c_m1; dg += c_m1; }
        if (gk & bp1) { rh += n1 * c_p1; dg += c_p1; }
        if (gk & bm2) { rh += n2 * c_m2; dg += c_m2; }
        if (gk & bp2) { rh += n3 * c_p2; dg += c_p2; }
        const long q = (long)vq + (so_q >> Buf<T>::shift);
        if (gk & bml) { const T a = al.am[r]; rh += a * phi[q - al.step]; dg += a; }
        if (gk & bpl) { const T a = al.ap[r]; rh += a * phi[q + al.step]; dg += a; }
    }
    const bool reg = live && (gk & MB_HELD) == 0u && !(fabs(dg) < EPS);
    typename Pair<T>::type bd;
    bd.x = reg ? dg : (T)0;
    bd.y = reg ? rh : old;
    return bd;
}

// x-lines (DIR 0): a block takes NL lines of one colour, k-major over the
// (j,k) half rows; rows are contiguous in memory.
template <typename T, int E>
__device__ __forceinline__ void dk_xsweep(const LevelDims &d, const Metrics<T> &m, T *__restrict__ phi,
                                          const T *__restrict__ rhs, const int8_t *__restrict__ igk,
                                          const LineSweep &g, const BcInfo &bc, const VBlock &vb)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ LineSlots L;
    using T2 = typename Pair<T>::type;
    T2 *sp = (T2 *)smem;
    const int NL = g.NL, Rp = g.Rp, R = g.R;
    T *acs = (T *)(sp + (size_t)NL * Rp);          // acx of rows 1..R, staged in LDS
    const int tile = xcd_remap(vb.x, g.nblocks);
    const int Jh = d.ny / 2;                       // slots per k: ceil((ny-1)/2)
    // loads that do not depend on the gather are issued first: the lane's
    // line coefficients and the ghost rows of the block's lines (only the
    // line itself writes them in this launch)
    const LineCoef<T, E> cf = ls_coef<T, E, 0>(m, g);
    T gh0 = (T)0, gh1 = (T)0;
    if (threadIdx.x < NL) {
        long ml = (long)tile * NL + threadIdx.x;
        int kk = 1 + (int)(ml / Jh), jr = (int)(ml % Jh);
        int j0 = (((1 + kk) & 1) == g.color) ? 1 : 2;
        int jj = j0 + 2 * jr;
        bool ok = (kk <= d.nz - 1) && (jj <= d.ny - 1);
        L.live[threadIdx.x] = ok;
        L.i[threadIdx.x] = 0;
        L.j[threadIdx.x] = ok ? jj : 1;
        L.k[threadIdx.x] = ok ? kk : 1;
        const long q0 = idx(d, 0, ok ? jj : 1, ok ? kk : 1);
        gh0 = phi[q0];
        gh1 = phi[q0 + R + 1];
    }
    for (int r = 1 + threadIdx.x; r <= R; r += blockDim.x) acs[r] = m.acx[r];
    __syncthreads();
    const long bytes = d.n << Buf<T>::shift;
    const __amdgpu_buffer_rsrc_t rphi = buf_rsrc(phi, bytes), rrhs = buf_rsrc(rhs, bytes), rigk = buf_rsrc(igk, d.n);
    const int o1 = (int)d.sy, o2 = (int)d.sz, sh = Buf<T>::shift;
    const int so_q = (o1 + o2) << sh, so_m1 = o2 << sh, so_p1 = (o2 + 2 * o1) << sh, so_m2 = o1 << sh,
              so_p2 = (o1 + 2 * o2) << sh, sb_q = o1 + o2;
    // gather: 8 groups of 32 threads, a group walks the rows of its lines
    const int tl = threadIdx.x & 31, tg = threadIdx.x >> 5;
    for (int s = tg; s < NL; s += 8) {
        const bool live = L.live[s];
        const int j = L.j[s], k = L.k[s];
        const T c_m1 = m.amy[j], c_p1 = m.apy[j], c_m2 = m.amz[k], c_p2 = m.apz[k];
        const T acy = m.acy[j], acz = m.acz[k];
        const int vq0 = (int)idx(d, 0, j, k) - o1 - o2;
        const AlongLine<T> al{m.amx, m.apx, 1};
#pragma unroll 4
        for (int r = 1 + tl; r <= R; r += 32) {
            const T diag = (acs[r] + acy) + acz;
            sp[lds_row(s, Rp, r - 1)] = sweep_row<T, 0>(rphi, rrhs, rigk, vq0 + r, so_q, so_m1, so_p1, so_m2, so_p2,
                                                        sb_q, c_m1, c_p1, c_m2, c_p2, diag, live, phi, al, r);
        }
    }
    if (threadIdx.x < NL) { L.g0[threadIdx.x] = (double)gh0; L.g1[threadIdx.x] = (double)gh1; }
    __syncthreads();
    ls_solve<T, E, 0>(sp, g, L, cf);
    __syncthreads();
    const T omega = (T)g.omega;
    const bool relax = g.omega != 1.0;
    for (int s = tg; s < NL; s += 8) {
        if (!L.live[s]) continue;
        const int j = L.j[s], k = L.k[s];
        const int vq0 = (int)idx(d, 0, j, k) - o1 - o2;
        const bool face = (j == 1) || (j == d.ny - 1) || (k == 1) || (k == d.nz - 1);
        for (int r = 1 + tl; r <= R; r += 32) {
            const int v = (vq0 + r) << sh;
            T x = sp[lds_row(s, Rp, r - 1)].y;
            if (relax) {
                const T o = Buf<T>::ld(rphi, v, so_q);
                x = g.relax_delta ? o + omega * (x - o) : o * ((T)1 - omega) + omega * x;
            }
            Buf<T>::st(rphi, x, v, so_q);
            if (g.post && (face || r == 1 || r == R)) bc_post<T>(d, phi, bc, r, j, k, x);
        }
    }
}

// y-lines (DIR 1) and z-lines (DIR 2): a block takes NL lines at the i of one
// colour inside a chunk of 2*NL consecutive i, at one k (y) or one j (z).
template <typename T, int E, int DIR>
__device__ __forceinline__ void dk_yzsweep(const LevelDims &d, const Metrics<T> &m, T *__restrict__ phi,
                                           const T *__restrict__ rhs, const int8_t *__restrict__ igk,
                                           const LineSweep &g, const BcInfo &bc, const VBlock &vb)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ LineSlots L;
    using T2 = typename Pair<T>::type;
    T2 *sp = (T2 *)smem;
    const int NL = g.NL, Rp = g.Rp, R = g.R;
    T *acs = (T *)(sp + (size_t)NL * Rp);          // ac along the line, rows 1..R, staged in LDS
    const int tile = xcd_remap(vb.x, g.nblocks);
    const int chunk = tile % g.chunks;
    const int other = 1 + tile / g.chunks;                     // k (y-lines) or j (z-lines)
    const int omax = (DIR == 1) ? d.nz - 1 : d.ny - 1;
    const int par = (g.color - other) & 1;
    const int oth = (other <= omax) ? other : 1;
    // the lane's line coefficients and the ghost rows of the block's lines
    // first (see dk_xsweep)
    const LineCoef<T, E> cf = ls_coef<T, E, DIR>(m, g);
    T gh0 = (T)0, gh1 = (T)0;
    if (threadIdx.x < NL) {
        int ii = 2 * NL * chunk + 2 * threadIdx.x + par;
        bool ok = (other <= omax) && ii >= 1 && ii <= d.nx - 1;
        L.live[threadIdx.x] = ok;
        L.i[threadIdx.x] = ok ? ii : 1;
        L.j[threadIdx.x] = (DIR == 2) ? oth : 0;
        L.k[threadIdx.x] = (DIR == 1) ? oth : 0;
        const long q0 = (DIR == 1) ? idx(d, ok ? ii : 1, 0, oth) : idx(d, ok ? ii : 1, oth, 0);
        const long st = (DIR == 1) ? d.sy : d.sz;
        gh0 = phi[q0];
        gh1 = phi[q0 + st * (R + 1)];
    }
    {
        const T *acl_g = (DIR == 1) ? m.acy : m.acz;
        for (int r = 1 + threadIdx.x; r <= R; r += blockDim.x) acs[r] = acl_g[r];
    }
    __syncthreads();
    const long bytes = d.n << Buf<T>::shift;
    const __amdgpu_buffer_rsrc_t rphi = buf_rsrc(phi, bytes), rrhs = buf_rsrc(rhs, bytes), rigk = buf_rsrc(igk, d.n);
    const int step = (DIR == 1) ? (int)d.sy : (int)d.sz;
    const int o1 = 1, o2 = (DIR == 1) ? (int)d.sz : (int)d.sy, sh = Buf<T>::shift;
    const int so_q = (o1 + o2) << sh, so_m1 = o2 << sh, so_p1 = (o2 + 2 * o1) << sh, so_m2 = o1 << sh,
              so_p2 = (o1 + 2 * o2) << sh, sb_q = o1 + o2;
    const T *m2 = (DIR == 1) ? m.amz : m.amy, *p2 = (DIR == 1) ? m.apz : m.apy;
    const T c_m2 = m2[oth], c_p2 = p2[oth];
    const T aco = (DIR == 1) ? m.acz[oth] : m.acy[oth];
    // gather: thread t -> line t % NL, rows t / NL + (256/NL) u
    const int s = threadIdx.x & (NL - 1);
    const int rg = threadIdx.x / NL, ng = 256 / NL;
    const bool live = L.live[s];
    const int i = L.i[s];
    const T c_m1 = m.amx[i], c_p1 = m.apx[i], acx = m.acx[i];
    const int vq0 = (int)((DIR == 1) ? idx(d, i, 0, oth) : idx(d, i, oth, 0)) - o1 - o2;
    const AlongLine<T> al{(DIR == 1) ? m.amy : m.amz, (DIR == 1) ? m.apy : m.apz, step};
#pragma unroll 4
    for (int r = 1 + rg; r <= R; r += ng) {
        const T diag = (DIR == 1) ? (acx + acs[r]) + aco : (acx + aco) + acs[r];
        sp[lds_row(s, Rp, r - 1)] = sweep_row<T, DIR>(rphi, rrhs, rigk, vq0 + step * r, so_q, so_m1, so_p1, so_m2,
                                                      so_p2, sb_q, c_m1, c_p1, c_m2, c_p2, diag, live, phi, al, r);
    }
    if (threadIdx.x < NL) { L.g0[threadIdx.x] = (double)gh0; L.g1[threadIdx.x] = (double)gh1; }
    __syncthreads();
    ls_solve<T, E, DIR>(sp, g, L, cf);
    __syncthreads();
    const T omega = (T)g.omega;
    const bool relax = g.omega != 1.0;
    if (live) {
        const bool face = (i == 1) || (i == d.nx - 1) || (oth == 1) || (oth == omax);
        for (int r = 1 + rg; r <= R; r += ng) {
            const int v = (vq0 + step * r) << sh;
            T x = sp[lds_row(s, Rp, r - 1)].y;
            if (relax) {
                const T o = Buf<T>::ld(rphi, v, so_q);
                x = g.relax_delta ? o + omega * (x - o) : o * ((T)1 - omega) + omega * x;
            }
            Buf<T>::st(rphi, x, v, so_q);
            if (g.post && (face || r == 1 || r == R))
                bc_post<T>(d, phi, bc, i, DIR == 1 ? r : oth, DIR == 2 ? r : oth, x);
        }
    }
}

template <typename T, int E>
__global__ void __launch_bounds__(256) k_xsweep(LevelDims d, Metrics<T> m, T *__restrict__ phi,
                                                const T *__restrict__ rhs, const int8_t *__restrict__ igk, LineSweep g,
                                                BcInfo bc)
{
    dk_xsweep<T, E>(d, m, phi, rhs, igk, g, bc, hw_block());
}
template <typename T, int E, int DIR>
__global__ void __launch_bounds__(256) k_yzsweep(LevelDims d, Metrics<T> m, T *__restrict__ phi,
                                                 const T *__restrict__ rhs, const int8_t *__restrict__ igk,
                                                 LineSweep g, BcInfo bc)
{
    dk_yzsweep<T, E, DIR>(d, m, phi, rhs, igk, g, bc, hw_block());
}

// ---------------------------------------------------------------------------
// Batched tdma (CORE_TDMA.F90:10-34) with the same in-wave solver, for
// nsys systems of n rows (a row-major stack of a, b, c, r).
template <typename T, int E>
__global__ void __launch_bounds__(64) k_tdma_batched(int nsys, int n, int LPL, const double *__restrict__ A,
                                                     const double *__restrict__ B, const double *__restrict__ Cc,
                                                     const double *__restrict__ R, double *__restrict__ U)
{
    const int lane = threadIdx.x;
    const int wl = 64 / LPL;
    const int p = lane % LPL;
    long sys = (long)blockIdx.x * wl + lane / LPL;
    bool live = sys < nsys;
    long base = (live ? sys : 0) * (long)n;
    T oa[E], ob[E], oc[E], od[E], x[E];
    for (int e = 0; e < E; ++e) {
        int r = p * E + e;
        T av = 0, bv = 1, cv = 0, dv = 0;
        if (live && r < n) {
            av = (r > 0) ? (T)A[base + r] : (T)0;
            bv = (T)B[base + r];
            cv = (r < n - 1) ? (T)Cc[base + r] : (T)0;
            dv = (T)R[base + r];
        }
        oa[e] = av; ob[e] = bv; oc[e] = cv; od[e] = dv;
    }
    line_solve<T, E>(oa, ob, oc, od, x, p, LPL);
    if (live)
        for (int e = 0; e < E; ++e) {
            int r = p * E + e;
            if (r < n) U[base + r] = (double)x[e];
        }
}

// ---------------------------------------------------------------------------
// GCM_ghostcell_pressure (GCM_GHOSTCELL_UPDATE.F90:361-739) as two Jacobi
// phases per sweep: all dead cells from the current field, then all ghost
// cells from the field holding the new dead values.  New values go to a
// list buffer first so every read of a phase sees the same field.
__global__ void k_gcm_dead(GcmDev g, const double *__restrict__ pres, const double *__restrict__ pprev,
                           const long *__restrict__ off8, int homo, double *__restrict__ outv)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nD || !g.dActive[n]) return;
    double dpdn = homo ? 0.0 : g.dDpdn[n];
    double pIP = 0.0;
    long base = g.dBase[n];
    if (g.dIdw[n]) {
        double pinv = 0.0, dist_int = 0.0;
        for (int r = 0; r < 8; ++r) {
            double w = g.dW[8 * (long)n + r];
            if (w != 0.0) {
                dist_int = dist_int + w;
                pinv = pinv + w * (pprev ? pprev[base + off8[r]] : 0.0);
            }
        }
        pIP = pinv / dist_int;
    } else {
        int self = g.dSelf[n];
        for (int r = 0; r < 8; ++r) {
            double cf = g.dCoef[8 * (long)n + r];
            pIP = pIP + (r == self ? cf * dpdn : cf * pres[base + off8[r]]);
        }
    }
    outv[n] = pIP + dpdn * g.dPl[n];
}

__global__ void k_gcm_ghost(GcmDev g, const double *__restrict__ pres, const double *__restrict__ r,
                            const long *__restrict__ off8, int homo, int mix, LevelDims d,
                            const double *__restrict__ dxcinv, const double *__restrict__ dycinv,
                            const double *__restrict__ dzcinv, const double *__restrict__ dxinv,
                            const double *__restrict__ dyinv, const double *__restrict__ dzinv,
                            const int *__restrict__ gijk, double *__restrict__ outv, double *__restrict__ reslap)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nG) return;
    double dpdn = homo ? 0.0 : g.gDpdn[n];
    long base = g.gBase[n];
    int self = g.gSelf[n];
    double pG = 0.0;
    for (int rr = 0; rr < 8; ++rr) {
        double cf = g.gCoef[8 * (long)n + rr];
        pG = pG + (rr == self ? cf * dpdn : cf * pres[base + off8[rr]]);
    }
    if (mix) {
        long q = g.gTarget[n];
        int iG = gijk[3 * n], jG = gijk[3 * n + 1], kG = gijk[3 * n + 2];
        double pc = pres[q];
        double dpdxW = (pc - pres[q - 1]) * dxcinv[iG];
        double dpdxE = (pres[q + 1] - pc) * dxcinv[iG + 1];
        double dpdyS = (pc - pres[q - d.sy]) * dycinv[jG];
        double dpdyN = (pres[q + d.sy] - pc) * dycinv[jG + 1];
        double dpdzB = (pc - pres[q - d.sz]) * dzcinv[kG];
        double dpdzF = (pres[q + d.sz] - pc) * dzcinv[kG + 1];
        double fl = g.gAlpha[n];
        double coeffG = -(dxcinv[iG] + dxcinv[iG + 1]) * dxinv[iG] - (dycinv[jG] + dycinv[jG + 1]) * dyinv[jG]
                        - (dzcinv[kG] + dzcinv[kG + 1]) * dzinv[kG];
        double resP = (dpdxE - dpdxW) * dxinv[iG] + (dpdyN - dpdyS) * dyinv[jG] + (dpdzF - dpdzB) * dzinv[kG]
                      - r[q];
        double pG2 = pc + resP / (-coeffG);
        reslap[n] = (-coeffG) * (pG2 - pG) * (1.0 - fl);
        outv[n] = (1.0 - fl) * pG + fl * pG2;
    } else {
        reslap[n] = 0.0;
        outv[n] = pG;
    }
}

// Level-scheduled exact GCM update: the ghost/dead relations of
// GCM_ghostcell_pressure (mix_GC_form = 0) form a dependency graph that is
// acyclic for image points outside their own stencil chains; evaluating its
// topological levels in order gives the fixed point the reference's
// list-order Gauss-Seidel sweeps converge to, in one pass.  order[lo:hi) is
// one level (entries < nD are dead cells, the rest ghost cells - nD); no
// entry of a level reads another entry of the same level.
// one relation of the topological schedule (record t)
__device__ __forceinline__ void gcm_node_at(const GcmNode *__restrict__ nodes, int t, double *__restrict__ pres,
                                            const double *__restrict__ pprev, const long *__restrict__ off8, int homo,
                                            double *__restrict__ gdelta)
{
    const GcmNode nd = nodes[t];
    const double dpdn = homo ? 0.0 : nd.dpdn;
    const long base = nd.base;
    double v = 0.0;
    if (nd.kind == 2) {
        double pinv = 0.0, dist_int = 0.0;
        for (int r = 0; r < 8; ++r) {
            const double w = nd.coef[r];
            if (w != 0.0) {
                dist_int = dist_int + w;
                pinv = pinv + w * (pprev ? pprev[base + off8[r]] : 0.0);
            }
        }
        v = pinv / dist_int;
    } else {
#pragma unroll
        for (int r = 0; r < 8; ++r) v = v + (r == nd.self ? nd.coef[r] * dpdn : nd.coef[r] * pres[base + off8[r]]);
    }
    if (nd.kind != 0) v = v + dpdn * nd.pl;
    if (gdelta) gdelta[t] = v - pres[nd.target];
    pres[nd.target] = v;
}

// one topological level (records lo..hi-1), one relation per thread
__device__ __forceinline__ void dk_gcm_level(const GcmNode *__restrict__ nodes, int lo, int hi,
                                             double *__restrict__ pres, const double *__restrict__ pprev,
                                             const long *__restrict__ off8, int homo, double *__restrict__ gdelta,
                                             const VBlock &vb)
{
    int t = lo + vb.x * blockDim.x + threadIdx.x;
    if (t < hi) gcm_node_at(nodes, t, pres, pprev, off8, homo, gdelta);
}
__global__ void __launch_bounds__(256) k_gcm_level(const GcmNode *__restrict__ nodes, int lo, int hi,
                                                   double *__restrict__ pres, const double *__restrict__ pprev,
                                                   const long *__restrict__ off8, int homo, double *__restrict__ gdelta)
{
    dk_gcm_level(nodes, lo, hi, pres, pprev, off8, homo, gdelta, hw_block());
}

// Mixed precision: move the changes the GCM pass made to its targets from
// the pressure into the FP32 correction (e += d, p -= d), so that p is the
// iterate x - x(1,1,1) and e the whole step G - p; e == nullptr moves them
// back into p (p += d) when the solve ends.  assign: e = d on the targets
// (the correction cycles with bodies as walls solve for the ghost cells too;
// the GCM relations, not the cycle, set a target's step)
template <typename TE>
__device__ __forceinline__ void dk_gcm_fold(const GcmNode *__restrict__ nodes, int n,
                                            const double *__restrict__ gdelta, TE *__restrict__ e,
                                            double *__restrict__ p, int assign, const VBlock &vb)
{
    const int t = vb.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const long q = nodes[t].target;
    if (e) {
        e[q] = assign ? (TE)gdelta[t] : (TE)((double)e[q] + gdelta[t]);
        p[q] = p[q] - gdelta[t];
    } else {
        p[q] = p[q] + gdelta[t];
    }
}
template <typename TE>
__global__ void __launch_bounds__(256) k_gcm_fold(const GcmNode *__restrict__ nodes, int n,
                                                  const double *__restrict__ gdelta, TE *__restrict__ e,
                                                  double *__restrict__ p, int assign)
{
    dk_gcm_fold<TE>(nodes, n, gdelta, e, p, assign, hw_block());
}

// write list values into the field, per-block max |new - old|
__global__ void k_gcm_scatter(int n, const long *__restrict__ target, const int *__restrict__ active,
                              const double *__restrict__ vals, double *__restrict__ pres,
                              double *__restrict__ partial)
{
    __shared__ double s[256];
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    double ch = 0.0;
    if (t < n && (!active || active[t])) {
        long q = target[t];
        ch = fabs(vals[t] - pres[q]);
        pres[q] = vals[t];
    }
    s[threadIdx.x] = ch;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// ---------------------------------------------------------------------------
// rhs_poisson (CORE_POISSON.F90:35-50): div on flag cells, per-block partial
// sums of div*dx*dy*dz and max |div/dtinv|.
// d: the layout of the face velocities and div; dm: that of the mask (the
// handle's padded layout; the allocation-free entry point takes fields in
// the unpadded Fortran layout)
__global__ void __launch_bounds__(256) k_rhs(LevelDims d, LevelDims dm, const double *__restrict__ fu,
                                             const double *__restrict__ fv,
                                             const double *__restrict__ fw, const int8_t *__restrict__ igk,
                                             const double *__restrict__ dxinv, const double *__restrict__ dyinv,
                                             const double *__restrict__ dzinv, const double *__restrict__ dx,
                                             const double *__restrict__ dy, const double *__restrict__ dz,
                                             double dtinv, double *__restrict__ div, double *__restrict__ partial)
{
    __shared__ double s1[256], s2[256];
    int j = 1 + blockIdx.x, k = 1 + blockIdx.y;
    double a1 = 0.0, a2 = 0.0;
    for (int i = 1 + threadIdx.x; i <= d.nx - 1; i += blockDim.x) {
        long q = idx(d, i, j, k);
        double v = dtinv * ((fu[q + 1] - fu[q]) * dxinv[i] + (fv[q + d.sy] - fv[q]) * dyinv[j]
                            + (fw[q + d.sz] - fw[q]) * dzinv[k]);
        v = igk[idx(dm, i, j, k)] ? 0.0 : v;
        div[q] = v;
        a1 += v * dx[i] * dy[j] * dz[k];
        a2 = fmax(a2, fabs(v / dtinv));
    }
    s1[threadIdx.x] = a1;
    s2[threadIdx.x] = a2;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) { s1[threadIdx.x] += s1[threadIdx.x + w]; s2[threadIdx.x] = fmax(s2[threadIdx.x], s2[threadIdx.x + w]); }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        long b = blockIdx.x + (long)gridDim.x * blockIdx.y;
        partial[2 * b] = s1[0];
        partial[2 * b + 1] = s2[0];
    }
}

// Block-wide reduction of NV values (sum, or max where MAXMASK has the bit);
// the result is valid in thread 0.  blockDim.x = 256.
template <int NV, unsigned MAXMASK>
__device__ __forceinline__ void block_reduce(double (&v)[NV], double *sh)
{
    for (int c = 0; c < NV; ++c) {
        double a = v[c];
        for (int o = 32; o > 0; o >>= 1) {
            double b = __shfl_down(a, o, 64);
            a = ((MAXMASK >> c) & 1u) ? fmax(a, b) : a + b;
        }
        v[c] = a;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int c = 0; c < NV; ++c) sh[wave * NV + c] = v[c];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int c = 0; c < NV; ++c) {
            double a = sh[c];
            for (int w = 1; w < 4; ++w) a = ((MAXMASK >> c) & 1u) ? fmax(a, sh[w * NV + c]) : a + sh[w * NV + c];
            v[c] = a;
        }
}

// ---------------------------------------------------------------------------
// Interior cells of a level in 4-cell chunks of the padded rows.  The row
// pitch is a multiple of 16 elements, so a chunk is 16-byte aligned in FP32
// and 32-byte aligned in FP64 and moves with vector loads.  Chunk c of row
// (j,k) holds i = 4c .. 4c+3; a block takes the chunks of its rows
// [row_begin, row_end) in turn, four cells per thread per pass.
struct Chunk {
    long q0;
    int i0, j, k;
    bool in[4];     // cell i0+e is interior (1 <= i <= nx-1)
    bool full;      // all four are
};
__device__ __forceinline__ int chunks_per_row(const LevelDims &d) { return ((d.nx - 1) >> 2) + 1; }
__device__ __forceinline__ Chunk chunk_at(const LevelDims &d, int r, int c)
{
    Chunk ch;
    ch.k = 1 + r / (d.ny - 1);
    ch.j = r - (ch.k - 1) * (d.ny - 1) + 1;
    ch.i0 = 4 * c;
    ch.q0 = (long)ch.j * d.sy + (long)ch.k * d.sz + ch.i0;
#pragma unroll
    for (int e = 0; e < 4; ++e) ch.in[e] = (ch.i0 + e >= 1) && (ch.i0 + e <= d.nx - 1);
    ch.full = (ch.i0 >= 1) && (ch.i0 + 3 <= d.nx - 1);
    return ch;
}
using D4 = V4<double>;
__device__ __forceinline__ D4 ld4(const double *p)
{
    const double2 a = *(const double2 *)p, b = *(const double2 *)(p + 2);
    return D4{{a.x, a.y, b.x, b.y}};
}
__device__ __forceinline__ void st4(double *p, const double (&x)[4], const Chunk &ch)
{
    if (ch.full) {
        *(double2 *)p = make_double2(x[0], x[1]);
        *(double2 *)(p + 2) = make_double2(x[2], x[3]);
    } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) if (ch.in[e]) p[e] = x[e];
    }
}
__device__ __forceinline__ void st4f(float *p, const float (&x)[4], const Chunk &ch)
{
    if (ch.full) {
        *(float4 *)p = make_float4(x[0], x[1], x[2], x[3]);
    } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) if (ch.in[e]) p[e] = x[e];
    }
}
__device__ __forceinline__ void ld4f(const float *p, float (&x)[4])
{
    const float4 a = *(const float4 *)p;
    x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w;
}
// four elements of a correction field (FP32 in the mixed path, FP64 in the
// FP64 correction path) to and from doubles
__device__ __forceinline__ void ld4e(const float *p, double (&x)[4])
{
    float f[4];
    ld4f(p, f);
#pragma unroll
    for (int e = 0; e < 4; ++e) x[e] = (double)f[e];
}
__device__ __forceinline__ void ld4e(const double *p, double (&x)[4])
{
    const D4 a = ld4(p);
#pragma unroll
    for (int e = 0; e < 4; ++e) x[e] = a.v[e];
}
__device__ __forceinline__ void st4e(float *p, const double (&x)[4], const Chunk &ch)
{
    float f[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) f[e] = (float)x[e];
    st4f(p, f, ch);
}
__device__ __forceinline__ void st4e(double *p, const double (&x)[4], const Chunk &ch) { st4(p, x, ch); }
// The x metrics (amx, apx, acx; the padded span of each array) staged in
// LDS by the block: xm must hold 3 * xmet_span(d) doubles.
__host__ __device__ inline int xmet_span(const LevelDims &d) { return (d.nx + 2 + 3) / 4 * 4 + 4; }
inline size_t xmet_bytes(const LevelDims &d) { return 3 * (size_t)xmet_span(d) * sizeof(double); }
__device__ __forceinline__ void stage_xmet(const LevelDims &d, const Metrics<double> &m, double *xm)
{
    const int sp = xmet_span(d);
    for (int i = threadIdx.x; i < sp; i += blockDim.x) {
        xm[i] = m.amx[i];
        xm[sp + i] = m.apx[i];
        xm[2 * sp + i] = m.acx[i];
    }
    __syncthreads();
}

// FP64 7-point residual div - A p of the four cells of a chunk (residual_mg
// rows, CORE_MG.F90:990-1030), 0 where igk is set; sums of res^2 and div^2
// of the unmasked interior cells added to s_res, s_div; returns the chunk's p
// xm: the x metrics staged in LDS by stage_xmet, or null (read from m).
// The caller's consecutive lanes hold consecutive chunks of a row, so the
// left and right neighbours of a chunk come from the adjacent lanes (loaded
// only at the wavefront's edges).
__device__ __forceinline__ D4 chunk_residual(const LevelDims &d, const Metrics<double> &m, const double *xm,
                                               const double *__restrict__ p, const double *__restrict__ div,
                                               const int8_t *__restrict__ igk, const Chunk &ch, double (&res)[4],
                                               double &s_res, double &s_div)
{
    const long bytes = d.n << 3;
    const __amdgpu_buffer_rsrc_t rp = buf_rsrc(p, bytes), rd = buf_rsrc(div, bytes), rg = buf_rsrc(igk, d.n);
    const int q0 = (int)ch.q0, v = q0 << 3, sy = (int)d.sy << 3, sz = (int)d.sz << 3;
    const V4<double> pc = bld4<double>(rp, v), pjm = bld4<double>(rp, v - sy), pjp = bld4<double>(rp, v + sy),
                     pkm = bld4<double>(rp, v - sz), pkp = bld4<double>(rp, v + sz), dv = bld4<double>(rd, v);
    const unsigned g = __builtin_amdgcn_raw_buffer_load_b32(rg, q0, 0, 0);
    // the left neighbour is only used when i0 >= 4 (cell 0 is never
    // interior), when the previous chunk of the row sits in the lane below;
    // the right one comes from the lane above unless this is the row's last
    // chunk
    const int lane = threadIdx.x & 63;
    double pl = __shfl_up(pc.v[3], 1, 64), pr = __shfl_down(pc.v[0], 1, 64);
    if (lane == 0) pl = Buf<double>::ld(rp, v - 8, 0);
    if (lane == 63 || ch.i0 + 4 > 4 * ((d.nx - 1) >> 2)) pr = Buf<double>::ld(rp, v + 32, 0);
    const double bmy = m.amy[ch.j], bpy = m.apy[ch.j], bmz = m.amz[ch.k], bpz = m.apz[ch.k];
    const double sjk = m.acy[ch.j] + m.acz[ch.k];
    const int vm = ch.i0 << 3;
    V4<double> amx, apx, acx;
    if (xm) {
        const int sp = xmet_span(d);
        amx = *(const V4<double> *)(xm + ch.i0);
        apx = *(const V4<double> *)(xm + sp + ch.i0);
        acx = *(const V4<double> *)(xm + 2 * sp + ch.i0);
    } else {
        amx = bld4<double>(met_rsrc(m.amx, d.nx), vm);
        apx = bld4<double>(met_rsrc(m.apx, d.nx), vm);
        acx = bld4<double>(met_rsrc(m.acx, d.nx), vm);
    }
    // branch-free: every element is computed and the masked ones selected
    // away, so no load is sunk into a per-element branch
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const double left = (e == 0) ? pl : pc.v[e - 1], right = (e == 3) ? pr : pc.v[e + 1];
        const double bc = acx.v[e] + sjk;
        const double rr = dv.v[e] - pc.v[e] * bc - left * amx.v[e] - right * apx.v[e] - pjm.v[e] * bmy -
                          pjp.v[e] * bpy - pkm.v[e] * bmz - pkp.v[e] * bpz;
        const bool on = ch.in[e] && ((g >> (8 * e)) & 0xffu) == 0u;
        res[e] = on ? rr : 0.0;
        s_res += on ? rr * rr : 0.0;
        s_div += on ? dv.v[e] * dv.v[e] : 0.0;
    }
    return pc;
}

// zero the cells i0..i0+3 of the chunk that lie in 0..nx (the interior and
// the two x ghosts); the row's last chunk also covers cell nx when it lies
// past the chunk
__device__ __forceinline__ void zero4(float *p) { *(float4 *)p = make_float4(0.f, 0.f, 0.f, 0.f); }
__device__ __forceinline__ void zero4(double *p)
{
    *(double2 *)p = make_double2(0.0, 0.0);
    *(double2 *)(p + 2) = make_double2(0.0, 0.0);
}
template <typename TE>
__device__ __forceinline__ void zero_row_cells(TE *p, const Chunk &ch, int nx)
{
    if (ch.i0 + 3 <= nx) {
        zero4(p);
    } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
            if (ch.i0 + e <= nx) p[e] = (TE)0;
    }
    if (ch.i0 == 4 * ((nx - 1) >> 2) && nx > ch.i0 + 3) p[nx - ch.i0] = (TE)0;
}

// ---------------------------------------------------------------------------
// Defect correction: the FP64 finest-level residual as the right-hand side
// of the correction equation (FP32 in the mixed path, FP64 in the FP64
// correction path: TE).  With partial: per-block sums of res^2 and div^2
// (calc_residual, components 0 and 1) and max|p| (component 3),
// component-major with one slot per block.
template <typename TE>
__device__ __forceinline__ void dk_resid64_to32(const LevelDims &d, const Metrics<double> &m, const double *xm,
                                                const double *__restrict__ p, const double *__restrict__ div,
                                                const int8_t *__restrict__ igk, TE *__restrict__ r32,
                                                TE *__restrict__ e32, double *__restrict__ partial,
                                                const VBlock &vb)
{
    __shared__ double sh[4 * 3];
    double v[3] = {0.0, 0.0, 0.0};
    // grid-stride over the chunks of all rows: the resident blocks sweep one
    // contiguous window of the field at a time
    const int cpr = chunks_per_row(d);
    const int nch = (d.ny - 1) * (d.nz - 1) * cpr;
    const float zero[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int t = vb.x * blockDim.x + threadIdx.x; t < nch; t += vb.nx * blockDim.x) {
        const int rr = t / cpr;
        const Chunk ch = chunk_at(d, rr, t - rr * cpr);
        double res[4];
        const D4 pc = chunk_residual(d, m, xm, p, div, igk, ch, res, v[0], v[1]);
#pragma unroll
        for (int e = 0; e < 4; ++e)
            if (ch.in[e]) v[2] = fmax(v[2], fabs(pc.v[e]));
        st4e(r32 + ch.q0, res, ch);
        if (e32) {
            // e starts at zero: its interior cells and the outer ghosts the
            // homogeneous boundary fill of its first smoothing would set to
            // +0 (x ghosts of this row; the y/z face rows next to it)
            zero_row_cells(e32 + ch.q0, ch, d.nx);
            if (ch.j == 1) zero_row_cells(e32 + ch.q0 - d.sy, ch, d.nx);
            if (ch.j == d.ny - 1) zero_row_cells(e32 + ch.q0 + d.sy, ch, d.nx);
            if (ch.k == 1) zero_row_cells(e32 + ch.q0 - d.sz, ch, d.nx);
            if (ch.k == d.nz - 1) zero_row_cells(e32 + ch.q0 + d.sz, ch, d.nx);
        }
    }
    if (!partial) return;
    block_reduce<3, 0x4u>(v, sh);
    if (threadIdx.x == 0) { partial[vb.x] = v[0]; partial[vb.nx + vb.x] = v[1]; partial[3 * vb.nx + vb.x] = v[2]; }
}
template <typename TE>
__global__ void __launch_bounds__(256) k_resid64_to32(LevelDims d, Metrics<double> m, const double *__restrict__ p,
                                                      const double *__restrict__ div, const int8_t *__restrict__ igk,
                                                      TE *__restrict__ r32, TE *__restrict__ e32,
                                                      double *__restrict__ partial)
{
    extern __shared__ double xm[];
    stage_xmet(d, m, xm);
    dk_resid64_to32<TE>(d, m, xm, p, div, igk, r32, e32, partial, hw_block());
}

// p += e over the whole padded field (e is zero where the V-cycle leaves it)
__global__ void k_add_correction(long n, double *__restrict__ p, const float *__restrict__ e)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; t < n; t += stride) p[t] += (double)e[t];
}

template <typename T>
__global__ void k_fill(long n, T *__restrict__ p, T v)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; t < n; t += stride) p[t] = v;
}

// ---------------------------------------------------------------------------
// Anderson acceleration of the outer fixed-point iteration (new; the
// reference iterates the plain map x -> G(x), one solve_poisson loop body).
// Type-II mixing with depth m <= 8 over the interior cells:
//   f_k = G_k - x_k,  dF = f_k - f_{k-1},  dG = G_k - G_{k-1} = f_k - c_{k-1}
//   gamma = argmin |f_k - dF gamma|,  c_k = dG gamma,  x_{k+1} = G_k - c_k
// where c_{k-1} is the previous mixing correction (x_k = G_{k-1} - c_{k-1}),
// so only x (FP64) and the small vectors f, c and the history columns (FP32)
// are stored.  The ring of columns, the Gram matrix and gamma stay on the
// device, so one outer iteration is a fixed launch sequence.

// G = p + e on interior cells (mixed precision: the FP32 V-cycle correction);
// e may be null (interior copy)
__device__ __forceinline__ void dk_add_into(const LevelDims &d, const double *__restrict__ p,
                                            const float *__restrict__ e, double *__restrict__ G, const VBlock &vb)
{
    const long nrow = (long)(d.ny - 1) * (d.nz - 1);
    const int r0 = (int)row_begin(nrow, vb), r1 = (int)row_end(nrow, vb);
    const int cpr = chunks_per_row(d);
    const int nch = (r1 - r0) * cpr;
    for (int t = threadIdx.x; t < nch; t += blockDim.x) {
        const int rr = t / cpr;
        const Chunk ch = chunk_at(d, r0 + rr, t - rr * cpr);
        const D4 pv = ld4(p + ch.q0);
        double out[4] = {pv.v[0], pv.v[1], pv.v[2], pv.v[3]};
        if (e) {
            float ev[4];
            ld4f(e + ch.q0, ev);
#pragma unroll
            for (int k = 0; k < 4; ++k) out[k] = pv.v[k] + (double)ev[k];
        }
        st4(G + ch.q0, out, ch);
    }
}

__global__ void __launch_bounds__(256) k_add_into(LevelDims d, const double *__restrict__ p,
                                                  const float *__restrict__ e, double *__restrict__ G)
{
    dk_add_into(d, p, e, G, hw_block());
}

// One pass over the interior after a map evaluation G: calc_residual sums of
// G (flag cells), max|f|, max|G|, the new history column and the dots
// dF_s . dF_new and dF_s . f for every live slot s.  Partials are stored
// component-major: partial[c * (blocks) + block], c < 4 + 2*8.
__device__ __forceinline__ void dk_aa_resid(const LevelDims &d, const Metrics<double> &m, const double *xm,
                                            const double *__restrict__ G, const double *__restrict__ div,
                                            const int8_t *__restrict__ igk, const double *__restrict__ x,
                                            float *__restrict__ fprev, const float *__restrict__ cprev,
                                            const AaCols &cols, const int *__restrict__ ring,
                                            double *__restrict__ partial, const VBlock &vb)
{
    __shared__ double sh[4 * ROW_COMPONENTS];
    const int have = ring[2];
    const int ns = have ? ring[3] : -1;                    // slot that takes the new column
    const unsigned live_mask = have ? (unsigned)ring[12] : 0u;   // slots in use after the insertion
    double v[ROW_COMPONENTS];
    for (int c = 0; c < ROW_COMPONENTS; ++c) v[c] = 0.0;
    const long nrow = (long)(d.ny - 1) * (d.nz - 1);
    const int r0 = (int)row_begin(nrow, vb), r1 = (int)row_end(nrow, vb);
    const int cpr = chunks_per_row(d);
    const int nch = (r1 - r0) * cpr;
    for (int t = threadIdx.x; t < nch; t += blockDim.x) {
        const int rr = t / cpr;
        const Chunk ch = chunk_at(d, r0 + rr, t - rr * cpr);
        const long q0 = ch.q0;
        double res[4];
        chunk_residual(d, m, xm, G, div, igk, ch, res, v[0], v[1]);
        const D4 gv = ld4(G + q0), xv = ld4(x + q0);
        float fp[4], cp[4];
        ld4f(fprev + q0, fp);
        ld4f(cprev + q0, cp);
        double f[4];
        float dfn[4], dgn[4], fn[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            f[e] = gv.v[e] - xv.v[e];
            fn[e] = (float)f[e];
            dfn[e] = (float)(f[e] - (double)fp[e]);
            dgn[e] = (float)(f[e] - (double)cp[e]);
            if (ch.in[e]) {
                v[2] = fmax(v[2], fabs(f[e]));
                v[3] = fmax(v[3], fabs(gv.v[e]));
            }
        }
        if (ns >= 0) {
            st4f(cols.dF[ns] + q0, dfn, ch);
            st4f(cols.dG[ns] + q0, dgn, ch);
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                // fixed trip count: v stays in registers
                if (!((live_mask >> s) & 1u)) continue;
                float cv[4];
                if (s == ns) { cv[0] = dfn[0]; cv[1] = dfn[1]; cv[2] = dfn[2]; cv[3] = dfn[3]; }
                else ld4f(cols.dF[s] + q0, cv);
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (ch.in[e]) {
                        v[4 + 2 * s] += (double)cv[e] * (double)dfn[e];
                        v[5 + 2 * s] += (double)cv[e] * f[e];
                    }
            }
        }
        st4f(fprev + q0, fn, ch);
    }
    block_reduce<ROW_COMPONENTS, 0xCu>(v, sh);
    if (threadIdx.x == 0)
        for (int c = 0; c < ROW_COMPONENTS; ++c) partial[(long)c * vb.nx + vb.x] = v[c];
}
__global__ void __launch_bounds__(256) k_aa_resid(LevelDims d, Metrics<double> m, const double *__restrict__ G,
                                                  const double *__restrict__ div, const int8_t *__restrict__ igk,
                                                  const double *__restrict__ x, float *__restrict__ fprev,
                                                  const float *__restrict__ cprev, AaCols cols,
                                                  const int *__restrict__ ring, double *__restrict__ partial)
{
    extern __shared__ double xm[];
    stage_xmet(d, m, xm);
    dk_aa_resid(d, m, xm, G, div, igk, x, fprev, cprev, cols, ring, partial, hw_block());
}

// The same history update for the mixed path, where the map's step is
// already in the FP32 correction: f = G - x = e - x(1,1,1) (e holds the V-cycle
// correction plus the GCM changes, p is the pinned iterate).  Writes max|f|
// (component 2) and the dots (components 4..) of the partials; components 0,
// 1 and 3 come from the residual of p (dk_resid64_to32).
template <typename TE>
__device__ __forceinline__ void dk_aa_fresid(const LevelDims &d, const TE *__restrict__ e,
                                             const double *__restrict__ scal, float *__restrict__ fprev,
                                             const float *__restrict__ cprev, const AaCols &cols,
                                             const int *__restrict__ ring, double *__restrict__ partial,
                                             const VBlock &vb)
{
    __shared__ double sh[4 * ROW_COMPONENTS];
    const int have = ring[2];
    const int ns = have ? ring[3] : -1;
    const unsigned live_mask = have ? (unsigned)ring[12] : 0u;
    const double corner = scal[92];
    double v[ROW_COMPONENTS];
    for (int c = 0; c < ROW_COMPONENTS; ++c) v[c] = 0.0;
    // grid-stride over the chunks of all rows: the resident blocks sweep one
    // contiguous window of the fields at a time
    const int cpr = chunks_per_row(d);
    const int nch = (d.ny - 1) * (d.nz - 1) * cpr;
    for (int t = vb.x * blockDim.x + threadIdx.x; t < nch; t += vb.nx * blockDim.x) {
        const int rr = t / cpr;
        const Chunk ch = chunk_at(d, rr, t - rr * cpr);
        const long q0 = ch.q0;
        double ev[4];
        float fp[4], cp[4];
        ld4e(e + q0, ev);
        ld4f(fprev + q0, fp);
        ld4f(cprev + q0, cp);
        double f[4];
        float dfn[4], dgn[4], fn[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            f[k] = ev[k] - corner;
            fn[k] = (float)f[k];
            dfn[k] = (float)(f[k] - (double)fp[k]);
            dgn[k] = (float)(f[k] - (double)cp[k]);
            if (ch.in[k]) v[2] = fmax(v[2], fabs(f[k]));
        }
        if (ns >= 0) {
            st4f(cols.dF[ns] + q0, dfn, ch);
            st4f(cols.dG[ns] + q0, dgn, ch);
#pragma unroll
            for (int sl = 0; sl < 8; ++sl) {
                if (!((live_mask >> sl) & 1u)) continue;
                float cv[4];
                if (sl == ns) { cv[0] = dfn[0]; cv[1] = dfn[1]; cv[2] = dfn[2]; cv[3] = dfn[3]; }
                else ld4f(cols.dF[sl] + q0, cv);
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (ch.in[k]) {
                        v[4 + 2 * sl] += (double)cv[k] * (double)dfn[k];
                        v[5 + 2 * sl] += (double)cv[k] * f[k];
                    }
            }
        }
        st4f(fprev + q0, fn, ch);
    }
    block_reduce<ROW_COMPONENTS, 0xCu>(v, sh);
    if (threadIdx.x == 0) {
        partial[2L * vb.nx + vb.x] = v[2];
        for (int c = 4; c < ROW_COMPONENTS; ++c) partial[(long)c * vb.nx + vb.x] = v[c];
    }
}
template <typename TE>
__global__ void __launch_bounds__(256) k_aa_fresid(LevelDims d, const TE *__restrict__ e,
                                                   const double *__restrict__ scal, float *__restrict__ fprev,
                                                   const float *__restrict__ cprev, AaCols cols,
                                                   const int *__restrict__ ring, double *__restrict__ partial)
{
    dk_aa_fresid<TE>(d, e, scal, fprev, cprev, cols, ring, partial, hw_block());
}

// Final reduction of component-major partials: block c reduces component c
// (sum, or max where the bit of maxmask is set) into out[c].
__device__ __forceinline__ void dk_reduce_cols(const double *__restrict__ partial, int nb, unsigned maxmask,
                                               double *__restrict__ out, const VBlock &vb)
{
    __shared__ double sh[4];
    const int c = vb.x;
    const bool mx = (maxmask >> c) & 1u;
    double a = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        double w = partial[(long)c * nb + b];
        a = mx ? fmax(a, w) : a + w;
    }
    for (int o = 32; o > 0; o >>= 1) {
        double w = __shfl_down(a, o, 64);
        a = mx ? fmax(a, w) : a + w;
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = sh[0];
        for (int w = 1; w < 4; ++w) t = mx ? fmax(t, sh[w]) : t + sh[w];
        out[c] = t;
    }
}

__global__ void __launch_bounds__(256) k_reduce_cols(const double *__restrict__ partial, int nb, unsigned maxmask,
                                                     double *__restrict__ out)
{
    dk_reduce_cols(partial, nb, maxmask, out, hw_block());
}

// Ring update, Gram update and the small least-squares solve (one thread,
// on an LDS copy of the state).
// scal layout: [0..4) residual sums and maxima, [4..20) dots by slot,
// [20..84) Gram by slot pairs, [84..92) gamma by slot, [92] the value of the
// next iterate at cell (1,1,1) (the mean-pinning reference).
// ring: [0] depth m, [1] columns in use K, [2] a previous evaluation exists,
// [3] slot of the next new column, [4..4+K) slots oldest..newest, [12] slots
// in use once the next column is inserted (bit mask).
// Least squares: the Gram matrix of the columns is scaled to unit diagonal
// and factorised by Cholesky; while a pivot falls below AA_PIVOT_TOL (the
// history is numerically dependent) the oldest column is dropped.
#define AA_PIVOT_TOL 1.0e-11
// The block (64 threads) stages ring[0..13) and scal[0..93) in LDS with one
// parallel load; thread 0 runs the update there (no dependent global
// accesses); the block writes the results back.
template <typename TE>
__device__ __forceinline__ void dk_aa_lsq(int *__restrict__ ring_g, double *__restrict__ scal_g, const AaCols &cols,
                                          const double *__restrict__ G, const TE *__restrict__ E, long q111,
                                          int reset)
{
    __shared__ double scal[93];
    __shared__ int ring[13];
    __shared__ double g111s, dg111[8];
    __shared__ double S[64], y[8], sc[8];
    __shared__ int ord[8];
    const int tid = threadIdx.x;
    for (int i = tid; i < 93; i += blockDim.x) scal[i] = scal_g[i];
    if (tid < 13) ring[tid] = ring_g[tid];
    // the map value at cell (1,1,1): G, or p + e in the mixed path; with a
    // Dirichlet pressure face the solution is unique and nothing is pinned
    // (q111 < 0: the pin value stays 0)
    if (tid == 0) g111s = (q111 >= 0) ? G[q111] + (E ? (double)E[q111] : 0.0) : 0.0;
    __syncthreads();
    double *gram = scal + 20, *gam = scal + 84;
    if (tid == 0) {
        const double g111 = g111s;
        const int M = ring[0];
        int K = ring[1];
        for (int s = 0; s < 8; ++s) gam[s] = 0.0;
        if (reset) {
            K = 0;
            ring[2] = 0;
            scal[92] = g111;
        } else {
            if (ring[2]) {
                const int ns = ring[3];
                if (K < M) {
                    ring[4 + K] = ns;
                    K++;
                } else {
                    for (int i = 0; i < M - 1; ++i) ring[4 + i] = ring[5 + i];
                    ring[4 + M - 1] = ns;
                }
                for (int i = 0; i < K; ++i) {
                    int s = ring[4 + i];
                    gram[s * 8 + ns] = gram[ns * 8 + s] = scal[4 + 2 * s];
                }
                // drop the oldest columns until the scaled Gram matrix factorises
                int first = 0;
                bool ok = false;
                while (K - first > 0 && !ok) {
                    const int k = K - first;
                    for (int i = 0; i < k; ++i) {
                        ord[i] = ring[4 + first + i];
                        double g = gram[ord[i] * 9];
                        sc[i] = (g > 0.0) ? 1.0 / sqrt(g) : 0.0;
                    }
                    ok = true;
                    for (int i = 0; i < k && ok; ++i)
                        for (int j = 0; j <= i; ++j) {
                            double v = gram[ord[i] * 8 + ord[j]] * sc[i] * sc[j];
                            for (int t = 0; t < j; ++t) v -= S[i * 8 + t] * S[j * 8 + t];
                            if (j == i) {
                                if (!(v > AA_PIVOT_TOL)) { ok = false; break; }
                                S[i * 8 + i] = sqrt(v);
                            } else {
                                S[i * 8 + j] = v / S[j * 8 + j];
                            }
                        }
                    if (!ok) first++;
                }
                if (first > 0) {
                    for (int i = 0; i < K - first; ++i) ring[4 + i] = ring[4 + first + i];
                    K -= first;
                }
                if (K > 0) {
                    // S S^T y = scaled b; gamma = scaled y
                    for (int i = 0; i < K; ++i) {
                        double v = scal[5 + 2 * ord[i]] * sc[i];
                        for (int t = 0; t < i; ++t) v -= S[i * 8 + t] * y[t];
                        y[i] = v / S[i * 8 + i];
                    }
                    for (int i = K - 1; i >= 0; --i) {
                        double v = y[i];
                        for (int t = i + 1; t < K; ++t) v -= S[t * 8 + i] * y[t];
                        y[i] = v / S[i * 8 + i];
                    }
                    for (int i = 0; i < K; ++i) gam[ord[i]] = y[i] * sc[i];
                }
            }
            ring[2] = 1;
        }
        ring[1] = K;
    }
    __syncthreads();
    // the correction at (1,1,1): slots with a coefficient (allocated columns)
    if (!reset && tid < 8) dg111[tid] = (gam[tid] != 0.0 && q111 >= 0) ? (double)cols.dG[tid][q111] : 0.0;
    __syncthreads();
    if (tid == 0) {
        int K = ring[1];
        const int M = ring[0];
        if (!reset) {
            double c = 0.0;
            for (int s = 0; s < 8; ++s)
                if (gam[s] != 0.0) c += gam[s] * dg111[s];
            scal[92] = g111s - c;
        }
        // slot of the next column and the slots in use after its insertion
        unsigned used = 0u;
        for (int i = 0; i < K; ++i) used |= 1u << ring[4 + i];
        int next;
        if (K < M) {
            next = 0;
            while ((used >> next) & 1u) next++;
            used |= 1u << next;
        } else {
            next = ring[4];
        }
        ring[1] = K;
        ring[3] = next;
        ring[12] = (int)used;
    }
    __syncthreads();
    for (int i = 20 + tid; i < 93; i += blockDim.x) scal_g[i] = scal[i];
    if (tid < 13) ring_g[tid] = ring[tid];
}

template <typename TE>
__global__ void k_aa_lsq(int *__restrict__ ring, double *__restrict__ scal, AaCols cols, const double *__restrict__ G,
                         const TE *__restrict__ E, long q111, int reset)
{
    dk_aa_lsq<TE>(ring, scal, cols, G, E, q111, reset);
}

// x = G - sum_s gamma_s dG_s on interior cells; c = sum_s gamma_s dG_s (FP32);
// p = x - x(1,1,1) (mean pinning, CORE_POISSON.F90:129-150).  G may alias p.
template <typename TE>
__device__ __forceinline__ void dk_aa_next(const LevelDims &d, const double *G, const TE *__restrict__ E,
                                           const AaCols &cols, const double *__restrict__ scal,
                                           double *__restrict__ x, float *__restrict__ c, double *p,
                                           const VBlock &vb)
{
    const double *gam = scal + 84;
    double gm[8];
    int act = 0;
#pragma unroll
    for (int s = 0; s < 8; ++s) { gm[s] = gam[s]; act |= (gm[s] != 0.0) << s; }
    const double corner = scal[92];
    // grid-stride over the chunks of all rows: the resident blocks sweep one
    // contiguous window of the fields at a time
    const int cpr = chunks_per_row(d);
    const int nch = (d.ny - 1) * (d.nz - 1) * cpr;
    for (int t = vb.x * blockDim.x + threadIdx.x; t < nch; t += vb.nx * blockDim.x) {
        const int rr = t / cpr;
        const Chunk ch = chunk_at(d, rr, t - rr * cpr);
        const long q0 = ch.q0;
        double cc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            if (!((act >> s) & 1)) continue;
            float cv[4];
            ld4f(cols.dG[s] + q0, cv);
#pragma unroll
            for (int e = 0; e < 4; ++e) cc[e] += gm[s] * (double)cv[e];
        }
        const D4 gv = ld4(G + q0);      // G may be p: read before p is written
        double ev[4] = {0.0, 0.0, 0.0, 0.0};
        if (E) ld4e(E + q0, ev);
        double xv[4], pv[4];
        float cf[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            xv[e] = (gv.v[e] + ev[e]) - cc[e];
            pv[e] = xv[e] - corner;
            cf[e] = (float)cc[e];
        }
        if (x) st4(x + q0, xv, ch);
        st4f(c + q0, cf, ch);
        st4(p + q0, pv, ch);
    }
}

template <typename TE>
__global__ void __launch_bounds__(256) k_aa_next(LevelDims d, const double *G, const TE *__restrict__ E,
                                                 AaCols cols, const double *__restrict__ scal,
                                                 double *__restrict__ x, float *__restrict__ c, double *p)
{
    dk_aa_next<TE>(d, G, E, cols, scal, x, c, p, hw_block());
}

// ===========================================================================
// Batched kernels (serving).  One launch runs one phase of the outer
// iteration for several requests: the blocks of entry e belong to request
// D[E[e].req] and see the block grid that request's own launch would have,
// so every request computes exactly what its single-grid launch computes,
// whatever else shares the launch.
__device__ __forceinline__ int bent_find(const BEnt *__restrict__ E, int ne)
{
    int lo = 0, hi = ne - 1;
    const int b = (int)blockIdx.x;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (E[mid].b0 <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}
__device__ __forceinline__ VBlock bent_block(const BEnt &e)
{
    const int lb = (int)blockIdx.x - e.b0;
    const int t = lb / e.gx;
    return VBlock{lb - t * e.gx, t % e.gy, t / e.gy, e.gx, e.gy, e.gz};
}
// false for the alignment padding between two entries (the whole block exits)
__device__ __forceinline__ bool bent_live(const BEnt &e)
{
    return (int)blockIdx.x - e.b0 < e.gx * e.gy * e.gz;
}

template <typename T> struct LevSel;
template <> struct LevSel<float> {
    __device__ static const Metrics<float> &m(const ReqLevel &L) { return L.m32; }
    __device__ static float *phi(const ReqLevel &L) { return L.phi32; }
    __device__ static float *rhs(const ReqLevel &L) { return L.rhs32; }
};
template <> struct LevSel<double> {
    __device__ static const Metrics<double> &m(const ReqLevel &L) { return L.m64; }
    __device__ static double *phi(const ReqLevel &L) { return L.phi64; }
    __device__ static double *rhs(const ReqLevel &L) { return L.rhs64; }
};

// fld 0: the level's phi; 1: p; 2: the map value G (aa_g)
template <typename T>
__global__ void __launch_bounds__(256) kb_bc_fill(const BEnt *__restrict__ E, int ne, const ReqDesc *__restrict__ D,
                                                  int lev, int fld, int bcsel)
{
    const BEnt e = E[bent_find(E, ne)];
    if (!bent_live(e)) return;
    const ReqDesc &R = D[e.req];
    const ReqLevel &L = R.L[lev];
    T *f = (fld == 0) ? LevSel<T>::phi(L) : (T *)((fld == 1) ? R.p : R.aa_g);
    dk_bc_fill<T>(L.d, f, R.bc[bcsel], bent_block(e));
}

__global__ void __launch_bounds__(256) kb_gcm_level(const BEnt *__restrict__ E, int ne, const ReqDesc *__restrict__ D,
                                                    int l)
{
    const BEnt e = E[bent_find(E, ne)];
    if (!bent_live(e)) return;
    const ReqDesc &R = D[e.req];
    dk_gcm_level(R.gnode, R.glevel[l], R.glevel[l + 1], R.p, R.pprev, R.off8, R.homo, R.gdelta,
                 bent_block(e));
}

__global__ void __launch_bounds__(256) kb_gcm_fold(const BEnt *__restrict__ E, int ne, const ReqDesc *__restrict__ D,
                                                   int back)
{
    const BEnt e = E[bent_find(E, ne)];
    if (!bent_live(e)) return;
    const ReqDesc &R = D[e.req];
    dk_gcm_fold<float>(R.gnode, R.ntgt, R.gdelta, back ? nullptr : R.L[0].phi32, R.p, R.walls, bent_block(e));
}

__global__ void __launch_bounds__(256) kb_resid64_to32(const BEnt *__restrict__ E, int ne,
                                                       const ReqDesc *__restrict__ D)
{
    const BEnt e = E[bent_find(E, ne)];
    if (!bent_live(e)) return;
    const ReqDesc &R = D[e.req];
    const ReqLevel &L = R.L[0];
    dk_resid64_to32<float>(L.d, L.m64, nullptr, R.p, R.div, L.igk, L.rhs32, L.phi32, R.partial, bent_block(e));
}

template <typename T, int DIR>
__global__ void __launch_bounds__(256) kb_sweep(const BEnt *__restrict__ E, int ne, const ReqDesc *__restrict__ D,
                                                int lev, int color, int bcsel)
{
    const BEnt e = E[bent_find(E, ne)];
    if (!bent_live(e)) return;
    const ReqDesc &R = D[e.req];
    const ReqLevel &L = R.L[lev];
    LineSweep g = L.sw[DIR];
    g.color = color;
    if (DIR == 0)
        dk_xsweep<T, 8>(L.d, LevSel<T>::m(L), LevSel<T>::phi(L), LevSel<T>::rhs(L), L.vrow, g, R.bc[bcsel],
                        bent_block(e));
    else
        dk_yzsweep<T, 8, DIR>(L.d, LevSel<T>::m(L), LevSel<T>::phi(L), LevSel<T>::rhs(L), L.vrow, g, R.bc[bcsel],
                              bent_block(e));
}

// level lev restricted into level lev+1
template <typename T>
__global__ void __launch_bounds__(256) kb_resrestrict(const BEnt *__restrict__ E, int ne,
                                                      const ReqDesc *__restrict__ D, int lev)
{
    const BEnt e = E[bent_find(E, ne)];
    if (!bent_live(e)) return;
    const ReqDesc &R = D[e.req];
    const ReqLevel &F = R.L[lev], &C = R.L[lev + 1];
    dk_resrestrict<T>(F.d, C.d, LevSel<T>::m(F), LevSel<T>::m(C), LevSel<T>::phi(F), LevSel<T>::rhs(F), F.vrow,
                      F.vibk, C.vrflag, LevSel<T>::rhs(C), LevSel<T>::phi(C), bent_block(e));
}

// level lev prolongated into level lev-1
template <typename T>
__global__ void __launch_bounds__(256) kb_prolong(const BEnt *__restrict__ E, int ne, const ReqDesc *__restrict__ D,
                                                  int lev)
{
    const BEnt e = E[bent_find(E, ne)];
    if (!bent_live(e)) return;
    const ReqDesc &R = D[e.req];
    const ReqLevel &F = R.L[lev - 1], &C = R.L[lev];
    dk_prolong<T>(F.d, C.d, LevSel<T>::m(F), LevSel<T>::m(C), LevSel<T>::phi(C), LevSel<T>::phi(F), F.vibk,
                  C.vibk, C.vpflag, R.walls, bent_block(e));
}

__global__ void __launch_bounds__(256) kb_aa_resid(const BEnt *__restrict__ E, int ne, const ReqDesc *__restrict__ D)
{
    const BEnt e = E[bent_find(E, ne)];
    if (!bent_live(e)) return;
    const ReqDesc &R = D[e.req];
    const ReqLevel &L = R.L[0];
    if (R.mixed)
        dk_aa_fresid<float>(L.d, L.phi32, R.scal, R.aa_fprev, R.aa_c, R.cols, R.ring, R.partial, bent_block(e));
    else
        dk_aa_resid(L.d, L.m64, nullptr, R.p, R.div, L.igk, R.aa_x, R.aa_fprev, R.aa_c, R.cols, R.ring, R.partial,
                    bent_block(e));
}

__global__ void __launch_bounds__(256) kb_reduce_cols(const BEnt *__restrict__ E, int ne,
                                                      const ReqDesc *__restrict__ D)
{
    const BEnt e = E[bent_find(E, ne)];
    if (!bent_live(e)) return;
    const ReqDesc &R = D[e.req];
    dk_reduce_cols(R.partial, R.nrb, 0xCu, R.scal, bent_block(e));
}

__global__ void kb_aa_lsq(const BEnt *__restrict__ E, int ne, const ReqDesc *__restrict__ D, int reset)
{
    const BEnt e = E[bent_find(E, ne)];
    if (!bent_live(e)) return;
    const ReqDesc &R = D[e.req];
    dk_aa_lsq<float>(R.ring, R.scal, R.cols, R.p, (reset || !R.mixed) ? nullptr : R.L[0].phi32, R.q111, reset);
}

// start: the first mixing step of a solve (G is the incoming pressure)
__global__ void __launch_bounds__(256) kb_aa_next(const BEnt *__restrict__ E, int ne, const ReqDesc *__restrict__ D,
                                                  int start)
{
    const BEnt e = E[bent_find(E, ne)];
    if (!bent_live(e)) return;
    const ReqDesc &R = D[e.req];
    dk_aa_next<float>(R.L[0].d, R.p, (start || !R.mixed) ? nullptr : R.L[0].phi32, R.cols, R.scal,
               R.mixed ? nullptr : R.aa_x, R.aa_c, R.p, bent_block(e));
}

// out[4*req + c] = scal[c], c < 4 (the stop-test sums of each request)
__global__ void kb_gather(const BEnt *__restrict__ E, int ne, const ReqDesc *__restrict__ D, double *__restrict__ out)
{
    const BEnt e = E[bent_find(E, ne)];
    if (!bent_live(e)) return;
    if (threadIdx.x < 4) out[4 * e.req + threadIdx.x] = D[e.req].scal[threadIdx.x];
}

// ===========================================================================
// launch wrappers
static inline dim3 grid1(long n, int bs) { return dim3((unsigned)((n + bs - 1) / bs)); }

template <typename T>
void bc_fill(hipStream_t st, const LevelDims &d, T *p, const BcInfo &bc)
{
    long n = (long)(d.ny - 1) * (d.nz - 1) + (long)(d.nx + 1) * (d.nz - 1) + (long)(d.nx + 1) * (d.ny - 1);
    hipLaunchKernelGGL(k_bc_fill<T>, grid1(n, 256), dim3(256), 0, st, d, p, bc);
}

template <typename T>
void residual(hipStream_t st, const LevelDims &d, const Metrics<T> &m, const T *phi, const T *rhs, T *res,
              const int8_t *igk)
{
    dim3 g((d.nx + 63) / 64, d.ny - 1, d.nz - 1);
    hipLaunchKernelGGL(k_residual<T>, g, dim3(64), 0, st, d, m, phi, rhs, res, igk);
}

template <typename T>
void resnorm(hipStream_t st, const LevelDims &d, const Metrics<T> &m, const T *phi, const T *rhs,
             const int8_t *igk, double *partial, double *out2)
{
    dim3 g(d.ny - 1, d.nz - 1);
    hipLaunchKernelGGL(k_resnorm<T>, g, dim3(256), 0, st, d, m, phi, rhs, igk, partial);
    long nb = (long)(d.ny - 1) * (d.nz - 1);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, st, partial, nb, 2, 0, 0, out2);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, st, partial, nb, 2, 1, 0, out2 + 1);
}

void maxdiff(hipStream_t st, const LevelDims &d, const double *p, const double *pt, const int8_t *iblank,
             double *partial, double *out)
{
    dim3 g(d.ny - 1, d.nz - 1);
    hipLaunchKernelGGL(k_maxdiff, g, dim3(256), 0, st, d, p, pt, iblank, partial);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, st, partial, (long)(d.ny - 1) * (d.nz - 1), 1, 0, 1, out);
}

void pin(hipStream_t st, const LevelDims &d, double *p, double *corner)
{
    hipLaunchKernelGGL(k_read_corner, dim3(1), dim3(1), 0, st, d, (const double *)p, corner);
    dim3 g((d.nx + 1 + 63) / 64, d.ny + 1, d.nz + 1);
    hipLaunchKernelGGL(k_pin, g, dim3(64), 0, st, d, p, (const double *)corner);
}

template <typename T>
void restrict_(hipStream_t st, const LevelDims &f, const LevelDims &c, const Metrics<T> &mf, const Metrics<T> &mc,
               const T *res_f, T *rhs_c, const int8_t *ibk_f, const int8_t *rflag_c)
{
    dim3 g((c.nx - 1 + 63) / 64, (c.ny - 1 + 3) / 4, c.nz - 1);
    hipLaunchKernelGGL(k_restrict<T>, g, dim3(64, 4), 0, st, f, c, mf, mc, res_f, rhs_c, ibk_f, rflag_c);
}

template <typename T>
void resrestrict(hipStream_t st, const LevelDims &f, const LevelDims &c, const Metrics<T> &mf, const Metrics<T> &mc,
                 const T *phi_f, const T *rhs_f, const int8_t *igk_f, const int8_t *ibk_f, const int8_t *rflag_c,
                 T *rhs_c, T *phi_c)
{
    dim3 g((c.nx - 1 + 63) / 64, (c.ny - 1 + 3) / 4, c.nz - 1);
    hipLaunchKernelGGL(k_resrestrict<T>, g, dim3(64, 4), 0, st, f, c, mf, mc, phi_f, rhs_f, igk_f, ibk_f, rflag_c,
                       rhs_c, phi_c);
}

dim3 prolong_grid(const LevelDims &f, const LevelDims &c)
{
    return dim3((prolong_ng(f.nx, c.ic) + 63) / 64, (prolong_ng(f.ny, c.jc) + 3) / 4, prolong_ng(f.nz, c.kc));
}

template <typename T>
void prolong(hipStream_t st, const LevelDims &f, const LevelDims &c, const Metrics<T> &mf, const Metrics<T> &mc,
             const T *phi_c, T *phi_f, const int8_t *ibk_f, const int8_t *ibk_c, const int8_t *pflag_c, int walls)
{
    hipLaunchKernelGGL(k_prolong<T>, prolong_grid(f, c), dim3(64, 4), 0, st, f, c, mf, mc, phi_c, phi_f, ibk_f, ibk_c,
                       pflag_c, walls);
}

// choose lanes per line and rows per lane for a line of L rows
static void line_shape(int L, int &LPL, int &E)
{
    int need = (L + 7) / 8;
    LPL = 1;
    while (LPL < need && LPL < 64) LPL <<= 1;
    E = (L + LPL - 1) / LPL;
}

template <typename T, int E>
static void launch_line_sweep(hipStream_t st, const LevelDims &d, const Metrics<T> &m, T *phi, const T *rhs,
                              const int8_t *igk, int dir, const LineSweep &g, const BcInfo &bc)
{
    const size_t sh = sweep_lds_bytes(g, sizeof(T));
    if (dir == 0)
        hipLaunchKernelGGL((k_xsweep<T, E>), dim3(g.nblocks), dim3(256), sh, st, d, m, phi, rhs, igk, g, bc);
    else if (dir == 1)
        hipLaunchKernelGGL((k_yzsweep<T, E, 1>), dim3(g.nblocks), dim3(256), sh, st, d, m, phi, rhs, igk, g, bc);
    else
        hipLaunchKernelGGL((k_yzsweep<T, E, 2>), dim3(g.nblocks), dim3(256), sh, st, d, m, phi, rhs, igk, g, bc);
}

// One colour of one direction of itsolv_mg (CORE_MG.F90:615-938) on a level.
// The ghost cells of the updated cells are rewritten by the sweep (bc_post),
// so a level needs one BC fill before its six half sweeps, not one per sweep.
size_t sweep_lds_bytes(const LineSweep &g, size_t elem)
{
    return ((size_t)g.NL * g.Rp * 2 * elem + (size_t)(g.R + 2) * elem + 15) / 16 * 16;
}

LineSweep sweep_shape(const LevelDims &d, int dir, int color, double omega)
{
    LineSweep g;
    g.color = color;
    g.omega = omega;
    g.post = 1;
    g.relax_delta = 0;
    const int naxis = (dir == 0) ? d.nx : (dir == 1) ? d.ny : d.nz;
    g.R = naxis - 1;
    const int E = (g.R > 512) ? 16 : 8;
    int lpl = 1, sh = 0;
    while (lpl * E < g.R) { lpl <<= 1; sh++; }
    g.LPL = lpl;
    g.lpl_shift = sh;
    g.Rp = (g.R + g.R / 16 + 1) | 1;
    // lines per block: 16 on large levels, fewer where 16 would leave the
    // chip with too few blocks to hide the gather's latency
    {
        const long lines = (dir == 0) ? (long)(d.ny / 2) * (d.nz - 1)
                                      : (long)(d.nx / 2) * ((dir == 1) ? d.nz - 1 : d.ny - 1);
        g.NL = (lines >= 4096) ? 16 : 4;
    }
    long nb;
    if (dir == 0) {
        long nl = (long)(d.ny / 2) * (d.nz - 1);
        nb = (nl + g.NL - 1) / g.NL;
        g.chunks = 0;
    } else {
        g.chunks = (d.nx + 2 * g.NL - 1) / (2 * g.NL);
        nb = (long)g.chunks * ((dir == 1) ? d.nz - 1 : d.ny - 1);
    }
    g.nblocks = (int)((nb + 7) / 8 * 8);     // the XCD remap needs a multiple of 8
    return g;
}

template <typename T>
void sweep(hipStream_t st, const LevelDims &d, const Metrics<T> &m, T *phi, const T *rhs, const int8_t *igk,
           int dir, int color, double omega, const BcInfo &bc)
{
    const LineSweep g = sweep_shape(d, dir, color, omega);
    if (g.R > 512) launch_line_sweep<T, 16>(st, d, m, phi, rhs, igk, dir, g, bc);
    else launch_line_sweep<T, 8>(st, d, m, phi, rhs, igk, dir, g, bc);
}

void sweep_itsolv(hipStream_t st, const LevelDims &d, const Metrics<double> &m, double *phi, const double *rhs,
                  const int8_t *igk, int dir, int color, double omega, const BcInfo &bc)
{
    LineSweep g = sweep_shape(d, dir, color, omega);
    g.post = 0;
    g.relax_delta = 1;
    if (g.R > 512) launch_line_sweep<double, 16>(st, d, m, phi, rhs, igk, dir, g, bc);
    else launch_line_sweep<double, 8>(st, d, m, phi, rhs, igk, dir, g, bc);
}

int tdma_batched(hipStream_t st, int nsys, int n, const double *a, const double *b, const double *c,
                 const double *r, double *u, int prec)
{
    int LPL, E;
    line_shape(n, LPL, E);
    if (E > 9) return -1;
    int wl = 64 / LPL;
    unsigned nb = (unsigned)((nsys + wl - 1) / wl);
#define TDMA_CASE(EE)                                                                                           \
    case EE:                                                                                                    \
        if (prec) hipLaunchKernelGGL((k_tdma_batched<float, EE>), dim3(nb), dim3(64), 0, st, nsys, n, LPL, a, b, c, r, u); \
        else hipLaunchKernelGGL((k_tdma_batched<double, EE>), dim3(nb), dim3(64), 0, st, nsys, n, LPL, a, b, c, r, u); \
        break;
    switch (E) {
        TDMA_CASE(1) TDMA_CASE(2) TDMA_CASE(3) TDMA_CASE(4) TDMA_CASE(5) TDMA_CASE(6) TDMA_CASE(7) TDMA_CASE(8)
        TDMA_CASE(9)
    }
#undef TDMA_CASE
    return 0;
}

void gcm_dead(hipStream_t st, const GcmDev &g, const double *pres, const double *pprev, const long *off8, int homo,
              double *outv)
{
    if (g.nD == 0) return;
    hipLaunchKernelGGL(k_gcm_dead, grid1(g.nD, 256), dim3(256), 0, st, g, pres, pprev, off8, homo, outv);
}

void gcm_ghost(hipStream_t st, const GcmDev &g, const double *pres, const double *r, const long *off8, int homo,
               int mix, const LevelDims &d, const double *const *met, const int *gijk, double *outv, double *reslap)
{
    if (g.nG == 0) return;
    hipLaunchKernelGGL(k_gcm_ghost, grid1(g.nG, 256), dim3(256), 0, st, g, pres, r, off8, homo, mix, d, met[0],
                       met[1], met[2], met[3], met[4], met[5], gijk, outv, reslap);
}

void gcm_scatter(hipStream_t st, int n, const long *target, const int *active, const double *vals, double *pres,
                 double *partial, double *out)
{
    if (n == 0) {
        hipLaunchKernelGGL(k_fill<double>, dim3(1), dim3(64), 0, st, 1L, out, 0.0);
        return;
    }
    unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_gcm_scatter, dim3(nb), dim3(256), 0, st, n, target, active, vals, pres, partial);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, st, partial, (long)nb, 1, 0, 1, out);
}

void gcm_levels(hipStream_t st, const GcmNode *nodes, const int *level_off, int nlevels, double *pres,
                const double *pprev, const long *off8, int homo, double *gdelta)
{
    for (int l = 0; l < nlevels; ++l) {
        int lo = level_off[l], hi = level_off[l + 1];
        if (hi <= lo) continue;
        hipLaunchKernelGGL(k_gcm_level, dim3((unsigned)((hi - lo + 255) / 256)), dim3(256), 0, st, nodes, lo, hi, pres,
                           pprev, off8, homo, gdelta);
    }
}

template <typename TE>
void gcm_fold(hipStream_t st, const GcmNode *nodes, int n, const double *gdelta, TE *e, double *p, int assign)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gcm_fold<TE>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, nodes, n, gdelta, e, p,
                       assign);
}

void rhs_poisson(hipStream_t st, const LevelDims &d, const double *fu, const double *fv, const double *fw,
                 const int8_t *igk, const double *const *met, double dtinv, double *div, double *partial,
                 double *out2)
{
    rhs_poisson2(st, d, d, fu, fv, fw, igk, met, dtinv, div, partial, out2);
}

void rhs_poisson2(hipStream_t st, const LevelDims &d, const LevelDims &dm, const double *fu, const double *fv,
                  const double *fw, const int8_t *igk, const double *const *met, double dtinv, double *div,
                  double *partial, double *out2)
{
    dim3 g(d.ny - 1, d.nz - 1);
    hipLaunchKernelGGL(k_rhs, g, dim3(256), 0, st, d, dm, fu, fv, fw, igk, met[0], met[1], met[2], met[3], met[4],
                       met[5], dtinv, div, partial);
    long nb = (long)(d.ny - 1) * (d.nz - 1);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, st, partial, nb, 2, 0, 0, out2);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, st, partial, nb, 2, 1, 1, out2 + 1);
}

// ---------------------------------------------------------------------------
// Direct solve of the coarsest level: e(q[t]) = sum_s M[t][s] rhs(q[s]) in
// FP64 over the n free cells (M row-major, n x n), one wave per row.
template <typename T>
__global__ void __launch_bounds__(256) k_coarse_direct(int n, const double *__restrict__ M,
                                                       const long *__restrict__ q, const T *__restrict__ rhs,
                                                       T *__restrict__ phi)
{
    const int lane = threadIdx.x & 63;
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= n) return;     // the whole wave leaves together
    const double *row = M + (long)t * n;
    double a = 0.0;
    for (int s = lane; s < n; s += 64) a += row[s] * (double)rhs[q[s]];
    for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
    if (lane == 0) phi[q[t]] = (T)a;
}

template <typename T>
void coarse_direct(hipStream_t st, int n, const double *M, const long *q, const T *rhs, T *phi)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_coarse_direct<T>, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, n, M, q, rhs, phi);
}
template void coarse_direct<float>(hipStream_t, int, const double *, const long *, const float *, float *);
template void coarse_direct<double>(hipStream_t, int, const double *, const long *, const double *, double *);

// ---------------------------------------------------------------------------
// Zero-gradient (outflow) outer faces.  enforce_global_mass_consv
// (CORE_SET_OUTER_BC.F90:515-735) removes the net mass flux of the face
// velocities by a uniform velocity on the zero-gradient faces;
// GCM_enforce_p_compatibility (GCM_GHOSTCELL_UPDATE.F90:865-1003) removes the
// net pressure-gradient flux of the pressure by a uniform gradient there.
// Both sum over the flag cells (one block per interior (j,k) row, partials
// reduced in row order) plus, with mix_GC_form = 1, terms of the ghost cells.
// zg: bit f set when outer face f (x1, x2, y1, y2, z1, z2) is zero-gradient.

// the net face mass flux of the flag cells of each row
__global__ void __launch_bounds__(256) k_massflux_rows(LevelDims d, const double *__restrict__ fu,
                                                       const double *__restrict__ fv, const double *__restrict__ fw,
                                                       const int8_t *__restrict__ igk, const double *__restrict__ dx,
                                                       const double *__restrict__ dy, const double *__restrict__ dz,
                                                       double *__restrict__ partial)
{
    __shared__ double s1[256];
    const int j = 1 + blockIdx.x, k = 1 + blockIdx.y;
    double a = 0.0;
    for (int i = 1 + threadIdx.x; i <= d.nx - 1; i += blockDim.x) {
        const long q = idx(d, i, j, k);
        if (igk[q]) continue;
        a += (-fu[q] + fu[q + 1]) * dy[j] * dz[k] + (-fv[q] + fv[q + d.sy]) * dx[i] * dz[k]
             + (-fw[q] + fw[q + d.sz]) * dx[i] * dy[j];
    }
    s1[threadIdx.x] = a;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s1[threadIdx.x] += s1[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x + (long)gridDim.x * blockIdx.y] = s1[0];
}

// correct_ghost_massflux (GCM_GHOSTCELL_UPDATE.F90:742-787): the ghost
// cells' face mass fluxes weighted by cos(pi/2 (1 - alphaGhost)), per-block
// partial sums; with div, also div = dtinv * (face divergence) * weight
__global__ void __launch_bounds__(256) k_ghost_massflux(LevelDims d, const GhostVel *__restrict__ g, int n,
                                                        const double *__restrict__ fu, const double *__restrict__ fv,
                                                        const double *__restrict__ fw, const double *__restrict__ dx,
                                                        const double *__restrict__ dy, const double *__restrict__ dz,
                                                        const double *__restrict__ dxinv,
                                                        const double *__restrict__ dyinv,
                                                        const double *__restrict__ dzinv, double dtinv,
                                                        double *__restrict__ div, double *__restrict__ partial)
{
    __shared__ double s1[256];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.0;
    if (t < n) {
        const GhostVel gc = g[t];
        const int i = gc.i, j = gc.j, k = gc.k;
        const long q = idx(d, i, j, k);
        const double w = cos(3.14159265358979323846 / 2.0 * (1.0 - gc.alpha));
        a = w * ((-fu[q] + fu[q + 1]) * dy[j] * dz[k] + (-fv[q] + fv[q + d.sy]) * dx[i] * dz[k]
                 + (-fw[q] + fw[q + d.sz]) * dx[i] * dy[j]);
        if (div) {
            const double v = dtinv * ((fu[q + 1] - fu[q]) * dxinv[i] + (fv[q + d.sy] - fv[q]) * dyinv[j]
                                      + (fw[q + d.sz] - fw[q]) * dzinv[k]);
            div[q] = v * w;
        }
    }
    s1[threadIdx.x] = a;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s1[threadIdx.x] += s1[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s1[0];
}

// the velocity correction on the zero-gradient faces: face velocities of the
// interior cells (face_u(1|nx, j, k), ...) and, when given, the outer-face
// velocities bcxu(0:1,0:ny+1,0:nz+1), bcyv(0:nx+1,0:1,0:nz+1),
// bczw(0:nx+1,0:ny+1,0:1) on 0..n of their plane.  Thread (a, b) of a 2-D
// grid covering the largest plane.
__global__ void k_outflow_velocity(LevelDims d, int zg, double corr, double *__restrict__ fu,
                                   double *__restrict__ fv, double *__restrict__ fw, double *__restrict__ bu,
                                   double *__restrict__ bv, double *__restrict__ bw)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y * blockDim.y + threadIdx.y;
    const int nx = d.nx, ny = d.ny, nz = d.nz, hx = nx + 2, hy = ny + 2;
    // x faces: (a, b) = (j, k)
    if (a >= 1 && a <= ny - 1 && b >= 1 && b <= nz - 1) {
        if (zg & 1) fu[idx(d, 1, a, b)] -= corr;
        if (zg & 2) fu[idx(d, nx, a, b)] += corr;
    }
    if (bu && a <= ny && b <= nz) {
        if (zg & 1) bu[0 + 2 * (a + (long)hy * b)] -= corr;
        if (zg & 2) bu[1 + 2 * (a + (long)hy * b)] += corr;
    }
    // y faces: (a, b) = (i, k)
    if (a >= 1 && a <= nx - 1 && b >= 1 && b <= nz - 1) {
        if (zg & 4) fv[idx(d, a, 1, b)] -= corr;
        if (zg & 8) fv[idx(d, a, ny, b)] += corr;
    }
    if (bv && a <= nx && b <= nz) {
        if (zg & 4) bv[a + (long)hx * (0 + 2 * b)] -= corr;
        if (zg & 8) bv[a + (long)hx * (1 + 2 * b)] += corr;
    }
    // z faces: (a, b) = (i, j)
    if (a >= 1 && a <= nx - 1 && b >= 1 && b <= ny - 1) {
        if (zg & 16) fw[idx(d, a, b, 1)] -= corr;
        if (zg & 32) fw[idx(d, a, b, nz)] += corr;
    }
    if (bw && a <= nx && b <= ny) {
        if (zg & 16) bw[a + (long)hx * (b + (long)hy * 0)] -= corr;
        if (zg & 32) bw[a + (long)hx * (b + (long)hy * 1)] += corr;
    }
}

// the pressure-gradient flux of the flag cells of each row
__global__ void __launch_bounds__(256) k_pflux_rows(LevelDims d, const double *__restrict__ p,
                                                    const int8_t *__restrict__ igk, const double *__restrict__ dxcinv,
                                                    const double *__restrict__ dycinv,
                                                    const double *__restrict__ dzcinv, const double *__restrict__ dx,
                                                    const double *__restrict__ dy, const double *__restrict__ dz,
                                                    double *__restrict__ partial)
{
    __shared__ double s1[256];
    const int j = 1 + blockIdx.x, k = 1 + blockIdx.y;
    double a = 0.0;
    for (int i = 1 + threadIdx.x; i <= d.nx - 1; i += blockDim.x) {
        const long q = idx(d, i, j, k);
        if (igk[q]) continue;
        const double pc = p[q];
        const double pgxw = (pc - p[q - 1]) * dxcinv[i], pgxe = (p[q + 1] - pc) * dxcinv[i + 1];
        const double pgys = (pc - p[q - d.sy]) * dycinv[j], pgyn = (p[q + d.sy] - pc) * dycinv[j + 1];
        const double pgzb = (pc - p[q - d.sz]) * dzcinv[k], pgzf = (p[q + d.sz] - pc) * dzcinv[k + 1];
        a += (-pgxw + pgxe) * dy[j] * dz[k] + (-pgys + pgyn) * dx[i] * dz[k] + (-pgzb + pgzf) * dx[i] * dy[j];
    }
    s1[threadIdx.x] = a;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s1[threadIdx.x] += s1[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x + (long)gridDim.x * blockIdx.y] = s1[0];
}

// the ghost cells' pressure-gradient flux less res_laplacian x volume
// (mix_GC_form = 1), per-block partial sums
__global__ void __launch_bounds__(256) k_pflux_ghost(LevelDims d, const GhostVel *__restrict__ g, int n,
                                                     const double *__restrict__ p, const double *__restrict__ reslap,
                                                     const double *__restrict__ dxcinv,
                                                     const double *__restrict__ dycinv,
                                                     const double *__restrict__ dzcinv, const double *__restrict__ dx,
                                                     const double *__restrict__ dy, const double *__restrict__ dz,
                                                     double *__restrict__ partial)
{
    __shared__ double s1[256];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.0;
    if (t < n) {
        const GhostVel gc = g[t];
        const int i = gc.i, j = gc.j, k = gc.k;
        const long q = idx(d, i, j, k);
        const double pc = p[q];
        const double pgxw = (pc - p[q - 1]) * dxcinv[i], pgxe = (p[q + 1] - pc) * dxcinv[i + 1];
        const double pgys = (pc - p[q - d.sy]) * dycinv[j], pgyn = (p[q + d.sy] - pc) * dycinv[j + 1];
        const double pgzb = (pc - p[q - d.sz]) * dzcinv[k], pgzf = (p[q + d.sz] - pc) * dzcinv[k + 1];
        a = ((-pgxw + pgxe) * dy[j] * dz[k] + (-pgys + pgyn) * dx[i] * dz[k] + (-pgzb + pgzf) * dx[i] * dy[j])
            - reslap[t] * dx[i] * dy[j] * dz[k];
    }
    s1[threadIdx.x] = a;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s1[threadIdx.x] += s1[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s1[0];
}

// pgrad on the zero-gradient faces += or -= -(flux[0] + flux[1]) / area over
// 0..n of the face plane (pgradx1(0:ny+1,0:nz+1) etc.)
__global__ void k_pgrad_correct(int nx, int ny, int nz, int zg, const double *__restrict__ flux, double area,
                                double *__restrict__ px1, double *__restrict__ px2, double *__restrict__ py1,
                                double *__restrict__ py2, double *__restrict__ pz1, double *__restrict__ pz2)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y * blockDim.y + threadIdx.y;
    const double corr = -(flux[0] + flux[1]) / area;
    if (a <= ny && b <= nz) {
        const long q = a + (long)(ny + 2) * b;
        if (zg & 1) px1[q] -= corr;
        if (zg & 2) px2[q] += corr;
    }
    if (a <= nx && b <= nz) {
        const long q = a + (long)(nx + 2) * b;
        if (zg & 4) py1[q] -= corr;
        if (zg & 8) py2[q] += corr;
    }
    if (a <= nx && b <= ny) {
        const long q = a + (long)(nx + 2) * b;
        if (zg & 16) pz1[q] -= corr;
        if (zg & 32) pz2[q] += corr;
    }
}

// out2[0] /= dtinv (the net mass flux sumd of rhs_poisson, CORE_POISSON.F90:
// 52-66, without the time-step factor)
__global__ void k_rhs_sumd(double *out2, double dtinv) { out2[0] = out2[0] / dtinv; }
void rhs_sumd(hipStream_t st, double *out2, double dtinv)
{
    hipLaunchKernelGGL(k_rhs_sumd, dim3(1), dim3(1), 0, st, out2, dtinv);
}

void massflux(hipStream_t st, const LevelDims &d, const double *fu, const double *fv, const double *fw,
              const int8_t *igk, const double *const *met, double *partial, double *out)
{
    dim3 g(d.ny - 1, d.nz - 1);
    hipLaunchKernelGGL(k_massflux_rows, g, dim3(256), 0, st, d, fu, fv, fw, igk, met[3], met[4], met[5], partial);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, st, partial, (long)(d.ny - 1) * (d.nz - 1), 1, 0, 0, out);
}

void ghost_massflux(hipStream_t st, const LevelDims &d, const GhostVel *g, int n, const double *fu, const double *fv,
                    const double *fw, const double *const *met, double dtinv, double *div, double *partial,
                    double *out)
{
    const int nb = (n + 255) / 256;
    if (n > 0)
        hipLaunchKernelGGL(k_ghost_massflux, dim3(nb), dim3(256), 0, st, d, g, n, fu, fv, fw, met[3], met[4], met[5],
                           met[0], met[1], met[2], dtinv, div, partial);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, st, partial, (long)nb, 1, 0, 0, out);
}

static dim3 plane_grid(const LevelDims &d)
{
    const int m = std::max(d.nx, std::max(d.ny, d.nz)) + 1;
    return dim3((unsigned)((m + 15) / 16), (unsigned)((m + 15) / 16));
}

void outflow_velocity(hipStream_t st, const LevelDims &d, int zg, double corr, double *fu, double *fv, double *fw,
                      double *bu, double *bv, double *bw)
{
    hipLaunchKernelGGL(k_outflow_velocity, plane_grid(d), dim3(16, 16), 0, st, d, zg, corr, fu, fv, fw, bu, bv, bw);
}

void pressure_flux(hipStream_t st, const LevelDims &d, const double *p, const int8_t *igk, const double *const *l1met,
                   const double *const *l1w, const GhostVel *g, int ng, const double *reslap, double *partial,
                   double *out2)
{
    dim3 gr(d.ny - 1, d.nz - 1);
    hipLaunchKernelGGL(k_pflux_rows, gr, dim3(256), 0, st, d, p, igk, l1met[0], l1met[1], l1met[2], l1w[3], l1w[4],
                       l1w[5], partial);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, st, partial, (long)(d.ny - 1) * (d.nz - 1), 1, 0, 0, out2);
    const int nb = (ng + 255) / 256;
    if (ng > 0 && reslap) {
        hipLaunchKernelGGL(k_pflux_ghost, dim3(nb), dim3(256), 0, st, d, g, ng, p, reslap, l1met[0], l1met[1],
                           l1met[2], l1w[3], l1w[4], l1w[5], partial);
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, st, partial, (long)nb, 1, 0, 0, out2 + 1);
    } else {
        hipLaunchKernelGGL(k_fill<double>, dim3(1), dim3(1), 0, st, 1L, out2 + 1, 0.0);
    }
}

void pgrad_correct(hipStream_t st, const LevelDims &d, int zg, const double *flux2, double area, double *const *pg)
{
    hipLaunchKernelGGL(k_pgrad_correct, plane_grid(d), dim3(16, 16), 0, st, d.nx, d.ny, d.nz, zg, flux2, area, pg[0],
                       pg[1], pg[2], pg[3], pg[4], pg[5]);
}

// ---------------------------------------------------------------------------
// Velocity projection (the calls after solve_poisson, CORE_TIME_STEP.F90:
// 211-222), FP64.  Every product and sum is rounded on its own (no fused
// multiply-add), in the reference's order of operations, so the results are
// the reference's to the bit.
//
// Face-interpolated pressure difference across cell i along one axis
// (correct_vel, CORE_FRAC_STEP_ROUTINES.F90:165-177):
//   pe = f(i+1) p(i+1) + (1 - f(i+1)) p(i),  pw = f(i) p(i) + (1 - f(i)) p(i-1),
//   result (pe - pw) * dinv(i).
__device__ __forceinline__ double face_grad(double f1, double f0, double pp, double pc, double pm, double dinv)
{
#pragma clang fp contract(off)
    const double pe = f1 * pp + (1.0 - f1) * pc;
    const double pw = f0 * pc + (1.0 - f0) * pm;
    return (pe - pw) * dinv;
}

// correct_vel (CORE_FRAC_STEP_ROUTINES.F90:134-210): on the interior cells,
// u -= dt pgx (1 - iblank) and likewise v, w.  One 4-cell chunk of an
// interior row per thread; blocks are numbered so that each XCD owns one
// contiguous eighth of the rows (the launch has a multiple of 8 blocks), so
// the j-1/j+1 and k-1/k+1 rows a block reads were fetched into the same L2.
__global__ void __launch_bounds__(256) k_correct_vel(LevelDims d, ProjMet m, const double *__restrict__ pp,
                                                     const int8_t *__restrict__ iblank, double dt,
                                                     double *__restrict__ u, double *__restrict__ v,
                                                     double *__restrict__ w)
{
#pragma clang fp contract(off)
    const int cpr = chunks_per_row(d);
    const int nch = (d.ny - 1) * (d.nz - 1) * cpr;
    {
        const int t = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
        if (t >= nch) return;
        const int rr = t / cpr;
        const Chunk ch = chunk_at(d, rr, t - rr * cpr);
        const long q0 = ch.q0;
        const int j = ch.j, k = ch.k;
        const D4 pc = ld4(pp + q0), pn = ld4(pp + q0 + d.sy), ps = ld4(pp + q0 - d.sy);
        const D4 pf = ld4(pp + q0 + d.sz), pb = ld4(pp + q0 - d.sz);
        const double pl = pp[q0 - 1], pr = pp[q0 + 4];
        const D4 uc = ld4(u + q0), vc = ld4(v + q0), wc = ld4(w + q0);
        const int ib4 = *(const int *)(iblank + q0);
        const double fy1 = m.fy[j + 1], fy0 = m.fy[j], fz1 = m.fz[k + 1], fz0 = m.fz[k];
        const double dyi = m.dyinv[j], dzi = m.dzinv[k];
        double ou[4], ov[4], ow[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int i = ch.i0 + e;
            const double pm = e == 0 ? pl : pc.v[e > 0 ? e - 1 : 0];
            const double pq = e == 3 ? pr : pc.v[e < 3 ? e + 1 : 3];
            const double pgx = face_grad(m.fx[i + 1], m.fx[i], pq, pc.v[e], pm, m.dxinv[i]);
            const double pgy = face_grad(fy1, fy0, pn.v[e], pc.v[e], ps.v[e], dyi);
            const double pgz = face_grad(fz1, fz0, pf.v[e], pc.v[e], pb.v[e], dzi);
            const double flag = 1.0 - (double)(int8_t)(ib4 >> (8 * e));
            ou[e] = uc.v[e] - dt * pgx * flag;
            ov[e] = vc.v[e] - dt * pgy * flag;
            ow[e] = wc.v[e] - dt * pgz * flag;
        }
        st4(u + q0, ou, ch);
        st4(v + q0, ov, ch);
        st4(w + q0, ow, ch);
    }
}

// correct_ghostcell_velocity (GCM_GHOSTCELL_UPDATE.F90:789-852): one thread
// per ghost cell, u -= dt pgx flag with flag = alphaGhost when mix_GC_form = 1
// and 0 otherwise.  Each ghost cell changes only its own u, v, w, and the
// list holds every cell once, so the list order does not matter.
__global__ void __launch_bounds__(256) k_correct_ghost_vel(LevelDims d, ProjMet m, const GhostVel *__restrict__ g,
                                                           int n, int mix, const double *__restrict__ pv, double dt,
                                                           double *__restrict__ u, double *__restrict__ v,
                                                           double *__restrict__ w)
{
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const GhostVel c = g[t];
    const long q = idx(d, c.i, c.j, c.k);
    const double flag = mix ? c.alpha : 0.0;
    const double pc = pv[q];
    const double pgx = face_grad(m.fx[c.i + 1], m.fx[c.i], pv[q + 1], pc, pv[q - 1], m.dxinv[c.i]);
    const double pgy = face_grad(m.fy[c.j + 1], m.fy[c.j], pv[q + d.sy], pc, pv[q - d.sy], m.dyinv[c.j]);
    const double pgz = face_grad(m.fz[c.k + 1], m.fz[c.k], pv[q + d.sz], pc, pv[q - d.sz], m.dzinv[c.k]);
    u[q] = u[q] - dt * pgx * flag;
    v[q] = v[q] - dt * pgy * flag;
    w[q] = w[q] - dt * pgz * flag;
}

// face_vel (CORE_FRAC_STEP_ROUTINES.F90:6-130) at one rank:
//   face_u(i,j,k) = fx(i) u(i) + (1 - fx(i)) u(i-1)   i = 2..nx-1, j < ny, k < nz
//   face_u(1,j,k) = bcxu(0,j,k),  face_u(nx,j,k) = bcxu(1,j,k)
// and likewise face_v (j = 2..ny-1; bcyv at j = 1, ny) and face_w (k =
// 2..nz-1; bczw at k = 1, nz).  Entries outside those ranges are left as
// they are.  One 4-cell chunk of the rows j = 1..ny, k = 1..nz (cells
// i = 1..nx) per thread, blocks in the XCD-contiguous order of k_correct_vel.
__global__ void __launch_bounds__(256) k_face_vel(LevelDims d, ProjMet m, const double *__restrict__ u,
                                                  const double *__restrict__ v, const double *__restrict__ w,
                                                  const double *__restrict__ bcu, const double *__restrict__ bcv,
                                                  const double *__restrict__ bcw, double *__restrict__ fu,
                                                  double *__restrict__ fv, double *__restrict__ fw)
{
#pragma clang fp contract(off)
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const long hx = nx + 2, hy = ny + 2;     // extents of the bc arrays (host layout)
    const int cpr = (nx >> 2) + 1;
    const long nch = (long)ny * nz * cpr;
    {
        const long t = (long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
        if (t >= nch) return;
        const long r = t / cpr;
        const int c = (int)(t - r * cpr);
        const int k = 1 + (int)(r / ny);
        const int j = 1 + (int)(r - (long)(k - 1) * ny);
        const int i0 = 4 * c;
        const long q0 = (long)i0 + (long)j * d.sy + (long)k * d.sz;
        const D4 uc = ld4(u + q0), vc = ld4(v + q0), vs = ld4(v + q0 - d.sy);
        const D4 wc = ld4(w + q0), wb = ld4(w + q0 - d.sz);
        const double ul = u[q0 - 1];
        const double fyj = m.fy[j], fzk = m.fz[k];
        double ou[4], ov[4], ow[4];
        int mu = 0, mv = 0, mw = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int i = i0 + e;
            const bool cell = i >= 1 && i <= nx;
            const double um = e == 0 ? ul : uc.v[e > 0 ? e - 1 : 0];
            if (cell && j <= ny - 1 && k <= nz - 1) {
                const long b = 2 * ((long)j + hy * k);
                ou[e] = i == 1 ? bcu[b] : i == nx ? bcu[b + 1] : m.fx[i] * uc.v[e] + (1.0 - m.fx[i]) * um;
                mu |= 1 << e;
            } else {
                ou[e] = 0.0;
            }
            if (cell && i <= nx - 1 && k <= nz - 1) {
                const long b = (long)i + hx * 2 * (long)k;
                ov[e] = j == 1 ? bcv[b] : j == ny ? bcv[b + hx] : fyj * vc.v[e] + (1.0 - fyj) * vs.v[e];
                mv |= 1 << e;
            } else {
                ov[e] = 0.0;
            }
            if (cell && i <= nx - 1 && j <= ny - 1) {
                const long b = (long)i + hx * (long)j;
                ow[e] = k == 1 ? bcw[b] : k == nz ? bcw[b + hx * hy] : fzk * wc.v[e] + (1.0 - fzk) * wb.v[e];
                mw |= 1 << e;
            } else {
                ow[e] = 0.0;
            }
        }
        double *dst[3] = {fu + q0, fv + q0, fw + q0};
        const int msk[3] = {mu, mv, mw};
        const double *val[3] = {ou, ov, ow};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (msk[a] == 15) {
                *(double2 *)dst[a] = make_double2(val[a][0], val[a][1]);
                *(double2 *)(dst[a] + 2) = make_double2(val[a][2], val[a][3]);
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (msk[a] >> e & 1) dst[a][e] = val[a][e];
            }
        }
    }
}

// update_pressure (CORE_FRAC_STEP_ROUTINES.F90:212-299): pPrime = 0 where
// pzero is set, then p = pPrime over the whole array.  Four entries per
// thread; the padded field length is a multiple of 16.
__global__ void __launch_bounds__(256) k_update_pressure(long n4, const int *__restrict__ pz4,
                                                         double *__restrict__ pp, double *__restrict__ p)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n4; t += (long)gridDim.x * blockDim.x) {
        const int z = pz4[t];
        const D4 a = ld4(pp + 4 * t);
        double o[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) o[e] = (z >> (8 * e)) & 0xff ? 0.0 : a.v[e];
        *(double2 *)(p + 4 * t) = make_double2(o[0], o[1]);
        *(double2 *)(p + 4 * t + 2) = make_double2(o[2], o[3]);
        if (z) {
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if ((z >> (8 * e)) & 0xff) pp[4 * t + e] = 0.0;
        }
    }
}

// blocks for one thread per item, rounded up to a multiple of 8 (xcd_remap
// is a bijection on such a grid)
static unsigned xcd_grid(long items)
{
    const long nb = (items + 255) / 256;
    return (unsigned)((nb + 7) / 8 * 8);
}

void correct_vel(hipStream_t st, const LevelDims &d, const ProjMet &m, const double *pp, const int8_t *iblank,
                 double dt, double *u, double *v, double *w)
{
    const long nch = (long)(d.ny - 1) * (d.nz - 1) * (((d.nx - 1) >> 2) + 1);
    hipLaunchKernelGGL(k_correct_vel, dim3(xcd_grid(nch)), dim3(256), 0, st, d, m, pp, iblank, dt, u, v, w);
}

void correct_ghost_vel(hipStream_t st, const LevelDims &d, const ProjMet &m, const GhostVel *g, int n, int mix,
                       const double *pv, double dt, double *u, double *v, double *w)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_correct_ghost_vel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d, m, g, n, mix, pv,
                       dt, u, v, w);
}

void face_vel(hipStream_t st, const LevelDims &d, const ProjMet &m, const double *u, const double *v, const double *w,
              const double *bc_u, const double *bc_v, const double *bc_w, double *fu, double *fv, double *fw)
{
    const long nch = (long)d.ny * d.nz * ((d.nx >> 2) + 1);
    hipLaunchKernelGGL(k_face_vel, dim3(xcd_grid(nch)), dim3(256), 0, st, d, m, u, v, w, bc_u, bc_v, bc_w, fu, fv,
                       fw);
}

void update_pressure(hipStream_t st, const LevelDims &d, const int8_t *pzero, double *pp, double *p)
{
    hipLaunchKernelGGL(k_update_pressure, dim3(ROW_BLOCKS), dim3(256), 0, st, d.n / 4, (const int *)pzero, pp, p);
}

// ---------------------------------------------------------------------------
// Advection-diffusion line solves (CORE_SOLVE_AD.F90), FP64.
//
// The nine coefficients of cell (i,j,k): the diffusion coefficients of
// set_solve_ad plus the Crank-Nicolson advection terms from the face
// velocities, central for 25 <= i <= 375 and first-order upwind elsewhere
// (the switch tests i on every axis).  CORE_SOLVE_AD.F90:212-297 (x-sweep;
// the y- and z-sweeps and calc_residual_ad repeat it).
// c = {amx, apx, acx, amy, apy, acy, amz, apz, acz}
__device__ __forceinline__ double sign1(double b) { return b >= 0.0 ? 1.0 : -1.0; }

__device__ __forceinline__ void ad_axis(double am0, double ap0, double lo, double hi, double f0, double f1,
                                        double dinv, double hd, bool central, double &am, double &ap, double &ac)
{
    am = am0;
    ap = ap0;
    ac = -(am0 + ap0);
    if (central) {
        am = am - hd * lo * (1.0 - f0) * dinv;
        ap = ap + hd * hi * f1 * dinv;
        ac = ac + hd * (hi * (1.0 - f1) - lo * f0) * dinv;
    } else {
        am = am - hd * (0.5 * (1.0 + sign1(lo))) * lo * dinv;
        ap = ap + hd * (0.5 * (1.0 - sign1(hi))) * hi * dinv;
        ac = ac + hd * (hi * (0.5 * (1.0 + sign1(hi))) - lo * (0.5 * (1.0 - sign1(lo)))) * dinv;
    }
}

__device__ __forceinline__ void ad_coef(const LevelDims &d, const AdMet &m, const double *__restrict__ fu,
                                        const double *__restrict__ fv, const double *__restrict__ fw, int i, int j,
                                        int k, long q, double (&c)[9])
{
    const bool central = i >= 25 && i <= 375;
    const double hd = m.half_dt;
    ad_axis(m.amx[i], m.apx[i], fu[q], fu[q + 1], m.pm.fx[i], m.pm.fx[i + 1], m.pm.dxinv[i], hd, central, c[0], c[1],
            c[2]);
    ad_axis(m.amy[j], m.apy[j], fv[q], fv[q + d.sy], m.pm.fy[j], m.pm.fy[j + 1], m.pm.dyinv[j], hd, central, c[3],
            c[4], c[5]);
    ad_axis(m.amz[k], m.apz[k], fw[q], fw[q + d.sz], m.pm.fz[k], m.pm.fz[k + 1], m.pm.dzinv[k], hd, central, c[6],
            c[7], c[8]);
}

// One half sweep of itsolv_ad: every line of direction DIR and one colour
// (x: (j+k)%2, y: (i+k)%2, z: (i+j)%2), NL lines per block, rows 0..n of a
// line with trivial rows at both ends.  Off-line neighbours go to the
// right-hand side, masked cells (igk: iblank or ghost) become identity rows,
// and after the solve var += omega (x - var) on rows 1..n-1 (x:
// CORE_SOLVE_AD.F90:206-346, y: :400-545, z: :600-772).
// Three phases:
//   build : the block's threads build the (a, b, c, d) rows of its lines
//           into LDS in an order that makes the global loads contiguous
//           (x-lines: consecutive threads take consecutive rows of one line;
//           y/z-lines: consecutive threads take the same row of the NL
//           lines, which are neighbours of one colour, two cells apart);
//   solve : one wavefront per line, E rows per lane in registers (the
//           partition solve of the pressure smoother), identity rows past n;
//   update: the relaxation in the build order.
template <int DIR, int E, int NL>
__global__ void __launch_bounds__(NL * 64) k_ad_sweep(LevelDims d, AdMet m, double *__restrict__ var,
                                                      const double *__restrict__ r, const double *__restrict__ fu,
                                                      const double *__restrict__ fv, const double *__restrict__ fw,
                                                      const int8_t *__restrict__ igk, int color)
{
    // a, b, c, d of every row, and for y/z-lines the old var (its re-read in
    // the update would be another strided pass): (DIR ? 5 : 4) * NL * RS doubles
    extern __shared__ double sm[];
    const int n = DIR == 0 ? d.nx : DIR == 1 ? d.ny : d.nz;
    const int RS = (n + 1) | 1;                      // LDS row stride of one line (odd, >= n + 1)
    double *sa_ = sm, *sb_ = sm + NL * RS, *sc_ = sm + 2 * NL * RS, *sd_ = sm + 3 * NL * RS, *sv_ = sm + 4 * NL * RS;
    const int na = DIR == 0 ? d.ny : d.nx;           // line coordinates (a, b): x (j, k), y (i, k), z (i, j)
    const int nb = DIR == 2 ? d.ny : d.nz;
    const int ha = na / 2;                           // slots per value of b
    const long st = DIR == 0 ? 1 : DIR == 1 ? d.sy : d.sz;
    const long sa = DIR == 0 ? d.sy : 1, sb = DIR == 2 ? d.sy : d.sz;     // off-line strides
    constexpr int ca = DIR == 0 ? 3 : 0, cb = DIR == 2 ? 3 : 6, cl = 3 * DIR;
    // the line of block-local index ml: its first cell offset, or -1.  x-lines:
    // consecutive slots.  y/z-lines (a = i): each value of b has its own run
    // of bpb blocks, and a block's NL lines a = 2 NL g + {0, 2, ..}
    // (+1 for the colour that starts at i = 1) lie in one aligned run of
    // 2 NL cells, so each row of the block reads whole 64-byte segments.
    // Blocks are renumbered so that each XCD runs one contiguous range (the
    // grid is a multiple of 8): the blocks one value of b apart, whose rows
    // are this block's off-line neighbours, then share its L2.
    const int bpb = (na / 2 + NL) / NL + 1;
    const int blk = xcd_remap(blockIdx.x, gridDim.x);
    auto line_q0 = [&](int ml, int &a, int &b) -> long {
        if (DIR == 0) {
            const long slot = (long)blk * NL + ml;
            b = 1 + (int)(slot / ha);
            const int a0 = (((1 + b) & 1) == color) ? 1 : 2;
            a = a0 + 2 * (int)(slot % ha);
        } else {
            b = 1 + blk / bpb;
            const int g = blk % bpb;
            const int odd = (((1 + b) & 1) == color) ? 1 : 0;      // lines at odd i
            a = 2 * (NL * g + ml) + odd;
            if (a < 1) return -1;
        }
        if (b > nb - 1 || a > na - 1) return -1;
        return DIR == 0 ? a * d.sy + b * d.sz : DIR == 1 ? a + b * d.sz : a + b * d.sy;
    };
    const int nrow = n + 1;
    // build
    for (int t = threadIdx.x; t < NL * nrow; t += NL * 64) {
        int ml, row;
        if (DIR == 0) { ml = t / nrow; row = t - ml * nrow; }
        else { ml = t % NL; row = t / NL; }
        int a, b;
        const long q0 = line_q0(ml, a, b);
        double A = 0.0, B = 1.0, Cc = 0.0, D = 0.0;
        if (q0 >= 0) {
            const long q = q0 + st * row;
            const double V = var[q];
            if (DIR != 0) sv_[ml * RS + row] = V;
            if (row == 0 || row == n) {
                D = V;
            } else {
                const int i = DIR == 0 ? row : a, j = DIR == 1 ? row : (DIR == 0 ? a : b), k = DIR == 2 ? row : b;
                double c[9];
                ad_coef(d, m, fu, fv, fw, i, j, k, q, c);
                const double rhs = r[q] - (var[q - sa] * c[ca] + var[q + sa] * c[ca + 1] + var[q - sb] * c[cb]
                                           + var[q + sb] * c[cb + 1]);
                const double rib = igk[q] ? 0.0 : 1.0;
                A = c[cl] * rib;
                Cc = c[cl + 1] * rib;
                B = 1.0 + (c[2] + c[5] + c[8]) * rib;
                D = rhs * rib + V * (1.0 - rib);
            }
        }
        const int o = ml * RS + row;
        sa_[o] = A; sb_[o] = B; sc_[o] = Cc; sd_[o] = D;
    }
    __syncthreads();
    // solve: wave w solves line w
    {
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        double oa[E], ob[E], oc[E], od[E], x[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int row = lane * E + e;
            const int o = w * RS + row;
            if (row < nrow) { oa[e] = sa_[o]; ob[e] = sb_[o]; oc[e] = sc_[o]; od[e] = sd_[o]; }
            else { oa[e] = 0.0; ob[e] = 1.0; oc[e] = 0.0; od[e] = 0.0; }
        }
        line_solve<double, E>(oa, ob, oc, od, x, lane, 64);
        __syncthreads();        // every build-phase read of sd_ is done before it is overwritten
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int row = lane * E + e;
            if (row < nrow) sd_[w * RS + row] = x[e];
        }
    }
    __syncthreads();
    // update
    for (int t = threadIdx.x; t < NL * nrow; t += NL * 64) {
        int ml, row;
        if (DIR == 0) { ml = t / nrow; row = t - ml * nrow; }
        else { ml = t % NL; row = t / NL; }
        if (row < 1 || row > n - 1) continue;
        int a, b;
        const long q0 = line_q0(ml, a, b);
        if (q0 < 0) continue;
        const long q = q0 + st * row;
        const double vq = DIR != 0 ? sv_[ml * RS + row] : var[q];
        var[q] = vq + m.omega * (sd_[ml * RS + row] - vq);
    }
}

// calc_residual_ad (CORE_SOLVE_AD.F90:780-964): per block the largest
// |res| with its smallest lexicographic cell index (i fastest, as the
// reference's loop finds the first maximum) and the sum of res^2.
__global__ void __launch_bounds__(256) k_ad_residual(LevelDims d, AdMet m, const double *__restrict__ var,
                                                     const double *__restrict__ r, const double *__restrict__ fu,
                                                     const double *__restrict__ fv, const double *__restrict__ fw,
                                                     const int8_t *__restrict__ igk, double *__restrict__ partial)
{
    __shared__ double sv[256], si[256], s2[256];
    const int ni = d.nx - 1, nj = d.ny - 1;
    const long ncell = (long)ni * nj * (d.nz - 1);
    double best = 0.0, bidx = 0.0, sum2 = 0.0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < ncell; t += (long)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(t % ni);
        const long rest = t / ni;
        const int j = 1 + (int)(rest % nj), k = 1 + (int)(rest / nj);
        const long q = idx(d, i, j, k);
        double c[9];
        ad_coef(d, m, fu, fv, fw, i, j, k, q, c);
        double res = r[q] - var[q] * (1.0 + c[2] + c[5] + c[8]) - var[q - 1] * c[0] - var[q + 1] * c[1]
                     - var[q - d.sy] * c[3] - var[q + d.sy] * c[4] - var[q - d.sz] * c[6] - var[q + d.sz] * c[7];
        res = igk[q] ? 0.0 : res;
        const double ar = fabs(res);
        if (ar > best) { best = ar; bidx = (double)t; }
        sum2 += res * res;
    }
    sv[threadIdx.x] = best;
    si[threadIdx.x] = best > 0.0 ? bidx : 1e300;
    s2[threadIdx.x] = sum2;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            const double v1 = sv[threadIdx.x], v2 = sv[threadIdx.x + w];
            const double i1 = si[threadIdx.x], i2 = si[threadIdx.x + w];
            if (v2 > v1 || (v2 == v1 && i2 < i1)) { sv[threadIdx.x] = v2; si[threadIdx.x] = i2; }
            s2[threadIdx.x] += s2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[3 * blockIdx.x] = sv[0];
        partial[3 * blockIdx.x + 1] = si[0];
        partial[3 * blockIdx.x + 2] = s2[0];
    }
}

// calc_residual / calc_residual2 with their loc output (CORE_POISSON.F90:
// 460-551, 679-724): per block the largest value and the first interior
// cell, in the reference's (k, j, i) loop order, where it occurs, and the sum
// of squares; k_ad_residual_final finishes the reduction.  mode 0: |7-point
// residual of var, r| on flag cells (mask = igk: 1 - flag); mode 1:
// |var - r| where iblank = 0 (mask = iblank; r is p_temp).
__global__ void __launch_bounds__(256) k_resloc(LevelDims d, Metrics<double> m, const double *__restrict__ var,
                                                const double *__restrict__ r, const int8_t *__restrict__ mask,
                                                int mode, double *__restrict__ partial)
{
    __shared__ double sv[256], si[256], s2[256];
    const int ni = d.nx - 1, nj = d.ny - 1;
    const long ncell = (long)ni * nj * (d.nz - 1);
    double best = 0.0, bidx = 0.0, sum2 = 0.0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < ncell; t += (long)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(t % ni);
        const long rest = t / ni;
        const int j = 1 + (int)(rest % nj), k = 1 + (int)(rest / nj);
        const long q = idx(d, i, j, k);
        double v;
        if (mode == 0) {
            const double bc = (m.acx[i] + m.acy[j]) + m.acz[k];
            v = r[q] - var[q] * bc - var[q - 1] * m.amx[i] - var[q + 1] * m.apx[i] - var[q - d.sy] * m.amy[j]
                - var[q + d.sy] * m.apy[j] - var[q - d.sz] * m.amz[k] - var[q + d.sz] * m.apz[k];
        } else {
            v = var[q] - r[q];
        }
        v = mask[q] ? 0.0 : v;
        const double av = fabs(v);
        if (av > best) { best = av; bidx = (double)t; }
        sum2 += v * v;
    }
    sv[threadIdx.x] = best;
    si[threadIdx.x] = best > 0.0 ? bidx : 1e300;
    s2[threadIdx.x] = sum2;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            const double v1 = sv[threadIdx.x], v2 = sv[threadIdx.x + w];
            const double i1 = si[threadIdx.x], i2 = si[threadIdx.x + w];
            if (v2 > v1 || (v2 == v1 && i2 < i1)) { sv[threadIdx.x] = v2; si[threadIdx.x] = i2; }
            s2[threadIdx.x] += s2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[3 * blockIdx.x] = sv[0];
        partial[3 * blockIdx.x + 1] = si[0];
        partial[3 * blockIdx.x + 2] = s2[0];
    }
}

__global__ void __launch_bounds__(256) k_ad_residual_final(const double *__restrict__ partial, int nb,
                                                           double *__restrict__ out3)
{
    __shared__ double sv[256], si[256], s2[256];
    double v = 0.0, ix = 1e300, q2 = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        const double v2 = partial[3 * b], i2 = partial[3 * b + 1];
        if (v2 > v || (v2 == v && i2 < ix)) { v = v2; ix = i2; }
        q2 += partial[3 * b + 2];
    }
    sv[threadIdx.x] = v;
    si[threadIdx.x] = ix;
    s2[threadIdx.x] = q2;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            const double v1 = sv[threadIdx.x], v2 = sv[threadIdx.x + w];
            const double i1 = si[threadIdx.x], i2 = si[threadIdx.x + w];
            if (v2 > v1 || (v2 == v1 && i2 < i1)) { sv[threadIdx.x] = v2; si[threadIdx.x] = i2; }
            s2[threadIdx.x] += s2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out3[0] = sv[0];
        out3[1] = sv[0] > 0.0 ? si[0] : -1.0;
        out3[2] = s2[0];
    }
}

template <int DIR>
static void ad_sweep_dir(hipStream_t st, const LevelDims &d, const AdMet &m, double *var, const double *r,
                         const double *fu, const double *fv, const double *fw, const int8_t *igk, int color, int n)
{
    const int na = DIR == 0 ? d.ny : d.nx, nb = DIR == 2 ? d.ny : d.nz;
    const long slots = (long)(na / 2) * (nb - 1);
    // x: consecutive slots; y/z: bpb aligned blocks per value of b (as in the kernel)
    const long b4 = DIR == 0 ? (slots + 3) / 4 : (long)((na / 2 + 4) / 4 + 1) * (nb - 1);
    const long b2 = DIR == 0 ? (slots + 1) / 2 : (long)((na / 2 + 2) / 2 + 1) * (nb - 1);
    const dim3 g4((unsigned)((b4 + 7) / 8 * 8)), g2((unsigned)((b2 + 7) / 8 * 8));   // xcd_remap needs a multiple of 8
    const size_t arrays = DIR == 0 ? 4 : 5, rs = (size_t)((n + 1) | 1);
    const size_t l4 = arrays * 4 * rs * sizeof(double), l2 = arrays * 2 * rs * sizeof(double);
    if (n + 1 <= 64 * 2)
        hipLaunchKernelGGL((k_ad_sweep<DIR, 2, 4>), g4, dim3(256), l4, st, d, m, var, r, fu, fv, fw, igk, color);
    else if (n + 1 <= 64 * 3)
        hipLaunchKernelGGL((k_ad_sweep<DIR, 3, 4>), g4, dim3(256), l4, st, d, m, var, r, fu, fv, fw, igk, color);
    else if (n + 1 <= 64 * 5)
        hipLaunchKernelGGL((k_ad_sweep<DIR, 5, 4>), g4, dim3(256), l4, st, d, m, var, r, fu, fv, fw, igk, color);
    else
        hipLaunchKernelGGL((k_ad_sweep<DIR, 9, 2>), g2, dim3(128), l2, st, d, m, var, r, fu, fv, fw, igk, color);
}

void ad_sweep(hipStream_t st, const LevelDims &d, const AdMet &m, double *var, const double *r, const double *fu,
              const double *fv, const double *fw, const int8_t *igk, int dir, int color)
{
    if (dir == 0) ad_sweep_dir<0>(st, d, m, var, r, fu, fv, fw, igk, color, d.nx);
    else if (dir == 1) ad_sweep_dir<1>(st, d, m, var, r, fu, fv, fw, igk, color, d.ny);
    else ad_sweep_dir<2>(st, d, m, var, r, fu, fv, fw, igk, color, d.nz);
}

void ad_residual(hipStream_t st, const LevelDims &d, const AdMet &m, const double *var, const double *r,
                 const double *fu, const double *fv, const double *fw, const int8_t *igk, double *partial,
                 double *out3)
{
    hipLaunchKernelGGL(k_ad_residual, dim3(ROW_BLOCKS), dim3(256), 0, st, d, m, var, r, fu, fv, fw, igk, partial);
    hipLaunchKernelGGL(k_ad_residual_final, dim3(1), dim3(256), 0, st, partial, ROW_BLOCKS, out3);
}

// ---------------------------------------------------------------------------
// rhs_advec_diff (CORE_SET_SOLVERS.F90:10-61), FP64: for each of (uOld, nlu),
// (vOld, nlv), (wOld, nlw), rhs_advec (:65-436) then rhs_diff (:440-526) on
// one interior cell, fused (each cell's result depends only on the old
// velocities, so the reference's three passes per component collapse into
// one).  Every product and sum is rounded on its own (no fused
// multiply-add), in the reference's order, so the results are the
// reference's to the bit.

// One axis of rhs_advec at a cell, from the component's values at n-2
// (clamped to 0), n-1, n, n+1 and n+2 (clamped to N) along the axis, the
// face velocities Fw (face n) and Fe (face n+1), the interpolation factors
// f(n), f(n+1), the upwind weights of index n and dinv(n):
//   central  ve = f(n+1) v(n+1) + (1 - f(n+1)) v(n),  vw = f(n) v(n) + (1 - f(n)) v(n-1)
//   upwind   second order (25 <= i <= 375, i the cell's x index on every
//            axis, as the reference tests it) or first order, the central
//            value on the inflow faces at n = 1 and n = N-1
//   term     (((1-alfa) ve + alfa ve_up) Fe - ((1-alfa) vw + alfa vw_up) Fw) dinv(n)
// SIGN(1, F) is +1 for F = -0.0 too (the Intel compilers' default).  The
// reference blends the two upwind candidates as ((1 + s) A + (1 - s) B) / 2
// with s = SIGN(1, F) = +-1; that is exactly A or B for finite operands
// below 2^1023 (the doubling, the zero product and the halving are exact),
// so the kernel selects instead; only the sign of a zero can differ.
__device__ __forceinline__ double advec_term(double vmm, double vm, double vc, double vp, double vpp, double Fw,
                                             double Fe, double f0, double f1, const UpwW &w, double dinv, bool lo,
                                             bool hi, bool central, double alfa)
{
#pragma clang fp contract(off)
    const bool pw = Fw >= 0.0, pe = Fe >= 0.0;
    const double ve = f1 * vp + (1.0 - f1) * vc;
    const double vw = f0 * vc + (1.0 - f0) * vm;
    double veu, vwu;
    if (hi && !pe) veu = ve;
    else if (central) veu = pe ? w.em1 * vc - w.em2 * vm : w.ep1 * vp - w.ep2 * vpp;
    else veu = pe ? vc : vp;
    if (lo && pw) vwu = vw;
    else if (central) vwu = pw ? w.wm1 * vm - w.wm2 * vmm : w.wp1 * vc - w.wp2 * vp;
    else vwu = pw ? vm : vc;
    return (((1.0 - alfa) * ve + alfa * veu) * Fe - ((1.0 - alfa) * vw + alfa * vwu) * Fw) * dinv;
}

// One thread per interior cell (i fastest); blocks in the XCD-contiguous
// order of the other stencil kernels (the launch has a multiple of 8
// blocks), so the rows j +- 2 and planes k +- 2 that a block reads are
// mostly in the same L2.  The components are done one after the other to
// keep the registers of one stencil live at a time.  Loads and stores go
// through raw buffer instructions with 32-bit byte offsets (a level-1 field
// is below 2 GiB), positive neighbour distances as scalar offsets.
__global__ void __launch_bounds__(256) k_rhs_advec_diff(LevelDims d, RhsAdMet m, const double *__restrict__ uo,
                                                        const double *__restrict__ vo, const double *__restrict__ wo,
                                                        const double *__restrict__ fu, const double *__restrict__ fv,
                                                        const double *__restrict__ fw, double *__restrict__ nlu,
                                                        double *__restrict__ nlv, double *__restrict__ nlw)
{
#pragma clang fp contract(off)
    using B = Buf<double>;
    const int nx = d.nx, ny = d.ny, nz = d.nz, ni = nx - 1;
    const long ncell = (long)ni * (ny - 1) * (nz - 1);
    const long t = (long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= ncell) return;
    const long r = t / ni;
    const int i = 1 + (int)(t - r * ni);
    const int k = 1 + (int)(r / (ny - 1));
    const int j = 1 + (int)(r - (long)(k - 1) * (ny - 1));
    const long fb = 8 * d.n;                              // bytes of one field
    const int sy = (int)(8 * d.sy), sz = (int)(8 * d.sz);  // neighbour distances in bytes
    const int v0 = (int)(8 * ((long)i + (long)j * d.sy + (long)k * d.sz));
    // neighbour offsets, the ones two away clamped as imm / ipp, jmm / jpp, kmm / kpp
    const int oxm = 8 * (max(i - 2, 0) - i), oxp = 8 * (min(i + 2, nx) - i);
    const int oym = (max(j - 2, 0) - j) * sy, oyp = (min(j + 2, ny) - j) * sy;
    const int ozm = (max(k - 2, 0) - k) * sz, ozp = (min(k + 2, nz) - k) * sz;
    const bool central = i >= 25 && i <= 375;
    const double alfa = m.alfa;
    // face velocities and metrics, shared by the three components
    const __amdgpu_buffer_rsrc_t rfu = buf_rsrc(fu, fb), rfv = buf_rsrc(fv, fb), rfw = buf_rsrc(fw, fb);
    const double Fuw = B::ld(rfu, v0, 0), Fue = B::ld(rfu, v0, 8), Fvs = B::ld(rfv, v0, 0), Fvn = B::ld(rfv, v0, sy);
    const double Fwb = B::ld(rfw, v0, 0), Fwf = B::ld(rfw, v0, sz);
    const double fx0 = m.pm.fx[i], fx1 = m.pm.fx[i + 1], fy0 = m.pm.fy[j], fy1 = m.pm.fy[j + 1];
    const double fz0 = m.pm.fz[k], fz1 = m.pm.fz[k + 1];
    const double dxi = m.pm.dxinv[i], dyi = m.pm.dyinv[j], dzi = m.pm.dzinv[k];
    const UpwW wx = m.wx[i], wy = m.wy[j], wz = m.wz[k];
    // rhs_diff's coefficients (CORE_SET_SOLVERS.F90:476-511)
    double nuE, nuW, nuN, nuS, nuF, nuB;
    {
        double c = m.nu, ve = c, vw = c, vn = c, vs = c, vf = c, vb = c;
        if (m.visc) {
            const __amdgpu_buffer_rsrc_t rv = buf_rsrc(m.visc, fb);
            c = B::ld(rv, v0, 0);
            ve = B::ld(rv, v0, 8); vw = B::ld(rv, v0 - 8, 0);
            vn = B::ld(rv, v0, sy); vs = B::ld(rv, v0 - sy, 0);
            vf = B::ld(rv, v0, sz); vb = B::ld(rv, v0 - sz, 0);
        }
        nuE = fx1 * ve + (1.0 - fx1) * c;
        nuW = fx0 * c + (1.0 - fx0) * vw;
        nuN = fy1 * vn + (1.0 - fy1) * c;
        nuS = fy0 * c + (1.0 - fy0) * vs;
        nuF = fz1 * vf + (1.0 - fz1) * c;
        nuB = fz0 * c + (1.0 - fz0) * vb;
    }
    double aamx = m.dxcinv[i] * dxi, aapx = m.dxcinv[i + 1] * dxi;
    aamx = aamx * nuW; aapx = aapx * nuE;
    const double aacx = -(aamx + aapx);
    double aamy = m.dycinv[j] * dyi, aapy = m.dycinv[j + 1] * dyi;
    aamy = aamy * nuS; aapy = aapy * nuN;
    const double aacy = -(aamy + aapy);
    double aamz = m.dzcinv[k] * dzi, aapz = m.dzcinv[k + 1] * dzi;
    aamz = aamz * nuB; aapz = aapz * nuF;
    const double aacz = -(aamz + aapz);
    const double hdt = 0.5 * m.dt;
    const double *vel[3] = {uo, vo, wo};
    double *nl[3] = {nlu, nlv, nlw};
#pragma unroll 1
    for (int c = 0; c < 3; ++c) {
        const __amdgpu_buffer_rsrc_t rv = buf_rsrc(vel[c], fb);
        const double vc = B::ld(rv, v0, 0);
        const double xm = B::ld(rv, v0 - 8, 0), xp = B::ld(rv, v0, 8);
        const double xmm = B::ld(rv, v0 + oxm, 0), xpp = B::ld(rv, v0 + oxp, 0);
        const double ym = B::ld(rv, v0 - sy, 0), yp = B::ld(rv, v0, sy);
        const double ymm = B::ld(rv, v0 + oym, 0), ypp = B::ld(rv, v0 + oyp, 0);
        const double zm = B::ld(rv, v0 - sz, 0), zp = B::ld(rv, v0, sz);
        const double zmm = B::ld(rv, v0 + ozm, 0), zpp = B::ld(rv, v0 + ozp, 0);
        double adv = advec_term(xmm, xm, vc, xp, xpp, Fuw, Fue, fx0, fx1, wx, dxi, i == 1, i == nx - 1, central, alfa);
        adv = adv + advec_term(ymm, ym, vc, yp, ypp, Fvs, Fvn, fy0, fy1, wy, dyi, j == 1, j == ny - 1, central, alfa);
        adv = adv + advec_term(zmm, zm, vc, zp, zpp, Fwb, Fwf, fz0, fz1, wz, dzi, k == 1, k == nz - 1, central, alfa);
        const double diffvel = aamx * xm + aacx * vc + aapx * xp + aamy * ym + aacy * vc + aapy * yp + aamz * zm +
                               aacz * vc + aapz * zp;
        B::st(buf_rsrc(nl[c], fb), vc + hdt * (-adv + diffvel), v0, 0);
    }
}

void rhs_advec_diff(hipStream_t st, const LevelDims &d, const RhsAdMet &m, const double *uo, const double *vo,
                    const double *wo, const double *fu, const double *fv, const double *fw, double *nlu, double *nlv,
                    double *nlw)
{
    const long ncell = (long)(d.nx - 1) * (d.ny - 1) * (d.nz - 1);
    if (ncell <= 0) return;
    const long nb = (ncell + 255) / 256;
    hipLaunchKernelGGL(k_rhs_advec_diff, dim3((unsigned)((nb + 7) / 8 * 8)), dim3(256), 0, st, d, m, uo, vo, wo, fu,
                       fv, fw, nlu, nlv, nlw);
}

void resloc(hipStream_t st, const LevelDims &d, const Metrics<double> &m, const double *var, const double *r,
            const int8_t *mask, int mode, double *partial, double *out3)
{
    hipLaunchKernelGGL(k_resloc, dim3(ROW_BLOCKS), dim3(256), 0, st, d, m, var, r, mask, mode, partial);
    hipLaunchKernelGGL(k_ad_residual_final, dim3(1), dim3(256), 0, st, partial, ROW_BLOCKS, out3);
}

void resid64_to32(hipStream_t st, const LevelDims &d, const Metrics<double> &m, const double *p, const double *div,
                  const int8_t *igk, float *r32, float *e32, double *partial, double *out2)
{
    hipLaunchKernelGGL(k_resid64_to32<float>, dim3(ROW_BLOCKS), dim3(256), xmet_bytes(d), st, d, m, p, div, igk, r32,
                       e32, out2 ? partial : nullptr);
    if (out2) hipLaunchKernelGGL(k_reduce_cols, dim3(2), dim3(256), 0, st, partial, ROW_BLOCKS, 0u, out2);
}
template <typename TE>
void resid64_to32_sums(hipStream_t st, const LevelDims &d, const Metrics<double> &m, const double *p,
                       const double *div, const int8_t *igk, TE *r32, TE *e32, double *partial)
{
    hipLaunchKernelGGL(k_resid64_to32<TE>, dim3(ROW_BLOCKS), dim3(256), xmet_bytes(d), st, d, m, p, div, igk, r32, e32,
                       partial);
}
void add_correction(hipStream_t st, long n, double *p, const float *e)
{
    hipLaunchKernelGGL(k_add_correction, dim3(2048), dim3(256), 0, st, n, p, e);
}


void add_into(hipStream_t st, const LevelDims &d, const double *p, const float *e, double *G)
{
    hipLaunchKernelGGL(k_add_into, dim3(ROW_BLOCKS), dim3(256), 0, st, d, p, e, G);
}

void aa_resid(hipStream_t st, const LevelDims &d, const Metrics<double> &m, const double *G, const double *div,
              const int8_t *igk, const double *x, float *fprev, const float *cprev, const AaCols &cols,
              const int *ring, double *partial, double *scal)
{
    hipLaunchKernelGGL(k_aa_resid, dim3(ROW_BLOCKS), dim3(256), xmet_bytes(d), st, d, m, G, div, igk, x, fprev, cprev, cols, ring,
                       partial);
    hipLaunchKernelGGL(k_reduce_cols, dim3(ROW_COMPONENTS), dim3(256), 0, st, partial, ROW_BLOCKS, 0xCu, scal);
}

template <typename TE>
void aa_fresid(hipStream_t st, const LevelDims &d, const TE *e, float *fprev, const float *cprev,
               const AaCols &cols, const int *ring, double *partial, double *scal)
{
    hipLaunchKernelGGL(k_aa_fresid<TE>, dim3(ROW_BLOCKS), dim3(256), 0, st, d, e, (const double *)scal, fprev, cprev,
                       cols, ring, partial);
    hipLaunchKernelGGL(k_reduce_cols, dim3(ROW_COMPONENTS), dim3(256), 0, st, partial, ROW_BLOCKS, 0xCu, scal);
}

template <typename TE>
void aa_lsq(hipStream_t st, int *ring, double *scal, const AaCols &cols, const double *G, const TE *E, long q111,
            int reset)
{
    hipLaunchKernelGGL(k_aa_lsq<TE>, dim3(1), dim3(64), 0, st, ring, scal, cols, G, E, q111, reset);
}

template <typename TE>
void aa_next(hipStream_t st, const LevelDims &d, const double *G, const TE *E, const AaCols &cols,
             const double *scal, double *x, float *c, double *p)
{
    hipLaunchKernelGGL(k_aa_next<TE>, dim3(ROW_BLOCKS), dim3(256), 0, st, d, G, E, cols, scal, x, c, p);
}

template <typename T>
void fill(hipStream_t st, long n, T *p, T v)
{
    hipLaunchKernelGGL(k_fill<T>, dim3(2048), dim3(256), 0, st, n, p, v);
}

// batched launches: E holds ne entries (ascending b0) covering nb blocks
void b_bc_fill(hipStream_t st, const BEnt *E, int ne, int nb, const ReqDesc *D, int prec, int lev, int fld, int bcsel)
{
    if (prec) hipLaunchKernelGGL(kb_bc_fill<float>, dim3(nb), dim3(256), 0, st, E, ne, D, lev, fld, bcsel);
    else hipLaunchKernelGGL(kb_bc_fill<double>, dim3(nb), dim3(256), 0, st, E, ne, D, lev, fld, bcsel);
}
void b_gcm_level(hipStream_t st, const BEnt *E, int ne, int nb, const ReqDesc *D, int l)
{
    hipLaunchKernelGGL(kb_gcm_level, dim3(nb), dim3(256), 0, st, E, ne, D, l);
}
void b_resid64_to32(hipStream_t st, const BEnt *E, int ne, int nb, const ReqDesc *D)
{
    hipLaunchKernelGGL(kb_resid64_to32, dim3(nb), dim3(256), 0, st, E, ne, D);
}
void b_sweep(hipStream_t st, const BEnt *E, int ne, int nb, size_t lds, const ReqDesc *D, int prec, int lev, int dir,
             int color, int bcsel)
{
#define BSW(T, DD) hipLaunchKernelGGL((kb_sweep<T, DD>), dim3(nb), dim3(256), lds, st, E, ne, D, lev, color, bcsel)
    if (prec) {
        if (dir == 0) BSW(float, 0); else if (dir == 1) BSW(float, 1); else BSW(float, 2);
    } else {
        if (dir == 0) BSW(double, 0); else if (dir == 1) BSW(double, 1); else BSW(double, 2);
    }
#undef BSW
}
void b_resrestrict(hipStream_t st, const BEnt *E, int ne, int nb, const ReqDesc *D, int prec, int lev)
{
    if (prec) hipLaunchKernelGGL(kb_resrestrict<float>, dim3(nb), dim3(64, 4), 0, st, E, ne, D, lev);
    else hipLaunchKernelGGL(kb_resrestrict<double>, dim3(nb), dim3(64, 4), 0, st, E, ne, D, lev);
}
void b_prolong(hipStream_t st, const BEnt *E, int ne, int nb, const ReqDesc *D, int prec, int lev)
{
    if (prec) hipLaunchKernelGGL(kb_prolong<float>, dim3(nb), dim3(64, 4), 0, st, E, ne, D, lev);
    else hipLaunchKernelGGL(kb_prolong<double>, dim3(nb), dim3(64, 4), 0, st, E, ne, D, lev);
}
void b_gcm_fold(hipStream_t st, const BEnt *E, int ne, int nb, const ReqDesc *D, int back)
{
    hipLaunchKernelGGL(kb_gcm_fold, dim3(nb), dim3(256), 0, st, E, ne, D, back);
}
void b_aa_resid(hipStream_t st, const BEnt *E, int ne, int nb, const ReqDesc *D)
{
    hipLaunchKernelGGL(kb_aa_resid, dim3(nb), dim3(256), 0, st, E, ne, D);
}
void b_reduce_cols(hipStream_t st, const BEnt *E, int ne, int nb, const ReqDesc *D)
{
    hipLaunchKernelGGL(kb_reduce_cols, dim3(nb), dim3(256), 0, st, E, ne, D);
}
void b_aa_lsq(hipStream_t st, const BEnt *E, int ne, int nb, const ReqDesc *D, int reset)
{
    hipLaunchKernelGGL(kb_aa_lsq, dim3(nb), dim3(64), 0, st, E, ne, D, reset);
}
void b_aa_next(hipStream_t st, const BEnt *E, int ne, int nb, const ReqDesc *D, int start)
{
    hipLaunchKernelGGL(kb_aa_next, dim3(nb), dim3(256), 0, st, E, ne, D, start);
}
void b_gather(hipStream_t st, const BEnt *E, int ne, int nb, const ReqDesc *D, double *out)
{
    hipLaunchKernelGGL(kb_gather, dim3(nb), dim3(64), 0, st, E, ne, D, out);
}

// ---------------------------------------------------------------------------
// solve_ad's outer velocity conditions and ghost-cell velocities, FP64 on
// level-1 fields (u, v, w padded; the nine face arrays bcxu, bcxv, bcxw,
// bcyu, ..., bczw in the host layout, VelBc).
//
// set_outer_velocity_bc (CORE_SET_OUTER_BC.F90:18-509): one thread per face
// entry; x faces over (j, k) in 0..ny+1 x 0..nz+1, y faces over i in 0..nx
// and k in 0..nz+1, z faces over i in 0..nx and j in 0..ny+1.  The value of
// each type, from the interior value vi next to the face and vp at the other
// end: Dirichlet val; zero gradient vi; pulsatile val sin(2 pi f t) (u of
// the x1 face: val t / 30); symmetry 0 for the normal component, vi for the
// others; periodic (vi + vp) / 2; user specified (x faces) the channel
// profile val 6 (y - y0)(y1 - y)/(y1 - y0)^2 for u, val for v and w.
__device__ __forceinline__ double vbc_value(const VelBc &b, int f, int c, double vi, double vp, double yj, double t)
{
    const int a = f >> 1, side = f & 1;
    const double val = b.val[6 * a + 2 * c + side], fr = b.freq[6 * a + 2 * c + side];
    switch (b.type[f]) {
    case 1: return val;
    case 2: return vi;
    case 3:
        if (f == 0 && c == 0) return val * t / 30.0;
        return val * sin(2.0 * b.pi * fr * t);
    case 4: return (c == a) ? 0.0 : vi;
    case 5: return 0.5 * (vi + vp);
    case 6:
        if (c == 0) return val * 6.0 * (yj - b.y0) * (b.y1 - yj) / ((b.y1 - b.y0) * (b.y1 - b.y0));
        return val;
    }
    return 0.0;
}

__global__ void __launch_bounds__(256) k_outer_velocity_bc(LevelDims d, VelBc b, const double *__restrict__ u,
                                                           const double *__restrict__ v,
                                                           const double *__restrict__ w,
                                                           const double *__restrict__ ynode, double t)
{
    long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const long hx = nx + 2, hy = ny + 2;
    const double *U[3] = {u, v, w};
    const long nX = hy * (nz + 2), nY = (long)(nx + 1) * (nz + 2), nZ = (long)(nx + 1) * hy;
    for (int side = 0; side < 2; ++side) {
        if (e < nX) {
            const int j = (int)(e % hy), k = (int)(e / hy);
            const int i = side ? nx - 1 : 1, ip = side ? 1 : nx - 1;
            for (int c = 0; c < 3; ++c)
                b.arr[c][side + 2 * (j + hy * k)] =
                    vbc_value(b, side, c, U[c][idx(d, i, j, k)], U[c][idx(d, ip, j, k)], ynode[j], t);
            return;
        }
        e -= nX;
    }
    for (int side = 0; side < 2; ++side) {
        if (e < nY) {
            const int i = (int)(e % (nx + 1)), k = (int)(e / (nx + 1));
            const int j = side ? ny - 1 : 1, jp = side ? 1 : ny - 1;
            for (int c = 0; c < 3; ++c)
                b.arr[3 + c][i + hx * (side + 2 * (long)k)] =
                    vbc_value(b, 2 + side, c, U[c][idx(d, i, j, k)], U[c][idx(d, i, jp, k)], 0.0, t);
            return;
        }
        e -= nY;
    }
    for (int side = 0; side < 2; ++side) {
        if (e < nZ) {
            const int i = (int)(e % (nx + 1)), j = (int)(e / (nx + 1));
            const int k = side ? nz - 1 : 1, kp = side ? 1 : nz - 1;
            for (int c = 0; c < 3; ++c)
                b.arr[6 + c][i + hx * (j + hy * side)] =
                    vbc_value(b, 4 + side, c, U[c][idx(d, i, j, k)], U[c][idx(d, i, j, kp)], 0.0, t);
            return;
        }
        e -= nZ;
    }
}

// set_outer_ghost_velocity (CORE_SET_OUTER_BC.F90:739-823): ghost = 2 bc -
// interior; axis 0 covers the whole x planes, axis 1 the y planes for i in
// 0..nx (after the x planes: the y ghosts at i = 0 and nx read x ghosts),
// axis 2 the z planes for i in 0..nx, j in 0..ny+1 (after both).
__global__ void __launch_bounds__(256) k_outer_ghost_velocity(LevelDims d, VelBc b, int axis, double *__restrict__ u,
                                                              double *__restrict__ v, double *__restrict__ w)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const long hx = nx + 2, hy = ny + 2;
    double *U[3] = {u, v, w};
    if (axis == 0) {
        if (e >= hy * (nz + 2)) return;
        const int j = (int)(e % hy), k = (int)(e / hy);
        for (int c = 0; c < 3; ++c) {
            U[c][idx(d, 0, j, k)] = 2.0 * b.arr[c][0 + 2 * (j + hy * k)] - U[c][idx(d, 1, j, k)];
            U[c][idx(d, nx, j, k)] = 2.0 * b.arr[c][1 + 2 * (j + hy * k)] - U[c][idx(d, nx - 1, j, k)];
        }
    } else if (axis == 1) {
        if (e >= (long)(nx + 1) * (nz + 2)) return;
        const int i = (int)(e % (nx + 1)), k = (int)(e / (nx + 1));
        for (int c = 0; c < 3; ++c) {
            U[c][idx(d, i, 0, k)] = 2.0 * b.arr[3 + c][i + hx * (0 + 2 * (long)k)] - U[c][idx(d, i, 1, k)];
            U[c][idx(d, i, ny, k)] = 2.0 * b.arr[3 + c][i + hx * (1 + 2 * (long)k)] - U[c][idx(d, i, ny - 1, k)];
        }
    } else {
        if (e >= (long)(nx + 1) * hy) return;
        const int i = (int)(e % (nx + 1)), j = (int)(e / (nx + 1));
        for (int c = 0; c < 3; ++c) {
            U[c][idx(d, i, j, 0)] = 2.0 * b.arr[6 + c][i + hx * (j + hy * 0)] - U[c][idx(d, i, j, 1)];
            U[c][idx(d, i, j, nz)] = 2.0 * b.arr[6 + c][i + hx * (j + hy * 1)] - U[c][idx(d, i, j, nz - 1)];
        }
    }
}

void outer_velocity_bc(hipStream_t st, const LevelDims &d, const VelBc &b, const double *u, const double *v,
                       const double *w, const double *ynode, double t)
{
    const long n = 2 * ((long)(d.ny + 2) * (d.nz + 2) + (long)(d.nx + 1) * (d.nz + 2) + (long)(d.nx + 1) * (d.ny + 2));
    hipLaunchKernelGGL(k_outer_velocity_bc, grid1(n, 256), dim3(256), 0, st, d, b, u, v, w, ynode, t);
}

void outer_ghost_velocity(hipStream_t st, const LevelDims &d, const VelBc &b, double *u, double *v, double *w)
{
    hipLaunchKernelGGL(k_outer_ghost_velocity, grid1((long)(d.ny + 2) * (d.nz + 2), 256), dim3(256), 0, st, d, b, 0,
                       u, v, w);
    hipLaunchKernelGGL(k_outer_ghost_velocity, grid1((long)(d.nx + 1) * (d.nz + 2), 256), dim3(256), 0, st, d, b, 1,
                       u, v, w);
    hipLaunchKernelGGL(k_outer_ghost_velocity, grid1((long)(d.nx + 1) * (d.ny + 2), 256), dim3(256), 0, st, d, b, 2,
                       u, v, w);
}

// GCM_ghostcell_velocity (GCM_GHOSTCELL_UPDATE.F90:1-359) as Jacobi steps:
// each step computes every active dead cell's velocities from the current
// field, writes them, then every ghost cell's from the field with the new
// dead values, and records max |du|+|dv|+|dw| of each part.  The reference
// sweeps the lists in order (Gauss-Seidel); both converge to the fixed point
// of the relations.
//   dead cell: IP = the inverse-distance average of the current velocities
//     (when the pressure relation takes that branch: same stencil, masks and
//     image point) or sum_r c_r u_r with the body-intercept velocity on the
//     self row; u = u_BI pl/(pl-1) - IP/(pl-1);
//   ghost cell: g = sum_r c_r u_r (self row: u_BI), u = (1-a) g + a t with
//     t = g, or with mix_GC_form = 1 the cell's advection-diffusion row
//     solved for it: t = (nl - sum_nb c_nb u_nb) / (1 + acx + acy + acz)
//     (ghostNScoeff, CORE_SOLVE_AD.F90:715-724, CORE_SET_SOLVERS.F90:46-58).
__global__ void __launch_bounds__(256) k_gcmv_dead(GcmDev g, GcmVel gv, const double *__restrict__ u,
                                                   const double *__restrict__ v, const double *__restrict__ w,
                                                   const long *__restrict__ off8, double *__restrict__ outv)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nD || !g.dActive[n]) return;
    const long base = g.dBase[n];
    const double *U[3] = {u, v, w};
    double ip[3] = {0.0, 0.0, 0.0};
    if (g.dIdw[n]) {
        double inv[3] = {0.0, 0.0, 0.0}, dist_int = 0.0;
        for (int r = 0; r < 8; ++r) {
            const double wt = g.dW[8 * (long)n + r];
            if (wt != 0.0) {
                dist_int = dist_int + wt;
                for (int c = 0; c < 3; ++c) inv[c] = inv[c] + wt * U[c][base + off8[r]];
            }
        }
        for (int c = 0; c < 3; ++c) ip[c] = inv[c] / dist_int;
    } else {
        const int self = g.dSelf[n];
        for (int r = 0; r < 8; ++r) {
            const double cf = gv.dCoef[8 * (long)n + r];
            for (int c = 0; c < 3; ++c)
                ip[c] = ip[c] + (r == self ? cf * gv.dBi[3 * (long)n + c] : cf * U[c][base + off8[r]]);
        }
    }
    for (int c = 0; c < 3; ++c) outv[3 * (long)n + c] = gv.dBi[3 * (long)n + c] * gv.biw - ip[c] * gv.ipw;
}

__global__ void __launch_bounds__(256) k_gcmv_ghost(GcmDev g, GcmVel gv, LevelDims d, AdMet am, int mix,
                                                    const double *__restrict__ u, const double *__restrict__ v,
                                                    const double *__restrict__ w, const double *__restrict__ nlu,
                                                    const double *__restrict__ nlv, const double *__restrict__ nlw,
                                                    const double *__restrict__ fu, const double *__restrict__ fv,
                                                    const double *__restrict__ fw, const int *__restrict__ gijk,
                                                    const long *__restrict__ off8, double *__restrict__ outv)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nG) return;
    const long base = g.gBase[n];
    const int self = g.gSelf[n];
    const double *U[3] = {u, v, w};
    double gg[3] = {0.0, 0.0, 0.0};
    for (int r = 0; r < 8; ++r) {
        const double cf = gv.gCoef[8 * (long)n + r];
        for (int c = 0; c < 3; ++c)
            gg[c] = gg[c] + (r == self ? cf * gv.gBi[3 * (long)n + c] : cf * U[c][base + off8[r]]);
    }
    const double fl = g.gAlpha[n];
    double t[3] = {gg[0], gg[1], gg[2]};
    if (mix) {
        const int i = gijk[3 * n], j = gijk[3 * n + 1], k = gijk[3 * n + 2];
        const long q = g.gTarget[n];
        double cf[9];
        ad_coef(d, am, fu, fv, fw, i, j, k, q, cf);
        const double bcx = 1.0 + (cf[2] + cf[5] + cf[8]);
        const double *NL[3] = {nlu, nlv, nlw};
        for (int c = 0; c < 3; ++c) {
            const double *X = U[c];
            t[c] = (NL[c][q] - cf[0] * X[q - 1] - cf[1] * X[q + 1] - cf[3] * X[q - d.sy] - cf[4] * X[q + d.sy] -
                    cf[6] * X[q - d.sz] - cf[7] * X[q + d.sz]) /
                   bcx;
        }
    }
    for (int c = 0; c < 3; ++c) outv[3 * (long)n + c] = (1.0 - fl) * gg[c] + fl * t[c];
}

// write the computed values at their targets; per-block max of |du|+|dv|+|dw|
__global__ void __launch_bounds__(256) k_gcmv_scatter(int n, const long *__restrict__ target,
                                                      const int *__restrict__ active, const double *__restrict__ val,
                                                      double *__restrict__ u, double *__restrict__ v,
                                                      double *__restrict__ w, double *__restrict__ partial)
{
    __shared__ double sm[256];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.0;
    if (t < n && (!active || active[t])) {
        const long q = target[t];
        double *U[3] = {u, v, w};
        for (int c = 0; c < 3; ++c) {
            const double nv = val[3 * (long)t + c];
            a += fabs(nv - U[c][q]);
            U[c][q] = nv;
        }
    }
    sm[threadIdx.x] = a;
    __syncthreads();
    for (int s2 = blockDim.x / 2; s2 > 0; s2 >>= 1) {
        if (threadIdx.x < s2) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s2]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sm[0];
}

void gcm_velocity_step(hipStream_t st, const GcmDev &g, const GcmVel &gv, const LevelDims &d, const AdMet &am,
                       int mix, int iopt, double *u, double *v, double *w, const double *nlu, const double *nlv,
                       const double *nlw, const double *fu, const double *fv, const double *fw, const int *gijk,
                       const long *off8, double *vals, double *partial, double *out2)
{
    const int nbD = (g.nD + 255) / 256, nbG = (g.nG + 255) / 256;
    if (g.nD > 0) {
        hipLaunchKernelGGL(k_gcmv_dead, dim3(nbD), dim3(256), 0, st, g, gv, u, v, w, off8, vals);
        hipLaunchKernelGGL(k_gcmv_scatter, dim3(nbD), dim3(256), 0, st, g.nD, g.dTarget, g.dActive, vals, u, v, w,
                           partial);
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, st, partial, (long)nbD, 1, 0, 1, out2);
    } else {
        hipLaunchKernelGGL(k_fill<double>, dim3(1), dim3(1), 0, st, 1L, out2, 0.0);
    }
    if (iopt != 1 && g.nG > 0) {
        hipLaunchKernelGGL(k_gcmv_ghost, dim3(nbG), dim3(256), 0, st, g, gv, d, am, mix, u, v, w, nlu, nlv, nlw, fu,
                           fv, fw, gijk, off8, vals);
        hipLaunchKernelGGL(k_gcmv_scatter, dim3(nbG), dim3(256), 0, st, g.nG, g.gTarget, nullptr, vals, u, v, w,
                           partial);
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, st, partial, (long)nbG, 1, 0, 1, out2 + 1);
    } else {
        hipLaunchKernelGGL(k_fill<double>, dim3(1), dim3(1), 0, st, 1L, out2 + 1, 0.0);
    }
}

// explicit instantiations
template void bc_fill<double>(hipStream_t, const LevelDims &, double *, const BcInfo &);
template void bc_fill<float>(hipStream_t, const LevelDims &, float *, const BcInfo &);
template void residual<double>(hipStream_t, const LevelDims &, const Metrics<double> &, const double *, const double *,
                               double *, const int8_t *);
template void residual<float>(hipStream_t, const LevelDims &, const Metrics<float> &, const float *, const float *,
                              float *, const int8_t *);
template void resnorm<double>(hipStream_t, const LevelDims &, const Metrics<double> &, const double *, const double *,
                              const int8_t *, double *, double *);
template void restrict_<double>(hipStream_t, const LevelDims &, const LevelDims &, const Metrics<double> &,
                                const Metrics<double> &, const double *, double *, const int8_t *, const int8_t *);
template void restrict_<float>(hipStream_t, const LevelDims &, const LevelDims &, const Metrics<float> &,
                               const Metrics<float> &, const float *, float *, const int8_t *, const int8_t *);
template void resrestrict<double>(hipStream_t, const LevelDims &, const LevelDims &, const Metrics<double> &,
                                  const Metrics<double> &, const double *, const double *, const int8_t *,
                                  const int8_t *, const int8_t *, double *, double *);
template void resrestrict<float>(hipStream_t, const LevelDims &, const LevelDims &, const Metrics<float> &,
                                 const Metrics<float> &, const float *, const float *, const int8_t *, const int8_t *,
                                 const int8_t *, float *, float *);
template void prolong<double>(hipStream_t, const LevelDims &, const LevelDims &, const Metrics<double> &,
                              const Metrics<double> &, const double *, double *, const int8_t *, const int8_t *,
                              const int8_t *, int);
template void prolong<float>(hipStream_t, const LevelDims &, const LevelDims &, const Metrics<float> &,
                             const Metrics<float> &, const float *, float *, const int8_t *, const int8_t *,
                             const int8_t *, int);
template void sweep<double>(hipStream_t, const LevelDims &, const Metrics<double> &, double *, const double *,
                            const int8_t *, int, int, double, const BcInfo &);
template void sweep<float>(hipStream_t, const LevelDims &, const Metrics<float> &, float *, const float *,
                           const int8_t *, int, int, double, const BcInfo &);
template void fill<double>(hipStream_t, long, double *, double);
template void resid64_to32_sums<float>(hipStream_t, const LevelDims &, const Metrics<double> &, const double *,
                                       const double *, const int8_t *, float *, float *, double *);
template void resid64_to32_sums<double>(hipStream_t, const LevelDims &, const Metrics<double> &, const double *,
                                        const double *, const int8_t *, double *, double *, double *);
template void gcm_fold<float>(hipStream_t, const GcmNode *, int, const double *, float *, double *, int);
template void gcm_fold<double>(hipStream_t, const GcmNode *, int, const double *, double *, double *, int);
template void aa_fresid<float>(hipStream_t, const LevelDims &, const float *, float *, const float *, const AaCols &,
                               const int *, double *, double *);
template void aa_fresid<double>(hipStream_t, const LevelDims &, const double *, float *, const float *, const AaCols &,
                                const int *, double *, double *);
template void aa_lsq<float>(hipStream_t, int *, double *, const AaCols &, const double *, const float *, long, int);
template void aa_lsq<double>(hipStream_t, int *, double *, const AaCols &, const double *, const double *, long, int);
template void aa_next<float>(hipStream_t, const LevelDims &, const double *, const float *, const AaCols &,
                             const double *, double *, float *, double *);
template void aa_next<double>(hipStream_t, const LevelDims &, const double *, const double *, const AaCols &,
                              const double *, double *, float *, double *);
template void fill<float>(hipStream_t, long, float *, float);

}  // namespace fc
