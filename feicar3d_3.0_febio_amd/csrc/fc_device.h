// fc_device.h -- device-side data structures of the gfx950 pressure-Poisson
// path, shared by the kernels (fc_kernels.hip) and the host orchestration
// (fc_poisson.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fc {

// One multigrid level as laid out in HBM: extents (nx+2, ny+2, nz+2) like
// the reference's mg_array(n) (CORE_MG.F90:584-597), i fastest, with the
// i-row padded to a multiple of 16 elements (pitch px) for aligned rows.
struct LevelDims {
    int nx, ny, nz;          // node counts (mgrid_I/J/K)
    int px;                  // padded i-extent (>= nx+2)
    long sy, sz;             // strides of j and k
    long n;                  // elements per field
    int ic, jc, kc;          // 1 if coarsened from the finer level (per axis)
};

// 1-D metrics of one level (double and float copies), index 0..n+1:
// node coordinates, centres, inverse spacings (CORE_MG.F90:180-330) and the
// line coefficients am(i)=dcinv(i)*dinv(i), ap(i)=dcinv(i+1)*dinv(i),
// ac(i)=-(am+ap) of itsolv_mg / residual_mg.
template <typename T>
struct Metrics {
    const T *x, *y, *z;
    const T *xc, *yc, *zc;
    const T *dxinv, *dyinv, *dzinv;
    const T *dxcinv, *dycinv, *dzcinv;
    const T *amx, *apx, *acx, *amy, *apy, *acy, *amz, *apz, *acz;
};

template <typename T>
struct LevelPtrs {
    T *phi, *rhs, *res;
};

// Outer pressure boundary data (CORE_SET_OUTER_BC.F90:831-1186).
struct BcInfo {
    int pbc[6];
    double ppp[6];
    const double *pgx1, *pgx2, *pgy1, *pgy2, *pgz1, *pgz2;  // level-1 sized
    double dxc1, dxcn, dyc1, dycn, dzc1, dzcn;            // dxc(1), dxc(nx) ...
    int homo;                                             // 1 on coarse levels
    int fnx, fny;                                          // level-1 node counts
    int per;        // bit 0: both x faces periodic, bit 1: both y faces periodic
};

// GCM lists, converted to linear offsets on the level-1 padded layout.
struct GcmDev {
    int nG, nD;
    const long *gTarget;      // ghost cell offset
    const long *gBase;        // host-cell (iCellIndex...) offset
    const int *gSelf;         // stencil row equal to the ghost cell, or -1
    const double *gCoef;      // 8 per ghost
    const double *gAlpha, *gDpdn;
    const long *dTarget, *dBase;
    const int *dSelf;
    const int *dActive;       // 0 for dead cells in the buffer rows (skipped)
    const int *dIdw;          // 1 if the inverse-distance branch applies
    const double *dCoef;      // 8 per dead cell
    const double *dW;         // 8 inverse-distance weights (0 = not valid)
    const double *dPl, *dDpdn;
};

// One relation of the exact GCM pass, stored at its position in the
// topological schedule so that a level's relations are contiguous records
// (GCM_GHOSTCELL_UPDATE.F90:361-739 per relation):
//   kind 0: ghost cell  v = sum_r (r == self ? c_r dpdn : c_r p[base+off8_r])
//   kind 1: dead cell   v = (the same sum with the dead-cell coefficients) + dpdn*pl
//   kind 2: dead cell, inverse distance  v = sum_{w_r != 0} w_r pprev[..] / sum w_r + dpdn*pl
struct GcmNode {
    int target;              // offset of the cell the relation sets
    int base;                // offset of the host cell (stencil corner)
    int self;                // stencil row equal to the target, or -1
    int kind;
    double coef[8];          // coefficients (kinds 0, 1) or weights (kind 2, 0 = not valid)
    double dpdn, pl;         // dp/dn and the probe length (dead cells)
};

// Level-1 metrics of the velocity projection: the face interpolation weights
// fx(i) = (x(i) - xc(i-1)) / (xc(i) - xc(i-1)) of metrics (CORE_GRID.F90:
// 184-205; 0 at i = 0 and n+1) and the inverse cell widths dxinv(i) =
// 1/(x(i+1) - x(i)) of make_grid (CORE_GRID.F90:120-139).  Each array is
// zero-padded past n+1 so that a 4-cell chunk may read beyond the end.
struct ProjMet {
    const double *fx, *fy, *fz;
    const double *dxinv, *dyinv, *dzinv;
};

// Level-1 data of the advection-diffusion line solves: the projection
// metrics, the diffusion coefficients of set_solve_ad (CORE_SET_SOLVERS.F90:
// 591-676; zero-padded like ProjMet), half_dt (dt/2 under Crank-Nicolson,
// else 0) and omega_ad.
struct AdMet {
    ProjMet pm;
    const double *amx, *apx, *amy, *apy, *amz, *apz;
    double half_dt, omega;
};

// Weights of the second-order upwind face values of rhs_advec
// (CORE_SET_SOLVERS.F90:105-121, 264-276, 355-369) at index n of one axis
// (dinv the inverse spacings, nm = max(n-2, 0), np = min(n+2, N)):
//   em1 = (dinv(n-1) + 2 dinv(n)) / (dinv(n-1) + dinv(n)),  em2 = dinv(n) / (dinv(n-1) + dinv(n))
//   ep1 = (dinv(np) + 2 dinv(n+1)) / (dinv(np) + dinv(n+1)), ep2 = dinv(n+1) / (dinv(np) + dinv(n+1))
//   wm1 = (dinv(nm) + 2 dinv(n-1)) / (dinv(nm) + dinv(n-1)), wm2 = dinv(n-1) / (dinv(nm) + dinv(n-1))
//   wp1 = (dinv(n+1) + 2 dinv(n)) / (dinv(n+1) + dinv(n)),  wp2 = dinv(n) / (dinv(n+1) + dinv(n))
struct UpwW {
    double em1, em2, ep1, ep2, wm1, wm2, wp1, wp2;
};

// Level-1 data of rhs_advec_diff: the projection metrics, the centre
// spacings' inverses dxcinv, dycinv, dzcinv (index 0..n), the upwind weights
// per axis, viscTot (null: the constant nu on every cell), alfa_upwind, dt.
struct RhsAdMet {
    ProjMet pm;
    const double *dxcinv, *dycinv, *dzcinv;
    const UpwW *wx, *wy, *wz;
    const double *visc;
    double nu, alfa, dt;
};

// Outer velocity conditions of solve_ad (set_outer_velocity_bc): per face
// its bc_vel type, the face values and frequencies (ux1, ux2, vx1, vx2, wx1,
// wx2, uy1, ..., wz2), the channel-profile bounds yOrigin and yout, pi, and
// the nine face arrays bcxu, bcxv, bcxw, bcyu, ..., bczw (host layout).
struct VelBc {
    int type[6];
    double val[18], freq[18];
    double y0, y1, pi;
    double *arr[9];
};

// The velocity relations of GCM_ghostcell_velocity, in the order of the GCM
// lists: coeffGCMD and coeffGCMDeadD (8 per relation), the body-intercept
// velocities (u, v, w per relation), and the weights pl/(pl-1), 1/(pl-1) of
// probeLengthNormalizedD.
struct GcmVel {
    const double *gCoef, *dCoef;
    const double *gBi, *dBi;
    double biw, ipw;
};

// One ghost cell of correct_ghostcell_velocity: its indices and alphaGhost.
struct GhostVel {
    int i, j, k, pad;
    double alpha;
};

// Anderson history columns (FP32), indexed by ring slot.
struct AaCols {
    float *dF[8];
    float *dG[8];
};

// Launch shape of one line-SOR half sweep (one colour of one direction).
struct LineSweep {
    int color;
    int R;          // interior rows per line
    int LPL;        // lanes per line (power of two, LPL*E >= R)
    int lpl_shift;  // log2(LPL)
    int Rp;         // LDS stride per line (odd)
    int NL;         // lines per block
    int nblocks;    // blocks launched (multiple of 8)
    int chunks;     // y/z: i-chunks of 2*NL per row of lines
    double omega;
    int post;       // 1: rewrite the outer ghost cells of updated cells (the
                    // multigrid sweeps); 0: leave them (itsolv)
    int relax_delta;  // relaxation: 0 = (1-omega) old + omega x (itsolv_mg),
                      // 1 = old + omega (x - old) (itsolv)
};

// ---------------------------------------------------------------------------
// Serving: the device view of one admitted request, read by the batched
// kernels, which run one phase of the outer iteration for many requests in
// one launch.  Everything a request's kernels need is reached from here.
constexpr int MAX_LEVELS = 10;     // max_grid_level (AMODULES.F90:550)

struct ReqLevel {
    LevelDims d;
    Metrics<float> m32;
    Metrics<double> m64;
    float *phi32, *rhs32;          // FP32 correction hierarchy (level 1: e, r)
    double *phi64, *rhs64;         // FP64 hierarchy (level 1: p, div)
    const int8_t *igk, *ibk, *rflag, *pflag;
    // the masks its V-cycles run with: igk, ibk, rflag, pflag, or the wall
    // masks of the correction cycles with bodies as walls (vrflag null)
    const int8_t *vrow, *vibk, *vrflag, *vpflag;
    LineSweep sw[3];               // half-sweep shapes by direction (colour set per launch)
};

struct ReqDesc {
    int nlev;
    int mixed;                     // 1: FP32 V-cycles on the correction (G = p + e), 0: FP64 (G is p)
    int walls;                     // 1: the correction cycles treat bodies as walls
    int homo;                      // is_PBC_homogeneous (GCM dpdn terms)
    int nrb;                       // blocks of the row-strided reductions
    long q111;                     // offset of cell (1,1,1)
    ReqLevel L[MAX_LEVELS];
    BcInfo bc[2];                  // [0] level-1 outer BC as given, [1] homogeneous
    GcmDev g;
    const GcmNode *gnode;          // relations in topological order
    const int *glevel;             // topological level offsets (device copy)
    double *gdelta;                // change of each target in the last pass (mixed path)
    int ntgt;                      // relations in gnode
    const long *off8;
    double *p, *div, *pprev;
    double *aa_x, *aa_g;
    float *aa_fprev, *aa_c;
    AaCols cols;
    int *ring;
    double *partial, *scal;
};

// One request's share of a batched launch: blocks [b0, b0 + gx*gy*gz) run
// request slot req with a gx x gy x gz block grid.
struct BEnt {
    int b0, gx, gy, gz, req;
};

}  // namespace fc
