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

// Anderson history columns (FP32), indexed by ring slot.
struct AaCols {
    float *dF[8];
    float *dG[8];
};

}  // namespace fc
