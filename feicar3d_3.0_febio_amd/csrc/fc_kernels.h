// fc_kernels.h -- launch wrappers of the gfx950 kernels (fc_kernels.hip).
#pragma once
#include "fc_device.h"

namespace fc {

template <typename T> void bc_fill(hipStream_t st, const LevelDims &d, T *p, const BcInfo &bc);
template <typename T>
void residual(hipStream_t st, const LevelDims &d, const Metrics<T> &m, const T *phi, const T *rhs, T *res,
              const int8_t *igk);
// out2[0] = sum (res*flag)^2, out2[1] = sum (rhs*flag)^2 over interior cells
template <typename T>
void resnorm(hipStream_t st, const LevelDims &d, const Metrics<T> &m, const T *phi, const T *rhs,
             const int8_t *igk, double *partial, double *out2);
void maxdiff(hipStream_t st, const LevelDims &d, const double *p, const double *pt, const int8_t *iblank,
             double *partial, double *out);
void pin(hipStream_t st, const LevelDims &d, double *p, double *corner);
template <typename T>
void restrict_(hipStream_t st, const LevelDims &f, const LevelDims &c, const Metrics<T> &mf, const Metrics<T> &mc,
               const T *res_f, T *rhs_c, const int8_t *ibk_f);
template <typename T>
void prolong(hipStream_t st, const LevelDims &f, const LevelDims &c, const Metrics<T> &mf, const Metrics<T> &mc,
             const T *phi_c, T *phi_f, const int8_t *ibk_f, const int8_t *ibk_c);
template <typename T>
void sweep(hipStream_t st, const LevelDims &d, const Metrics<T> &m, T *phi, const T *rhs, const int8_t *igk,
           int dir, int color, double omega);
int tdma_batched(hipStream_t st, int nsys, int n, const double *a, const double *b, const double *c,
                 const double *r, double *u, int prec);
void gcm_dead(hipStream_t st, const GcmDev &g, const double *pres, const double *pprev, const long *off8, int homo,
              double *outv);
// met = {dxcinv, dycinv, dzcinv, dxinv, dyinv, dzinv} of level 1
void gcm_ghost(hipStream_t st, const GcmDev &g, const double *pres, const double *r, const long *off8, int homo,
               int mix, const LevelDims &d, const double *const *met, const int *gijk, double *outv, double *reslap);
// one exact pass over the topological levels (level_off has nlevels+1
// entries on the host); returns the number of partial blocks used
long gcm_levels(hipStream_t st, const GcmDev &g, const int *order, const int *level_off, int nlevels, double *pres,
                const double *pprev, const long *off8, int homo, double *partial, double *out);
void gcm_scatter(hipStream_t st, int n, const long *target, const int *active, const double *vals, double *pres,
                 double *partial, double *out);
// met = {dxinv, dyinv, dzinv, dx, dy, dz}; out2 = {sumd*dtinv, resDiv}
void rhs_poisson(hipStream_t st, const LevelDims &d, const double *fu, const double *fv, const double *fw,
                 const int8_t *igk, const double *const *met, double dtinv, double *div, double *partial,
                 double *out2);
void resid64_to32(hipStream_t st, const LevelDims &d, const Metrics<double> &m, const double *p, const double *div,
                  const int8_t *igk, float *r32, double *partial, double *out2);
void add_correction(hipStream_t st, long n, double *p, const float *e);
template <typename T> void fill(hipStream_t st, long n, T *p, T v);
// Anderson acceleration (history depth <= 8)
void aa_update(hipStream_t st, long n, const double *g, const double *x, float *fprev, double *gprev, float *dF,
               float *dG, int have_prev, double *partial, double *out2);
void aa_dots(hipStream_t st, long n, const float *const *cols_dev, int m, const float *v1, const float *v2,
             double *partial, double *out);
void aa_combine(hipStream_t st, long n, const float *const *cols_dev, int m, const double *gam, const double *g,
                double *x, double *p);

}  // namespace fc
