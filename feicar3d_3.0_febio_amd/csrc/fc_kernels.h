// fc_kernels.h -- launch wrappers of the gfx950 kernels (fc_kernels.hip).
#pragma once
#include "fc_device.h"

namespace fc {

// Row-strided reduction kernels run ROW_BLOCKS blocks and store up to
// ROW_COMPONENTS partial sums per block, component-major; the reduction
// scratch must hold ROW_BLOCKS * ROW_COMPONENTS doubles.
constexpr int ROW_BLOCKS = 2048;
constexpr int ROW_COMPONENTS = 20;

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
// out3 = {max |v|, its first interior cell in (k, j, i) order as the index
// (k-1)(nx-1)(ny-1) + (j-1)(nx-1) + i-1 (-1 if v is zero everywhere), sum v^2};
// mode 0: v = the flag-cell 7-point residual of var, r (mask = igk);
// mode 1: v = var - r where mask (iblank) = 0.  partial: 3*ROW_BLOCKS.
void resloc(hipStream_t st, const LevelDims &d, const Metrics<double> &m, const double *var, const double *r,
            const int8_t *mask, int mode, double *partial, double *out3);
void pin(hipStream_t st, const LevelDims &d, double *p, double *corner);
template <typename T>
void restrict_(hipStream_t st, const LevelDims &f, const LevelDims &c, const Metrics<T> &mf, const Metrics<T> &mc,
               const T *res_f, T *rhs_c, const int8_t *ibk_f, const int8_t *rflag_c);
// residual of level f restricted into the rhs of level c, coarse phi zeroed;
// rflag_c / pflag_c: the precomputed ibk flags of the coarse level
template <typename T>
void resrestrict(hipStream_t st, const LevelDims &f, const LevelDims &c, const Metrics<T> &mf, const Metrics<T> &mc,
                 const T *phi_f, const T *rhs_f, const int8_t *igk_f, const int8_t *ibk_f, const int8_t *rflag_c,
                 T *rhs_c, T *phi_c);
// prolong's block grid (64 x 4 threads, one thread per group of fine cells
// that share a coarse interpolation block)
dim3 prolong_grid(const LevelDims &f, const LevelDims &c);
// walls != 0: the correction cycles with bodies as walls (ibk_f / ibk_c are
// the wall masks, pflag_c marks interpolation blocks with a held coarse cell)
template <typename T>
void prolong(hipStream_t st, const LevelDims &f, const LevelDims &c, const Metrics<T> &mf, const Metrics<T> &mc,
             const T *phi_c, T *phi_f, const int8_t *ibk_f, const int8_t *ibk_c, const int8_t *pflag_c, int walls);
// one colour of one direction; also rewrites the ghost cells of the cells it
// updates (the BC of bc), so ghosts stay consistent with the interior
template <typename T>
void sweep(hipStream_t st, const LevelDims &d, const Metrics<T> &m, T *phi, const T *rhs, const int8_t *igk,
           int dir, int color, double omega, const BcInfo &bc);
// one colour of one direction of itsolv (CORE_POISSON.F90:205-456): the
// relaxation old + omega (x - old), ghost cells left as they are
void sweep_itsolv(hipStream_t st, const LevelDims &d, const Metrics<double> &m, double *phi, const double *rhs,
                  const int8_t *igk, int dir, int color, double omega, const BcInfo &bc);
// the launch shape sweep() uses for this level, direction and colour
LineSweep sweep_shape(const LevelDims &d, int dir, int color, double omega);
// dynamic LDS of one half-sweep block: the (b, d) pairs of its lines and the
// staged diagonal metric along the lines
size_t sweep_lds_bytes(const LineSweep &g, size_t elem);
int tdma_batched(hipStream_t st, int nsys, int n, const double *a, const double *b, const double *c,
                 const double *r, double *u, int prec);
void gcm_dead(hipStream_t st, const GcmDev &g, const double *pres, const double *pprev, const long *off8, int homo,
              double *outv);
// met = {dxcinv, dycinv, dzcinv, dxinv, dyinv, dzinv} of level 1
void gcm_ghost(hipStream_t st, const GcmDev &g, const double *pres, const double *r, const long *off8, int homo,
               int mix, const LevelDims &d, const double *const *met, const int *gijk, double *outv, double *reslap);
// one exact pass over the topological levels (level_off: nlevels+1 host
// entries); gdelta (may be null) receives new - old of each target, indexed
// like the schedule records
void gcm_levels(hipStream_t st, const GcmNode *nodes, const int *level_off, int nlevels, double *pres,
                const double *pprev, const long *off8, int homo, double *gdelta);
// for the targets q of nodes[t], t < n: e(q) += gdelta[t] (assign: e(q) =
// gdelta[t]), p(q) -= gdelta[t]; with e == nullptr: p(q) += gdelta[t]
template <typename TE>
void gcm_fold(hipStream_t st, const GcmNode *nodes, int n, const double *gdelta, TE *e, double *p, int assign);
void gcm_scatter(hipStream_t st, int n, const long *target, const int *active, const double *vals, double *pres,
                 double *partial, double *out);
// met = {dxinv, dyinv, dzinv, dx, dy, dz}; out2 = {sumd*dtinv, resDiv};
// partial: 2 doubles per interior (j,k) row
void rhs_poisson(hipStream_t st, const LevelDims &d, const double *fu, const double *fv, const double *fw,
                 const int8_t *igk, const double *const *met, double dtinv, double *div, double *partial,
                 double *out2);
// the same with the fields (fu, fv, fw, div) in layout d and the mask in dm
void rhs_poisson2(hipStream_t st, const LevelDims &d, const LevelDims &dm, const double *fu, const double *fv,
                  const double *fw, const int8_t *igk, const double *const *met, double dtinv, double *div,
                  double *partial, double *out2);
// out2[0] /= dtinv
void rhs_sumd(hipStream_t st, double *out2, double dtinv);
// set_outer_velocity_bc into b.arr from u, v, w (ynode: node y of level 1)
void outer_velocity_bc(hipStream_t st, const LevelDims &d, const VelBc &b, const double *u, const double *v,
                       const double *w, const double *ynode, double t);
// set_outer_ghost_velocity from b.arr
void outer_ghost_velocity(hipStream_t st, const LevelDims &d, const VelBc &b, double *u, double *v, double *w);
// one Jacobi step of GCM_ghostcell_velocity (dead cells, then with iopt != 1
// ghost cells); out2 = {max change of the dead cells, of the ghost cells};
// vals: 3 max(nG, nD) doubles
void gcm_velocity_step(hipStream_t st, const GcmDev &g, const GcmVel &gv, const LevelDims &d, const AdMet &am,
                       int mix, int iopt, double *u, double *v, double *w, const double *nlu, const double *nlv,
                       const double *nlw, const double *fu, const double *fv, const double *fw, const int *gijk,
                       const long *off8, double *vals, double *partial, double *out2);
// phi(q[t]) = sum_s M[t][s] rhs(q[s]) (FP64 sums), t, s < n
template <typename T>
void coarse_direct(hipStream_t st, int n, const double *M, const long *q, const T *rhs, T *phi);
// Zero-gradient outer faces (zg: bit f for face x1, x2, y1, y2, z1, z2).
// out = the flag cells' net face mass flux (enforce_global_mass_consv);
// met = {dxinv, dyinv, dzinv, dx, dy, dz} of level 1 (l1rhs)
void massflux(hipStream_t st, const LevelDims &d, const double *fu, const double *fv, const double *fw,
              const int8_t *igk, const double *const *met, double *partial, double *out);
// out = the weighted ghost-cell fluxes of correct_ghost_massflux; div (may
// be null) receives their weighted divergence
void ghost_massflux(hipStream_t st, const LevelDims &d, const GhostVel *g, int n, const double *fu, const double *fv,
                    const double *fw, const double *const *met, double dtinv, double *div, double *partial,
                    double *out);
// the uniform velocity correction of the zero-gradient faces (face
// velocities; bu/bv/bw = bcxu/bcyv/bczw may be null)
void outflow_velocity(hipStream_t st, const LevelDims &d, int zg, double corr, double *fu, double *fv, double *fw,
                      double *bu, double *bv, double *bw);
// out2 = {pressure-gradient flux of the flag cells, of the ghost cells less
// res_laplacian (reslap null: 0)} (GCM_enforce_p_compatibility);
// l1met = {dxcinv, dycinv, dzcinv, ...}, l1w = {dxinv, dyinv, dzinv, dx, dy, dz}
void pressure_flux(hipStream_t st, const LevelDims &d, const double *p, const int8_t *igk, const double *const *l1met,
                   const double *const *l1w, const GhostVel *g, int ng, const double *reslap, double *partial,
                   double *out2);
// pgrad of the zero-gradient faces corrected by -(flux2[0] + flux2[1]) / area
void pgrad_correct(hipStream_t st, const LevelDims &d, int zg, const double *flux2, double area, double *const *pg);
// out2 may be null (no sums); e32 may be null (else zeroed)
// Velocity projection after the pressure solve (CORE_TIME_STEP.F90:211-222),
// FP64 on level-1 fields: correct_vel, correct_ghostcell_velocity, face_vel,
// update_pressure.  bc_u/bc_v/bc_w are bcxu(0:1,0:ny+1,0:nz+1),
// bcyv(0:nx+1,0:1,0:nz+1), bczw(0:nx+1,0:ny+1,0:1) in the host layout; pzero
// marks the entries update_pressure sets to zero.
void correct_vel(hipStream_t st, const LevelDims &d, const ProjMet &m, const double *pp, const int8_t *iblank,
                 double dt, double *u, double *v, double *w);
void correct_ghost_vel(hipStream_t st, const LevelDims &d, const ProjMet &m, const GhostVel *g, int n, int mix,
                       const double *pv, double dt, double *u, double *v, double *w);
void face_vel(hipStream_t st, const LevelDims &d, const ProjMet &m, const double *u, const double *v, const double *w,
              const double *bc_u, const double *bc_v, const double *bc_w, double *fu, double *fv, double *fw);
void update_pressure(hipStream_t st, const LevelDims &d, const int8_t *pzero, double *pp, double *p);

// Advection-diffusion line solves (CORE_SOLVE_AD.F90): one half sweep of
// itsolv_ad (lines of direction dir and one colour), and calc_residual_ad
// (out3 = max |res|, its lexicographic cell index, sum res^2).
void ad_sweep(hipStream_t st, const LevelDims &d, const AdMet &m, double *var, const double *r, const double *fu,
              const double *fv, const double *fw, const int8_t *igk, int dir, int color);
// rhs_advec_diff (CORE_SET_SOLVERS.F90:10-61) on the interior cells of level 1
void rhs_advec_diff(hipStream_t st, const LevelDims &d, const RhsAdMet &m, const double *uo, const double *vo,
                    const double *wo, const double *fu, const double *fv, const double *fw, double *nlu, double *nlv,
                    double *nlw);
void ad_residual(hipStream_t st, const LevelDims &d, const AdMet &m, const double *var, const double *r,
                 const double *fu, const double *fv, const double *fw, const int8_t *igk, double *partial,
                 double *out3);

void resid64_to32(hipStream_t st, const LevelDims &d, const Metrics<double> &m, const double *p, const double *div,
                  const int8_t *igk, float *r32, float *e32, double *partial, double *out2);
// the same, with the calc_residual sums (components 0, 1) and max|p|
// (component 3) left in partial for a later reduce (aa_fresid)
// (TE: the correction's element type, float in the mixed path, double in
// the FP64 correction path)
template <typename TE>
void resid64_to32_sums(hipStream_t st, const LevelDims &d, const Metrics<double> &m, const double *p,
                       const double *div, const int8_t *igk, TE *r32, TE *e32, double *partial);
void add_correction(hipStream_t st, long n, double *p, const float *e);
template <typename T> void fill(hipStream_t st, long n, T *p, T v);
// Anderson acceleration of the outer loop (history depth <= 8); ring[0] =
// depth, ring[1] = columns in use, ring[2] = a previous evaluation exists,
// ring[4..12) = slots oldest..newest; scal layout in k_aa_lsq.
void add_into(hipStream_t st, const LevelDims &d, const double *p, const float *e, double *G);
void aa_resid(hipStream_t st, const LevelDims &d, const Metrics<double> &m, const double *G, const double *div,
              const int8_t *igk, const double *x, float *fprev, const float *cprev, const AaCols &cols,
              const int *ring, double *partial, double *scal);
// mixed path: the history update from the map's step held in e (f = e -
// x(1,1,1)), completing the partials resid64_to32_sums left, then the
// reduction into scal
template <typename TE>
void aa_fresid(hipStream_t st, const LevelDims &d, const TE *e, float *fprev, const float *cprev,
               const AaCols &cols, const int *ring, double *partial, double *scal);
// G (+ E when not null) is the map value
template <typename TE>
void aa_lsq(hipStream_t st, int *ring, double *scal, const AaCols &cols, const double *G, const TE *E, long q111,
            int reset);
// x = G (+ E) - dG gamma (x may be null), c = dG gamma, p = x - x(1,1,1)
template <typename TE>
void aa_next(hipStream_t st, const LevelDims &d, const double *G, const TE *E, const AaCols &cols,
             const double *scal, double *x, float *c, double *p);

// Batched launches (serving).  E: device array of ne entries, ascending b0,
// covering nb blocks; each entry names a request slot of D and the block
// grid its own launch would use.  prec: 0 FP64, 1 FP32 hierarchy; bcsel:
// index into ReqDesc::bc.
void b_bc_fill(hipStream_t st, const BEnt *E, int ne, int nb, const ReqDesc *D, int prec, int lev, int fld, int bcsel);
void b_gcm_level(hipStream_t st, const BEnt *E, int ne, int nb, const ReqDesc *D, int l);
void b_resid64_to32(hipStream_t st, const BEnt *E, int ne, int nb, const ReqDesc *D);
void b_sweep(hipStream_t st, const BEnt *E, int ne, int nb, size_t lds, const ReqDesc *D, int prec, int lev, int dir,
             int color, int bcsel);
void b_resrestrict(hipStream_t st, const BEnt *E, int ne, int nb, const ReqDesc *D, int prec, int lev);
void b_prolong(hipStream_t st, const BEnt *E, int ne, int nb, const ReqDesc *D, int prec, int lev);
void b_gcm_fold(hipStream_t st, const BEnt *E, int ne, int nb, const ReqDesc *D, int back);
void b_aa_resid(hipStream_t st, const BEnt *E, int ne, int nb, const ReqDesc *D);
void b_reduce_cols(hipStream_t st, const BEnt *E, int ne, int nb, const ReqDesc *D);
void b_aa_lsq(hipStream_t st, const BEnt *E, int ne, int nb, const ReqDesc *D, int reset);
void b_aa_next(hipStream_t st, const BEnt *E, int ne, int nb, const ReqDesc *D, int start);
void b_gather(hipStream_t st, const BEnt *E, int ne, int nb, const ReqDesc *D, double *out);

}  // namespace fc
