/*
 * fc_geom_oracle.h -- CPU FP64 restatement of FEICAR3D's immersed-boundary
 * geometry refresh for triangulated bodies (one time step of a moving body):
 * calculate_arclength_norm_ds, set_iblank_body_fast with
 * calc_bodyIntercept_Unstruc, and GCM_set_internal_boundary (smooth_iblank,
 * the 2-D smoothing passes, GCM_define_ghostcells, the ghost-cell relations
 * with GCM_Calc_VanMatrixDN, GCM_set_deadcells).  TEST INFRASTRUCTURE ONLY:
 * used by tests/ and the cpu_baseline leg of bench.py as the checker; the
 * product library never links or calls it.
 *
 * Same data contract as include/fc_geometry.h.
 */
#ifndef FC_GEOM_ORACLE_H
#define FC_GEOM_ORACLE_H
#include "../include/fc_geometry.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct og_state og_state;

og_state *og_create(const fc_grid *g, const fc_geom_params *p);
void og_destroy(og_state *s);
int  og_set_body(og_state *s, int ibody, int npts, int nelem, const double *x, const double *y,
                 const double *z, const int *tri);
int  og_move_body(og_state *s, int ibody, const double *x, const double *y, const double *z);
int  og_refresh(og_state *s, fc_geom_stats *st);
int  og_get_counts(const og_state *s, int *nGhost, int *nDead);
int  og_get_masks(const og_state *s, int8_t *iblank, int8_t *ghost, int8_t *dead, int8_t *bodyNum,
                  int8_t *ps1, int8_t *ps2);
int  og_get_ghosts(const og_state *s, fc_geom_ghosts *g);
int  og_get_dead(const og_state *s, fc_geom_dead *d);
int  og_get_surface(const og_state *s, int ibody, fc_geom_surface *f);
int  og_get_dcip(const og_state *s, double *xyzDCIP);
/* marker velocities of a body, current and previous step */
int  og_set_marker_velocity(og_state *s, int ibody, const double *u, const double *v, const double *w,
                            const double *uOld, const double *vOld, const double *wOld);
/* GCM_SetBodyInterceptValues (dpdn only when ntime > 1) */
int  og_body_intercept_values(og_state *s, double dt, int ntime);
int  og_get_bi_values(const og_state *s, double *uG, double *vG, double *wG, double *dpdnG, double *uD,
                      double *vD, double *wD, double *dpdnD);
/* compute_marker_stress without the FSI blending: per body the nine totals
 * cxt cyt czt cmxt cmyt cmzt cxs cys czs into totals[9*ibody ..] */
int  og_marker_stress(og_state *s, const double *p, const double *u, const double *v, const double *w,
                      double reinv, double *totals);
int  og_get_marker_forces(const og_state *s, int ibody, fc_geom_forces *f);
/* one body intercept (calc_bodyIntercept_Unstruc) of the cell (i,j,k) */
int  og_body_intercept(const og_state *s, int ibody, int i, int j, int k, double *bi, int *elem,
                       double *dist, double *dotNorm);
/* GCM_Calc_VanMatrixDN for one relation */
int  og_van_matrix(const og_state *s, int iG, int jG, int kG, int iC, int jC, int kC,
                   const double *eval, const double *bi, const double *bin, double *cD, double *cN);

#ifdef __cplusplus
}
#endif
#endif /* FC_GEOM_ORACLE_H */
