/*
 * fc_oracle.h -- CPU FP64 restatement of FEICAR3D's pressure-Poisson hot
 * path.  TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py as the checker.  The product library
 * never links or calls it.
 *
 * Same data contract as include/fc_poisson.h (Fortran column-major arrays,
 * 1-based GCM indices).  ordering selects the line/GCM update order:
 *   OR_ORDER_REFERENCE  lexicographic line Gauss-Seidel exactly as
 *                       itsolv_mg (CORE_MG.F90:650-925) and list-order
 *                       Gauss-Seidel GCM updates (GCM_GHOSTCELL_UPDATE.F90:
 *                       425-688);
 *   OR_ORDER_COLORED    the GPU ordering: per direction two checkerboard
 *                       colours of lines, outer-BC fill before each colour;
 *                       GCM dead cells then ghost cells, each as a Jacobi
 *                       step.  Same fixed point, different path.
 */
#ifndef FC_ORACLE_H
#define FC_ORACLE_H
#include "../include/fc_poisson.h"

#ifdef __cplusplus
extern "C" {
#endif

#define OR_ORDER_REFERENCE 0
#define OR_ORDER_COLORED   1

typedef struct or_state or_state;

or_state *or_create(const fc_grid *g, const fc_pson_params *p, int ordering);
void   or_destroy(or_state *s);
void   or_set_params(or_state *s, const fc_pson_params *p, int ordering);
int    or_levels(const or_state *s, int *nlev, int *nxl, int *nyl, int *nzl);
int    or_set_masks(or_state *s, const int8_t *iblank, const int8_t *ghost, const int8_t *dead);
int    or_set_gcm(or_state *s,
                  int nG, const int *iG, const int *jG, const int *kG,
                  const int *iC, const int *jC, const int *kC,
                  const double *coeffN, const double *alpha, const double *dpdnG,
                  int nD, const int *iD, const int *jD, const int *kD,
                  const int *iDC, const int *jDC, const int *kDC,
                  const double *coeffDN, const double *probeLenD,
                  const double *dpdnD, const double *ipD);
int    or_set_bc(or_state *s, const double *pgx1, const double *pgx2,
                 const double *pgy1, const double *pgy2,
                 const double *pgz1, const double *pgz2, double outflow_area);
int    or_get_bc(or_state *s, double *pgx1, double *pgx2, double *pgy1,
                 double *pgy2, double *pgz1, double *pgz2);

int    or_solve_poisson(or_state *s, double *pPrime, const double *div,
                        const double *p_prev, double *res_laplacian, fc_pson_stats *st);
int    or_mg_solve(or_state *s, double *var, const double *rr, fc_pson_stats *st);
int    or_itsolv(or_state *s, double *var, const double *r);
int    or_ghostcell_pressure(or_state *s, double *pres, const double *r,
                             const double *p_prev, double *res_laplacian, fc_pson_stats *st);
double or_calc_residual(or_state *s, const double *var, const double *r);
double or_calc_residual2(or_state *s, const double *var, const double *p_temp);
double or_rel_residual(or_state *s, const double *var, const double *r);
int    or_rhs_poisson(or_state *s, double *fu, double *fv, double *fw,
                      double dtinv, double *div, double *sumd, double *resDiv);
void   or_enforce_p_compatibility(or_state *s, const double *pres);

/* velocity projection after the solve (CORE_TIME_STEP.F90:211-222) */
int    or_set_vel_bc(or_state *s, const double *bcxu, const double *bcyv, const double *bczw);
int    or_get_vel_bc(or_state *s, double *bcxu, double *bcyv, double *bczw);
double or_outflow_area(or_state *s);
int    or_correct_vel(or_state *s, const double *pPrime, double dt, double *u, double *v, double *w);
int    or_correct_ghostcell_velocity(or_state *s, const double *pVar, double dt_fac,
                                     double *u, double *v, double *w);
int    or_face_vel(or_state *s, const double *u, const double *v, const double *w,
                   double *fu, double *fv, double *fw);
int    or_update_pressure(or_state *s, double *pPrime, double *p);

/* advection-diffusion line solves (CORE_SOLVE_AD.F90, CORE_SET_SOLVERS.F90) */
int    or_set_solve_ad(or_state *s, double reinv, double dt, int crank_nicolson, double omega_ad);
void   or_get_solve_ad(or_state *s, double *amx, double *apx, double *amy, double *apy,
                       double *amz, double *apz);
int    or_itsolv_ad(or_state *s, double *var, const double *r, const double *fu,
                    const double *fv, const double *fw);
double or_calc_residual_ad(or_state *s, const double *var, const double *r, const double *fu,
                           const double *fv, const double *fw, int *loc, double *res2);
/* rhs_advec_diff (CORE_SET_SOLVERS.F90:10-61) from uOld, vOld, wOld and the
 * face velocities; visc = viscTot (NULL: the constant reinv of
 * or_set_solve_ad); writes the interior cells of nlu, nlv, nlw */
int    or_rhs_advec_diff(or_state *s, const double *uo, const double *vo, const double *wo, const double *fu,
                         const double *fv, const double *fw, const double *visc, double alfa_upwind, double dt,
                         double *nlu, double *nlv, double *nlw);

/* solve_ad's loop (CORE_SOLVE_AD.F90:11-127) and what it calls besides
 * itsolv_ad: set_outer_velocity_bc / set_outer_ghost_velocity
 * (CORE_SET_OUTER_BC.F90:18-509, 739-823), GCM_ghostcell_velocity
 * (GCM_GHOSTCELL_UPDATE.F90:1-359).  vals/freq: ux1, ux2, vx1, vx2, wx1, wx2,
 * uy1, ..., wz2; out9: bcxu, bcxv, bcxw, bcyu, bcyv, bcyw, bczu, bczv, bczw
 * (NULL skipped); ubi: u, v, w per ghost cell (uBodyIntercept ...). */
int    or_set_outer_velocity(or_state *s, const double *vals, const double *freq, double y_origin,
                             double y_out);
int    or_set_outer_velocity_bc(or_state *s, const double *u, const double *v, const double *w, double time);
int    or_set_outer_ghost_velocity(or_state *s, double *u, double *v, double *w);
int    or_get_outer_velocity_bc(or_state *s, double *const *out9);
int    or_set_gcm_vel(or_state *s, const double *coeffD, const double *coeffDeadD, const double *ubi,
                      const double *ubiDead, double probe_norm_d);
int    or_ghostcell_velocity(or_state *s, int iOpt, double *u, double *v, double *w, const double *nlu,
                             const double *nlv, const double *nlw, const double *fu, const double *fv,
                             const double *fw, double restol_gcv, int *sweeps, double *resmax);
int    or_solve_ad(or_state *s, double *u, double *v, double *w, const double *nlu, const double *nlv,
                   const double *nlw, double *fu, double *fv, double *fw, double time, int itermax_ad,
                   double restol_ad, double restol_gcv, int *iters, double *res, int *loc);

/* level operations on the internal hierarchy (0 = phi, 1 = rhs, 2 = res,
 * 3 = ibk, 4 = igk) */
void  *or_level_ptr(or_state *s, int lev, int which);
void   or_level_metric(or_state *s, int lev, int which, double **ptr);
void   or_itsolv_mg(or_state *s, int lev);
void   or_sweep(or_state *s, int lev, int dir, int color);
void   or_bc_fill(or_state *s, int lev);
void   or_residual_mg(or_state *s, int lev);
void   or_mg_inject(or_state *s, int lev);
void   or_mg_interp(or_state *s, int lev);
void   or_vcycle(or_state *s);

/* tdma(a,b,c,r,u,n1,n2) on 0-based arrays of length n2+1 */
void   or_tdma(const double *a, const double *b, const double *c, const double *r,
               double *u, int n1, int n2);

#ifdef __cplusplus
}
#endif
#endif
