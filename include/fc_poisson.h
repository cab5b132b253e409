/*
 * fc_poisson.h -- C-ABI of the MI355X (gfx950) pressure-Poisson hot path of
 * FEICAR3D: solve_poisson -> mg_solve V-cycle with the itsolv_mg line-SOR
 * (tdma) smoother on the iblank-masked Cartesian grid, plus the
 * GCM_ghostcell_pressure ghost/dead-cell update it calls every outer
 * iteration.
 *
 * Every entry point takes plain pointers and sizes.  Field arrays are the
 * reference's Fortran module arrays, passed as-is: column-major, i fastest,
 * extents (0:nx+1, 0:ny+1, 0:nz+1) at one flow rank (AMODULES.F90:276-297 with
 * yb1=zb1=0, yb2=ny+1, zb2=nz+1, CORE_PARALLEL.F90:170-185).  Masks are int8
 * (INT_K=1, AMODULES.F90:5).  GCM list indices are the reference's Fortran
 * cell indices (i,j,k as in iGhost/iCellIndex); the field arrays start at
 * index 0, so an index is also the 0-based offset along its axis.
 *
 * Return codes: 0 = converged, 1 = an iteration cap was hit (the reference
 * prints and continues, CORE_POISSON.F90:195-198), <0 = error; the text of
 * the last error of the calling thread is fc_pson_last_error().
 *
 * A handle owns one HIP stream and its device-resident multigrid hierarchy;
 * calls on one handle are serialised by the caller (the reference is
 * single-threaded per rank and not reentrant).  Only one flow rank is
 * supported: the whole domain lives on one GPU.
 */
#ifndef FC_POISSON_H
#define FC_POISSON_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reference parameter codes (AMODULES.F90:15-44). */
#define FC_BC_TYPE_DIRICHLET      1
#define FC_BC_TYPE_ZERO_GRADIENT  2
#define FC_BC_TYPE_PERIODIC       5
#define FC_IT_SOLVER_TYPE_LSOR    1
#define FC_IT_SOLVER_TYPE_MG      3
#define FC_PBC_DIRICHLET          1
#define FC_PBC_NEUMANN            2

/* Precision modes (new; the reference is FP64 only). */
#define FC_PREC_FP64   0   /* plain FP64 drop-in, reference V-cycle shape     */
#define FC_PREC_MIXED  1   /* FP32 V-cycles inside FP64 defect correction     */

/* Outer-loop stopping rule. */
#define FC_OUTER_REFERENCE 0 /* max|p - p_temp| <= restol_gcp (CORE_POISSON.F90:126) */
#define FC_OUTER_RELRES    1 /* flag-cell relative residual <= rel_tol and the
                                mean-free pressure change <= restol_gcp        */

/* Grid: node coordinates x(0:nx+1), y(0:ny+1), z(0:nz+1) as produced by
 * make_grid (CORE_GRID.F90:1-60).  Cell centres and inverse spacings are
 * derived inside with the formulas of CORE_GRID.F90:106-150. */
typedef struct fc_grid {
    int nx, ny, nz;              /* node counts (reference nx, ny, nz)        */
    const double *x, *y, *z;     /* length nx+2, ny+2, nz+2                   */
} fc_grid;

/* Hot-path parameters read from input.dat (MAIN.F90:214-254), plus the
 * fields that are new in this implementation (precision, rel_tol, ...). */
typedef struct fc_pson_params {
    int    it_solver_type;       /* FC_IT_SOLVER_TYPE_MG or _LSOR            */
    double omega_pson;           /* line-SOR relaxation factor               */
    int    itermax_pson;         /* max outer iterations                     */
    double restol_pson;          /* absolute RMS tolerance of mg_solve       */
    int    iterInter;            /* V-cycles per mg_solve = iterInter + 1    */
    int    mgrids_x, mgrids_y, mgrids_z;  /* 0 = automatic level count      */
    double restol_gcp;           /* GCM pressure and outer-loop tolerance    */
    int    itermax_gc;           /* max GCM sweeps per call                  */
    int    mix_GC_form;          /* 1 = hybrid ghost update (alphaGhost)     */
    int    is_PBC_homogeneous;   /* 1 = dp/dn = 0 on immersed bodies         */
    int    pbc[6];               /* pbcx1,pbcx2,pbcy1,pbcy2,pbcz1,pbcz2      */
    double ppp[6];               /* pppx1 ... pppz2 (Dirichlet values)       */
    int    bc_vel[6];            /* bcx1 ... bcz2 velocity BC types.  2 (zero
                                    gradient): outflow face (rhs_poisson's
                                    mass correction, the pressure
                                    compatibility correction of the pgrad
                                    planes).  5 on both faces of x or of y:
                                    periodic pressure ghosts (CORE_SET_OUTER_
                                    BC.F90:1007-1026); on both z faces an
                                    error, as in the reference              */
    int    infoconv;             /* 1 = print MG convergence history         */
    /* new fields */
    int    precision;            /* FC_PREC_FP64 or FC_PREC_MIXED            */
    int    outer_rule;           /* FC_OUTER_REFERENCE or FC_OUTER_RELRES    */
    double rel_tol;              /* relative-residual target (FC_OUTER_RELRES
                                    and the mixed-precision inner target)    */
    int    fp32_cycles;          /* FP32 V-cycles per FP64 refinement step   */
    int    aa_depth;             /* Anderson history depth for FC_OUTER_RELRES
                                    (1..8; 0 = 8)                            */
    int    launch_mode;          /* 0 = one HIP graph per accelerated outer
                                    iteration when no host decision is
                                    needed; 1 = launch every kernel from the
                                    host                                     */
    int    cycle;                /* V-cycle shape: 1 (and 0) = the
                                    reference's (no post-smoothing on the
                                    finest level, CORE_MG.F90:1618-1766);
                                    2 = plus one post-smoothing pass on the
                                    finest level                             */
    int    vc_walls;             /* correction V-cycles of the accelerated
                                    loop (FC_OUTER_RELRES with one V-cycle
                                    per iteration and the exact GCM pass):
                                    1 = immersed bodies are zero-flux walls
                                    of the fluid cells (ghost cells included)
                                    on every level; 0 = masked cells are held
                                    (the reference's igk identity rows).  The
                                    converged pressure is the same; 1 needs
                                    fewer outer iterations                   */
    int    coarse_solve;         /* 1 = FP64 direct solve on the coarsest
                                    level of the wall correction V-cycles
                                    (vc_walls >= 1; when that level has at
                                    most FC_COARSE_DIRECT_MAX free cells):
                                    the dense inverse, formed once on the
                                    host, applied on the device; groups of
                                    free cells without a Dirichlet face get
                                    their right-hand side's volume-weighted
                                    mean removed.  0 = one line-SOR pass
                                    there, as the reference (CORE_MG.F90:
                                    1618-1704).  Not in batched serving.     */
    int    reserved[3];
} fc_pson_params;

/* Largest coarsest level (free cells) the direct solve takes. */
#define FC_COARSE_DIRECT_MAX 1024

/* Per-call statistics (the reference prints these; here they are returned). */
typedef struct fc_pson_stats {
    int    outer_iters;          /* solve_poisson outer iterations           */
    int    vcycles_fp64;         /* executed FP64 V-cycles                   */
    int    vcycles_fp32;         /* executed FP32 V-cycles                   */
    int    refine_steps;         /* FP64 residual/update pairs (mixed mode)  */
    int    gcm_sweeps;           /* executed GCM dead+ghost sweeps           */
    int    converged;            /* 1 if the outer stopping rule was met     */
    double maxres;               /* final calc_residual RMS (node-normalised) */
    double maxres2;              /* final max|p - p_temp| over non-iblank    */
    double rel_resid;            /* final flag-cell ||div-Ap||/||div||       */
    double gcm_change;           /* final GCM max change                     */
    double time_ms;              /* wall time of the call on the stream      */
    double reserved[6];
} fc_pson_stats;

typedef struct fc_pson fc_pson;

/* Library/version and error text. */
const char *fc_pson_version(void);
const char *fc_pson_last_error(void);

/* Fills *p with the defaults this library uses when a caller has no
 * input.dat (recorded with every benchmark). */
void fc_pson_default_params(fc_pson_params *p);

/* MG_initial (CORE_MG.F90:22-599): builds the level hierarchy and allocates
 * it in HBM.  Returns 0 or <0. */
int fc_pson_create(const fc_grid *g, const fc_pson_params *p, fc_pson **out);

/* Replaces the parameters of an existing handle (same grid). */
int fc_pson_set_params(fc_pson *h, const fc_pson_params *p);

/* Frees the hierarchy and the stream. */
int fc_pson_destroy(fc_pson *h);

/* Level hierarchy as built by MG_initial: node counts per level. */
int fc_pson_levels(const fc_pson *h, int *nlev, int *nxl, int *nyl, int *nzl);

/* Masks iblank, ghostCellMark, dead_cell (int8, (nx+2)(ny+2)(nz+2)) and the
 * per-level masks of mg_prepare (CORE_MG.F90:1365-1555), which is run here
 * every time geometry changes (the reference runs it only for moving
 * bodies, CORE_POISSON.F90:121-123).  dead_cell may be NULL (all zero). */
int fc_pson_set_masks(fc_pson *h, const int8_t *iblank,
                      const int8_t *ghostCellMark, const int8_t *dead_cell);

/* Ghost- and dead-cell lists of GCM_set_internal_boundary /
 * GCM_set_deadcells (GCM_SET_BOUNDARY.F90:236-600, 880-1491), consumed by
 * GCM_ghostcell_pressure (GCM_GHOSTCELL_UPDATE.F90:361-739).  Indices are
 * the Fortran array indices; coefficient arrays are coeff(8,n) column-major
 * (iRow fastest); xyzDCIP is the image point of each dead cell (3 per dead
 * cell, x,y,z), i.e. xyzDCIP(iDead(n),jDead(n),kDead(n),0:2). */
int fc_pson_set_gcm(fc_pson *h,
                    int nGhost, const int *iGhost, const int *jGhost, const int *kGhost,
                    const int *iCellIndex, const int *jCellIndex, const int *kCellIndex,
                    const double *coeffGCMN, const double *alphaGhost,
                    const double *dpdnBodyIntercept,
                    int nDead, const int *iDead, const int *jDead, const int *kDead,
                    const int *iDeadCellIndex, const int *jDeadCellIndex, const int *kDeadCellIndex,
                    const double *coeffGCMDeadN, const double *probeLengthDead,
                    const double *dpdnInterceptDead, const double *xyzDCIP);

/* Outer pressure-gradient BC arrays pgradx1(0:ny+1,0:nz+1), pgradx2,
 * pgrady1(0:nx+1,0:nz+1), pgrady2, pgradz1(0:nx+1,0:ny+1), pgradz2 and
 * outflow_area (CORE_SET_OUTER_BC.F90:831-1042, GCM_GHOSTCELL_UPDATE.F90:
 * 865-1003).  NULL arrays mean zero. */
int fc_pson_set_bc(fc_pson *h, const double *pgradx1, const double *pgradx2,
                   const double *pgrady1, const double *pgrady2,
                   const double *pgradz1, const double *pgradz2,
                   double outflow_area);
/* Reads back the pgrad arrays (GCM_enforce_p_compatibility updates them on
 * zero-gradient faces). */
int fc_pson_get_bc(fc_pson *h, double *pgradx1, double *pgradx2,
                   double *pgrady1, double *pgrady2,
                   double *pgradz1, double *pgradz2);

/* solve_poisson() (CORE_POISSON.F90:83-200): in/out pPrime, in div and the
 * previous-step pressure p (read by the dead-cell fallback,
 * GCM_GHOSTCELL_UPDATE.F90:479,501).  res_laplacian (nGhost) is written when
 * mix_GC_form=1 and may be NULL.  Host arrays; copies in and out. */
int fc_pson_solve(fc_pson *h, double *pPrime, const double *div,
                  const double *p_prev, double *res_laplacian,
                  fc_pson_stats *st);

/* mg_solve(var, rr) (CORE_MG.F90:1572-1848): V-cycles on the current masks
 * with ghost values held fixed. */
int fc_pson_mg_solve(fc_pson *h, double *var, const double *rr, fc_pson_stats *st);

/* itsolv(var, r) (CORE_POISSON.F90:205-456): one plain line-SOR pass on the
 * finest level (x-, y- then z-lines, each direction in the GPU's two
 * checkerboard colours), relaxation var + omega (x - var) on interior rows,
 * masked cells and the ghost layers left as they are.  solve_poisson runs it
 * as its inner solve when it_solver_type = 1 (FP64 only). */
int fc_pson_itsolv(fc_pson *h, double *var, const double *r);

/* GCM_ghostcell_pressure(pres, r) (GCM_GHOSTCELL_UPDATE.F90:361-739). */
int fc_pson_ghostcell_pressure(fc_pson *h, double *pres, const double *r,
                               const double *p_prev, double *res_laplacian,
                               fc_pson_stats *st);

/* calc_residual(var, r, resm, loc) (CORE_POISSON.F90:460-551) -> RMS over
 * flag cells with the node-count normalisation; calc_residual2(var, r, resm,
 * loc) (:679-724) -> max|var - p_temp| over iblank = 0 (the reference reads
 * its module array p_temp there, passed here).  loc (int[3], may be NULL):
 * the 1-based (i, j, k) of the first cell in the reference's k-j-i loop
 * order where the largest |residual| (|var - p_temp|) occurs, (0,0,0) if all
 * are zero.  Like every host-array entry point of a handle, they stage the
 * arrays through the handle's resident buffers (pressure, rhs, p_temp),
 * which a later fc_pson_solve_resident would then see. */
int fc_pson_calc_residual(fc_pson *h, const double *var, const double *r, double *resm, int *loc);
int fc_pson_calc_residual2(fc_pson *h, const double *var, const double *p_temp, double *resm, int *loc);

/* rhs_poisson(sumd) (CORE_POISSON.F90:9-79): div = dtinv*(div of face
 * velocities)*flag, sumd = the net mass flux, resDiv = max|div|/dtinv.
 * face_u(0:nx+1,...) etc. have the field extents.  With zero-gradient
 * outer faces (bc_vel = 2) enforce_global_mass_consv (CORE_SET_OUTER_BC.F90:
 * 515-735) runs first, as in the reference: the face velocities of those
 * faces are corrected in place by a uniform velocity, the correction is
 * also applied to the handle's outer-face velocities (fc_vel_set_bc) and
 * the fluid area of those faces becomes the handle's outflow_area.  With
 * mix_GC_form = 1 the ghost-cell mass fluxes are added
 * (correct_ghost_massflux, GCM_GHOSTCELL_UPDATE.F90:742-787). */
int fc_pson_rhs(fc_pson *h, double *face_u, double *face_v,
                double *face_w, double dtinv, double *div, double *sumd,
                double *resDiv);

/* The outflow area and the velocity correction of the last fc_pson_rhs
 * (enforce_global_mass_consv's outflow_area and correction_vel, which the
 * reference also applies to bcxu/bcyv/bczw). */
int fc_pson_get_outflow(fc_pson *h, double *outflow_area, double *correction_vel);

/* rhs_poisson without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops).  No
 * counterpart in the reference.  The same divergence as fc_pson_rhs, on
 * device arrays in the unpadded Fortran layout (0:nx+1, 0:ny+1, 0:nz+1):
 * face_u/v/w are read, div is written on its interior cells (others are not
 * touched), sums[0..1] receive sumd and resDiv (device memory).  The kernels'
 * reduction partials live in a static device array of the library (4 MiB per
 * GPU, held until the process ends): grids with more than
 * FC_NOALLOC_MAX_ROWS interior (j,k) rows are refused with -1.  Calls that
 * overlap on one GPU share that array and race; calls queued on one stream
 * do not.  stream: a hipStream_t, or NULL for the handle's stream.  The
 * work is queued only (no synchronisation).  0 or <0. */
#define FC_NOALLOC_MAX_ROWS 262144
int fc_pson_rhs_noalloc(fc_pson *h, const double *face_u, const double *face_v, const double *face_w,
                        double dtinv, double *div, double *sums, void *stream);

/* tdma(a,b,c,r,u,n1,n2) (CORE_TDMA.F90:10-34), batched on the GPU: nsys
 * independent systems of n rows each, stored one after another; row 0 has
 * no a and row n-1 has no c.  Solved in the precision given by prec. */
int fc_tdma_batched(int nsys, int n, const double *a, const double *b,
                    const double *c, const double *r, double *u, int prec);

/* ---- device-resident path (serving, benchmarks) ------------------------ */
/* Copies pPrime, div, p_prev into the handle's HBM buffers. */
int fc_pson_upload(fc_pson *h, const double *pPrime, const double *div, const double *p_prev);
/* solve_poisson on the resident buffers; nothing crosses PCIe except the
 * stats scalars. */
int fc_pson_solve_resident(fc_pson *h, fc_pson_stats *st);
/* Copies the resident pressure (and res_laplacian if non-NULL) out. */
int fc_pson_download(fc_pson *h, double *pPrime, double *res_laplacian);
/* Bytes of HBM the handle holds. */
size_t fc_pson_device_bytes(const fc_pson *h);
/* Sets the resident pressure to zero (the initial guess of a new request). */
int fc_pson_reset(fc_pson *h);
/* Finest-level line-sweep timing: when enabled, every finest-level half
 * sweep (one colour of one direction of itsolv_mg) is bracketed by HIP
 * events on the handle's stream; _read returns the summed milliseconds and
 * the number of launches since enabling. */
int fc_pson_timing(fc_pson *h, int enable);
int fc_pson_timing_read(fc_pson *h, double *sweep_ms, long *launches);

/* ---- velocity projection after the solve --------------------------------
 * The calls that follow solve_poisson in a time step (CORE_TIME_STEP.F90:
 * 211-222), in FP64 with the reference's rounding of every operation, so the
 * results equal the reference's to the bit.  Arrays have the field extents
 * (0:nx+1, 0:ny+1, 0:nz+1), Fortran order; the handle's masks
 * (fc_pson_set_masks) and ghost list (fc_pson_set_gcm) are used.  Entries a
 * routine does not write keep the caller's values.  The host-pointer calls
 * use the handle's resident pressure buffer as scratch. */
/* The outer-face velocities face_vel copies onto the domain faces:
 * bcxu(0:1,0:ny+1,0:nz+1), bcyv(0:nx+1,0:1,0:nz+1), bczw(0:nx+1,0:ny+1,0:1)
 * (CORE_MEMORY_ALLOCATE.F90:66-80); NULL = zeros. */
int fc_vel_set_bc(fc_pson *h, const double *bcxu, const double *bcyv, const double *bczw);
/* Reads back bcxu/bcyv/bczw (rhs_poisson's enforce_global_mass_consv,
 * CORE_SET_OUTER_BC.F90:664-735, corrects them on zero-gradient faces);
 * NULL pointers are skipped. */
int fc_vel_get_bc(fc_pson *h, double *bcxu, double *bcyv, double *bczw);
/* correct_vel() (CORE_FRAC_STEP_ROUTINES.F90:134-210): u, v, w -= dt * grad
 * pPrime * (1 - iblank) on the interior cells. */
int fc_correct_vel(fc_pson *h, const double *pPrime, double dt, double *u, double *v, double *w);
/* correct_ghostcell_velocity(pVar) (GCM_GHOSTCELL_UPDATE.F90:789-852): the
 * ghost cells' u, v, w -= dt_fac * grad pVar * alphaGhost when mix_GC_form =
 * 1 (flag 0 otherwise).  dt_fac is dt, or 2*rk_bet(rk_step)*dt under RK3.
 * Each ghost cell must be listed once. */
int fc_correct_ghostcell_velocity(fc_pson *h, const double *pVar, double dt_fac, double *u, double *v,
                                  double *w);
/* face_vel(u_vel, v_vel, w_vel) (CORE_FRAC_STEP_ROUTINES.F90:6-130): face
 * velocities by the fx/fy/fz interpolation, outer faces from the arrays of
 * fc_vel_set_bc.  face_u/v/w are in/out. */
int fc_face_vel(fc_pson *h, const double *u, const double *v, const double *w, double *face_u,
                double *face_v, double *face_w);
/* update_pressure() (CORE_FRAC_STEP_ROUTINES.F90:212-299): pPrime = 0 on
 * solid cells that are not dead-cell candidates (dead_cell /= -1) and on the
 * domain-edge lines, then p = pPrime.  pPrime in/out, p out. */
int fc_update_pressure(fc_pson *h, double *pPrime, double *p);
/* Device-resident form: velocities and face velocities stay in HBM between
 * steps (NULL pointers are skipped). */
int fc_vel_upload(fc_pson *h, const double *u, const double *v, const double *w, const double *face_u,
                  const double *face_v, const double *face_w);
int fc_vel_download(fc_pson *h, double *u, double *v, double *w, double *face_u, double *face_v,
                    double *face_w);
/* CORE_TIME_STEP.F90:211-222 on the resident buffers: correct_vel,
 * correct_ghostcell_velocity(pPrime), face_vel, update_pressure.  The
 * resident pressure is pPrime; update_pressure writes the resident p that
 * the next solve reads as p_prev. */
int fc_project_resident(fc_pson *h, double dt, double dt_fac);
/* face_vel(u, v, w) (CORE_FRAC_STEP_ROUTINES.F90:6-130) on the resident
 * velocities into the resident face velocities. */
int fc_face_vel_resident(fc_pson *h);
/* rhs_poisson(sumd) (CORE_POISSON.F90:9-79, with enforce_global_mass_consv
 * on zero-gradient faces, which corrects the resident face velocities) from
 * the resident face velocities into the resident div that
 * fc_pson_solve_resident solves for. */
int fc_pson_rhs_resident(fc_pson *h, double dtinv, double *sumd, double *resDiv);
/* Copies one resident field into a dense (nx+2)(ny+2)(nz+2) FP64 array in
 * device memory (Fortran order; the resident fields are row-padded): which
 * = 0..5 u, v, w, face_u, face_v, face_w; 6 p (update_pressure's, the next
 * solve's p_prev); 7 pPrime.  For device consumers such as
 * fc_geom_marker_stress_device. */
int fc_pson_resident_copy(fc_pson *h, int which, double *dst_device);
/* With fc_pson_timing enabled: milliseconds of the four kernels of the last
 * fc_project_resident call (HIP events on the handle's stream). */
int fc_project_timing_read(fc_pson *h, double *ms4);
/* Copies the resident p (p_prev of the next solve) out. */
int fc_pson_download_p(fc_pson *h, double *p);

/* ---- advection-diffusion line solves ------------------------------------
 * The momentum line solves of solve_ad (CORE_SOLVE_AD.F90), FP64, one flow
 * rank.  Lines are solved in the GPU order (per direction the lines of one
 * checkerboard colour, then the other: x (j+k)%2, y (i+k)%2, z (i+j)%2);
 * the reference's lexicographic Gauss-Seidel order reaches the same fixed
 * point by a different path.  Lines may hold at most 576 rows (n <= 575).
 * The handle's masks give the identity rows (iblank or ghostCellMark). */
/* set_solve_ad() (CORE_SET_SOLVERS.F90:591-676) with nu = reinv, plus the
 * two scalars itsolv_ad reads: half_dt = dt/2 when crank_nicolson is
 * non-zero (advec_scheme == CRANK_NICOLSON2, CORE_SOLVE_AD.F90:196-200),
 * else 0, and omega_ad. */
int fc_set_solve_ad(fc_pson *h, double reinv, double dt, int crank_nicolson, double omega_ad);
/* itsolv_ad(var, r) (CORE_SOLVE_AD.F90:131-776), nsweep times: x-, y- and
 * z-line solves of one velocity component with the Crank-Nicolson advection
 * coefficients of face_u/v/w.  var in/out.  The host arrays are staged
 * through the resident buffers: this call overwrites the resident u, the
 * resident face velocities and the resident nlu (fc_vel_upload,
 * fc_ad_upload), so a host that mixes it with the resident path uploads
 * those again afterwards.  fc_calc_residual_ad does the same. */
int fc_itsolv_ad(fc_pson *h, double *var, const double *r, const double *face_u, const double *face_v,
                 const double *face_w, int nsweep);
/* calc_residual_ad(var, r, resm, loc) (CORE_SOLVE_AD.F90:780-964): max |res|
 * over the cells with iblank = ghostCellMark = 0, its first location in the
 * reference's loop order (i, j, k array indices; 0 0 0 if the residual is
 * zero everywhere) and the sum of res^2 (loc and res2 may be NULL). */
int fc_calc_residual_ad(fc_pson *h, const double *var, const double *r, const double *face_u,
                        const double *face_v, const double *face_w, double *resm, int *loc, double *res2);
/* Device-resident form: the right-hand sides nlu, nlv, nlw stay in HBM
 * (NULL skipped); the velocities and face velocities are those of
 * fc_vel_upload.  comp: 0 = u with nlu, 1 = v with nlv, 2 = w with nlw. */
int fc_ad_upload(fc_pson *h, const double *nlu, const double *nlv, const double *nlw);
/* Host copies of the resident nlu, nlv, nlw (NULL skipped). */
int fc_ad_download(fc_pson *h, double *nlu, double *nlv, double *nlw);
int fc_itsolv_ad_resident(fc_pson *h, int comp, int nsweep);
int fc_calc_residual_ad_resident(fc_pson *h, int comp, double *resm, int *loc, double *res2);
/* With fc_pson_timing enabled: milliseconds of the last
 * fc_itsolv_ad_resident or fc_rhs_advec_diff_resident call (HIP events on
 * the handle's stream). */
int fc_ad_timing_read(fc_pson *h, double *ms);

/* rhs_advec_diff() (CORE_SET_SOLVERS.F90:10-61): the explicit right-hand
 * side of the advection-diffusion step,
 *   nl = vel + dt/2 (-advection(vel) + div(viscTot grad vel))
 * for (uOld, nlu), (vOld, nlv), (wOld, nlw) on the interior cells, with the
 * reference's hybrid central / upwind advection (rhs_advec :65-436, blend
 * alfa_upwind; second-order upwind for 25 <= i <= 375, first order
 * elsewhere, i tested on every axis) and its diffusion (rhs_diff :440-526).
 * Other entries of nlu, nlv, nlw keep their values.  ghostNScoeff(8:10)
 * (:46-58) are nlu, nlv, nlw at the ghost cells, which
 * fc_ghostcell_velocity reads from the resident nlu, nlv, nlw.  Needs
 * fc_set_solve_ad; viscTot is the constant reinv given there unless
 * fc_set_visc_tot set an array (NULL returns to the constant).  The
 * resident form reads the resident u, v, w as uOld, vOld, wOld (uOld = u at
 * the start of a step, BOUNDARY_MOVE.F90:211) and the resident face
 * velocities, and writes the resident nlu, nlv, nlw; the host form stages
 * its arrays through them. */
int fc_set_visc_tot(fc_pson *h, const double *viscTot);
int fc_rhs_advec_diff_resident(fc_pson *h, double alfa_upwind, double dt);
int fc_rhs_advec_diff(fc_pson *h, const double *uOld, const double *vOld, const double *wOld, const double *face_u,
                      const double *face_v, const double *face_w, double alfa_upwind, double dt, double *nlu,
                      double *nlv, double *nlw);

/* ---- solve_ad's loop (CORE_SOLVE_AD.F90:11-127) ----------------------
 * The outer-face velocities are nine arrays: bcxu, bcxv, bcxw
 * (0:1, 0:ny+1, 0:nz+1), bcyu, bcyv, bcyw (0:nx+1, 0:1, 0:nz+1), bczu,
 * bczv, bczw (0:nx+1, 0:ny+1, 0:1); bcxu, bcyv, bczw are those of
 * fc_vel_set_bc / face_vel.  The calls work on the resident u, v, w and face
 * velocities (fc_vel_upload) and nlu, nlv, nlw (fc_ad_upload). */
/* The inputs of set_outer_velocity_bc: the face values ux1, ux2, vx1, vx2,
 * wx1, wx2, uy1, ..., wz2 (18; NULL = zeros), the frequencies of the
 * pulsatile faces in the same order, and yOrigin / yout of the
 * user-specified channel profile. */
int fc_vel_set_outer(fc_pson *h, const double *vals, const double *freq, double y_origin, double y_out);
/* set_outer_velocity_bc() (CORE_SET_OUTER_BC.F90:18-509) at time `time`,
 * per face by bc_vel: 1 Dirichlet, 2 zero gradient, 3 pulsatile, 4
 * symmetry, 5 periodic, 6 user specified (x faces only; the reference's y
 * and z branches write the x arrays); 7 (shear) is an error, as the
 * reference stops there. */
int fc_outer_velocity_bc(fc_pson *h, double time);
/* set_outer_ghost_velocity() (CORE_SET_OUTER_BC.F90:739-823): ghost = 2 bc -
 * interior on the resident u, v, w. */
int fc_outer_ghost_velocity(fc_pson *h);
/* Host copies of all nine outer-face arrays (in9/out9 in the order above;
 * NULL entries: set zero / skipped). */
int fc_vel_set_outer_all(fc_pson *h, const double *const *in9);
int fc_vel_get_outer_all(fc_pson *h, double *const *out9);
/* The velocity relations of GCM_ghostcell_velocity, in the order of the
 * lists of fc_pson_set_gcm: coeffGCMD (nGhost x 8), coeffGCMDeadD (nDead x
 * 8), the body-intercept velocities uBodyIntercept, vBodyIntercept,
 * wBodyIntercept (nGhost x 3) and of the dead cells (nDead x 3), and
 * probeLengthNormalizedD (GCM_GHOSTCELL_UPDATE.F90:72-73). */
int fc_set_gcm_vel(fc_pson *h, const double *coeffGCMD, const double *coeffGCMDeadD,
                   const double *uvwBodyIntercept, const double *uvwBodyInterceptDead,
                   double probeLengthNormalizedD);
/* GCM_ghostcell_velocity(iOpt, u, v, w) (GCM_GHOSTCELL_UPDATE.F90:1-359) on
 * the resident velocities, iOpt 1 (dead cells) or 2 (dead and ghost cells),
 * until the largest change of a step is at most restol_gcv or itermax_gc
 * steps (return 1; the reference stops there).  With mix_GC_form = 1 the
 * ghost value blends in the cell's advection-diffusion row (needs
 * fc_set_solve_ad and the resident nlu, nlv, nlw).  The steps are Jacobi
 * steps (the reference sweeps the lists in order): same fixed point. */
int fc_ghostcell_velocity(fc_pson *h, int iOpt, double restol_gcv, int *sweeps, double *resmax);
/* solve_ad() (CORE_SOLVE_AD.F90:11-127): each pass runs itsolv_ad on u, v,
 * w with nlu, nlv, nlw, GCM_ghostcell_velocity(2), face_vel,
 * enforce_global_mass_consv, set_outer_velocity_bc(time),
 * set_outer_ghost_velocity and calc_residual_ad of the three components,
 * until the largest residual is at most restol_ad (return 0) or itermax_ad
 * passes (return 1).  iters, res and loc (i, j, k) report the passes, the
 * final residual and its location (NULL skipped).  Needs fc_set_solve_ad
 * and, with ghost or dead cells, fc_set_gcm_vel. */
int fc_solve_ad_resident(fc_pson *h, double time, int itermax_ad, double restol_ad, double restol_gcv,
                         int *iters, double *res, int *loc);
/* The same on host arrays: u, v, w and the face velocities in/out, nlu,
 * nlv, nlw in (staged through the resident buffers, which it overwrites). */
int fc_solve_ad(fc_pson *h, double *u, double *v, double *w, const double *nlu, const double *nlv,
                const double *nlw, double *face_u, double *face_v, double *face_w, double time,
                int itermax_ad, double restol_ad, double restol_gcv, int *iters, double *res, int *loc);

/* ---- single-operation entry points on one MG level (itsolv_mg,
 * residual_mg, mg_inject, mg_interp, set_outer_*): each runs exactly one
 * of the reference's level operations on the handle's resident hierarchy.
 * which: 0 = phi, 1 = rhs, 2 = res (FP64 hierarchy), 3 = ibk, 4 = igk
 * (int8), 10 = phi, 11 = rhs, 12 = res (FP32 hierarchy of the mixed path),
 * 5 = the wall mask, 6 = the wall prolongation flags (int8, read only) of
 * the correction cycles with bodies as walls (fc_pson_params.vc_walls).
 * prec: FC_PREC_FP64 or FC_PREC_MIXED, plus FC_LVL_WALLS to run the
 * operation as those cycles do (wall faces in the sweeps and residual, no
 * restriction volume branch, renormalised prolongation). */
#define FC_LVL_WALLS 16
int fc_lvl_write(fc_pson *h, int lev, int which, const void *host);
int fc_lvl_read(fc_pson *h, int lev, int which, void *host);
/* itsolv_mg (CORE_MG.F90:615-938) in the GPU ordering: x, y, z line sweeps,
 * each as two checkerboard colours, outer-BC fill before every colour. */
int fc_lvl_smooth(fc_pson *h, int lev, int prec);
/* One directional half sweep: dir 0/1/2 = x/y/z, color 0/1. */
int fc_lvl_sweep(fc_pson *h, int lev, int dir, int color, int prec);
/* set_outer_pressure_bc + set_outer_ghost_pressure (CORE_SET_OUTER_BC.F90:
 * 831-1186) on phi of level lev. */
int fc_lvl_bc(fc_pson *h, int lev, int prec);
/* res = rhs copy + residual_mg (CORE_MG.F90:956-1043, 1672-1681). */
int fc_lvl_residual(fc_pson *h, int lev, int prec);
/* mg_inject res(lev) -> rhs(lev+1) (CORE_MG.F90:1061-1183). */
int fc_lvl_restrict(fc_pson *h, int lev, int prec);
/* mg_interp phi(lev) added into phi(lev-1) (CORE_MG.F90:1200-1351). */
int fc_lvl_prolong(fc_pson *h, int lev, int prec);
/* New (replaces the single line-SOR pass of the coarsest level,
 * CORE_MG.F90:1618-1704, when coarse_solve = 1): phi = the FP64 direct
 * solution of the wall operator on the coarsest level for its rhs, then the
 * coarse boundary fill; -2 when the direct solve is not in use. */
int fc_lvl_coarse_solve(fc_pson *h, int prec);

/* ---- request server ------------------------------------------------------
 * New: the reference has no server.  It runs solve_poisson (CORE_POISSON.F90:
 * 83-200) for many independent masked grids at once.  Requests are admitted
 * in arrival order while a slot is free and the device memory pool can hold
 * their hierarchy; every fc_serve_step runs one accelerated outer iteration
 * (the loop body of solve_poisson with Anderson mixing, as fc_pson_solve with
 * outer_rule FC_OUTER_RELRES) for all admitted requests together, one phase
 * at a time, each phase one kernel launch over all of them.  A request
 * leaves as soon as it meets its own stop test, and waiting requests join at
 * the next step.  Each request computes exactly what fc_pson_solve computes
 * for it alone with the same parameters (its kernels see the same block
 * grid and its reductions run in the same order), so its result does not
 * depend on what else is in the batch.
 *
 * Return codes as above, plus -6 = the request does not fit the pool. */
typedef struct fc_serve fc_serve;

typedef struct {
    size_t pool_bytes;   /* device memory for resident requests           */
    int max_active;      /* requests solved together, at most 256         */
    int admit_threads;   /* host threads building admitted requests' handles
                            (0 = 8)                                        */
    int reserved[6];
} fc_serve_config;

/* One request: the arguments of fc_pson_create / _set_masks / _set_gcm /
 * _set_bc / _solve for one grid.  The params must describe an outer loop
 * with no host decision inside an iteration: outer_rule FC_OUTER_RELRES,
 * restol_pson <= 0, iterInter = 0, fp32_cycles <= 1, mix_GC_form = 0.
 * pgrad[f] may be NULL (zero plane).  The server starts from pPrime = 0 and
 * p_prev = 0 and writes the answer into pPrime.  Every pointer must stay
 * valid and unchanged until fc_serve_poll reports the request. */
typedef struct {
    fc_grid grid;
    fc_pson_params params;
    const int8_t *iblank, *ghost, *dead;
    int nGhost;
    const int *iGhost, *jGhost, *kGhost, *iCellIndex, *jCellIndex, *kCellIndex;
    const double *coeffGCMN, *alphaGhost, *dpdnBodyIntercept;
    int nDead;
    const int *iDead, *jDead, *kDead, *iDeadCellIndex, *jDeadCellIndex, *kDeadCellIndex;
    const double *coeffGCMDeadN, *probeLengthDead, *dpdnInterceptDead, *xyzDCIP;
    const double *pgrad[6];
    const double *div;
    double *pPrime;
} fc_serve_request;

typedef struct {
    int active;                /* requests in the batch                        */
    int queued;                /* submitted, waiting for a slot or memory      */
    int finished;              /* done, not yet polled                         */
    int launches;              /* kernel launches of the last step             */
    long long steps;           /* steps run so far                             */
    size_t pool_used, pool_capacity;
} fc_serve_status;

int fc_serve_create(const fc_serve_config *cfg, fc_serve **out);
/* Waits for the device, releases every request (unpolled results are
 * dropped) and the pool. */
int fc_serve_destroy(fc_serve *s);
/* Queues a request; *id numbers requests from 1 in submission order. */
int fc_serve_submit(fc_serve *s, const fc_serve_request *rq, long long *id);
/* Admits what fits, runs one outer iteration of every admitted request and
 * retires the requests that finished (answer copied into their pPrime). */
int fc_serve_step(fc_serve *s, fc_serve_status *status);
/* Pops one finished request: returns 1 and its id, its return code (0, 1 or
 * <0 as fc_pson_solve) and stats (time_ms = submission to retirement, host
 * clock), or 0 when none is waiting. */
int fc_serve_poll(fc_serve *s, long long *id, int *rc, fc_pson_stats *stats);

#ifdef __cplusplus
}
#endif
#endif /* FC_POISSON_H */
