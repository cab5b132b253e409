/*
 * fc_geometry.h -- C-ABI of the MI355X (gfx950) immersed-boundary geometry
 * refresh of FEICAR3D: what one time step of a moving triangulated body
 * recomputes before the flow solve (CORE_TIME_STEP.F90:100-118):
 *   calculate_arclength_norm_ds   BOUNDARY_SET_ARCLENGTH_NORM.F90:2-189
 *   set_iblank_body_fast          BOUNDARY_IBLANK_FAST.F90:2-854
 *     calc_bodyIntercept_Unstruc  BOUNDARY_IBLANK_FAST.F90:857-1216
 *   GCM_set_internal_boundary     GCM_SET_BOUNDARY.F90:2-604
 *     smooth_iblank               GCM_SET_BOUNDARY.F90:1618-1697
 *     GCM_define_ghostcells       GCM_SET_BOUNDARY.F90:800-876
 *     GCM_Calc_VanMatrixDN        GCM_SET_BOUNDARY.F90:607-797
 *     GCM_set_deadcells           GCM_SET_BOUNDARY.F90:880-1491
 * Bodies are the reference's unstructured surfaces (canonical_body_type
 * UNSTRUCTURED_SURFACE): marker coordinates xBodyMarker(1:nPts) ... and the
 * connectivity triElemNeig(1:3, 1:nElem) with 1-based marker numbers, both
 * as the reference stores them (AMODULES.F90:343-392).  Grids, index bases
 * and return codes are those of fc_poisson.h (one flow rank).  Everything
 * stays on the device between calls; the getters copy results to the host.
 */
#ifndef FC_GEOMETRY_H
#define FC_GEOMETRY_H

#include "fc_poisson.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FC_GEOM_MAX_BODIES 8
/* unstruc_surface_type (AMODULES.F90:63-64) */
#define FC_SURFACE_SOLID_BODY 1
#define FC_SURFACE_MEMBRANE   2

typedef struct fc_geom fc_geom;

typedef struct fc_geom_params {
    int nBody;                                 /* bodies, 1..FC_GEOM_MAX_BODIES */
    int unstruc_surface_type[FC_GEOM_MAX_BODIES]; /* SOLID_BODY or MEMBRANE  */
    int body_dim[FC_GEOM_MAX_BODIES];          /* 2: z of intercepts held at the cell's */
    double membrane_tkns;                      /* membrane thickness (BOUNDARY_MARKER_INIT.F90:428) */
    double probeLengthNormalized;              /* ghost-cell image points   */
    double probeLengthNormalizedD;             /* dead-cell image points    */
    int reserved[8];
} fc_geom_params;

typedef struct fc_geom_stats {
    int nGhost, nDead;
    int n_undecided;        /* cells classified by a body intercept        */
    int n_smoothed;         /* cells the smoothing passes turned solid      */
    int smooth_rounds;      /* dependency rounds of the smoothing passes    */
    int n_unselected_side;  /* intercepts whose outside-triangle side test
                               selected no side (the reference reads an
                               unset variable there; side 1 is taken)      */
    double time_ms;         /* device time of the refresh (HIP events)      */
    double scan_ms;         /* of which the closest-marker searches of the
                               undecided, ghost and dead cells (HIP events) */
    double scan_pairs;      /* markers those searches compared (cell-marker
                               distances computed)                          */
    int reserved[4];
} fc_geom_stats;

/* Ghost-cell results (GCM_arrays, AMODULES.F90:427-501), nGhost entries
 * each; a NULL pointer skips that array.  Index arrays are 1-based cell
 * indices as the reference stores them; coefficients are (8, nGhost)
 * column-major. */
typedef struct fc_geom_ghosts {
    int *iGhost, *jGhost, *kGhost;
    int *iCellIndex, *jCellIndex, *kCellIndex;
    int *closestElementGC;
    double *xBodyIntercept, *yBodyIntercept, *zBodyIntercept;
    double *xBodyInterceptNorm, *yBodyInterceptNorm, *zBodyInterceptNorm;
    double *probeLength;
    double *xImagePoint, *yImagePoint, *zImagePoint;
    double *coeffGCMD, *coeffGCMN;
    double *alphaGhost;
} fc_geom_ghosts;

typedef struct fc_geom_dead {
    int *iDead, *jDead, *kDead;
    int *iDeadCellIndex, *jDeadCellIndex, *kDeadCellIndex;
    int *closestElementDead;
    double *xBodyInterceptDead, *yBodyInterceptDead, *zBodyInterceptDead;
    double *xNormInterceptDead, *yNormInterceptDead, *zNormInterceptDead;
    double *probeLengthDead;
    double *coeffGCMDeadD, *coeffGCMDeadN;
} fc_geom_dead;

/* Surface quantities of one body (nElem / nPts entries). */
typedef struct fc_geom_surface {
    double *triElemNormX, *triElemNormY, *triElemNormZ, *triElemArea;
    double *triElemCentX, *triElemCentY, *triElemCentZ;
    double *xNormBodyMarker, *yNormBodyMarker, *zNormBodyMarker;
} fc_geom_surface;

/* Marker forces of one body after compute_marker_stress (nPts entries
 * each; AMODULES.F90:393-405). */
typedef struct fc_geom_forces {
    double *xMarkerForce, *yMarkerForce, *zMarkerForce;
    double *xMarkerShear, *yMarkerShear, *zMarkerShear;
    double *MarkerArea;
    double *xMarkerStress, *yMarkerStress, *zMarkerStress;
} fc_geom_forces;

/* Grid metrics of CORE_GRID.F90:106-150 on the device; no bodies yet. */
int fc_geom_create(const fc_grid *g, const fc_geom_params *p, fc_geom **out);
int fc_geom_destroy(fc_geom *h);
/* Body ibody (1-based): markers and connectivity (BOUNDARY_MARKER_INIT.F90
 * reads them; triElemNeig holds 1-based marker numbers, oriented as the
 * reference leaves them after determine_norm_dir_unstruc, normals into the
 * body).  Builds the marker -> element incidence once. */
int fc_geom_set_body(fc_geom *h, int ibody, int nPts, int nElem, const double *x, const double *y,
                     const double *z, const int *triElemNeig);
/* New marker positions of a moving body (move_boundary, BOUNDARY_MOVE.F90);
 * same connectivity. */
int fc_geom_move_body(fc_geom *h, int ibody, const double *x, const double *y, const double *z);
/* calculate_arclength_norm_ds, set_iblank_body_fast and
 * GCM_set_internal_boundary for the current marker positions, on the
 * device.  Returns -5 when an image point has no host cell (the reference
 * prints and dumps) or a dead cell has no body. */
int fc_geom_refresh(fc_geom *h, fc_geom_stats *st);
int fc_geom_get_counts(fc_geom *h, int *nGhost, int *nDead);
/* iblank, ghostCellMark, dead_cell, bodyNum, ps1, ps2: int8 arrays
 * (0:nx+1, 0:ny+1, 0:nz+1); NULL skips one. */
int fc_geom_get_masks(fc_geom *h, int8_t *iblank, int8_t *ghostCellMark, int8_t *dead_cell,
                      int8_t *bodyNum, int8_t *ps1, int8_t *ps2);
int fc_geom_get_ghosts(fc_geom *h, const fc_geom_ghosts *g);
int fc_geom_get_dead(fc_geom *h, const fc_geom_dead *d);
int fc_geom_get_surface(fc_geom *h, int ibody, const fc_geom_surface *f);
/* xyzDCIP(0:nx+1, 0:ny+1, 0:nz+1, 0:2) (image points of ghost and dead cells) */
int fc_geom_get_dcip(fc_geom *h, double *xyzDCIP);
/* xyzDCIP at the dead cells only, in dead-list order (x, y, z per cell:
 * 3*nDead values), the form fc_pson_set_gcm takes; no counterpart in the
 * reference (it reads the grid array at iDead, jDead, kDead). */
int fc_geom_get_dead_ip(fc_geom *h, double *xyz);

/* Marker velocities of body ibody: u/v/wBodyMarker and the previous step's
 * u/v/wBodyMarkerOld (BOUNDARY_MARKER_VEL.F90 sets them), nPts each. */
int fc_geom_set_marker_velocity(fc_geom *h, int ibody, const double *u, const double *v, const double *w,
                                const double *uOld, const double *vOld, const double *wOld);
/* GCM_SetBodyInterceptValues (GCM_BODYINTERCEPT_VALUES.F90:22-178, with
 * GCM_calc_BIVelocity_Unstruc :198-388) after fc_geom_refresh: the body
 * velocity at every ghost and dead cell's body intercept, linear over the
 * closest element, and dp/dn = -(du/dt . n) there when ntime > 1. */
int fc_geom_body_intercept_values(fc_geom *h, double dt, int ntime);
/* u/v/wBodyIntercept, dpdnBodyIntercept (nGhost), u/v/wBodyInterceptDead,
 * dpdnInterceptDead (nDead); NULL skips one. */
int fc_geom_get_bi_values(fc_geom *h, double *uG, double *vG, double *wG, double *dpdnG, double *uD, double *vD,
                          double *wD, double *dpdnD);
/* compute_marker_stress (UTIL_DRAG_LIFT.F90:29-342, with
 * marker_interpolation :375-686) after fc_geom_body_intercept_values: the
 * pressure and viscous force of every surface element from the flow fields
 * p, u, v, w ((0:nx+1, 0:ny+1, 0:nz+1), host arrays), summed per body into
 * totals[9*(ibody-1) .. +8] = cxt cyt czt cmxt cmyt cmzt cxs cys czs (drag,
 * lift, side force, moments, viscous parts), and per marker (one third of
 * each element's force and area).  The FSI blending with the previous
 * step's marker forces (fsi_filter_f) is left to the caller. */
int fc_geom_marker_stress(fc_geom *h, const double *p, const double *u, const double *v, const double *w,
                          double reinv, double *totals);
/* The same with p, u, v, w already on the device (dense (nx+2)(ny+2)(nz+2)
 * FP64 arrays on the handle's device, Fortran order). */
int fc_geom_marker_stress_device(fc_geom *h, const double *p, const double *u, const double *v, const double *w,
                                 double reinv, double *totals);
/* The same on the resident fields of a flow handle of the same grid and
 * device (the p of update_pressure, u, v, w of fc_project_resident), copied
 * device to device into this handle. */
int fc_geom_marker_stress_resident(fc_geom *h, fc_pson *flow, double reinv, double *totals);
int fc_geom_get_marker_forces(fc_geom *h, int ibody, const fc_geom_forces *f);
/* Device time (HIP events on the handle's stream) of the last
 * fc_geom_body_intercept_values (bi_ms) and of the last marker-stress call:
 * the per-element interpolation (interp_ms) and the body totals and marker
 * forces (sum_ms).  No counterpart in the reference (a measurement hook). */
int fc_geom_get_stress_timing(fc_geom *h, double *bi_ms, double *interp_ms, double *sum_ms);

#ifdef __cplusplus
}
#endif
#endif /* FC_GEOMETRY_H */
