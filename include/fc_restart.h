/*
 * fc_restart.h -- the restart file of FEICAR3D, written and read from the
 * device-resident state of one flow rank:
 *   write_restart  MAIN.F90:1049-1122  (called every nrestart steps and at
 *                                      the end, CORE_TIME_STEP.F90:484-487)
 *   read_restart   MAIN.F90:983-1047   (nread == 1, MAIN.F90:935-938)
 * The file is the reference's 'restart_NNNN.dat' (NNNN = lProc, I4.4): a
 * Fortran unformatted sequential file as the reference's compiler writes it
 * (little-endian, 4-byte record markers; a logical record longer than
 * 2147483639 bytes is split into subrecords, the leading marker negative
 * when another subrecord follows, the trailing one negative when one
 * precedes), holding
 *   record 1: time (real*8), ntime (integer*4)
 *   record 2: u, v, w, p, face_u, face_v, face_w, pPrime
 *             (0:nx+1, 0:ny+1, 0:nz+1) real*8 each;
 *             bcxu, bcxv, bcxw, bcxp (0:1, 0:ny+1, 0:nz+1);
 *             bcyu, bcyv, bcyw, bcyp (0:nx+1, 0:1, 0:nz+1);
 *             bczu, bczv, bczw, bczp (0:nx+1, 0:ny+1, 0:1);
 *             pgradx1, pgradx2 (0:ny+1, 0:nz+1); pgrady1, pgrady2
 *             (0:nx+1, 0:nz+1); pgradz1, pgradz2 (0:nx+1, 0:ny+1);
 *             iblank, ghostCellMark (0:nx+1, 0:ny+1, 0:nz+1) integer*1
 *   one record per body: x/y/zBodyMarker, u/v/wBodyMarker,
 *             x/y/zMarkerForce (1:nPts) real*8; triElemNeig(1:3, 1:nElem)
 *             integer*4; xcent, ycent, zcent, angx, angy, angz, vxcent,
 *             vycent, vzcent, angvx, angvy, angvz real*8.
 * Array bounds are those of one rank holding the whole domain
 * (CORE_MEMORY_ALLOCATE.F90:33-86 with yb1 = zb1 = 0, yb2 = ny+1,
 * zb2 = nz+1).  Return codes: 0 = done, -1 = bad argument, -3 = HIP error,
 * -4 = the file's records do not have the sizes these dimensions give,
 * -7 = the file could not be opened, written or read; the text is
 * fc_pson_last_error().
 */
#ifndef FC_RESTART_H
#define FC_RESTART_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The flow state of record 2 (AMODULES.F90 flow_arrays, boundary_arrays).
 * With on_device = 1 the 26 real*8 and 2 integer*1 arrays are device
 * pointers (for instance the resident buffers of the flow path); they are
 * staged through pinned host memory in chunks, the copies of one chunk
 * overlapping the file I/O of the previous one.  With on_device = 0 they
 * are host arrays and no HIP call is made. */
typedef struct fc_restart_flow {
    double *u, *v, *w, *p, *face_u, *face_v, *face_w, *pPrime;
    double *bcxu, *bcxv, *bcxw, *bcxp;
    double *bcyu, *bcyv, *bcyw, *bcyp;
    double *bczu, *bczv, *bczw, *bczp;
    double *pgradx1, *pgradx2, *pgrady1, *pgrady2, *pgradz1, *pgradz2;
    int8_t *iblank, *ghostCellMark;
    int on_device;
} fc_restart_flow;

/* One body's record (host arrays; AMODULES.F90 boundary_arrays,
 * unstructured_surface_arrays). */
typedef struct fc_restart_body {
    int nPts, nElem;
    double *xBodyMarker, *yBodyMarker, *zBodyMarker;
    double *uBodyMarker, *vBodyMarker, *wBodyMarker;
    double *xMarkerForce, *yMarkerForce, *zMarkerForce;
    int *triElemNeig;                      /* (3, nElem), 1-based markers */
    double cent[3];                        /* xcent, ycent, zcent        */
    double ang[3];                         /* angx, angy, angz           */
    double vcent[3];                       /* vxcent, vycent, vzcent     */
    double angv[3];                        /* angvx, angvy, angvz        */
} fc_restart_body;

/* 'restart_'//NNNN//'.dat' with NNNN = lProc in I4.4 (MAIN.F90:1068-1069);
 * out holds at least 17 bytes. */
int fc_restart_name(int lProc, char *out, int out_len);

/* write_restart: time and ntime, the flow state, then nBody body records. */
int fc_restart_write(const char *path, int nx, int ny, int nz, double time, int ntime, const fc_restart_flow *f,
                     int nBody, const fc_restart_body *bodies);

/* read_restart: fills time, ntime_start (ntime as written), the flow arrays
 * and the bodies' arrays and scalars (nPts and nElem given by the caller,
 * as the reference takes them from its body input).  The reference then
 * copies x/y/zMarkerForce into x/y/zMarkerForceOld; that is the caller's. */
int fc_restart_read(const char *path, int nx, int ny, int nz, double *time, int *ntime_start,
                    const fc_restart_flow *f, int nBody, fc_restart_body *bodies);

/* Largest user-data length of one subrecord, 2147483639 by default (the
 * 4-byte record markers' limit).  A smaller value (at least 8) makes small
 * files exercise the subrecord framing, as gfortran's
 * -fmax-subrecord-length does; no counterpart in the reference.  Returns
 * the previous value. */
long long fc_restart_set_max_subrecord(long long bytes);

#ifdef __cplusplus
}
#endif
#endif /* FC_RESTART_H */
