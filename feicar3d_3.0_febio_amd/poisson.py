"""Host-side mirror of the reference's pressure-Poisson interface over the HIP
library (ctypes over include/fc_poisson.h).

Method names follow the reference subroutines: solve_poisson
(CORE_POISSON.F90:83), mg_solve (CORE_MG.F90:1572), ghostcell_pressure
(GCM_GHOSTCELL_UPDATE.F90:361), calc_residual / calc_residual2
(CORE_POISSON.F90:460, 679), rhs_poisson (CORE_POISSON.F90:9) and the level
operations itsolv_mg / residual_mg / mg_inject / mg_interp.  Non-zero return
codes follow the reference: 1 = an iteration cap was hit (the reference
prints and continues); errors raise FcError.  There is no CPU fallback: the
HIP library must be built and a GPU present.
"""
import ctypes as C
import os

import numpy as np

from . import capi


class FcError(RuntimeError):
    pass


_lib = None


def library():
    """Loads csrc/_build/libfc_poisson.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(capi.LIB_PATH):
        raise FcError("HIP library not built: %s (run __graft_entry__.build())" % capi.LIB_PATH)
    L = C.CDLL(capi.LIB_PATH)
    vp = C.c_void_p
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    bp = C.POINTER(C.c_int8)
    i = C.c_int
    L.fc_pson_version.restype = C.c_char_p
    L.fc_pson_last_error.restype = C.c_char_p
    L.fc_pson_default_params.argtypes = [C.POINTER(capi.FcParams)]
    L.fc_pson_create.argtypes = [C.POINTER(capi.FcGrid), C.POINTER(capi.FcParams), C.POINTER(vp)]
    L.fc_pson_set_params.argtypes = [vp, C.POINTER(capi.FcParams)]
    L.fc_pson_destroy.argtypes = [vp]
    L.fc_pson_levels.argtypes = [vp, ip, ip, ip, ip]
    L.fc_pson_set_masks.argtypes = [vp, bp, bp, bp]
    L.fc_pson_set_gcm.argtypes = [vp, i, ip, ip, ip, ip, ip, ip, dp, dp, dp,
                                  i, ip, ip, ip, ip, ip, ip, dp, dp, dp, dp]
    L.fc_pson_set_bc.argtypes = [vp, dp, dp, dp, dp, dp, dp, C.c_double]
    L.fc_pson_get_bc.argtypes = [vp, dp, dp, dp, dp, dp, dp]
    L.fc_pson_solve.argtypes = [vp, dp, dp, dp, dp, C.POINTER(capi.FcStats)]
    L.fc_pson_mg_solve.argtypes = [vp, dp, dp, C.POINTER(capi.FcStats)]
    L.fc_pson_itsolv.argtypes = [vp, dp, dp]
    L.fc_pson_ghostcell_pressure.argtypes = [vp, dp, dp, dp, dp, C.POINTER(capi.FcStats)]
    L.fc_pson_calc_residual.argtypes = [vp, dp, dp, dp, ip]
    L.fc_pson_calc_residual2.argtypes = [vp, dp, dp, dp, ip]
    L.fc_pson_rhs.argtypes = [vp, dp, dp, dp, C.c_double, dp, dp, dp]
    L.fc_pson_rhs_noalloc.argtypes = [vp, vp, vp, vp, C.c_double, vp, vp, vp]
    L.fc_tdma_batched.argtypes = [i, i, dp, dp, dp, dp, dp, i]
    L.fc_pson_upload.argtypes = [vp, dp, dp, dp]
    L.fc_pson_solve_resident.argtypes = [vp, C.POINTER(capi.FcStats)]
    L.fc_pson_download.argtypes = [vp, dp, dp]
    L.fc_pson_device_bytes.restype = C.c_size_t
    L.fc_pson_device_bytes.argtypes = [vp]
    L.fc_pson_reset.argtypes = [vp]
    L.fc_pson_timing.argtypes = [vp, i]
    L.fc_pson_timing_read.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_long)]
    L.fc_vel_set_bc.argtypes = [vp, dp, dp, dp]
    L.fc_vel_get_bc.argtypes = [vp, dp, dp, dp]
    L.fc_pson_get_outflow.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.fc_vel_set_outer.argtypes = [vp, dp, dp, C.c_double, C.c_double]
    L.fc_outer_velocity_bc.argtypes = [vp, C.c_double]
    L.fc_outer_ghost_velocity.argtypes = [vp]
    L.fc_vel_set_outer_all.argtypes = [vp, C.POINTER(dp)]
    L.fc_vel_get_outer_all.argtypes = [vp, C.POINTER(dp)]
    L.fc_set_gcm_vel.argtypes = [vp, dp, dp, dp, dp, C.c_double]
    L.fc_ghostcell_velocity.argtypes = [vp, i, C.c_double, ip, C.POINTER(C.c_double)]
    L.fc_solve_ad_resident.argtypes = [vp, C.c_double, i, C.c_double, C.c_double, ip, C.POINTER(C.c_double), ip]
    L.fc_solve_ad.argtypes = [vp, dp, dp, dp, dp, dp, dp, dp, dp, dp, C.c_double, i, C.c_double, C.c_double, ip,
                              C.POINTER(C.c_double), ip]
    L.fc_correct_vel.argtypes = [vp, dp, C.c_double, dp, dp, dp]
    L.fc_correct_ghostcell_velocity.argtypes = [vp, dp, C.c_double, dp, dp, dp]
    L.fc_face_vel.argtypes = [vp, dp, dp, dp, dp, dp, dp]
    L.fc_update_pressure.argtypes = [vp, dp, dp]
    L.fc_vel_upload.argtypes = [vp, dp, dp, dp, dp, dp, dp]
    L.fc_vel_download.argtypes = [vp, dp, dp, dp, dp, dp, dp]
    L.fc_project_resident.argtypes = [vp, C.c_double, C.c_double]
    L.fc_face_vel_resident.argtypes = [vp]
    L.fc_pson_rhs_resident.argtypes = [vp, C.c_double, dp, dp]
    L.fc_pson_resident_copy.argtypes = [vp, i, vp]
    L.fc_project_timing_read.argtypes = [vp, dp]
    L.fc_pson_download_p.argtypes = [vp, dp]
    L.fc_set_solve_ad.argtypes = [vp, C.c_double, C.c_double, i, C.c_double]
    L.fc_set_visc_tot.argtypes = [vp, dp]
    L.fc_ad_download.argtypes = [vp, dp, dp, dp]
    L.fc_rhs_advec_diff_resident.argtypes = [vp, C.c_double, C.c_double]
    L.fc_rhs_advec_diff.argtypes = [vp, dp, dp, dp, dp, dp, dp, C.c_double, C.c_double, dp, dp, dp]
    L.fc_itsolv_ad.argtypes = [vp, dp, dp, dp, dp, dp, i]
    L.fc_calc_residual_ad.argtypes = [vp, dp, dp, dp, dp, dp, dp, ip, dp]
    L.fc_ad_upload.argtypes = [vp, dp, dp, dp]
    L.fc_itsolv_ad_resident.argtypes = [vp, i, i]
    L.fc_calc_residual_ad_resident.argtypes = [vp, i, dp, ip, dp]
    L.fc_ad_timing_read.argtypes = [vp, dp]
    L.fc_lvl_write.argtypes = [vp, i, i, vp]
    L.fc_lvl_read.argtypes = [vp, i, i, vp]
    L.fc_lvl_smooth.argtypes = [vp, i, i]
    L.fc_lvl_sweep.argtypes = [vp, i, i, i, i]
    L.fc_lvl_coarse_solve.argtypes = [vp, i]
    for f in ("fc_lvl_bc", "fc_lvl_residual", "fc_lvl_restrict", "fc_lvl_prolong"):
        getattr(L, f).argtypes = [vp, i, i]
    L.fc_serve_create.argtypes = [C.POINTER(capi.FcServeConfig), C.POINTER(vp)]
    L.fc_serve_destroy.argtypes = [vp]
    L.fc_serve_submit.argtypes = [vp, C.POINTER(capi.FcServeRequest), C.POINTER(C.c_longlong)]
    L.fc_serve_step.argtypes = [vp, C.POINTER(capi.FcServeStatus)]
    L.fc_serve_poll.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_int), C.POINTER(capi.FcStats)]
    _lib = L
    return L


def _check(rc, what):
    if rc < 0:
        raise FcError("%s failed (%d): %s" % (what, rc, library().fc_pson_last_error().decode()))
    return rc


def tdma_batched(a, b, c, r, prec=capi.PREC_FP64):
    """Batched tdma (CORE_TDMA.F90:10-34): arrays of shape (nsys, n)."""
    L = library()
    a, b, c, r = [np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c, r)]
    nsys, n = r.shape
    u = np.zeros_like(r)
    _check(L.fc_tdma_batched(nsys, n, capi.dptr(a.ravel()), capi.dptr(b.ravel()), capi.dptr(c.ravel()),
                             capi.dptr(r.ravel()), capi.dptr(u.ravel()), prec), "fc_tdma_batched")
    return u


class PoissonSolver:
    """One handle: the device-resident multigrid hierarchy of one grid."""

    def __init__(self, x, y, z, params=None):
        self.L = library()
        self.x, self.y, self.z = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z)]
        self.nodes = (self.x.size - 2, self.y.size - 2, self.z.size - 2)
        self.params = params if params is not None else capi.default_params()
        g = capi.FcGrid(self.nodes[0], self.nodes[1], self.nodes[2],
                        capi.dptr(self.x), capi.dptr(self.y), capi.dptr(self.z))
        h = C.c_void_p()
        _check(self.L.fc_pson_create(C.byref(g), C.byref(self.params), C.byref(h)), "fc_pson_create")
        self.h = h
        nlev = C.c_int()
        a, b, c = (C.c_int * 10)(), (C.c_int * 10)(), (C.c_int * 10)()
        self.L.fc_pson_levels(self.h, C.byref(nlev), a, b, c)
        self.nlev = nlev.value
        self.level_nodes = [(a[k], b[k], c[k]) for k in range(self.nlev)]
        self.nG = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.fc_pson_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape(self):
        nx, ny, nz = self.nodes
        return (nx + 2, ny + 2, nz + 2)

    def set_params(self, params):
        self.params = params
        _check(self.L.fc_pson_set_params(self.h, C.byref(params)), "fc_pson_set_params")

    def device_bytes(self):
        return int(self.L.fc_pson_device_bytes(self.h))

    def set_masks(self, iblank, ghost, dead=None):
        ib = np.asfortranarray(iblank, dtype=np.int8)
        gh = np.asfortranarray(ghost, dtype=np.int8)
        de = None if dead is None else np.asfortranarray(dead, dtype=np.int8)
        sh = self.shape
        _check(self.L.fc_pson_set_masks(self.h, capi.bptr(ib, sh, "iblank"), capi.bptr(gh, sh, "ghostCellMark"),
                                        capi.bptr(de, sh, "dead_cell")), "fc_pson_set_masks")

    def set_gcm(self, g):
        g = {k: np.ascontiguousarray(v) for k, v in g.items()}
        self._g = g
        self.nG = g["iGhost"].size
        _check(self.L.fc_pson_set_gcm(
            self.h, g["iGhost"].size, capi.iptr(g["iGhost"]), capi.iptr(g["jGhost"]), capi.iptr(g["kGhost"]),
            capi.iptr(g["iCellIndex"]), capi.iptr(g["jCellIndex"]), capi.iptr(g["kCellIndex"]),
            capi.dptr(g["coeffGCMN"]), capi.dptr(g["alphaGhost"]), capi.dptr(g["dpdnBodyIntercept"]),
            g["iDead"].size, capi.iptr(g["iDead"]), capi.iptr(g["jDead"]), capi.iptr(g["kDead"]),
            capi.iptr(g["iDeadCellIndex"]), capi.iptr(g["jDeadCellIndex"]), capi.iptr(g["kDeadCellIndex"]),
            capi.dptr(g["coeffGCMDeadN"]), capi.dptr(g["probeLengthDead"]), capi.dptr(g["dpdnInterceptDead"]),
            capi.dptr(g["xyzDCIP"])), "fc_pson_set_gcm")

    def set_workload(self, w):
        self.set_masks(w.iblank, w.ghost, w.dead)
        self.set_gcm(w.gcm)

    def set_bc(self, pgx1=None, pgx2=None, pgy1=None, pgy2=None, pgz1=None, pgz2=None, outflow_area=1.0):
        arrs = [None if a is None else np.asfortranarray(a, dtype=np.float64) for a in (pgx1, pgx2, pgy1, pgy2, pgz1, pgz2)]
        self._bc = arrs
        nx, ny, nz = self.nodes
        shp = [(ny + 2, nz + 2)] * 2 + [(nx + 2, nz + 2)] * 2 + [(nx + 2, ny + 2)] * 2
        names = ("pgradx1", "pgradx2", "pgrady1", "pgrady2", "pgradz1", "pgradz2")
        _check(self.L.fc_pson_set_bc(self.h, *[capi.plane_ptr(a, s, n) for a, s, n in zip(arrs, shp, names)],
                                     outflow_area), "fc_pson_set_bc")

    def get_bc(self):
        """the six pgrad planes (GCM_enforce_p_compatibility corrects the
        zero-gradient ones during solve_poisson)"""
        nx, ny, nz = self.nodes
        shp = [(ny + 2, nz + 2)] * 2 + [(nx + 2, nz + 2)] * 2 + [(nx + 2, ny + 2)] * 2
        names = ("pgradx1", "pgradx2", "pgrady1", "pgrady2", "pgradz1", "pgradz2")
        out = [np.zeros(s, order="F") for s in shp]
        _check(self.L.fc_pson_get_bc(self.h, *[capi.plane_ptr(a, s, n) for a, s, n in zip(out, shp, names)]),
               "fc_pson_get_bc")
        return out

    def get_vel_bc(self):
        """bcxu, bcyv, bczw as rhs_poisson left them"""
        nx, ny, nz = self.nodes
        out = [np.zeros(s, order="F") for s in ((2, ny + 2, nz + 2), (nx + 2, 2, nz + 2), (nx + 2, ny + 2, 2))]
        _check(self.L.fc_vel_get_bc(self.h, *[capi.dptr(a) for a in out]), "fc_vel_get_bc")
        return out

    # the reference's subroutines ------------------------------------------
    def solve_poisson(self, pPrime, div, p_prev=None):
        """solve_poisson(): pPrime updated in place; returns (rc, stats, res_laplacian)."""
        st = capi.FcStats()
        reslap = np.zeros(max(self.nG, 1))
        rc = _check(self.L.fc_pson_solve(self.h, self._fp(pPrime, "pPrime"), self._fp(div, "div"),
                                         self._fp(p_prev, "p_prev"), capi.dptr(reslap), C.byref(st)), "solve_poisson")
        return rc, st, reslap[:self.nG]

    def mg_solve(self, var, rr):
        st = capi.FcStats()
        rc = _check(self.L.fc_pson_mg_solve(self.h, self._fp(var, "var"), self._fp(rr, "rr"), C.byref(st)),
                    "mg_solve")
        return rc, st

    def itsolv(self, var, r):
        """itsolv (CORE_POISSON.F90:205-456): one plain line-SOR pass, var
        updated in place"""
        _check(self.L.fc_pson_itsolv(self.h, self._fp(var, "var"), self._fp(r, "r")), "itsolv")

    def ghostcell_pressure(self, pres, r, p_prev=None):
        st = capi.FcStats()
        reslap = np.zeros(max(self.nG, 1))
        rc = _check(self.L.fc_pson_ghostcell_pressure(self.h, self._fp(pres, "pres"), self._fp(r, "r"),
                                                      self._fp(p_prev, "p_prev"), capi.dptr(reslap), C.byref(st)),
                    "GCM_ghostcell_pressure")
        return rc, st, reslap[:self.nG]

    def calc_residual(self, var, r, with_loc=False):
        """calc_residual(var, r, resm, loc): resm, and with with_loc
        (resm, (i, j, k)) -- the 1-based cell of the largest |residual|"""
        out = C.c_double()
        loc = (C.c_int * 3)()
        _check(self.L.fc_pson_calc_residual(self.h, self._fp(var, "var"), self._fp(r, "r"), C.byref(out), loc),
               "calc_residual")
        return (out.value, tuple(loc)) if with_loc else out.value

    def calc_residual2(self, var, p_temp, with_loc=False):
        """calc_residual2: max |var - p_temp| over iblank = 0 (and its cell)"""
        out = C.c_double()
        loc = (C.c_int * 3)()
        _check(self.L.fc_pson_calc_residual2(self.h, self._fp(var, "var"), self._fp(p_temp, "p_temp"),
                                             C.byref(out), loc), "calc_residual2")
        return (out.value, tuple(loc)) if with_loc else out.value

    def rhs_poisson(self, fu, fv, fw, dtinv=1.0):
        """div, sumd, resDiv.  With zero-gradient faces fu, fv, fw (Fortran
        float64 fields) are corrected in place (enforce_global_mass_consv)."""
        div = np.zeros(self.shape, order="F")
        s = C.c_double()
        rd = C.c_double()
        _check(self.L.fc_pson_rhs(self.h, self._fp(fu, "face_u"), self._fp(fv, "face_v"), self._fp(fw, "face_w"),
                                  dtinv, capi.dptr(div), C.byref(s), C.byref(rd)), "rhs_poisson")
        return div, s.value, rd.value

    def get_outflow(self):
        """(outflow_area, correction_vel) of the last rhs_poisson"""
        a, c = C.c_double(), C.c_double()
        _check(self.L.fc_pson_get_outflow(self.h, C.byref(a), C.byref(c)), "get_outflow")
        return a.value, c.value

    def rhs_poisson_noalloc(self, fu_d, fv_d, fw_d, div_d, sums_d, dtinv=1.0, stream=None):
        """fc_pson_rhs_noalloc on device arrays given by address (unpadded
        Fortran layout; sums_d: 2 doubles); queued on stream (None: the
        handle's stream), not synchronised; returns the entry point's code
        (raises on errors)"""
        return _check(self.L.fc_pson_rhs_noalloc(self.h, fu_d, fv_d, fw_d, dtinv, div_d, sums_d, stream),
                      "rhs_poisson_noalloc")

    # device-resident path ---------------------------------------------------
    def upload(self, pPrime=None, div=None, p_prev=None):
        _check(self.L.fc_pson_upload(self.h, self._fp(pPrime, "pPrime"), self._fp(div, "div"),
                                     self._fp(p_prev, "p_prev")), "upload")

    def solve_resident(self):
        st = capi.FcStats()
        rc = _check(self.L.fc_pson_solve_resident(self.h, C.byref(st)), "solve_resident")
        return rc, st

    def reset(self):
        _check(self.L.fc_pson_reset(self.h), "reset")

    def timing(self, enable=True):
        _check(self.L.fc_pson_timing(self.h, 1 if enable else 0), "timing")

    def timing_read(self):
        ms = C.c_double()
        n = C.c_long()
        _check(self.L.fc_pson_timing_read(self.h, C.byref(ms), C.byref(n)), "timing_read")
        return ms.value, n.value

    def download(self):
        p = np.zeros(self.shape, order="F")
        _check(self.L.fc_pson_download(self.h, capi.dptr(p), None), "download")
        return p

    # velocity projection (CORE_TIME_STEP.F90:211-222) ------------------------
    def _fp(self, a, name):
        """pointer to a full field (0:nx+1, 0:ny+1, 0:nz+1), or None"""
        if a is None:
            return None
        try:
            return capi.dptr(a, self.shape, name)
        except ValueError as e:
            raise FcError(str(e)) from None

    def _field(self, a, name):
        if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.shape == self.shape
                and a.flags.f_contiguous):
            raise FcError("%s must be a Fortran-ordered float64 array of shape %s" % (name, self.shape))
        return a

    def set_vel_bc(self, bcxu=None, bcyv=None, bczw=None):
        """bcxu(0:1,0:ny+1,0:nz+1), bcyv(0:nx+1,0:1,0:nz+1), bczw(0:nx+1,0:ny+1,0:1)."""
        arrs = [None if a is None else np.asfortranarray(a, dtype=np.float64) for a in (bcxu, bcyv, bczw)]
        _check(self.L.fc_vel_set_bc(self.h, *[capi.dptr(a) for a in arrs]), "fc_vel_set_bc")

    def correct_vel(self, pPrime, dt, u, v, w):
        """correct_vel() (CORE_FRAC_STEP_ROUTINES.F90:134-210): u, v, w in place."""
        for a, n in ((pPrime, "pPrime"), (u, "u"), (v, "v"), (w, "w")):
            self._field(a, n)
        _check(self.L.fc_correct_vel(self.h, capi.dptr(pPrime), dt, capi.dptr(u), capi.dptr(v), capi.dptr(w)),
               "correct_vel")

    def correct_ghostcell_velocity(self, pVar, dt_fac, u, v, w):
        """correct_ghostcell_velocity(pVar) (GCM_GHOSTCELL_UPDATE.F90:789-852): u, v, w in place."""
        for a, n in ((pVar, "pVar"), (u, "u"), (v, "v"), (w, "w")):
            self._field(a, n)
        _check(self.L.fc_correct_ghostcell_velocity(self.h, capi.dptr(pVar), dt_fac, capi.dptr(u), capi.dptr(v),
                                                    capi.dptr(w)), "correct_ghostcell_velocity")

    def face_vel(self, u, v, w, face_u, face_v, face_w):
        """face_vel(u_vel, v_vel, w_vel) (CORE_FRAC_STEP_ROUTINES.F90:6-130): face arrays in place."""
        for a, n in ((u, "u"), (v, "v"), (w, "w"), (face_u, "face_u"), (face_v, "face_v"), (face_w, "face_w")):
            self._field(a, n)
        _check(self.L.fc_face_vel(self.h, capi.dptr(u), capi.dptr(v), capi.dptr(w), capi.dptr(face_u),
                                  capi.dptr(face_v), capi.dptr(face_w)), "face_vel")

    def update_pressure(self, pPrime, p):
        """update_pressure() (CORE_FRAC_STEP_ROUTINES.F90:212-299): pPrime and p in place."""
        self._field(pPrime, "pPrime")
        self._field(p, "p")
        _check(self.L.fc_update_pressure(self.h, capi.dptr(pPrime), capi.dptr(p)), "update_pressure")

    def vel_upload(self, u=None, v=None, w=None, face_u=None, face_v=None, face_w=None):
        arrs = [None if a is None else self._field(a, "velocity") for a in (u, v, w, face_u, face_v, face_w)]
        _check(self.L.fc_vel_upload(self.h, *[capi.dptr(a) for a in arrs]), "vel_upload")

    def vel_download(self):
        out = [np.zeros(self.shape, order="F") for _ in range(6)]
        _check(self.L.fc_vel_download(self.h, *[capi.dptr(a) for a in out]), "vel_download")
        return out

    def project_resident(self, dt, dt_fac=None):
        """correct_vel, correct_ghostcell_velocity, face_vel, update_pressure
        on the resident buffers (CORE_TIME_STEP.F90:211-222)."""
        _check(self.L.fc_project_resident(self.h, dt, dt if dt_fac is None else dt_fac), "project_resident")

    def face_vel_resident(self):
        """face_vel on the resident velocities into the resident face
        velocities"""
        _check(self.L.fc_face_vel_resident(self.h), "face_vel_resident")

    def rhs_resident(self, dtinv):
        """rhs_poisson from the resident face velocities into the resident
        div: (sumd, resDiv)"""
        s, rd = C.c_double(), C.c_double()
        _check(self.L.fc_pson_rhs_resident(self.h, dtinv, C.byref(s), C.byref(rd)), "rhs_resident")
        return s.value, rd.value

    def resident_copy(self, which, dst):
        """one resident field (0..5 u, v, w, face_u, face_v, face_w; 6 p;
        7 pPrime) into dst, a contiguous float64 torch tensor on the device of
        shape (nz+2, ny+2, nx+2)"""
        if (not dst.is_cuda or str(dst.dtype) != "torch.float64" or not dst.is_contiguous()
                or tuple(dst.shape) != self.shape[::-1]):
            raise FcError("resident_copy: contiguous float64 device tensor of shape %s" % (self.shape[::-1],))
        _check(self.L.fc_pson_resident_copy(self.h, which, C.c_void_p(dst.data_ptr())), "resident_copy")

    def project_timing(self):
        ms = (C.c_double * 4)()
        _check(self.L.fc_project_timing_read(self.h, ms), "project_timing_read")
        return list(ms)

    def download_p(self):
        p = np.zeros(self.shape, order="F")
        _check(self.L.fc_pson_download_p(self.h, capi.dptr(p)), "download_p")
        return p

    # advection-diffusion line solves (CORE_SOLVE_AD.F90) -------------------
    def set_solve_ad(self, reinv, dt, crank_nicolson=1, omega_ad=1.0):
        """set_solve_ad() (CORE_SET_SOLVERS.F90:591-676) plus half_dt and omega_ad."""
        _check(self.L.fc_set_solve_ad(self.h, reinv, dt, crank_nicolson, omega_ad), "set_solve_ad")

    def itsolv_ad(self, var, r, face_u, face_v, face_w, nsweep=1):
        """itsolv_ad(var, r) (CORE_SOLVE_AD.F90:131-776), nsweep times; var in place."""
        for a, n in ((var, "var"), (r, "r"), (face_u, "face_u"), (face_v, "face_v"), (face_w, "face_w")):
            self._field(a, n)
        _check(self.L.fc_itsolv_ad(self.h, capi.dptr(var), capi.dptr(r), capi.dptr(face_u), capi.dptr(face_v),
                                   capi.dptr(face_w), nsweep), "itsolv_ad")

    def calc_residual_ad(self, var, r, face_u, face_v, face_w):
        """calc_residual_ad (CORE_SOLVE_AD.F90:780-964): (resm, loc, sum res^2)."""
        for a, n in ((var, "var"), (r, "r"), (face_u, "face_u"), (face_v, "face_v"), (face_w, "face_w")):
            self._field(a, n)
        resm, r2, loc = C.c_double(), C.c_double(), (C.c_int * 3)()
        _check(self.L.fc_calc_residual_ad(self.h, capi.dptr(var), capi.dptr(r), capi.dptr(face_u),
                                          capi.dptr(face_v), capi.dptr(face_w), C.byref(resm), loc, C.byref(r2)),
               "calc_residual_ad")
        return resm.value, tuple(loc), r2.value

    def ad_upload(self, nlu=None, nlv=None, nlw=None):
        arrs = [None if a is None else self._field(a, "rhs") for a in (nlu, nlv, nlw)]
        _check(self.L.fc_ad_upload(self.h, *[capi.dptr(a) for a in arrs]), "ad_upload")

    # solve_ad's loop (CORE_SOLVE_AD.F90:11-127) ------------------------------
    def _outer_shapes(self):
        nx, ny, nz = self.nodes
        return [(2, ny + 2, nz + 2)] * 3 + [(nx + 2, 2, nz + 2)] * 3 + [(nx + 2, ny + 2, 2)] * 3

    def vel_set_outer(self, vals=None, freq=None, y_origin=0.0, y_out=1.0):
        """face values and frequencies ux1, ux2, vx1, vx2, wx1, wx2, uy1, ...,
        wz2 and the channel-profile bounds of set_outer_velocity_bc"""
        a = [None if x is None else np.ascontiguousarray(x, dtype=np.float64) for x in (vals, freq)]
        for x, n in zip(a, ("vals", "freq")):
            if x is not None and x.shape != (18,):
                raise FcError("%s must hold 18 values" % n)
        _check(self.L.fc_vel_set_outer(self.h, capi.dptr(a[0]), capi.dptr(a[1]), y_origin, y_out), "vel_set_outer")

    def outer_velocity_bc(self, time=0.0):
        _check(self.L.fc_outer_velocity_bc(self.h, time), "set_outer_velocity_bc")

    def outer_ghost_velocity(self):
        _check(self.L.fc_outer_ghost_velocity(self.h), "set_outer_ghost_velocity")

    def vel_set_outer_all(self, arrs):
        """the nine outer-face arrays bcxu, bcxv, bcxw, bcyu, ..., bczw (None = zeros)"""
        shp = self._outer_shapes()
        keep = []
        for a, sh in zip(arrs, shp):
            if a is None:
                keep.append(None)
                continue
            a = np.asfortranarray(a, dtype=np.float64)
            if a.shape != sh:
                raise FcError("outer-face array of shape %s, expected %s" % (a.shape, sh))
            keep.append(a)
        ptrs = (C.POINTER(C.c_double) * 9)(*[capi.dptr(a) if a is not None else None for a in keep])
        _check(self.L.fc_vel_set_outer_all(self.h, ptrs), "vel_set_outer_all")

    def vel_get_outer_all(self):
        out = [np.zeros(sh, order="F") for sh in self._outer_shapes()]
        ptrs = (C.POINTER(C.c_double) * 9)(*[capi.dptr(a) for a in out])
        _check(self.L.fc_vel_get_outer_all(self.h, ptrs), "vel_get_outer_all")
        return out

    def set_gcm_vel(self, coeffD, coeffDeadD, ubi, ubiDead, probe_norm_d):
        """coeffGCMD (nGhost, 8), coeffGCMDeadD (nDead, 8), body-intercept
        velocities (nGhost, 3) and (nDead, 3), probeLengthNormalizedD"""
        self._gv = [np.ascontiguousarray(a, dtype=np.float64) for a in (coeffD, coeffDeadD, ubi, ubiDead)]
        _check(self.L.fc_set_gcm_vel(self.h, *[capi.dptr(a) for a in self._gv], probe_norm_d), "set_gcm_vel")

    def ghostcell_velocity(self, iopt=2, restol_gcv=1e-10):
        """GCM_ghostcell_velocity on the resident velocities: (rc, sweeps, last change)"""
        sw, rm = C.c_int(), C.c_double()
        rc = _check(self.L.fc_ghostcell_velocity(self.h, iopt, restol_gcv, C.byref(sw), C.byref(rm)),
                    "ghostcell_velocity")
        return rc, sw.value, rm.value

    def ghostcell_velocity_host(self, u, v, w, iopt=2, restol_gcv=1e-10):
        """GCM_ghostcell_velocity(iOpt, u, v, w) on host arrays (in place),
        staged through the resident velocities: (rc, sweeps, last change)"""
        self.vel_upload(u, v, w)
        out = self.ghostcell_velocity(iopt, restol_gcv)
        uu, vv, ww = self.vel_download()[:3]
        u[...], v[...], w[...] = uu, vv, ww
        return out

    def solve_ad_resident(self, time=0.0, itermax_ad=200, restol_ad=1e-10, restol_gcv=1e-12):
        it, res, loc = C.c_int(), C.c_double(), (C.c_int * 3)()
        rc = _check(self.L.fc_solve_ad_resident(self.h, time, itermax_ad, restol_ad, restol_gcv, C.byref(it),
                                                C.byref(res), loc), "solve_ad")
        return rc, it.value, res.value, tuple(loc)

    def solve_ad(self, u, v, w, nlu, nlv, nlw, fu, fv, fw, time=0.0, itermax_ad=200, restol_ad=1e-10,
                 restol_gcv=1e-12):
        """solve_ad() on host arrays: u, v, w and the face velocities in place"""
        for a, n in ((u, "u"), (v, "v"), (w, "w"), (nlu, "nlu"), (nlv, "nlv"), (nlw, "nlw"), (fu, "face_u"),
                     (fv, "face_v"), (fw, "face_w")):
            self._field(a, n)
        it, res, loc = C.c_int(), C.c_double(), (C.c_int * 3)()
        d = capi.dptr
        rc = _check(self.L.fc_solve_ad(self.h, d(u), d(v), d(w), d(nlu), d(nlv), d(nlw), d(fu), d(fv), d(fw), time,
                                       itermax_ad, restol_ad, restol_gcv, C.byref(it), C.byref(res), loc), "solve_ad")
        return rc, it.value, res.value, tuple(loc)

    def set_visc_tot(self, visc=None):
        """viscTot of rhs_advec_diff (None: the reinv of set_solve_ad)"""
        if visc is not None:
            self._field(visc, "viscTot")
        _check(self.L.fc_set_visc_tot(self.h, capi.dptr(visc)), "set_visc_tot")

    def rhs_advec_diff(self, uo, vo, wo, fu, fv, fw, alfa_upwind, dt, nlu, nlv, nlw):
        """rhs_advec_diff() (CORE_SET_SOLVERS.F90:10-61): the interior of nlu,
        nlv, nlw in place"""
        for a, n in ((uo, "uOld"), (vo, "vOld"), (wo, "wOld"), (fu, "face_u"), (fv, "face_v"), (fw, "face_w"),
                     (nlu, "nlu"), (nlv, "nlv"), (nlw, "nlw")):
            self._field(a, n)
        d = capi.dptr
        _check(self.L.fc_rhs_advec_diff(self.h, d(uo), d(vo), d(wo), d(fu), d(fv), d(fw), alfa_upwind, dt, d(nlu),
                                        d(nlv), d(nlw)), "rhs_advec_diff")

    def ad_download(self):
        """host copies of the resident nlu, nlv, nlw"""
        out = [np.zeros(self.shape, order="F") for _ in range(3)]
        _check(self.L.fc_ad_download(self.h, *[capi.dptr(a) for a in out]), "ad_download")
        return out

    def rhs_advec_diff_resident(self, alfa_upwind, dt):
        _check(self.L.fc_rhs_advec_diff_resident(self.h, alfa_upwind, dt), "rhs_advec_diff_resident")

    def itsolv_ad_resident(self, comp, nsweep=1):
        _check(self.L.fc_itsolv_ad_resident(self.h, comp, nsweep), "itsolv_ad_resident")

    def calc_residual_ad_resident(self, comp):
        resm, r2, loc = C.c_double(), C.c_double(), (C.c_int * 3)()
        _check(self.L.fc_calc_residual_ad_resident(self.h, comp, C.byref(resm), loc, C.byref(r2)),
               "calc_residual_ad_resident")
        return resm.value, tuple(loc), r2.value

    def ad_timing(self):
        ms = C.c_double()
        _check(self.L.fc_ad_timing_read(self.h, C.byref(ms)), "ad_timing_read")
        return ms.value

    # level operations (itsolv_mg, residual_mg, mg_inject, mg_interp, BC) ---
    def level_shape(self, lev):
        nx, ny, nz = self.level_nodes[lev - 1]
        return (nx + 2, ny + 2, nz + 2)

    def lvl_write(self, lev, which, arr):
        dt = np.int8 if which in (3, 4) else (np.float32 if which >= 10 else np.float64)
        a = np.asfortranarray(arr, dtype=dt)
        assert a.shape == self.level_shape(lev)
        _check(self.L.fc_lvl_write(self.h, lev, which, a.ctypes.data_as(C.c_void_p)), "fc_lvl_write")

    def lvl_read(self, lev, which):
        dt = np.int8 if which in (3, 4, 5, 6) else (np.float32 if which >= 10 else np.float64)
        a = np.zeros(self.level_shape(lev), dtype=dt, order="F")
        _check(self.L.fc_lvl_read(self.h, lev, which, a.ctypes.data_as(C.c_void_p)), "fc_lvl_read")
        return a

    def itsolv_mg(self, lev, prec=capi.PREC_FP64):
        _check(self.L.fc_lvl_smooth(self.h, lev, prec), "itsolv_mg")

    def sweep(self, lev, d, color, prec=capi.PREC_FP64):
        _check(self.L.fc_lvl_sweep(self.h, lev, d, color, prec), "sweep")

    def bc_fill(self, lev, prec=capi.PREC_FP64):
        _check(self.L.fc_lvl_bc(self.h, lev, prec), "set_outer_pressure_bc")

    def residual_mg(self, lev, prec=capi.PREC_FP64):
        _check(self.L.fc_lvl_residual(self.h, lev, prec), "residual_mg")

    def mg_inject(self, lev, prec=capi.PREC_FP64):
        _check(self.L.fc_lvl_restrict(self.h, lev, prec), "mg_inject")

    def mg_interp(self, lev, prec=capi.PREC_FP64):
        _check(self.L.fc_lvl_prolong(self.h, lev, prec), "mg_interp")

    def coarse_solve(self, prec=capi.PREC_FP64):
        """the FP64 direct solve of the coarsest level (coarse_solve = 1 with
        the wall correction cycles)"""
        _check(self.L.fc_lvl_coarse_solve(self.h, prec), "coarse_solve")


class Server:
    """Request server (fc_serve_*): many independent solves on one GPU with
    continuous batching.  submit() keeps the request's arrays alive until
    its result is polled."""

    def __init__(self, pool_bytes, max_active, admit_threads=0):
        self.L = library()
        cfg = capi.FcServeConfig(pool_bytes, max_active, admit_threads)
        h = C.c_void_p()
        _check(self.L.fc_serve_create(C.byref(cfg), C.byref(h)), "fc_serve_create")
        self.h = h
        self._keep = {}

    def close(self):
        if getattr(self, "h", None):
            self.L.fc_serve_destroy(self.h)
            self.h = None
            self._keep = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, w, params, out=None, planes=None):
        """Queue workload w (geometry.Workload) with params; the answer goes
        into out (a zeroed Fortran field when None).  Returns (id, out)."""
        x, y, z = [np.ascontiguousarray(a, dtype=np.float64) for a in w.grid]
        ib = np.asfortranarray(w.iblank, dtype=np.int8)
        gh = np.asfortranarray(w.ghost, dtype=np.int8)
        de = np.asfortranarray(w.dead, dtype=np.int8)
        g = {k: np.ascontiguousarray(v) for k, v in w.gcm.items()}
        div = np.asfortranarray(w.div, dtype=np.float64)
        if out is None:
            out = np.zeros(div.shape, order="F")
        pl = [None] * 6 if planes is None else [None if a is None else np.asfortranarray(a, dtype=np.float64)
                                                for a in planes]
        rq = capi.FcServeRequest()
        rq.grid = capi.FcGrid(x.size - 2, y.size - 2, z.size - 2, capi.dptr(x), capi.dptr(y), capi.dptr(z))
        rq.params = params
        rq.iblank, rq.ghost, rq.dead = capi.bptr(ib), capi.bptr(gh), capi.bptr(de)
        rq.nGhost = g["iGhost"].size
        for f in ("iGhost", "jGhost", "kGhost", "iCellIndex", "jCellIndex", "kCellIndex"):
            setattr(rq, f, capi.iptr(g[f]))
        for f in ("coeffGCMN", "alphaGhost", "dpdnBodyIntercept"):
            setattr(rq, f, capi.dptr(g[f]))
        rq.nDead = g["iDead"].size
        for f in ("iDead", "jDead", "kDead", "iDeadCellIndex", "jDeadCellIndex", "kDeadCellIndex"):
            setattr(rq, f, capi.iptr(g[f]))
        for f in ("coeffGCMDeadN", "probeLengthDead", "dpdnInterceptDead", "xyzDCIP"):
            setattr(rq, f, capi.dptr(g[f]))
        for f in range(6):
            rq.pgrad[f] = capi.plane_ptr(pl[f])
        rq.div = capi.dptr(div)
        rq.pPrime = capi.dptr(out)
        rid = C.c_longlong()
        _check(self.L.fc_serve_submit(self.h, C.byref(rq), C.byref(rid)), "fc_serve_submit")
        self._keep[rid.value] = (x, y, z, ib, gh, de, g, div, out, pl, rq)
        return rid.value, out

    def step(self):
        st = capi.FcServeStatus()
        _check(self.L.fc_serve_step(self.h, C.byref(st)), "fc_serve_step")
        return st

    def poll(self):
        """One finished request as (id, rc, stats, out), or None."""
        rid, rc, st = C.c_longlong(), C.c_int(), capi.FcStats()
        if not self.L.fc_serve_poll(self.h, C.byref(rid), C.byref(rc), C.byref(st)):
            return None
        keep = self._keep.pop(rid.value, None)
        return rid.value, rc.value, st, (keep[8] if keep else None)

    def drain(self, max_steps=100000):
        """Steps until every submitted request is finished; returns the polled results by id."""
        res = {}
        for _ in range(max_steps):
            st = self.step()
            while True:
                r = self.poll()
                if r is None:
                    break
                res[r[0]] = r
            if st.active == 0 and st.queued == 0:
                break
        return res
