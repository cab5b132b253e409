"""Host-side mirror of the immersed-boundary geometry refresh over the HIP
library (ctypes over include/fc_geometry.h).

GeometryRefresh drives what one time step of a moving body recomputes
(CORE_TIME_STEP.F90:100-118): calculate_arclength_norm_ds
(BOUNDARY_SET_ARCLENGTH_NORM.F90:2), set_iblank_body_fast
(BOUNDARY_IBLANK_FAST.F90:2) and GCM_set_internal_boundary
(GCM_SET_BOUNDARY.F90:2) with smooth_iblank, GCM_define_ghostcells,
GCM_Calc_VanMatrixDN and GCM_set_deadcells.  Results come back in the
reference's array names.  There is no CPU fallback: the HIP library must be
built and a GPU present.
"""
import ctypes as C

import numpy as np

from . import capi

MAX_BODIES = 8


class FcGeomParams(C.Structure):
    _fields_ = [("nBody", C.c_int), ("unstruc_surface_type", C.c_int * MAX_BODIES),
                ("body_dim", C.c_int * MAX_BODIES), ("membrane_tkns", C.c_double),
                ("probeLengthNormalized", C.c_double), ("probeLengthNormalizedD", C.c_double),
                ("reserved", C.c_int * 8)]


class FcGeomStats(C.Structure):
    _fields_ = [("nGhost", C.c_int), ("nDead", C.c_int), ("n_undecided", C.c_int), ("n_smoothed", C.c_int),
                ("smooth_rounds", C.c_int), ("n_unselected_side", C.c_int), ("time_ms", C.c_double),
                ("scan_ms", C.c_double), ("scan_pairs", C.c_double), ("reserved", C.c_int * 4)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

GHOST_INT = ["iGhost", "jGhost", "kGhost", "iCellIndex", "jCellIndex", "kCellIndex", "closestElementGC"]
GHOST_DBL = ["xBodyIntercept", "yBodyIntercept", "zBodyIntercept", "xBodyInterceptNorm", "yBodyInterceptNorm",
             "zBodyInterceptNorm", "probeLength", "xImagePoint", "yImagePoint", "zImagePoint"]
GHOST_COEF = ["coeffGCMD", "coeffGCMN"]
DEAD_INT = ["iDead", "jDead", "kDead", "iDeadCellIndex", "jDeadCellIndex", "kDeadCellIndex", "closestElementDead"]
DEAD_DBL = ["xBodyInterceptDead", "yBodyInterceptDead", "zBodyInterceptDead", "xNormInterceptDead",
            "yNormInterceptDead", "zNormInterceptDead", "probeLengthDead"]
DEAD_COEF = ["coeffGCMDeadD", "coeffGCMDeadN"]
SURF_ELEM = ["triElemNormX", "triElemNormY", "triElemNormZ", "triElemArea", "triElemCentX", "triElemCentY",
             "triElemCentZ"]
SURF_VERT = ["xNormBodyMarker", "yNormBodyMarker", "zNormBodyMarker"]
FORCE_NAMES = ["xMarkerForce", "yMarkerForce", "zMarkerForce", "xMarkerShear", "yMarkerShear", "zMarkerShear",
               "MarkerArea", "xMarkerStress", "yMarkerStress", "zMarkerStress"]
TOTAL_NAMES = ["cxt", "cyt", "czt", "cmxt", "cmyt", "cmzt", "cxs", "cys", "czs"]
BI_NAMES = ["uBodyIntercept", "vBodyIntercept", "wBodyIntercept", "dpdnBodyIntercept", "uBodyInterceptDead",
            "vBodyInterceptDead", "wBodyInterceptDead", "dpdnInterceptDead"]


class FcGeomGhosts(C.Structure):
    _fields_ = [(f, _ip) for f in GHOST_INT] + [(f, _dp) for f in GHOST_DBL + GHOST_COEF + ["alphaGhost"]]


class FcGeomDead(C.Structure):
    _fields_ = [(f, _ip) for f in DEAD_INT] + [(f, _dp) for f in DEAD_DBL + DEAD_COEF]


class FcGeomSurface(C.Structure):
    _fields_ = [(f, _dp) for f in SURF_ELEM + SURF_VERT]


class FcGeomForces(C.Structure):
    _fields_ = [(f, _dp) for f in FORCE_NAMES]


def collect_stress(get_bi, get_forces, nG, nD, surfaces, totals):
    """The body-intercept values and marker forces of one backend, named as
    the reference names them; totals: the nine per-body sums."""
    out = {}
    arrs = [np.zeros(max(nG if k < 4 else nD, 1)) for k in range(8)]
    get_bi(*[a.ctypes.data_as(_dp) for a in arrs])
    for k, (name, a) in enumerate(zip(BI_NAMES, arrs)):
        out[name] = a[:nG if k < 4 else nD]
    out["forces"] = []
    for b, s in enumerate(surfaces):
        f = FcGeomForces()
        d = {}
        for name in FORCE_NAMES:
            d[name] = np.zeros(s.npts)
            setattr(f, name, d[name].ctypes.data_as(_dp))
        get_forces(b + 1, C.byref(f))
        d["totals"] = np.array(totals[9 * b:9 * b + 9])
        out["forces"].append(d)
    return out


def collect_bi(get_bi, nG, nD):
    """The body-intercept values of one backend, named as the reference
    names them."""
    arrs = [np.zeros(max(nG if k < 4 else nD, 1)) for k in range(8)]
    get_bi(*[a.ctypes.data_as(_dp) for a in arrs])
    return {name: a[:nG if k < 4 else nD] for k, (name, a) in enumerate(zip(BI_NAMES, arrs))}


HANDOFF_GHOST_INT = ["iGhost", "jGhost", "kGhost", "iCellIndex", "jCellIndex", "kCellIndex"]
HANDOFF_DEAD_INT = ["iDead", "jDead", "kDead", "iDeadCellIndex", "jDeadCellIndex", "kDeadCellIndex"]


def collect_handoff(get_counts, get_masks, get_ghosts, get_dead, get_dead_ip, shape):
    """What the flow solver takes from a refresh (masks, the GCM lists and
    coefficients, the dead cells' image points), nothing else."""
    nG, nD = C.c_int(), C.c_int()
    get_counts(C.byref(nG), C.byref(nD))
    nG, nD = nG.value, nD.value
    out = dict(nGhost=nG, nDead=nD)
    masks = [np.zeros(shape, dtype=np.int8, order="F") for _ in range(3)]
    get_masks(*[m.ctypes.data_as(C.POINTER(C.c_int8)) for m in masks], None, None, None)
    out["iblank"], out["ghostCellMark"], out["dead_cell"] = masks
    g = FcGeomGhosts()
    for f in HANDOFF_GHOST_INT:
        out[f] = np.zeros(max(nG, 1), dtype=np.int32)
        setattr(g, f, out[f].ctypes.data_as(_ip))
    out["alphaGhost"] = np.zeros(max(nG, 1))
    g.alphaGhost = out["alphaGhost"].ctypes.data_as(_dp)
    for f in GHOST_COEF:
        out[f] = np.zeros((max(nG, 1), 8))
        setattr(g, f, out[f].ctypes.data_as(_dp))
    get_ghosts(C.byref(g))
    d = FcGeomDead()
    for f in HANDOFF_DEAD_INT:
        out[f] = np.zeros(max(nD, 1), dtype=np.int32)
        setattr(d, f, out[f].ctypes.data_as(_ip))
    out["probeLengthDead"] = np.zeros(max(nD, 1))
    d.probeLengthDead = out["probeLengthDead"].ctypes.data_as(_dp)
    for f in DEAD_COEF:
        out[f] = np.zeros((max(nD, 1), 8))
        setattr(d, f, out[f].ctypes.data_as(_dp))
    get_dead(C.byref(d))
    ip = np.zeros((max(nD, 1), 3))
    if nD:
        get_dead_ip(ip)
    out["xyzDCIP_dead"] = ip[:nD]
    for f in HANDOFF_GHOST_INT + GHOST_COEF + ["alphaGhost"]:
        out[f] = out[f][:nG]
    for f in HANDOFF_DEAD_INT + DEAD_COEF + ["probeLengthDead"]:
        out[f] = out[f][:nD]
    return out


def make_params(surfaces, membrane_tkns=0.0, probe=2.0, probeD=2.0):
    if not 1 <= len(surfaces) <= MAX_BODIES:
        raise ValueError("1 to %d bodies" % MAX_BODIES)
    p = FcGeomParams()
    p.nBody = len(surfaces)
    for b, s in enumerate(surfaces):
        p.unstruc_surface_type[b] = s.kind
        p.body_dim[b] = s.dim
    p.membrane_tkns = membrane_tkns
    p.probeLengthNormalized = probe
    p.probeLengthNormalizedD = probeD
    return p


def collect(get_counts, get_masks, get_ghosts, get_dead, get_surface, get_dcip, shape, surfaces):
    """Reads every result of a refresh through the getters of one backend
    (the HIP library or the oracle) into numpy arrays named as the
    reference names them."""
    nG, nD = C.c_int(), C.c_int()
    get_counts(C.byref(nG), C.byref(nD))
    nG, nD = nG.value, nD.value
    out = dict(nGhost=nG, nDead=nD)
    masks = [np.zeros(shape, dtype=np.int8, order="F") for _ in range(6)]
    get_masks(*[m.ctypes.data_as(C.POINTER(C.c_int8)) for m in masks])
    for name, m in zip(["iblank", "ghostCellMark", "dead_cell", "bodyNum", "ps1", "ps2"], masks):
        out[name] = m
    g = FcGeomGhosts()
    for f in GHOST_INT:
        out[f] = np.zeros(max(nG, 1), dtype=np.int32)
        setattr(g, f, out[f].ctypes.data_as(_ip))
    for f in GHOST_DBL + ["alphaGhost"]:
        out[f] = np.zeros(max(nG, 1))
        setattr(g, f, out[f].ctypes.data_as(_dp))
    for f in GHOST_COEF:
        out[f] = np.zeros((max(nG, 1), 8))
        setattr(g, f, out[f].ctypes.data_as(_dp))
    get_ghosts(C.byref(g))
    d = FcGeomDead()
    for f in DEAD_INT:
        out[f] = np.zeros(max(nD, 1), dtype=np.int32)
        setattr(d, f, out[f].ctypes.data_as(_ip))
    for f in DEAD_DBL:
        out[f] = np.zeros(max(nD, 1))
        setattr(d, f, out[f].ctypes.data_as(_dp))
    for f in DEAD_COEF:
        out[f] = np.zeros((max(nD, 1), 8))
        setattr(d, f, out[f].ctypes.data_as(_dp))
    get_dead(C.byref(d))
    for f in GHOST_INT + GHOST_DBL + GHOST_COEF + ["alphaGhost"]:
        out[f] = out[f][:nG]
    for f in DEAD_INT + DEAD_DBL + DEAD_COEF:
        out[f] = out[f][:nD]
    out["surfaces"] = []
    for b, s in enumerate(surfaces):
        f = FcGeomSurface()
        sd = {}
        for name in SURF_ELEM:
            sd[name] = np.zeros(s.nelem)
            setattr(f, name, sd[name].ctypes.data_as(_dp))
        for name in SURF_VERT:
            sd[name] = np.zeros(s.npts)
            setattr(f, name, sd[name].ctypes.data_as(_dp))
        get_surface(b + 1, C.byref(f))
        out["surfaces"].append(sd)
    dcip = np.zeros(shape + (3,), order="F")
    get_dcip(dcip.ctypes.data_as(_dp))
    out["xyzDCIP"] = dcip
    return out


_lib = None


def library():
    global _lib
    if _lib is not None:
        return _lib
    from .poisson import library as pson_library
    L = pson_library()
    vp = C.c_void_p
    i = C.c_int
    L.fc_geom_create.argtypes = [C.POINTER(capi.FcGrid), C.POINTER(FcGeomParams), C.POINTER(vp)]
    L.fc_geom_destroy.argtypes = [vp]
    L.fc_geom_set_body.argtypes = [vp, i, i, i, _dp, _dp, _dp, _ip]
    L.fc_geom_move_body.argtypes = [vp, i, _dp, _dp, _dp]
    L.fc_geom_refresh.argtypes = [vp, C.POINTER(FcGeomStats)]
    L.fc_geom_get_counts.argtypes = [vp, _ip, _ip]
    L.fc_geom_get_masks.argtypes = [vp] + [C.POINTER(C.c_int8)] * 6
    L.fc_geom_get_ghosts.argtypes = [vp, C.POINTER(FcGeomGhosts)]
    L.fc_geom_get_dead.argtypes = [vp, C.POINTER(FcGeomDead)]
    L.fc_geom_get_surface.argtypes = [vp, i, C.POINTER(FcGeomSurface)]
    L.fc_geom_get_dcip.argtypes = [vp, _dp]
    L.fc_geom_get_dead_ip.argtypes = [vp, _dp]
    L.fc_geom_set_marker_velocity.argtypes = [vp, i] + [_dp] * 6
    L.fc_geom_body_intercept_values.argtypes = [vp, C.c_double, i]
    L.fc_geom_get_bi_values.argtypes = [vp] + [_dp] * 8
    L.fc_geom_marker_stress.argtypes = [vp, _dp, _dp, _dp, _dp, C.c_double, _dp]
    L.fc_geom_marker_stress_device.argtypes = [vp] + [vp] * 4 + [C.c_double, _dp]
    L.fc_geom_get_marker_forces.argtypes = [vp, i, C.POINTER(FcGeomForces)]
    L.fc_geom_get_stress_timing.argtypes = [vp, _dp, _dp, _dp]
    L.fc_geom_marker_stress_resident.argtypes = [vp, vp, C.c_double, _dp]
    _lib = L
    return L


def _check(L, rc, what):
    if rc < 0:
        from .poisson import FcError
        raise FcError("%s failed (%d): %s" % (what, rc, L.fc_pson_last_error().decode()))
    return rc


class GeometryRefresh:
    """Device-resident geometry of a grid and its triangulated bodies."""

    def __init__(self, x, y, z, surfaces, membrane_tkns=0.0, probe=2.0, probeD=2.0):
        self.L = library()
        self.grid_arrays = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z)]
        g = capi.FcGrid()
        g.nx, g.ny, g.nz = (a.size - 2 for a in self.grid_arrays)
        g.x, g.y, g.z = (a.ctypes.data_as(_dp) for a in self.grid_arrays)
        self.shape = (g.nx + 2, g.ny + 2, g.nz + 2)
        self.surfaces = list(surfaces)
        p = make_params(self.surfaces, membrane_tkns, probe, probeD)
        h = C.c_void_p()
        _check(self.L, self.L.fc_geom_create(C.byref(g), C.byref(p), C.byref(h)), "fc_geom_create")
        self.h = h
        for b, s in enumerate(self.surfaces):
            _check(self.L, self.L.fc_geom_set_body(self.h, b + 1, s.npts, s.nelem, s.x.ctypes.data_as(_dp),
                                                   s.y.ctypes.data_as(_dp), s.z.ctypes.data_as(_dp),
                                                   s.tri.ctypes.data_as(_ip)), "fc_geom_set_body")

    def move_body(self, ibody, surface):
        """BOUNDARY_MOVE.F90's new marker positions of body ibody (1-based)."""
        s = surface
        self.surfaces[ibody - 1] = s
        _check(self.L, self.L.fc_geom_move_body(self.h, ibody, s.x.ctypes.data_as(_dp), s.y.ctypes.data_as(_dp),
                                                s.z.ctypes.data_as(_dp)), "fc_geom_move_body")

    def refresh(self):
        st = FcGeomStats()
        _check(self.L, self.L.fc_geom_refresh(self.h, C.byref(st)), "fc_geom_refresh")
        return st

    def results(self):
        L, h = self.L, self.h
        return collect(lambda a, b: L.fc_geom_get_counts(h, a, b),
                       lambda *m: L.fc_geom_get_masks(h, *m),
                       lambda g: L.fc_geom_get_ghosts(h, g),
                       lambda d: L.fc_geom_get_dead(h, d),
                       lambda b, f: L.fc_geom_get_surface(h, b, f),
                       lambda p: L.fc_geom_get_dcip(h, p), self.shape, self.surfaces)

    def handoff(self):
        """The masks, GCM lists and coefficients and the dead cells' image
        points of the last refresh: what the flow solver takes."""
        L, h = self.L, self.h
        return collect_handoff(lambda a, b: L.fc_geom_get_counts(h, a, b),
                               lambda *m: _check(L, L.fc_geom_get_masks(h, *m), "fc_geom_get_masks"),
                               lambda g: _check(L, L.fc_geom_get_ghosts(h, g), "fc_geom_get_ghosts"),
                               lambda d: _check(L, L.fc_geom_get_dead(h, d), "fc_geom_get_dead"),
                               lambda ip: _check(L, L.fc_geom_get_dead_ip(h, ip.ctypes.data_as(_dp)),
                                                 "fc_geom_get_dead_ip"), self.shape)

    def set_marker_velocity(self, ibody, u, v, w, uOld, vOld, wOld):
        """u/v/wBodyMarker and u/v/wBodyMarkerOld of body ibody (1-based)."""
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (u, v, w, uOld, vOld, wOld)]
        _check(self.L, self.L.fc_geom_set_marker_velocity(self.h, ibody, *[x.ctypes.data_as(_dp) for x in a]),
               "fc_geom_set_marker_velocity")

    def body_intercept_values(self, dt, ntime):
        """GCM_SetBodyInterceptValues (GCM_BODYINTERCEPT_VALUES.F90:22)."""
        _check(self.L, self.L.fc_geom_body_intercept_values(self.h, dt, ntime), "fc_geom_body_intercept_values")

    def marker_stress(self, p, u, v, w, reinv):
        """compute_marker_stress (UTIL_DRAG_LIFT.F90:29) on Fortran-ordered
        host fields; returns the nine totals per body (cxt ... czs)."""
        fields = []
        for a in (p, u, v, w):
            capi._check_array(a, np.float64, self.shape, "flow field")
            fields.append(a.ctypes.data_as(_dp))
        tot = np.zeros(9 * len(self.surfaces))
        _check(self.L, self.L.fc_geom_marker_stress(self.h, *fields, reinv, tot.ctypes.data_as(_dp)),
               "fc_geom_marker_stress")
        self._totals = tot
        return tot.reshape(-1, 9)

    def marker_stress_device(self, p, u, v, w, reinv):
        """The same on device-resident fields: torch FP64 tensors on the
        handle's device, contiguous with shape (nz+2, ny+2, nx+2) (the
        Fortran (0:nx+1, 0:ny+1, 0:nz+1) layout)."""
        ptrs = []
        for a in (p, u, v, w):
            if (not a.is_cuda or str(a.dtype) != "torch.float64" or not a.is_contiguous()
                    or tuple(a.shape) != self.shape[::-1]):
                raise ValueError("flow field: contiguous float64 device tensor of shape %s" % (self.shape[::-1],))
            ptrs.append(C.c_void_p(a.data_ptr()))
        tot = np.zeros(9 * len(self.surfaces))
        _check(self.L, self.L.fc_geom_marker_stress_device(self.h, *ptrs, reinv, tot.ctypes.data_as(_dp)),
               "fc_geom_marker_stress_device")
        self._totals = tot
        return tot.reshape(-1, 9)

    def marker_stress_resident(self, flow, reinv):
        """The same on the resident p, u, v, w of a PoissonSolver of this
        grid (after its fc_project_resident)."""
        tot = np.zeros(9 * len(self.surfaces))
        _check(self.L, self.L.fc_geom_marker_stress_resident(self.h, flow.h, reinv, tot.ctypes.data_as(_dp)),
               "fc_geom_marker_stress_resident")
        self._totals = tot
        return tot.reshape(-1, 9)

    def stress_timing(self):
        """(bi_ms, interp_ms, sum_ms): device times of the last
        body-intercept values and marker-stress calls."""
        v = [C.c_double() for _ in range(3)]
        _check(self.L, self.L.fc_geom_get_stress_timing(self.h, *[C.byref(a) for a in v]),
               "fc_geom_get_stress_timing")
        return tuple(a.value for a in v)

    def bi_values(self):
        """u/v/wBodyIntercept, dpdnBodyIntercept and the dead cells' of the
        last fc_geom_body_intercept_values."""
        L, h = self.L, self.h
        nG, nD = C.c_int(), C.c_int()
        L.fc_geom_get_counts(h, C.byref(nG), C.byref(nD))
        return collect_bi(lambda *a: _check(L, L.fc_geom_get_bi_values(h, *a), "fc_geom_get_bi_values"),
                          nG.value, nD.value)

    def stress_results(self):
        L, h = self.L, self.h
        nG, nD = C.c_int(), C.c_int()
        L.fc_geom_get_counts(h, C.byref(nG), C.byref(nD))
        return collect_stress(lambda *a: L.fc_geom_get_bi_values(h, *a),
                              lambda b, f: L.fc_geom_get_marker_forces(h, b, f), nG.value, nD.value,
                              self.surfaces, getattr(self, "_totals", np.zeros(9 * len(self.surfaces))))

    def close(self):
        if self.h:
            self.L.fc_geom_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
