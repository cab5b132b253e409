"""ctypes mirror of include/fc_poisson.h (structs, constants, library loading).

The Python layer is host plumbing over the C-ABI; every compute call goes
through the HIP library ``csrc/_build/libfc_poisson.so``.
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "csrc", "_build", "libfc_poisson.so")

# AMODULES.F90:15-44
BC_TYPE_DIRICHLET = 1
BC_TYPE_ZERO_GRADIENT = 2
BC_TYPE_PERIODIC = 5
IT_SOLVER_TYPE_LSOR = 1
IT_SOLVER_TYPE_MG = 3
PBC_DIRICHLET = 1
PBC_NEUMANN = 2
PREC_FP64 = 0
PREC_MIXED = 1
OUTER_REFERENCE = 0
OUTER_RELRES = 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_bp = C.POINTER(C.c_int8)


class FcGrid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("x", _dp), ("y", _dp), ("z", _dp)]


class FcParams(C.Structure):
    _fields_ = [("it_solver_type", C.c_int), ("omega_pson", C.c_double),
                ("itermax_pson", C.c_int), ("restol_pson", C.c_double),
                ("iterInter", C.c_int), ("mgrids_x", C.c_int), ("mgrids_y", C.c_int),
                ("mgrids_z", C.c_int), ("restol_gcp", C.c_double), ("itermax_gc", C.c_int),
                ("mix_GC_form", C.c_int), ("is_PBC_homogeneous", C.c_int),
                ("pbc", C.c_int * 6), ("ppp", C.c_double * 6), ("bc_vel", C.c_int * 6),
                ("infoconv", C.c_int), ("precision", C.c_int), ("outer_rule", C.c_int),
                ("rel_tol", C.c_double), ("fp32_cycles", C.c_int), ("aa_depth", C.c_int),
                ("launch_mode", C.c_int), ("cycle", C.c_int), ("vc_walls", C.c_int), ("coarse_solve", C.c_int),
                ("reserved", C.c_int * 3)]


class FcStats(C.Structure):
    _fields_ = [("outer_iters", C.c_int), ("vcycles_fp64", C.c_int), ("vcycles_fp32", C.c_int),
                ("refine_steps", C.c_int), ("gcm_sweeps", C.c_int), ("converged", C.c_int),
                ("maxres", C.c_double), ("maxres2", C.c_double), ("rel_resid", C.c_double),
                ("gcm_change", C.c_double), ("time_ms", C.c_double), ("reserved", C.c_double * 6)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


class FcServeConfig(C.Structure):
    _fields_ = [("pool_bytes", C.c_size_t), ("max_active", C.c_int), ("admit_threads", C.c_int),
                ("reserved", C.c_int * 6)]


class FcServeRequest(C.Structure):
    _fields_ = [("grid", FcGrid), ("params", FcParams),
                ("iblank", _bp), ("ghost", _bp), ("dead", _bp),
                ("nGhost", C.c_int),
                ("iGhost", _ip), ("jGhost", _ip), ("kGhost", _ip),
                ("iCellIndex", _ip), ("jCellIndex", _ip), ("kCellIndex", _ip),
                ("coeffGCMN", _dp), ("alphaGhost", _dp), ("dpdnBodyIntercept", _dp),
                ("nDead", C.c_int),
                ("iDead", _ip), ("jDead", _ip), ("kDead", _ip),
                ("iDeadCellIndex", _ip), ("jDeadCellIndex", _ip), ("kDeadCellIndex", _ip),
                ("coeffGCMDeadN", _dp), ("probeLengthDead", _dp), ("dpdnInterceptDead", _dp), ("xyzDCIP", _dp),
                ("pgrad", _dp * 6), ("div", _dp), ("pPrime", _dp)]


class FcServeStatus(C.Structure):
    _fields_ = [("active", C.c_int), ("queued", C.c_int), ("finished", C.c_int), ("launches", C.c_int),
                ("steps", C.c_longlong), ("pool_used", C.c_size_t), ("pool_capacity", C.c_size_t)]


def default_params(**kw):
    """Parameters used when no input.dat exists (the reference ships none);
    every benchmark records them.  Keyword arguments override fields."""
    p = FcParams()
    p.it_solver_type = IT_SOLVER_TYPE_MG
    p.omega_pson = 1.0
    p.itermax_pson = 200
    p.restol_pson = 1e-10
    p.iterInter = 1
    p.mgrids_x = p.mgrids_y = p.mgrids_z = 0
    p.restol_gcp = 1e-10
    p.itermax_gc = 10
    p.mix_GC_form = 0
    p.is_PBC_homogeneous = 1
    for f in range(6):
        p.pbc[f] = PBC_NEUMANN
        p.ppp[f] = 0.0
        p.bc_vel[f] = BC_TYPE_DIRICHLET
    p.infoconv = 0
    p.precision = PREC_FP64
    p.outer_rule = OUTER_REFERENCE
    p.rel_tol = 1e-8
    p.fp32_cycles = 1
    p.vc_walls = 1
    p.coarse_solve = 0
    for k, v in kw.items():
        if k in ("pbc", "ppp", "bc_vel"):
            for f in range(6):
                getattr(p, k)[f] = v[f]
        else:
            setattr(p, k, v)
    return p


def params_dict(p):
    d = {}
    for f, _ in FcParams._fields_:
        if f == "reserved":
            continue
        v = getattr(p, f)
        d[f] = list(v) if f in ("pbc", "ppp", "bc_vel") else v
    return d


def _check_array(a, dtype, shape, what, fortran=None):
    """Raise ValueError unless a has the dtype, the shape (when given) and the
    memory order the C side reads: Fortran order for 3-D fields, C order for
    lists and per-relation tables such as coeffGCMN(8, nGhost).  The C side reads exactly the extents the
    handle was created with, so a wrong shape would be read out of bounds."""
    if not isinstance(a, np.ndarray) or a.dtype != dtype:
        raise ValueError("%s must be a numpy array of %s" % (what, np.dtype(dtype).name))
    if shape is not None and tuple(a.shape) != tuple(shape):
        raise ValueError("%s has shape %s, expected %s" % (what, tuple(a.shape), tuple(shape)))
    if fortran is None:
        fortran = a.ndim == 3
    if fortran and not a.flags["F_CONTIGUOUS"]:
        raise ValueError("%s must be Fortran-ordered" % what)
    if not fortran and not a.flags["C_CONTIGUOUS"]:
        raise ValueError("%s must be contiguous" % what)


def dptr(a, shape=None, what="array"):
    """float64 pointer; shape, when given, is checked."""
    if a is None:
        return None
    _check_array(a, np.float64, shape, what)
    return a.ctypes.data_as(_dp)


def plane_ptr(a, shape=None, what="plane"):
    """A boundary plane such as pgradx1(0:ny+1,0:nz+1): 2-D, Fortran-ordered."""
    if a is None:
        return None
    if not isinstance(a, np.ndarray) or a.ndim != 2:
        raise ValueError("%s must be a 2-D array" % what)
    _check_array(a, np.float64, shape, what, fortran=True)
    return a.ctypes.data_as(_dp)


def iptr(a, shape=None, what="index list"):
    if a is None:
        return None
    _check_array(a, np.int32, shape, what)
    return a.ctypes.data_as(_ip)


def bptr(a, shape=None, what="mask"):
    if a is None:
        return None
    _check_array(a, np.int8, shape, what)
    return a.ctypes.data_as(_bp)


def field(shape_nodes, dtype=np.float64):
    """Zeroed Fortran-ordered field (0:nx+1, 0:ny+1, 0:nz+1)."""
    nx, ny, nz = shape_nodes
    return np.zeros((nx + 2, ny + 2, nz + 2), dtype=dtype, order="F")
