"""Python binding of the restart file (include/fc_restart.h): write_restart
and read_restart of MAIN.F90:983-1122 for one flow rank, from numpy arrays
or device-resident arrays.

FLOW_NAMES is the order of record 2; every name is the reference's array.
Grid-shaped arrays are (nx+2, ny+2, nz+2) Fortran-ordered numpy arrays or,
on the device, contiguous torch tensors of the reversed shape or any object
whose data_ptr() holds the Fortran-ordered array."""
import ctypes as C

import numpy as np

FLOW_NAMES = ["u", "v", "w", "p", "face_u", "face_v", "face_w", "pPrime",
              "bcxu", "bcxv", "bcxw", "bcxp", "bcyu", "bcyv", "bcyw", "bcyp",
              "bczu", "bczv", "bczw", "bczp", "pgradx1", "pgradx2", "pgrady1", "pgrady2", "pgradz1", "pgradz2",
              "iblank", "ghostCellMark"]
BODY_ARRAYS = ["xBodyMarker", "yBodyMarker", "zBodyMarker", "uBodyMarker", "vBodyMarker", "wBodyMarker",
               "xMarkerForce", "yMarkerForce", "zMarkerForce"]
BODY_SCALARS = ["xcent", "ycent", "zcent", "angx", "angy", "angz", "vxcent", "vycent", "vzcent",
                "angvx", "angvy", "angvz"]

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class FcRestartFlow(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in FLOW_NAMES] + [("on_device", C.c_int)]


class FcRestartBody(C.Structure):
    _fields_ = ([("nPts", C.c_int), ("nElem", C.c_int)] + [(n, _dp) for n in BODY_ARRAYS]
                + [("triElemNeig", _ip), ("cent", C.c_double * 3), ("ang", C.c_double * 3),
                   ("vcent", C.c_double * 3), ("angv", C.c_double * 3)])


def flow_shapes(nx, ny, nz):
    """Fortran shapes of the record-2 arrays (CORE_MEMORY_ALLOCATE.F90:33-86,
    one rank)."""
    X, Y, Z = nx + 2, ny + 2, nz + 2
    s = {}
    for n in FLOW_NAMES[:8] + ["iblank", "ghostCellMark"]:
        s[n] = (X, Y, Z)
    for n in ("bcxu", "bcxv", "bcxw", "bcxp"):
        s[n] = (2, Y, Z)
    for n in ("bcyu", "bcyv", "bcyw", "bcyp"):
        s[n] = (X, 2, Z)
    for n in ("bczu", "bczv", "bczw", "bczp"):
        s[n] = (X, Y, 2)
    s["pgradx1"] = s["pgradx2"] = (Y, Z)
    s["pgrady1"] = s["pgrady2"] = (X, Z)
    s["pgradz1"] = s["pgradz2"] = (X, Y)
    return s


_lib = None


def library():
    global _lib
    if _lib is not None:
        return _lib
    from .poisson import library as pson_library
    L = pson_library()
    L.fc_restart_name.argtypes = [C.c_int, C.c_char_p, C.c_int]
    L.fc_restart_write.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                   C.POINTER(FcRestartFlow), C.c_int, C.POINTER(FcRestartBody)]
    L.fc_restart_read.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, _dp, _ip,
                                  C.POINTER(FcRestartFlow), C.c_int, C.POINTER(FcRestartBody)]
    L.fc_restart_set_max_subrecord.argtypes = [C.c_longlong]
    L.fc_restart_set_max_subrecord.restype = C.c_longlong
    _lib = L
    return L


def _check(L, rc, what):
    if rc < 0:
        from .poisson import FcError
        raise FcError("%s failed (%d): %s" % (what, rc, L.fc_pson_last_error().decode()))


def restart_name(lproc):
    L = library()
    buf = C.create_string_buffer(32)
    _check(L, L.fc_restart_name(lproc, buf, 32), "fc_restart_name")
    return buf.value.decode()


def _flow_struct(nx, ny, nz, flow, on_device):
    shapes = flow_shapes(nx, ny, nz)
    f = FcRestartFlow()
    f.on_device = 1 if on_device else 0
    for n in FLOW_NAMES:
        a = flow[n]
        want = np.int8 if n in ("iblank", "ghostCellMark") else np.float64
        if on_device:
            # a device array: a torch tensor (checked), or any object with
            # data_ptr() holding the Fortran-ordered array (its size is the
            # caller's to get right)
            if hasattr(a, "is_cuda"):
                import torch
                tw = torch.int8 if want is np.int8 else torch.float64
                if not a.is_cuda or a.dtype != tw or not a.is_contiguous() or tuple(a.shape) != shapes[n][::-1]:
                    raise ValueError("%s: contiguous %s device tensor of shape %s" % (n, tw, shapes[n][::-1]))
            elif not hasattr(a, "data_ptr"):
                raise ValueError("%s: a device array (data_ptr())" % n)
            elif hasattr(a, "nbytes") and a.nbytes != int(np.prod(shapes[n])) * np.dtype(want).itemsize:
                raise ValueError("%s: %d bytes expected" % (n, int(np.prod(shapes[n])) * np.dtype(want).itemsize))
            setattr(f, n, a.data_ptr())
        else:
            if a.dtype != want or a.shape != shapes[n] or not a.flags.f_contiguous:
                raise ValueError("%s: Fortran-ordered %s array of shape %s" % (n, np.dtype(want), shapes[n]))
            setattr(f, n, a.ctypes.data)
    return f


def _body_structs(bodies):
    arr = (FcRestartBody * max(len(bodies), 1))()
    for k, b in enumerate(bodies):
        s = arr[k]
        s.nPts = b["xBodyMarker"].size
        s.nElem = b["triElemNeig"].shape[1]
        for n in BODY_ARRAYS:
            a = b[n]
            if a.dtype != np.float64 or a.shape != (s.nPts,) or not a.flags.c_contiguous:
                raise ValueError("%s: float64 array of %d markers" % (n, s.nPts))
            setattr(s, n, a.ctypes.data_as(_dp))
        t = b["triElemNeig"]
        if t.dtype != np.int32 or t.shape != (3, s.nElem) or not t.flags.f_contiguous:
            raise ValueError("triElemNeig: Fortran-ordered int32 (3, nElem)")
        s.triElemNeig = t.ctypes.data_as(_ip)
        sc = [float(b.get(n, 0.0)) for n in BODY_SCALARS]
        s.cent[:], s.ang[:], s.vcent[:], s.angv[:] = sc[0:3], sc[3:6], sc[6:9], sc[9:12]
    return arr


def write_restart(path, nx, ny, nz, time, ntime, flow, bodies=(), on_device=False):
    """write_restart: flow maps FLOW_NAMES to arrays; bodies is a list of
    dicts with BODY_ARRAYS, triElemNeig and BODY_SCALARS."""
    L = library()
    f = _flow_struct(nx, ny, nz, flow, on_device)
    b = _body_structs(list(bodies))
    _check(L, L.fc_restart_write(str(path).encode(), nx, ny, nz, time, ntime, C.byref(f), len(bodies), b),
           "fc_restart_write")


def read_restart(path, nx, ny, nz, flow, bodies=(), on_device=False):
    """read_restart into the given arrays (and the bodies' arrays); returns
    (time, ntime_start) and sets the bodies' scalars in their dicts."""
    L = library()
    f = _flow_struct(nx, ny, nz, flow, on_device)
    bodies = list(bodies)
    b = _body_structs(bodies)
    t, nt = C.c_double(), C.c_int()
    _check(L, L.fc_restart_read(str(path).encode(), nx, ny, nz, C.byref(t), C.byref(nt), C.byref(f), len(bodies), b),
           "fc_restart_read")
    for k, d in enumerate(bodies):
        sc = list(b[k].cent) + list(b[k].ang) + list(b[k].vcent) + list(b[k].angv)
        for n, v in zip(BODY_SCALARS, sc):
            d[n] = v
    return t.value, nt.value
