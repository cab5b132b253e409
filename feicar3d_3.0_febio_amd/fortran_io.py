"""Binary case/result files of the Fortran driver (fortran/fc_driver.f90):
little-endian stream files in the order the driver reads them."""
import numpy as np

from . import capi


def write_case(path, w, params, p_prev=None, p0=None):
    x, y, z = w.grid
    nx, ny, nz = x.size - 2, y.size - 2, z.size - 2
    g = w.gcm
    with open(path, "wb") as f:
        np.array([nx, ny, nz], dtype="<i4").tofile(f)
        for a in (x, y, z):
            np.asarray(a, dtype="<f8").tofile(f)
        f.write(bytes(params))
        for m in (w.iblank, w.ghost, w.dead):
            np.asarray(m, dtype=np.int8).ravel(order="F").tofile(f)
        nG = g["iGhost"].size
        np.array([nG], dtype="<i4").tofile(f)
        if nG:
            for k in ("iGhost", "jGhost", "kGhost", "iCellIndex", "jCellIndex", "kCellIndex"):
                np.asarray(g[k], dtype="<i4").tofile(f)
            for k in ("coeffGCMN", "alphaGhost", "dpdnBodyIntercept"):
                np.asarray(g[k], dtype="<f8").ravel().tofile(f)
        nD = g["iDead"].size
        np.array([nD], dtype="<i4").tofile(f)
        if nD:
            for k in ("iDead", "jDead", "kDead", "iDeadCellIndex", "jDeadCellIndex", "kDeadCellIndex"):
                np.asarray(g[k], dtype="<i4").tofile(f)
            for k in ("coeffGCMDeadN", "probeLengthDead", "dpdnInterceptDead", "xyzDCIP"):
                np.asarray(g[k], dtype="<f8").ravel().tofile(f)
        shp = w.div.shape
        for a in (w.div, np.zeros(shp) if p_prev is None else p_prev, np.zeros(shp) if p0 is None else p0):
            np.asarray(a, dtype="<f8").ravel(order="F").tofile(f)


def read_result(path, shape):
    n = int(np.prod(shape))
    raw = open(path, "rb").read()
    p = np.frombuffer(raw[:8 * n], dtype="<f8").reshape(shape, order="F").copy(order="F")
    st = capi.FcStats.from_buffer_copy(raw[8 * n:8 * n + C_sizeof_stats()])
    return p, st


def C_sizeof_stats():
    import ctypes
    return ctypes.sizeof(capi.FcStats)
