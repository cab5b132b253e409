"""ctypes wrapper of the geometry-refresh restatement in the CPU oracle
(oracle/_build/libfc_oracle.so, og_* in fc_geom_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/ and the cpu_baseline leg of
bench.py as the checker, never by the product path.
"""
import ctypes as C

import numpy as np

import fc_oracle

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_bp = C.POINTER(C.c_int8)
_set = False


def _lib(capi, geom):
    global _set
    L = fc_oracle.lib(capi)
    if not _set:
        vp = C.c_void_p
        i = C.c_int
        L.og_create.restype = vp
        L.og_create.argtypes = [C.POINTER(capi.FcGrid), C.POINTER(geom.FcGeomParams)]
        L.og_destroy.argtypes = [vp]
        L.og_set_body.argtypes = [vp, i, i, i, _dp, _dp, _dp, _ip]
        L.og_move_body.argtypes = [vp, i, _dp, _dp, _dp]
        L.og_refresh.argtypes = [vp, C.POINTER(geom.FcGeomStats)]
        L.og_get_counts.argtypes = [vp, _ip, _ip]
        L.og_get_masks.argtypes = [vp] + [_bp] * 6
        L.og_get_ghosts.argtypes = [vp, C.POINTER(geom.FcGeomGhosts)]
        L.og_get_dead.argtypes = [vp, C.POINTER(geom.FcGeomDead)]
        L.og_get_surface.argtypes = [vp, i, C.POINTER(geom.FcGeomSurface)]
        L.og_get_dcip.argtypes = [vp, _dp]
        L.og_body_intercept.argtypes = [vp, i, i, i, i, _dp, _ip, _dp, _dp]
        L.og_van_matrix.argtypes = [vp, i, i, i, i, i, i, _dp, _dp, _dp, _dp, _dp]
        L.og_set_marker_velocity.argtypes = [vp, i] + [_dp] * 6
        L.og_body_intercept_values.argtypes = [vp, C.c_double, i]
        L.og_get_bi_values.argtypes = [vp] + [_dp] * 8
        L.og_marker_stress.argtypes = [vp, _dp, _dp, _dp, _dp, C.c_double, _dp]
        L.og_get_marker_forces.argtypes = [vp, i, C.POINTER(geom.FcGeomForces)]
        _set = True
    return L


class GeomOracle:
    """Same methods as geom.GeometryRefresh, on the CPU restatement."""

    def __init__(self, capi, geom, x, y, z, surfaces, membrane_tkns=0.0, probe=2.0, probeD=2.0):
        self.geom = geom
        self.L = _lib(capi, geom)
        self.grid_arrays = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z)]
        g = capi.FcGrid()
        g.nx, g.ny, g.nz = (a.size - 2 for a in self.grid_arrays)
        g.x, g.y, g.z = (a.ctypes.data_as(_dp) for a in self.grid_arrays)
        self.shape = (g.nx + 2, g.ny + 2, g.nz + 2)
        self.surfaces = list(surfaces)
        p = geom.make_params(self.surfaces, membrane_tkns, probe, probeD)
        self.h = self.L.og_create(C.byref(g), C.byref(p))
        if not self.h:
            raise ValueError("og_create refused the arguments")
        for b, s in enumerate(self.surfaces):
            rc = self.L.og_set_body(self.h, b + 1, s.npts, s.nelem, s.x.ctypes.data_as(_dp), s.y.ctypes.data_as(_dp),
                                    s.z.ctypes.data_as(_dp), s.tri.ctypes.data_as(_ip))
            if rc != 0:
                raise ValueError("og_set_body: %d" % rc)

    def move_body(self, ibody, s):
        self.surfaces[ibody - 1] = s
        self.L.og_move_body(self.h, ibody, s.x.ctypes.data_as(_dp), s.y.ctypes.data_as(_dp), s.z.ctypes.data_as(_dp))

    def refresh(self):
        st = self.geom.FcGeomStats()
        rc = self.L.og_refresh(self.h, C.byref(st))
        if rc < 0:
            raise ValueError("og_refresh: %d" % rc)
        return st

    def results(self):
        L, h = self.L, self.h
        return self.geom.collect(lambda a, b: L.og_get_counts(h, a, b),
                                 lambda *m: L.og_get_masks(h, *m),
                                 lambda g: L.og_get_ghosts(h, g),
                                 lambda d: L.og_get_dead(h, d),
                                 lambda b, f: L.og_get_surface(h, b, f),
                                 lambda p: L.og_get_dcip(h, p), self.shape, self.surfaces)

    def handoff(self):
        r = self.results()
        out = {k: r[k] for k in ["nGhost", "nDead", "iblank", "ghostCellMark", "dead_cell", "alphaGhost",
                                 "probeLengthDead"] + self.geom.HANDOFF_GHOST_INT + self.geom.HANDOFF_DEAD_INT
               + self.geom.GHOST_COEF + self.geom.DEAD_COEF}
        out["xyzDCIP_dead"] = np.ascontiguousarray(r["xyzDCIP"][r["iDead"], r["jDead"], r["kDead"], :])
        return out

    def set_marker_velocity(self, ibody, u, v, w, uOld, vOld, wOld):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (u, v, w, uOld, vOld, wOld)]
        self.L.og_set_marker_velocity(self.h, ibody, *[x.ctypes.data_as(_dp) for x in a])

    def body_intercept_values(self, dt, ntime):
        rc = self.L.og_body_intercept_values(self.h, dt, ntime)
        if rc < 0:
            raise ValueError("og_body_intercept_values: %d" % rc)

    def marker_stress(self, p, u, v, w, reinv):
        f = [np.asfortranarray(a, dtype=np.float64) for a in (p, u, v, w)]
        tot = np.zeros(9 * len(self.surfaces))
        rc = self.L.og_marker_stress(self.h, *[a.ctypes.data_as(_dp) for a in f], reinv, tot.ctypes.data_as(_dp))
        if rc < 0:
            raise ValueError("og_marker_stress: %d" % rc)
        self._totals = tot
        return tot.reshape(-1, 9)

    def bi_values(self):
        L, h = self.L, self.h
        nG, nD = C.c_int(), C.c_int()
        L.og_get_counts(h, C.byref(nG), C.byref(nD))
        return self.geom.collect_bi(lambda *a: L.og_get_bi_values(h, *a), nG.value, nD.value)

    def stress_results(self):
        L, h = self.L, self.h
        nG, nD = C.c_int(), C.c_int()
        L.og_get_counts(h, C.byref(nG), C.byref(nD))
        return self.geom.collect_stress(lambda *a: L.og_get_bi_values(h, *a),
                                        lambda b, f: L.og_get_marker_forces(h, b, f), nG.value, nD.value,
                                        self.surfaces, self._totals)

    def body_intercept(self, ibody, i, j, k):
        bi = np.zeros(3)
        el, dist, dn = C.c_int(), C.c_double(), C.c_double()
        self.L.og_body_intercept(self.h, ibody, i, j, k, bi.ctypes.data_as(_dp), C.byref(el), C.byref(dist),
                                 C.byref(dn))
        return bi, el.value, dist.value, dn.value

    def van_matrix(self, iG, jG, kG, iC, jC, kC, ev, bi, bn):
        cD, cN = np.zeros(8), np.zeros(8)
        args = [np.ascontiguousarray(a, dtype=np.float64) for a in (ev, bi, bn)]
        self.L.og_van_matrix(self.h, iG, jG, kG, iC, jC, kC, *[a.ctypes.data_as(_dp) for a in args],
                             cD.ctypes.data_as(_dp), cN.ctypes.data_as(_dp))
        return cD, cN

    def close(self):
        if self.h:
            self.L.og_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
