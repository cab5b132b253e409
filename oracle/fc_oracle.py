"""ctypes wrapper of the CPU oracle (oracle/_build/libfc_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker, never by the product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libfc_oracle.so")

ORDER_REFERENCE = 0
ORDER_COLORED = 1


def build(force=False):
    if force or not os.path.exists(LIB):
        subprocess.check_call(["make", "-s", "-C", HERE] + (["-B"] if force else []))
    return LIB


_lib = None


def lib(capi):
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB)
        vp = C.c_void_p
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        bp = C.POINTER(C.c_int8)
        L.or_create.restype = vp
        L.or_create.argtypes = [C.POINTER(capi.FcGrid), C.POINTER(capi.FcParams), C.c_int]
        L.or_destroy.argtypes = [vp]
        L.or_set_params.argtypes = [vp, C.POINTER(capi.FcParams), C.c_int]
        L.or_levels.argtypes = [vp, ip, ip, ip, ip]
        L.or_set_masks.argtypes = [vp, bp, bp, bp]
        L.or_set_gcm.argtypes = [vp, C.c_int, ip, ip, ip, ip, ip, ip, dp, dp, dp,
                                 C.c_int, ip, ip, ip, ip, ip, ip, dp, dp, dp, dp]
        L.or_set_bc.argtypes = [vp, dp, dp, dp, dp, dp, dp, C.c_double]
        L.or_solve_poisson.argtypes = [vp, dp, dp, dp, dp, C.POINTER(capi.FcStats)]
        L.or_mg_solve.argtypes = [vp, dp, dp, C.POINTER(capi.FcStats)]
        L.or_itsolv.argtypes = [vp, dp, dp]
        L.or_ghostcell_pressure.argtypes = [vp, dp, dp, dp, dp, C.POINTER(capi.FcStats)]
        L.or_calc_residual.restype = C.c_double
        L.or_calc_residual.argtypes = [vp, dp, dp]
        L.or_calc_residual2.restype = C.c_double
        L.or_calc_residual2.argtypes = [vp, dp, dp]
        L.or_rel_residual.restype = C.c_double
        L.or_rel_residual.argtypes = [vp, dp, dp]
        L.or_rhs_poisson.argtypes = [vp, dp, dp, dp, C.c_double, dp, dp, dp]
        L.or_enforce_p_compatibility.argtypes = [vp, dp]
        L.or_get_bc.argtypes = [vp, dp, dp, dp, dp, dp, dp]
        L.or_get_vel_bc.argtypes = [vp, dp, dp, dp]
        L.or_outflow_area.restype = C.c_double
        L.or_outflow_area.argtypes = [vp]
        L.or_level_ptr.restype = vp
        L.or_level_ptr.argtypes = [vp, C.c_int, C.c_int]
        L.or_level_metric.argtypes = [vp, C.c_int, C.c_int, C.POINTER(dp)]
        for f in ("or_itsolv_mg", "or_bc_fill", "or_residual_mg", "or_mg_inject", "or_mg_interp"):
            getattr(L, f).argtypes = [vp, C.c_int]
        L.or_sweep.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.or_vcycle.argtypes = [vp]
        L.or_tdma.argtypes = [dp, dp, dp, dp, dp, C.c_int, C.c_int]
        L.or_set_vel_bc.argtypes = [vp, dp, dp, dp]
        L.or_correct_vel.argtypes = [vp, dp, C.c_double, dp, dp, dp]
        L.or_correct_ghostcell_velocity.argtypes = [vp, dp, C.c_double, dp, dp, dp]
        L.or_face_vel.argtypes = [vp, dp, dp, dp, dp, dp, dp]
        L.or_update_pressure.argtypes = [vp, dp, dp]
        L.or_set_solve_ad.argtypes = [vp, C.c_double, C.c_double, C.c_int, C.c_double]
        L.or_get_solve_ad.argtypes = [vp, dp, dp, dp, dp, dp, dp]
        L.or_itsolv_ad.argtypes = [vp, dp, dp, dp, dp, dp]
        L.or_calc_residual_ad.restype = C.c_double
        L.or_calc_residual_ad.argtypes = [vp, dp, dp, dp, dp, dp, ip, dp]
        L.or_rhs_advec_diff.argtypes = [vp, dp, dp, dp, dp, dp, dp, dp, C.c_double, C.c_double, dp, dp, dp]
        L.or_set_outer_velocity.argtypes = [vp, dp, dp, C.c_double, C.c_double]
        L.or_set_outer_velocity_bc.argtypes = [vp, dp, dp, dp, C.c_double]
        L.or_set_outer_ghost_velocity.argtypes = [vp, dp, dp, dp]
        L.or_get_outer_velocity_bc.argtypes = [vp, C.POINTER(dp)]
        L.or_set_gcm_vel.argtypes = [vp, dp, dp, dp, dp, C.c_double]
        L.or_ghostcell_velocity.argtypes = [vp, C.c_int, dp, dp, dp, dp, dp, dp, dp, dp, dp, C.c_double, ip, dp]
        L.or_solve_ad.argtypes = [vp, dp, dp, dp, dp, dp, dp, dp, dp, dp, C.c_double, C.c_int, C.c_double,
                                  C.c_double, ip, dp, ip]
        _lib = L
    return _lib


class Oracle:
    """FP64 CPU restatement of solve_poisson / mg_solve / GCM update."""

    def __init__(self, capi, x, y, z, params, ordering=ORDER_REFERENCE):
        self.capi = capi
        self.L = lib(capi)
        self.x, self.y, self.z = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z)]
        self.nodes = (self.x.size - 2, self.y.size - 2, self.z.size - 2)
        g = capi.FcGrid(self.nodes[0], self.nodes[1], self.nodes[2],
                        capi.dptr(self.x), capi.dptr(self.y), capi.dptr(self.z))
        self.params = params
        self.h = self.L.or_create(C.byref(g), C.byref(params), ordering)
        nlev = C.c_int()
        a, b, c = (C.c_int * 10)(), (C.c_int * 10)(), (C.c_int * 10)()
        self.L.or_levels(self.h, C.byref(nlev), a, b, c)
        self.nlev = nlev.value
        self.level_nodes = [(a[i], b[i], c[i]) for i in range(self.nlev)]

    def close(self):
        if self.h:
            self.L.or_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params, ordering):
        self.params = params
        self.L.or_set_params(self.h, C.byref(params), ordering)

    def set_masks(self, iblank, ghost, dead=None):
        capi = self.capi
        self._masks = (np.asfortranarray(iblank, dtype=np.int8), np.asfortranarray(ghost, dtype=np.int8),
                       None if dead is None else np.asfortranarray(dead, dtype=np.int8))
        self.L.or_set_masks(self.h, capi.bptr(self._masks[0]), capi.bptr(self._masks[1]),
                            capi.bptr(self._masks[2]) if dead is not None else None)

    def set_gcm(self, g):
        capi = self.capi
        self._g = {k: np.ascontiguousarray(v) for k, v in g.items()}
        g = self._g
        nG, nD = g["iGhost"].size, g["iDead"].size
        self.L.or_set_gcm(self.h, nG, capi.iptr(g["iGhost"]), capi.iptr(g["jGhost"]), capi.iptr(g["kGhost"]),
                          capi.iptr(g["iCellIndex"]), capi.iptr(g["jCellIndex"]), capi.iptr(g["kCellIndex"]),
                          capi.dptr(g["coeffGCMN"]), capi.dptr(g["alphaGhost"]), capi.dptr(g["dpdnBodyIntercept"]),
                          nD, capi.iptr(g["iDead"]), capi.iptr(g["jDead"]), capi.iptr(g["kDead"]),
                          capi.iptr(g["iDeadCellIndex"]), capi.iptr(g["jDeadCellIndex"]),
                          capi.iptr(g["kDeadCellIndex"]), capi.dptr(g["coeffGCMDeadN"]),
                          capi.dptr(g["probeLengthDead"]), capi.dptr(g["dpdnInterceptDead"]),
                          capi.dptr(g["xyzDCIP"]))

    def set_workload(self, w):
        self.set_masks(w.iblank, w.ghost, w.dead)
        self.set_gcm(w.gcm)

    def set_bc(self, pgx1=None, pgx2=None, pgy1=None, pgy2=None, pgz1=None, pgz2=None, outflow_area=1.0):
        capi = self.capi
        self._bc = [None if a is None else np.asfortranarray(a, dtype=np.float64)
                    for a in (pgx1, pgx2, pgy1, pgy2, pgz1, pgz2)]
        self.L.or_set_bc(self.h, *[capi.plane_ptr(a) for a in self._bc], outflow_area)

    def get_bc(self):
        """the six pgrad planes as the oracle holds them"""
        nx, ny, nz = self.nodes
        shp = [(ny + 2, nz + 2)] * 2 + [(nx + 2, nz + 2)] * 2 + [(nx + 2, ny + 2)] * 2
        out = [np.zeros(s, order="F") for s in shp]
        self.L.or_get_bc(self.h, *[self.capi.plane_ptr(a) for a in out])
        return out

    def get_vel_bc(self):
        """bcxu(0:1,0:ny+1,0:nz+1), bcyv(0:nx+1,0:1,0:nz+1), bczw(0:nx+1,0:ny+1,0:1)"""
        nx, ny, nz = self.nodes
        out = [np.zeros((2, ny + 2, nz + 2), order="F"), np.zeros((nx + 2, 2, nz + 2), order="F"),
               np.zeros((nx + 2, ny + 2, 2), order="F")]
        self.L.or_get_vel_bc(self.h, *[self.capi.dptr(a) for a in out])
        return out

    def outflow_area(self):
        return self.L.or_outflow_area(self.h)

    def enforce_p_compatibility(self, pres):
        """GCM_enforce_p_compatibility(pres): adjusts the pgrad planes of the
        zero-gradient faces (uses res_laplacian of the last GCM update)"""
        self.L.or_enforce_p_compatibility(self.h, self.capi.dptr(pres))

    def solve_poisson(self, pPrime, div, p_prev=None):
        capi = self.capi
        st = capi.FcStats()
        nG = self._g["iGhost"].size if hasattr(self, "_g") else 0
        reslap = np.zeros(max(nG, 1))
        rc = self.L.or_solve_poisson(self.h, capi.dptr(pPrime), capi.dptr(div),
                                     capi.dptr(p_prev) if p_prev is not None else None,
                                     capi.dptr(reslap), C.byref(st))
        return rc, st, reslap[:nG]

    def mg_solve(self, var, rr):
        st = self.capi.FcStats()
        rc = self.L.or_mg_solve(self.h, self.capi.dptr(var), self.capi.dptr(rr), C.byref(st))
        return rc, st

    def itsolv(self, var, r):
        return self.L.or_itsolv(self.h, self.capi.dptr(var), self.capi.dptr(r))

    def ghostcell_pressure(self, pres, r, p_prev=None):
        capi = self.capi
        st = capi.FcStats()
        nG = self._g["iGhost"].size
        reslap = np.zeros(max(nG, 1))
        rc = self.L.or_ghostcell_pressure(self.h, capi.dptr(pres), capi.dptr(r),
                                          capi.dptr(p_prev) if p_prev is not None else None,
                                          capi.dptr(reslap), C.byref(st))
        return rc, st, reslap[:nG]

    def calc_residual(self, var, r):
        return self.L.or_calc_residual(self.h, self.capi.dptr(var), self.capi.dptr(r))

    def calc_residual2(self, var, p_temp):
        return self.L.or_calc_residual2(self.h, self.capi.dptr(var), self.capi.dptr(p_temp))

    def rel_residual(self, var, r):
        return self.L.or_rel_residual(self.h, self.capi.dptr(var), self.capi.dptr(r))

    def rhs_poisson(self, fu, fv, fw, dtinv=1.0):
        capi = self.capi
        div = np.zeros_like(fu)
        sumd = C.c_double()
        rd = C.c_double()
        rc = self.L.or_rhs_poisson(self.h, capi.dptr(fu), capi.dptr(fv), capi.dptr(fw), dtinv,
                                   capi.dptr(div), C.byref(sumd), C.byref(rd))
        return rc, div, sumd.value, rd.value

    # velocity projection (CORE_TIME_STEP.F90:211-222); arrays in place ------
    def set_vel_bc(self, bcxu=None, bcyv=None, bczw=None):
        self._vbc = [None if a is None else np.asfortranarray(a, dtype=np.float64) for a in (bcxu, bcyv, bczw)]
        self.L.or_set_vel_bc(self.h, *[self.capi.dptr(a) for a in self._vbc])

    def correct_vel(self, pPrime, dt, u, v, w):
        d = self.capi.dptr
        self.L.or_correct_vel(self.h, d(pPrime), dt, d(u), d(v), d(w))

    def correct_ghostcell_velocity(self, pVar, dt_fac, u, v, w):
        d = self.capi.dptr
        self.L.or_correct_ghostcell_velocity(self.h, d(pVar), dt_fac, d(u), d(v), d(w))

    def face_vel(self, u, v, w, fu, fv, fw):
        d = self.capi.dptr
        self.L.or_face_vel(self.h, d(u), d(v), d(w), d(fu), d(fv), d(fw))

    def update_pressure(self, pPrime, p):
        d = self.capi.dptr
        self.L.or_update_pressure(self.h, d(pPrime), d(p))

    # advection-diffusion line solves (CORE_SOLVE_AD.F90) ---------------
    def set_solve_ad(self, reinv, dt, crank_nicolson=1, omega_ad=1.0):
        self.L.or_set_solve_ad(self.h, reinv, dt, crank_nicolson, omega_ad)

    def get_solve_ad(self):
        nx, ny, nz = self.nodes
        out = [np.zeros(n + 2) for n in (nx, nx, ny, ny, nz, nz)]
        self.L.or_get_solve_ad(self.h, *[self.capi.dptr(a) for a in out])
        return out

    def itsolv_ad(self, var, r, fu, fv, fw):
        d = self.capi.dptr
        self.L.or_itsolv_ad(self.h, d(var), d(r), d(fu), d(fv), d(fw))

    def calc_residual_ad(self, var, r, fu, fv, fw):
        d = self.capi.dptr
        loc = (C.c_int * 3)()
        r2 = C.c_double()
        resm = self.L.or_calc_residual_ad(self.h, d(var), d(r), d(fu), d(fv), d(fw), loc, C.byref(r2))
        return resm, tuple(loc), r2.value

    def rhs_advec_diff(self, uo, vo, wo, fu, fv, fw, alfa_upwind, dt, nlu, nlv, nlw, visc=None):
        """rhs_advec_diff (CORE_SET_SOLVERS.F90:10-61): writes the interior
        of nlu, nlv, nlw; visc = viscTot (None: the reinv of set_solve_ad)"""
        d = self.capi.dptr
        return self.L.or_rhs_advec_diff(self.h, d(uo), d(vo), d(wo), d(fu), d(fv), d(fw), d(visc), alfa_upwind, dt,
                                        d(nlu), d(nlv), d(nlw))

    # solve_ad's loop (CORE_SOLVE_AD.F90:11-127) ---------------------------
    def set_outer_velocity(self, vals=None, freq=None, y_origin=0.0, y_out=1.0):
        """face values and frequencies ux1, ux2, vx1, vx2, wx1, wx2, uy1, ...,
        wz2 and the channel profile bounds of set_outer_velocity_bc"""
        d = self.capi.dptr
        self._vals = None if vals is None else np.ascontiguousarray(vals, dtype=np.float64)
        self._freq = None if freq is None else np.ascontiguousarray(freq, dtype=np.float64)
        return self.L.or_set_outer_velocity(self.h, d(self._vals), d(self._freq), y_origin, y_out)

    def set_outer_velocity_bc(self, u, v, w, time=0.0):
        d = self.capi.dptr
        return self.L.or_set_outer_velocity_bc(self.h, d(u), d(v), d(w), time)

    def set_outer_ghost_velocity(self, u, v, w):
        d = self.capi.dptr
        return self.L.or_set_outer_ghost_velocity(self.h, d(u), d(v), d(w))

    def get_outer_velocity_bc(self):
        """bcxu, bcxv, bcxw (2, ny+2, nz+2), bcyu, bcyv, bcyw (nx+2, 2, nz+2),
        bczu, bczv, bczw (nx+2, ny+2, 2)"""
        nx, ny, nz = self.nodes
        shp = [(2, ny + 2, nz + 2)] * 3 + [(nx + 2, 2, nz + 2)] * 3 + [(nx + 2, ny + 2, 2)] * 3
        out = [np.zeros(sh, order="F") for sh in shp]
        arr = (C.POINTER(C.c_double) * 9)(*[self.capi.dptr(a) for a in out])
        self.L.or_get_outer_velocity_bc(self.h, arr)
        return out

    def set_gcm_vel(self, coeffD, coeffDeadD, ubi, ubiDead, probe_norm_d):
        """coeffGCMD (nGhost, 8), coeffGCMDeadD (nDead, 8), body-intercept
        velocities (nGhost, 3) and (nDead, 3), probeLengthNormalizedD"""
        self._gv = [np.ascontiguousarray(a, dtype=np.float64) for a in (coeffD, coeffDeadD, ubi, ubiDead)]
        d = self.capi.dptr
        return self.L.or_set_gcm_vel(self.h, *[d(a) for a in self._gv], probe_norm_d)

    def ghostcell_velocity(self, u, v, w, nlu=None, nlv=None, nlw=None, fu=None, fv=None, fw=None,
                           restol_gcv=1e-10, iopt=2):
        d = self.capi.dptr
        sw = C.c_int()
        rm = C.c_double()
        rc = self.L.or_ghostcell_velocity(self.h, iopt, d(u), d(v), d(w), d(nlu), d(nlv), d(nlw), d(fu), d(fv),
                                          d(fw), restol_gcv, C.byref(sw), C.byref(rm))
        return rc, sw.value, rm.value

    def solve_ad(self, u, v, w, nlu, nlv, nlw, fu, fv, fw, time=0.0, itermax_ad=200, restol_ad=1e-10,
                 restol_gcv=1e-12):
        d = self.capi.dptr
        it = C.c_int()
        res = C.c_double()
        loc = (C.c_int * 3)()
        rc = self.L.or_solve_ad(self.h, d(u), d(v), d(w), d(nlu), d(nlv), d(nlw), d(fu), d(fv), d(fw), time,
                                itermax_ad, restol_ad, restol_gcv, C.byref(it), C.byref(res), loc)
        return rc, it.value, res.value, tuple(loc)

    # level access ------------------------------------------------------
    def level(self, lev, which):
        """numpy view (Fortran order) of phi/rhs/res/ibk/igk of level lev."""
        nx, ny, nz = self.level_nodes[lev - 1]
        ptr = self.L.or_level_ptr(self.h, lev, which)
        dt = np.float64 if which < 3 else np.int8
        n = (nx + 2) * (ny + 2) * (nz + 2)
        buf = (C.c_double * n).from_address(ptr) if which < 3 else (C.c_int8 * n).from_address(ptr)
        return np.frombuffer(buf, dtype=dt).reshape((nx + 2, ny + 2, nz + 2), order="F")

    def level_metric(self, lev, which):
        """x,y,z,xc,yc,zc,dxinv,dyinv,dzinv,dxcinv,dycinv,dzcinv (0..11)"""
        nx, ny, nz = self.level_nodes[lev - 1]
        n = (nx, ny, nz)[which % 3] + 2
        p = C.POINTER(C.c_double)()
        self.L.or_level_metric(self.h, lev, which, C.byref(p))
        return np.ctypeslib.as_array(p, shape=(n,)).copy()

    def itsolv_mg(self, lev):
        self.L.or_itsolv_mg(self.h, lev)

    def sweep(self, lev, d, color):
        self.L.or_sweep(self.h, lev, d, color)

    def bc_fill(self, lev):
        self.L.or_bc_fill(self.h, lev)

    def residual_mg(self, lev):
        self.L.or_residual_mg(self.h, lev)

    def mg_inject(self, lev):
        self.L.or_mg_inject(self.h, lev)

    def mg_interp(self, lev):
        self.L.or_mg_interp(self.h, lev)

    def vcycle(self):
        self.L.or_vcycle(self.h)


def tdma(capi, a, b, c, r, n1, n2):
    L = lib(capi)
    a, b, c, r = [np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c, r)]
    u = np.zeros_like(r)
    L.or_tdma(capi.dptr(a), capi.dptr(b), capi.dptr(c), capi.dptr(r), capi.dptr(u), n1, n2)
    return u
