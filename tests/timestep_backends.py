"""Backends of feicar3d_amd.timestep.FractionalStep for the tests: the HIP
library's PoissonSolver / GeometryRefresh, and the CPU oracle's Oracle /
GeomOracle behind the same method names (the oracle is the checker here,
never the product)."""
import numpy as np

import fc_geom_oracle
import fc_oracle

VALS = np.array([1.0, 1.0, 0.0, 0.0, 0.0, 0.0] * 3)     # u = 1, v = w = 0 on every face


class OracleFlow(fc_oracle.Oracle):
    def rhs_poisson(self, fu, fv, fw, dtinv=1.0):
        rc, div, sumd, rd = super().rhs_poisson(fu, fv, fw, dtinv)
        assert rc >= 0
        return div, sumd, rd

    def ghostcell_velocity_host(self, u, v, w, iopt=2, restol_gcv=1e-10):
        return self.ghostcell_velocity(u, v, w, restol_gcv=restol_gcv, iopt=iopt)

    def vel_set_outer(self, vals=None, freq=None, y_origin=0.0, y_out=1.0):
        return self.set_outer_velocity(vals, freq, y_origin, y_out)


def params(capi, gpu):
    """all velocity faces Dirichlet (u = 1), pressure Dirichlet (0) on the
    x = 1 face and Neumann elsewhere.  The oracle and gpu = "fp64" / "mixed"
    (FP32 V-cycles in FP64 defect correction) run the reference's outer loop
    to its fixed point; "accelerated" / "accelerated_mixed" run the
    Anderson-accelerated loop to a relative residual of 1e-12 / 1e-11 (the
    Dirichlet face: nothing is pinned)."""
    pbc = [capi.PBC_NEUMANN, capi.PBC_DIRICHLET] + [capi.PBC_NEUMANN] * 4
    mixed = gpu in ("mixed", "accelerated_mixed")
    prec = capi.PREC_MIXED if mixed else capi.PREC_FP64
    if gpu in ("accelerated", "accelerated_mixed"):
        return capi.default_params(bc_vel=[1] * 6, pbc=pbc, restol_pson=0.0, iterInter=0, itermax_pson=400,
                                   restol_gcp=1e-12, itermax_gc=200, outer_rule=capi.OUTER_RELRES,
                                   rel_tol=1e-11 if mixed else 1e-12, precision=prec)
    return capi.default_params(bc_vel=[1] * 6, pbc=pbc, restol_pson=1e-14, iterInter=1, itermax_pson=4000,
                               restol_gcp=1e-13, itermax_gc=200, precision=prec)


def make(capi, geom, ts, x, y, z, bodies, tk, gpu, gpu_mod=None, dt=0.01, reinv=1.0 / 200.0, resident=False):
    """A FractionalStep on the GPU (gpu=True, gpu_mod = the poisson module)
    or on the oracle, with the flow at u = 1."""
    sh = (x.size, y.size, z.size)
    p = params(capi, gpu)
    if gpu:
        flow = gpu_mod.PoissonSolver(x, y, z, p)
        g = geom.GeometryRefresh(x, y, z, bodies, membrane_tkns=tk)
    else:
        flow = OracleFlow(capi, x, y, z, p, ordering=fc_oracle.ORDER_REFERENCE)
        g = fc_geom_oracle.GeomOracle(capi, geom, x, y, z, bodies, membrane_tkns=tk)
    flow.vel_set_outer(VALS)
    vbc = (np.ones((2, sh[1], sh[2])), np.zeros((sh[0], 2, sh[2])), np.zeros((sh[0], sh[1], 2)))
    return ts.FractionalStep(flow, g, sh, reinv, dt, 2.0, itermax_ad=400, restol_ad=1e-13, restol_gcv=1e-14,
                             u0=(1.0, 0.0, 0.0), vel_bc=vbc, resident=resident)
