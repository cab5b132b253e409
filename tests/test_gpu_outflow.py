"""Zero-gradient (outflow) outer faces on the GPU against the oracle:
rhs_poisson with enforce_global_mass_consv (CORE_SET_OUTER_BC.F90:515-735)
and correct_ghost_massflux (GCM_GHOSTCELL_UPDATE.F90:742-787), and
solve_poisson with GCM_enforce_p_compatibility (GCM_GHOSTCELL_UPDATE.F90:
865-1003) adjusting the pgrad planes of the outflow face in every outer
iteration.

The GPU sums the net fluxes as a tree, the oracle in the reference's loop
order, so the corrections agree to rounding (1e-12 relative) rather than to
the bit.  The solve has a unique fixed point (pressure up to a constant and
the outflow pgrad plane), which both reach: 1e-9 relative L2 after mean
removal on the pressure, 1e-9 on the plane, in FP64 and with Anderson; 1e-8
with mixed precision."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _flag_interior(w):
    f = (w.iblank == 0) & (w.ghost == 0)
    m = np.zeros_like(f)
    m[1:-2, 1:-2, 1:-2] = True
    return f & m


def _mean_free_err(a, b, mask):
    x = a[mask] - a[mask].mean()
    y = b[mask] - b[mask].mean()
    return np.linalg.norm(x - y) / np.linalg.norm(y)


def _planes(nx, ny, nz, face, value):
    pl = [np.zeros(s, order="F") for s in [(ny + 2, nz + 2)] * 2 + [(nx + 2, nz + 2)] * 2 + [(nx + 2, ny + 2)] * 2]
    pl[face][:] = value
    return pl


def _vel_planes(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    return [np.asfortranarray(rng.standard_normal(s))
            for s in ((2, ny + 2, nz + 2), (nx + 2, 2, nz + 2), (nx + 2, ny + 2, 2))]


@pytest.fixture(scope="module")
def c1(geo):
    return geo.config_c1(16)


@pytest.mark.parametrize("mix", [0, 1])
def test_rhs_outflow(capi, geo, oracle, gpu, c1, mix):
    """Outflow faces at x = 1 and z = 0 (the sphere reaches neither)."""
    w = c1
    nx, ny, nz = (a.size - 2 for a in w.grid)
    bcv = [capi.BC_TYPE_DIRICHLET] * 6
    bcv[1] = capi.BC_TYPE_ZERO_GRADIENT
    bcv[4] = capi.BC_TYPE_ZERO_GRADIENT
    p = capi.default_params(bc_vel=bcv, mix_GC_form=mix)
    vb = _vel_planes(nx, ny, nz, 3)
    fu, fv, fw = geo.face_velocities(nx, ny, nz, seed=11)
    o = oracle.Oracle(capi, *w.grid, p)
    o.set_workload(w)
    o.set_vel_bc(*vb)
    fo = [fu.copy(order="F"), fv.copy(order="F"), fw.copy(order="F")]
    rc, d_o, s_o, r_o = o.rhs_poisson(*fo, 2.0)
    vb_o = o.get_vel_bc()
    area_o = o.outflow_area()
    o.close()
    s = gpu.PoissonSolver(*w.grid, p)
    s.set_workload(w)
    s.set_vel_bc(*vb)
    fg = [fu.copy(order="F"), fv.copy(order="F"), fw.copy(order="F")]
    d_g, s_g, r_g = s.rhs_poisson(*fg, 2.0)
    vb_g = s.get_vel_bc()
    area_g, corr_g = s.get_outflow()
    s.close()
    print("mix %d: corr %.6e sumd %.3e / %.3e" % (mix, corr_g, s_g, s_o), flush=True)
    assert rc == 0
    assert area_g == area_o and abs(area_g - 2.0) <= 1e-14
    # the correction is the only change of the face velocities and planes
    corr_o = fo[0][nx, 1, 1] - fu[nx, 1, 1]
    assert abs(corr_g - corr_o) <= 1e-12 * abs(corr_o)
    for a, b, c in zip(fg, fo, (fu, fv, fw)):
        assert np.max(np.abs(a - b)) <= 1e-12 * abs(corr_o)
        assert np.array_equal(a != c, b != c)
    for a, b in zip(vb_g, vb_o):
        assert np.max(np.abs(a - b)) <= 1e-12 * abs(corr_o)
    assert np.max(np.abs(d_g - d_o)) <= 1e-12 * np.max(np.abs(d_o))
    scale = np.max(np.abs(d_o)) * 1e-12
    assert abs(s_g - s_o) <= scale and abs(s_o) <= scale
    assert abs(r_g - r_o) <= 1e-12 * r_o


@pytest.mark.parametrize("rule,prec", [(0, 0), (1, 0), (1, 1)])
@pytest.mark.parametrize("mix", [0, 1])
def test_solve_outflow(capi, geo, oracle, gpu, c1, mix, rule, prec):
    """solve_poisson with an outflow face at x = 1: an incompatible initial
    pgrad plane (0.3 there) is corrected every outer iteration until the
    Neumann problem is compatible; the GPU (reference outer loop, or
    Anderson with the relative-residual stop, FP64 or mixed) reaches the
    oracle's pressure and plane."""
    w = c1
    nx, ny, nz = (a.size - 2 for a in w.grid)
    bcv = [capi.BC_TYPE_DIRICHLET] * 6
    bcv[1] = capi.BC_TYPE_ZERO_GRADIENT
    q = capi.default_params(restol_pson=1e-13, iterInter=1, itermax_pson=1500, restol_gcp=1e-12, itermax_gc=100,
                            bc_vel=bcv, mix_GC_form=mix)
    o = oracle.Oracle(capi, *w.grid, q, ordering=oracle.ORDER_REFERENCE)
    o.set_workload(w)
    o.set_bc(*_planes(nx, ny, nz, 1, 0.3))
    fu, fv, fw = geo.face_velocities(nx, ny, nz, seed=5)
    rc, div, _, _ = o.rhs_poisson(fu.copy(order="F"), fv.copy(order="F"), fw.copy(order="F"), 1.0)
    a = np.zeros_like(w.div)
    rc_o, st_o, _ = o.solve_poisson(a, div)
    pg_o = o.get_bc()
    o.close()
    assert rc_o == 0, st_o.as_dict()
    if rule == 0:
        p = capi.default_params(restol_pson=1e-12, iterInter=1, itermax_pson=1500, restol_gcp=1e-11, itermax_gc=100,
                                bc_vel=bcv, mix_GC_form=mix)
    else:
        p = capi.default_params(restol_pson=0.0, iterInter=0, itermax_pson=600, restol_gcp=1e-11, itermax_gc=100,
                                bc_vel=bcv, mix_GC_form=mix, precision=prec, outer_rule=capi.OUTER_RELRES,
                                rel_tol=1e-11 if prec == 0 else 1e-10, vc_walls=0 if prec == 0 else 1)
    s = gpu.PoissonSolver(*w.grid, p)
    s.set_workload(w)
    s.set_bc(*_planes(nx, ny, nz, 1, 0.3))
    d_g, _, _ = s.rhs_poisson(fu.copy(order="F"), fv.copy(order="F"), fw.copy(order="F"), 1.0)
    assert np.max(np.abs(d_g - div)) <= 1e-12 * np.max(np.abs(div))
    b = np.zeros_like(w.div)
    rc, st, _ = s.solve_poisson(b, div)
    pg_g = s.get_bc()
    s.close()
    print("mix %d rule %d prec %d: outer %d (oracle %d), rel %.2e" % (mix, rule, prec, st.outer_iters,
                                                                     st_o.outer_iters, st.rel_resid), flush=True)
    assert rc == 0 and st.converged == 1, st.as_dict()
    tol = 1e-9 if prec == 0 else 1e-8
    f = _flag_interior(w)
    assert _mean_free_err(b, a, f) < tol
    # the plane moved from 0.3 and both reach the same value on 0..n
    x2o, x2g = pg_o[1][0:ny + 1, 0:nz + 1], pg_g[1][0:ny + 1, 0:nz + 1]
    assert abs(x2o[0, 0] - 0.3) > 1e-3
    assert np.max(np.abs(x2g - x2o)) <= tol * max(1.0, np.max(np.abs(x2o)))
    for k in (0, 2, 3, 4, 5):
        assert np.array_equal(pg_g[k], pg_o[k])
