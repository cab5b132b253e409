"""Periodic x and y outer faces on the GPU against the oracle
(set_outer_pressure_bc's periodic branch, CORE_SET_OUTER_BC.F90:1007-1026;
the oracle's fill is pinned in test_oracle_pins.py): the boundary fill to the
bit, half sweeps in the GPU ordering to 1e-12, and solve_poisson to the
oracle's fixed point (1e-10 relative L2 after mean removal in FP64, 1e-8
with mixed precision).  Periodic z is an error, as in the reference."""
import numpy as np
import pytest

from conftest import no_edges

pytestmark = pytest.mark.gpu


def _params(capi, axes=(0, 1), **kw):
    bcv = [capi.BC_TYPE_DIRICHLET] * 6
    for a in axes:
        bcv[2 * a] = bcv[2 * a + 1] = capi.BC_TYPE_PERIODIC
    return capi.default_params(bc_vel=bcv, **kw)


def _rand(shape, seed):
    return np.asfortranarray(np.random.default_rng(seed).standard_normal(shape))


@pytest.fixture(scope="module")
def small(geo):
    return geo.config_c2(24)


def _pair(capi, oracle, gpu, w, p):
    o = oracle.Oracle(capi, *w.grid, p, ordering=oracle.ORDER_COLORED)
    o.set_workload(w)
    s = gpu.PoissonSolver(*w.grid, p)
    s.set_workload(w)
    return o, s


@pytest.mark.parametrize("axes", [(0,), (1,), (0, 1)])
@pytest.mark.parametrize("lev", [1, 2])
def test_periodic_bc_fill(capi, oracle, gpu, small, lev, axes):
    o, s = _pair(capi, oracle, gpu, small, _params(capi, axes))
    phi = _rand(s.level_shape(lev), 10 + lev)
    s.lvl_write(lev, 0, phi)
    o.level(lev, 0)[:] = phi
    s.bc_fill(lev)
    o.bc_fill(lev)
    g, r = s.lvl_read(lev, 0), o.level(lev, 0)
    m = no_edges(g)
    assert np.array_equal(g[m], r[m])
    s.close()
    o.close()


@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("lev", [1, 2])
def test_periodic_sweep(capi, oracle, gpu, small, lev, d):
    o, s = _pair(capi, oracle, gpu, small, _params(capi, (0, 1), omega_pson=1.2))
    phi = _rand(s.level_shape(lev), 70 + lev)
    rhs = _rand(s.level_shape(lev), 80 + lev)
    for which, a in ((0, phi), (1, rhs)):
        s.lvl_write(lev, which, a)
        o.level(lev, which)[:] = a
    for color in (0, 1):
        s.sweep(lev, d, color)
        o.sweep(lev, d, color)
        g, r = s.lvl_read(lev, 0), o.level(lev, 0)
        m = no_edges(g)
        err = np.max(np.abs(g[m] - r[m])) / np.max(np.abs(r[m]))
        assert err < 1e-12, (lev, d, color, err)
    s.close()
    o.close()


def _box(geo, nx, ny, nz):
    w = geo.Workload()
    w.grid = (geo.tanh_axis(nx + 1, [0.5], 1.0 / 256.0, width=0.25), geo.uniform_axis(ny + 1),
              geo.uniform_axis(nz + 1))
    sh = tuple(a.size for a in w.grid)
    w.iblank = np.zeros(sh, np.int8, order="F")
    w.ghost = np.zeros(sh, np.int8, order="F")
    w.dead = np.zeros(sh, np.int8, order="F")
    w.gcm = geo.empty_gcm()
    return w


def _compatible_rhs(w, seed):
    """a seeded right-hand side with zero volume-weighted sum (solvable with
    periodic and zero-flux faces)"""
    x, y, z = w.grid
    sh = tuple(a.size for a in w.grid)
    nx, ny, nz = (a - 2 for a in sh)
    dx, dy, dz = (np.diff(a[1:-1]) for a in (x, y, z))
    vol = dx[:, None, None] * dy[None, :, None] * dz[None, None, :]
    div = np.zeros(sh, order="F")
    r = np.random.default_rng(seed).standard_normal((nx - 1, ny - 1, nz - 1))
    div[1:nx, 1:ny, 1:nz] = r - np.sum(vol * r) / np.sum(vol)
    return div


@pytest.mark.parametrize("rule,prec", [(0, 0), (1, 0), (1, 1)])
def test_periodic_solve(capi, geo, oracle, gpu, rule, prec):
    """A 48 x 24 x 20 box (x stretched) periodic in x and y, zero-flux z
    faces: the GPU solve (reference outer loop, or Anderson FP64 / mixed)
    reaches the oracle's fixed point."""
    w = _box(geo, 48, 24, 20)
    div = _compatible_rhs(w, 3)
    q = _params(capi, (0, 1), restol_pson=1e-14, iterInter=1, itermax_pson=3000, restol_gcp=1e-13, itermax_gc=10)
    o = oracle.Oracle(capi, *w.grid, q, ordering=oracle.ORDER_REFERENCE)
    o.set_workload(w)
    a = np.zeros_like(div)
    rc_o, st_o, _ = o.solve_poisson(a, div)
    o.close()
    assert rc_o == 0, st_o.as_dict()
    if rule == 0:
        p = _params(capi, (0, 1), restol_pson=1e-13, iterInter=1, itermax_pson=3000, restol_gcp=1e-13,
                    itermax_gc=10)
    else:
        p = _params(capi, (0, 1), restol_pson=0.0, iterInter=0, itermax_pson=600, restol_gcp=1e-11, itermax_gc=10,
                    precision=prec, outer_rule=capi.OUTER_RELRES, rel_tol=1e-12 if prec == 0 else 1e-10,
                    vc_walls=1)
    s = gpu.PoissonSolver(*w.grid, p)
    s.set_workload(w)
    b = np.zeros_like(div)
    rc, st, _ = s.solve_poisson(b, div)
    s.close()
    nx, ny, nz = (t - 2 for t in div.shape)
    f = np.zeros(div.shape, bool)
    f[1:nx, 1:ny, 1:nz] = True
    x = b[f] - b[f].mean()
    y = a[f] - a[f].mean()
    err = np.linalg.norm(x - y) / np.linalg.norm(y)
    print("rule %d prec %d: outer %d (oracle %d), rel %.2e, vs oracle %.2e" % (rule, prec, st.outer_iters,
                                                                             st_o.outer_iters, st.rel_resid, err),
          flush=True)
    assert rc == 0 and st.converged == 1, st.as_dict()
    assert err < (1e-10 if prec == 0 else 1e-8)


def test_periodic_z_is_an_error(capi, gpu, small):
    s = gpu.PoissonSolver(*small.grid, _params(capi, (2,)))
    s.set_workload(small)
    b = np.zeros_like(small.div)
    with pytest.raises(gpu.FcError, match="periodic condition in z"):
        s.solve_poisson(b, small.div)
    s.close()
