"""GPU parity of the advection-diffusion line solves (CORE_SOLVE_AD.F90:
itsolv_ad, calc_residual_ad; set_solve_ad CORE_SET_SOLVERS.F90:591-676)
through the C-ABI, against the oracle in the GPU's line order
(checkerboard colours per direction).  The GPU solves each line by a
partition method and the oracle by the reference's tdma, so the sweeps agree
to round-off (1e-12 relative); the fixed point is also checked against the
oracle in the reference's own lexicographic order.

The workloads reach every variant of the GPU line kernel in every direction:
lines of up to 128 rows (2 rows per lane), 129-192 (3), 193-320 (5) and
321-576 (9, two lines per block); on the stretched axes the cell Courant
number |u| dt / dx of the finest cells is 3.5 to 10 for |u| = 3, where the
Crank-Nicolson rows are no longer diagonally dominant and the unpivoted eliminations (the reference's tdma,
the GPU's partition solve) are compared at their weakest."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REINV, DT = 1.0 / 150.0, 0.004   # cell Courant numbers up to ~4 on the finest cells


def _rand(sh, rng, scale=1.0):
    return np.asfortranarray(scale * rng.standard_normal(sh))


def _empty_workload(geo, nx, ny, nz, stretch=(True, False, False)):
    """a box without bodies; stretched axes cluster toward the centre down to
    a spacing of 1/1024"""
    w = geo.Workload()
    w.grid = tuple(geo.tanh_axis(n + 1, [0.5], 1.0 / 1024.0, width=0.25) if st else geo.uniform_axis(n + 1)
                   for n, st in zip((nx, ny, nz), stretch))
    sh = tuple(a.size for a in w.grid)       # nodes + 2 per axis
    w.iblank = np.zeros(sh, np.int8, order="F")
    w.ghost = np.zeros(sh, np.int8, order="F")
    w.dead = np.zeros(sh, np.int8, order="F")
    w.gcm = geo.empty_gcm()
    return w


# cells per axis; a line along an axis of n cells has n + 2 rows
_BOXES = {
    "long_400": (400, 10, 9),          # x: 9 rows per lane
    "x140_z200": (140, 12, 200),       # x: 3, z: 5
    "x200_y140": (200, 140, 12),       # x: 5, y: 3
    "y200_z140": (12, 200, 140),       # y: 5, z: 3
    "y400": (10, 400, 9),              # y: 9
    "z400": (10, 9, 400),              # z: 9
}


@pytest.fixture(scope="module", params=["c2_40", "c1_16"] + list(_BOXES))
def wl(request, geo):
    if request.param == "c2_40":
        return geo.config_c2(40)
    if request.param == "c1_16":
        return geo.config_c1(16)
    n = _BOXES[request.param]
    return _empty_workload(geo, *n, stretch=tuple(m >= 100 for m in n))


def _pair(capi, oracle, gpu, w, cn, omega, ordering=None):
    o = oracle.Oracle(capi, *w.grid, capi.default_params(),
                      ordering=oracle.ORDER_COLORED if ordering is None else ordering)
    o.set_workload(w)
    o.set_solve_ad(REINV, DT, cn, omega)
    s = gpu.PoissonSolver(*w.grid, capi.default_params())
    s.set_workload(w)
    s.set_solve_ad(REINV, DT, cn, omega)
    return o, s


@pytest.mark.parametrize("cn,omega", [(1, 1.0), (0, 1.0), (1, 0.8)])
def test_itsolv_ad_matches_oracle(capi, geo, oracle, gpu, wl, cn, omega):
    w = wl
    print("workload", w.shape, "cn", cn, "omega", omega, flush=True)
    o, s = _pair(capi, oracle, gpu, w, cn, omega)
    rng = np.random.default_rng(4)
    sh = w.shape
    var = _rand(sh, rng)
    r = _rand(sh, rng)
    fu, fv, fw = (_rand(sh, rng) for _ in range(3))
    for nsweep in (1, 2):
        a = var.copy(order="F")
        b = var.copy(order="F")
        for _ in range(nsweep):
            o.itsolv_ad(a, r, fu, fv, fw)
        s.itsolv_ad(b, r, fu, fv, fw, nsweep=nsweep)
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(a)), nsweep
        assert not np.array_equal(a, var)
    # calc_residual_ad on the random start (an O(1) residual: location and
    # value are decided well above round-off) and after the sweeps (absolute
    # agreement at the round-off of the operands)
    ra = o.calc_residual_ad(var, r, fu, fv, fw)
    rb = s.calc_residual_ad(var, r, fu, fv, fw)
    assert abs(ra[0] - rb[0]) <= 1e-12 * ra[0] and ra[1] == rb[1]
    assert abs(ra[2] - rb[2]) <= 1e-10 * ra[2]
    ra = o.calc_residual_ad(a, r, fu, fv, fw)
    rb = s.calc_residual_ad(b, r, fu, fv, fw)
    assert abs(ra[0] - rb[0]) <= 1e-10 * ra[0] + 1e-12 * np.max(np.abs(r))
    s.close()


def test_ad_fixed_point_matches_reference_order(capi, geo, oracle, gpu):
    """Resident sweeps on the GPU, to calc_residual_ad < 1e-12, reach the
    oracle's fixed point in the reference's own line order."""
    w = geo.config_c2(40)
    o, s = _pair(capi, oracle, gpu, w, 1, 1.0, ordering=oracle.ORDER_REFERENCE)
    rng = np.random.default_rng(8)
    sh = w.shape
    u0 = _rand(sh, rng, 0.1)
    r = _rand(sh, rng)
    fu, fv, fw = (_rand(sh, rng) for _ in range(3))
    a = u0.copy(order="F")
    for it in range(200):
        o.itsolv_ad(a, r, fu, fv, fw)
        if o.calc_residual_ad(a, r, fu, fv, fw)[0] < 1e-12:
            break
    s.vel_upload(np.zeros_like(u0), u0, np.zeros_like(u0), fu, fv, fw)
    s.ad_upload(nlv=r)
    for it2 in range(200):
        s.itsolv_ad_resident(1, 1)
        resm = s.calc_residual_ad_resident(1)[0]
        if resm < 1e-12:
            break
    print("oracle sweeps", it + 1, "gpu sweeps", it2 + 1, flush=True)
    assert resm < 1e-12
    b = s.vel_download()[1]
    assert np.max(np.abs(a - b)) <= 1e-10 * np.max(np.abs(a))
    s.close()


def test_ad_call_checks(capi, geo, gpu):
    """itsolv_ad before set_solve_ad, and lines beyond 576 rows, are errors
    (negative return codes raised as FcError) found before any kernel runs."""
    w = geo.config_c1(16)
    s = gpu.PoissonSolver(*w.grid, capi.default_params())
    s.set_workload(w)
    z = np.zeros(w.shape, order="F")
    with pytest.raises(gpu.FcError, match="set_solve_ad"):
        s.itsolv_ad(z.copy(order="F"), z, z, z, z)
    s.close()
    long = _empty_workload(geo, 600, 6, 6)
    s = gpu.PoissonSolver(*long.grid, capi.default_params())
    s.set_workload(long)
    with pytest.raises(gpu.FcError, match="576 rows"):
        s.set_solve_ad(REINV, DT, 1, 1.0)
    s.close()
