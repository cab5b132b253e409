"""GPU parity of rhs_advec_diff (CORE_SET_SOLVERS.F90:10-61) through the
C-ABI against the oracle, which equals the reference's own compiled
CORE_SET_SOLVERS.F90 bit for bit (tests/test_rhs_ad_oracle.py).  The kernel
rounds every product and sum on its own in the reference's order, so the bar
is bit-exact: every entry of nlu, nlv, nlw, the untouched ones included.

The workloads cover the stretched C2 grid with bodies, the C1 sphere, a box
longer than 375 cells in x (the hybrid switch's upper end: first-order
upwind again beyond i = 375) and an odd-sized uniform box; face velocities
of both signs with zeros (SIGN(1, 0) = +1), alfa_upwind 0, 0.6 and 1, and
viscTot both constant (reinv) and a random array."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REINV, DT = 1.0 / 200.0, 0.007


def _box(geo, nx, ny, nz, stretch):
    w = geo.Workload()
    w.grid = tuple(geo.tanh_axis(n + 1, [0.5], 1.0 / 2048.0, width=0.2) if st else geo.uniform_axis(n + 1)
                   for n, st in zip((nx, ny, nz), stretch))
    sh = tuple(a.size for a in w.grid)
    w.iblank = np.zeros(sh, np.int8, order="F")
    w.ghost = np.zeros(sh, np.int8, order="F")
    w.dead = np.zeros(sh, np.int8, order="F")
    w.gcm = geo.empty_gcm()
    return w


@pytest.fixture(scope="module", params=["c2_40", "c1_16", "long_400", "odd_37"])
def wl(request, geo):
    if request.param == "c2_40":
        return geo.config_c2(40)
    if request.param == "c1_16":
        return geo.config_c1(16)
    if request.param == "long_400":
        return _box(geo, 400, 7, 6, (True, False, True))
    return _box(geo, 37, 23, 19, (False, False, False))


def _fields(sh, seed):
    rng = np.random.default_rng(seed)
    U = [np.asfortranarray(rng.standard_normal(sh)) for _ in range(3)]
    FU = []
    for _ in range(3):
        f = rng.standard_normal(sh)
        f[rng.random(sh) < 0.1] = 0.0
        FU.append(np.asfortranarray(f))
    NL = [np.asfortranarray(rng.standard_normal(sh)) for _ in range(3)]
    visc = np.asfortranarray(REINV * (1.0 + rng.random(sh)))
    return U, FU, NL, visc


@pytest.mark.parametrize("alfa,use_visc", [(0.0, False), (0.6, True), (1.0, False), (1.0, True)])
def test_rhs_advec_diff_matches_oracle(capi, geo, oracle, gpu, wl, alfa, use_visc):
    w = wl
    print("workload", w.shape, "alfa", alfa, "visc array", use_visc, flush=True)
    U, FU, NL0, visc = _fields(w.shape, 9)
    o = oracle.Oracle(capi, *w.grid, capi.default_params())
    o.set_workload(w)
    o.set_solve_ad(REINV, DT, 1, 1.0)
    want = [a.copy(order="F") for a in NL0]
    assert o.rhs_advec_diff(*U, *FU, alfa, DT, *want, visc=visc if use_visc else None) == 0
    o.close()
    s = gpu.PoissonSolver(*w.grid, capi.default_params())
    s.set_workload(w)
    s.set_solve_ad(REINV, DT, 1, 1.0)
    s.set_visc_tot(visc if use_visc else None)
    got = [a.copy(order="F") for a in NL0]
    s.rhs_advec_diff(*U, *FU, alfa, DT, *got)
    s.close()
    for c in range(3):
        d = np.abs(got[c] - want[c])
        assert np.array_equal(got[c], want[c]), (c, d.max(), np.unravel_index(np.argmax(d), d.shape))


def test_rhs_advec_diff_resident_feeds_solve_ad(capi, geo, oracle, gpu):
    """The resident form reads the resident u, v, w and face velocities and
    leaves nlu, nlv, nlw where itsolv_ad and the mix_GC_form = 1 velocity
    relations read them: the same bits as the host form, and viscTot falls
    back to reinv after set_visc_tot(None)."""
    w = geo.config_c2(40)
    U, FU, NL0, visc = _fields(w.shape, 21)
    s = gpu.PoissonSolver(*w.grid, capi.default_params())
    s.set_workload(w)
    s.set_solve_ad(REINV, DT, 1, 1.0)
    s.set_visc_tot(visc)
    s.set_visc_tot(None)
    host = [a.copy(order="F") for a in NL0]
    s.rhs_advec_diff(*U, *FU, 0.8, DT, *host)
    s.vel_upload(*U, *FU)
    s.ad_upload(*NL0)
    s.rhs_advec_diff_resident(0.8, DT)
    res = s.ad_download()
    s.close()
    for c in range(3):
        assert np.array_equal(res[c], host[c])
