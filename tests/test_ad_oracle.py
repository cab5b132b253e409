"""Pins of the oracle's advection-diffusion line solves (CORE_SOLVE_AD.F90:
itsolv_ad :131-776, calc_residual_ad :780-964; set_solve_ad
CORE_SET_SOLVERS.F90:591-676).

* set_solve_ad: bit for bit against the reference's own compiled
  CORE_SET_SOLVERS.F90 (oracle/ref.mk, harness mode 3).
* itsolv_ad's operator, closed form: on any stretching, the diffusion part
  annihilates a field linear in xc, yc, zc, and with constant face
  velocities (U, V, W) the Crank-Nicolson advection part adds
  half_dt * U * a per axis on the central cells (25 <= i <= 375) and
  half_dt * U * a * dxc(i) / dx(i) (U > 0) or dxc(i+1) / dx(i) (U < 0) on
  the upwind cells.  With r set to that, the linear field is the fixed
  point: one sweep leaves it in place, sweeps from a random start converge
  to it in both line orders, and calc_residual_ad vanishes there.  A swapped
  coefficient, a wrong branch or a sign error moves the fixed point.

CORE_SOLVE_AD.F90 itself does not compile with flang (the Intel-only
'PE15.5' edit descriptor at :120), so itsolv_ad has no bit-level reference
run; these closed forms pin it."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import _ref_mg  # noqa: E402

REINV, DT = 1.0 / 150.0, 0.02
UVW = (0.8, -0.45, 0.3)
ABC = (1.3, -0.7, 0.55)


@pytest.fixture(scope="module")
def w(geo):
    return geo.config_c2(40)


def _centres(x):
    xc = 0.5 * (x[1:] + x[:-1])
    return np.concatenate([xc, x[-1:]])


def _setup(capi, oracle, w, ordering, cn=1, omega=1.0):
    o = oracle.Oracle(capi, *w.grid, capi.default_params(), ordering=ordering)
    o.set_workload(w)
    o.set_solve_ad(REINV, DT, cn, omega)
    return o


def linear_case(w, cn=1):
    """(u_lin, r, faces) with r such that u_lin is the fixed point."""
    x, y, z = w.grid
    xc, yc, zc = (_centres(t) for t in w.grid)
    a, b, c = ABC
    U, V, W = UVW
    sh = w.shape
    u = np.asfortranarray(0.2 + a * xc[:, None, None] + b * yc[None, :, None] + c * zc[None, None, :])
    faces = [np.asfortranarray(np.full(sh, v)) for v in (U, V, W)]
    hd = 0.5 * DT if cn else 0.0
    nx, ny, nz = x.size - 2, y.size - 2, z.size - 2
    i = np.arange(sh[0])

    def upwind_factor(t, n, vel):
        tc = _centres(t)
        d = np.zeros(n + 2); d[:n + 1] = np.diff(t)
        dc = np.zeros(n + 2); dc[1:n + 1] = np.diff(tc[:n + 1])
        dcp = np.zeros(n + 2); dcp[:n] = dc[1:n + 1]
        with np.errstate(divide="ignore", invalid="ignore"):
            return (dc if vel > 0 else dcp) / np.where(d == 0, 1, d)

    central = ((i >= 25) & (i <= 375))[:, None, None]
    fx_up = upwind_factor(x, nx, U)[:, None, None]
    fy_up = upwind_factor(y, ny, V)[None, :, None]
    fz_up = upwind_factor(z, nz, W)[None, None, :]
    adv = np.where(central, U * a + V * b + W * c, U * a * fx_up + V * b * fy_up + W * c * fz_up)
    r = np.asfortranarray(u + hd * adv)
    return u, r, faces


def _interior(sh):
    m = np.zeros(sh, bool)
    m[1:-2, 1:-2, 1:-2] = True
    return m


@pytest.mark.skipif(not _ref_mg.available(), reason="reference tree and harness absent")
def test_set_solve_ad_matches_reference(capi, oracle, w, tmp_path):
    got = _ref_mg.run_ad_coefficients(tmp_path, *w.grid, capi.default_params(), w.iblank, w.ghost, w.dead,
                                      REINV, DT)
    o = _setup(capi, oracle, w, oracle.ORDER_REFERENCE)
    mine = o.get_solve_ad()
    for g, m in zip(got, mine):
        assert np.array_equal(g, m)
    assert np.count_nonzero(mine[0]) > 0


@pytest.mark.parametrize("cn", [1, 0])
@pytest.mark.parametrize("order", ["reference", "colored"])
def test_itsolv_ad_linear_fixed_point(capi, oracle, w, cn, order):
    o = _setup(capi, oracle, w, oracle.ORDER_REFERENCE if order == "reference" else oracle.ORDER_COLORED, cn)
    u, r, (fu, fv, fw) = linear_case(w, cn)
    m = _interior(w.shape)
    resm, loc, _ = o.calc_residual_ad(u, r, fu, fv, fw)
    assert resm < 1e-11
    v = u.copy(order="F")
    o.itsolv_ad(v, r, fu, fv, fw)
    assert np.max(np.abs(v - u)) < 1e-11
    # from a random start on the unmasked cells (masked rows are identities
    # and boundary layers trivial rows: they hold the linear values)
    flag = m & (w.iblank == 0) & (w.ghost == 0)
    rng = np.random.default_rng(3)
    v = u.copy(order="F")
    v[flag] += 0.1 * rng.standard_normal(np.count_nonzero(flag))
    for it in range(60):
        o.itsolv_ad(v, r, fu, fv, fw)
        resm, loc, _ = o.calc_residual_ad(v, r, fu, fv, fw)
        if resm < 1e-12:
            break
    assert resm < 1e-12, (it, resm, loc)
    assert np.max(np.abs(v[flag] - u[flag])) < 1e-10


def test_upwind_branch_differs(capi, oracle, w):
    """The fixed point above uses both advection branches (i < 25 upwind,
    i >= 25 central): the central-branch right-hand side on the upwind cells
    is no fixed point."""
    o = _setup(capi, oracle, w, oracle.ORDER_REFERENCE)
    u, r, (fu, fv, fw) = linear_case(w)
    a, b, c = ABC
    U, V, W = UVW
    r2 = np.asfortranarray(u + 0.5 * DT * (U * a + V * b + W * c))
    resm, loc, _ = o.calc_residual_ad(u, r2, fu, fv, fw)
    assert resm > 1e-6 and loc[0] < 25
