"""Pins of the oracle's rhs_advec_diff (CORE_SET_SOLVERS.F90:10-61: rhs_advec
:65-436, rhs_diff :440-526), the explicit right-hand side of the
advection-diffusion step.

* Bit for bit against the reference's own compiled CORE_SET_SOLVERS.F90
  (oracle/ref.mk, harness mode 4) on stretched grids with random velocities,
  face velocities of both signs (and zeros), a random viscTot and several
  alfa_upwind, including a grid longer than 375 cells in x (the hybrid
  switch's upper end), and ghostNScoeff(8:10) = nlu, nlv, nlw at the ghost
  cells.
* Advection, closed form: on a uniform grid the three face values of the
  scheme are exact for a quadratic up to a known offset: central
  (fx = 1/2) c (x_f^2 + h^2/4), second-order upwind (weights 3/2, -1/2)
  c (x_f^2 - 3 h^2/4), first-order upwind the upwind cell's value, with the
  inflow faces at n = 1 and n = N-1 taking the central value.  With face
  velocities of random sign the advection term is
  ((1-alfa) g_e + alfa g_e^up) F_e - ((1-alfa) g_w + alfa g_w^up) F_w) / h
  per axis, the upwind form chosen by i on every axis (second order for
  25 <= i <= 375).  A swapped weight, a wrong upwind cell or a dropped
  offset moves the result by O(h^2 F / h).
* Diffusion, closed form on stretched grids: with a velocity linear in
  (xc, yc, zc) and a viscosity linear in space the face viscosities are exact
  and div(nu grad u) = sum a_d n_d; with zero face velocities
  nl = u + dt/2 (a . n)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import _ref_mg  # noqa: E402


def _F(a):
    return np.asfortranarray(a, dtype=np.float64)


def _grid_state(capi, oracle, x, y, z, reinv=0.0, dt=0.01):
    o = oracle.Oracle(capi, x, y, z, capi.default_params())
    sh = (x.size, y.size, z.size)
    o.set_masks(np.zeros(sh, np.int8, order="F"), np.zeros(sh, np.int8, order="F"), np.zeros(sh, np.int8, order="F"))
    o.set_solve_ad(reinv, dt, 1, 1.0)
    return o


def _run(o, U, FU, alfa, dt, visc=None, NL=None):
    sh = U[0].shape
    NL = [_F(np.zeros(sh)) for _ in range(3)] if NL is None else NL
    assert o.rhs_advec_diff(*U, *FU, alfa, dt, *NL, visc=visc) == 0
    return NL


# ---------------------------------------------------------------------------
# closed forms

def _face_values(c, xc, xf_e, xf_w, h, Fe, Fw, n, N):
    """central and upwind face values of c x^2 at the east / west faces of
    cell n (arrays over the cells of one axis)"""
    ge = c * (xf_e ** 2 + h * h / 4.0)
    gw = c * (xf_w ** 2 + h * h / 4.0)
    up2e = c * (xf_e ** 2 - 0.75 * h * h)
    up2w = c * (xf_w ** 2 - 0.75 * h * h)
    up1e = np.where(Fe >= 0, c * xc[n] ** 2, c * xc[np.minimum(n + 1, N)] ** 2)
    up1w = np.where(Fw >= 0, c * xc[n - 1] ** 2, c * xc[n] ** 2)
    # inflow faces at the ends take the central value
    spe = (n == N - 1) & (Fe < 0)
    spw = (n == 1) & (Fw >= 0)
    up2e, up1e = np.where(spe, ge, up2e), np.where(spe, ge, up1e)
    up2w, up1w = np.where(spw, gw, up2w), np.where(spw, gw, up1w)
    return ge, gw, up2e, up2w, up1e, up1w


@pytest.mark.parametrize("alfa", [0.0, 0.35, 1.0])
def test_advection_closed_form_uniform(capi, geo, oracle, alfa):
    N = (33, 13, 11)                           # node counts: 32 x 12 x 10 cells (i spans both sides of 25)
    x, y, z = (geo.uniform_axis(n, 0.0, 0.1 * n) for n in N)
    h = [a[2] - a[1] for a in (x, y, z)]
    xc = [geo.metrics(a)[0] for a in (x, y, z)]
    sh = (x.size, y.size, z.size)
    cq = (0.7, -1.3, 0.45)
    X, Y, Z = np.meshgrid(xc[0], xc[1], xc[2], indexing="ij")
    rng = np.random.default_rng(5)
    U = [_F(cq[0] * X ** 2 + cq[1] * Y ** 2 + cq[2] * Z ** 2 + s) for s in (0.0, 0.4, -0.2)]
    FU = [_F(rng.uniform(0.2, 1.0, sh) * rng.choice([-1.0, 1.0], sh)) for _ in range(3)]
    dt = 0.02
    o = _grid_state(capi, oracle, x, y, z)
    NL = _run(o, U, FU, alfa, dt)
    o.close()
    nn = [np.arange(1, n) for n in N]
    I, J, K = np.meshgrid(*nn, indexing="ij")
    idx = (I, J, K)
    central_band = (I >= 25) & (I <= 375)
    for c in range(3):
        adv = np.zeros(I.shape)
        for a, (st, F) in enumerate(zip(((1, 0, 0), (0, 1, 0), (0, 0, 1)), FU)):
            n = idx[a]
            Fw = F[I, J, K]
            Fe = F[I + st[0], J + st[1], K + st[2]]
            nodes = (x, y, z)[a]
            xf_w, xf_e = nodes[n], nodes[n + 1]
            ge, gw, u2e, u2w, u1e, u1w = _face_values(cq[a], xc[a], xf_e, xf_w, h[a], Fe, Fw, n, N[a])
            # the other two axes' parts of the field are constant on the faces of this axis
            rest = U[c][I, J, K] - cq[a] * xc[a][n] ** 2
            ue = np.where(central_band, u2e, u1e) + rest
            uw = np.where(central_band, u2w, u1w) + rest
            adv += (((1 - alfa) * (ge + rest) + alfa * ue) * Fe - ((1 - alfa) * (gw + rest) + alfa * uw) * Fw) / h[a]
        want = U[c][I, J, K] - 0.5 * dt * adv
        got = NL[c][1:N[0], 1:N[1], 1:N[2]]
        assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want)), c


def test_advection_closed_form_detects_a_wrong_branch(capi, geo, oracle):
    """the closed form separates the schemes: with face velocities that
    change from face to face, central and upwind results differ on every
    cell by far more than the tolerance used above"""
    N = (33, 9, 9)
    x, y, z = (geo.uniform_axis(n, 0.0, 0.1 * n) for n in N)
    xc = [geo.metrics(a)[0] for a in (x, y, z)]
    X = np.meshgrid(xc[0], xc[1], xc[2], indexing="ij")[0]
    sh = X.shape
    U = [_F(0.7 * X ** 2) for _ in range(3)]
    XF = np.broadcast_to(x[:, None, None], sh)          # face_u(i) sits on node x(i)
    FU = [_F(0.5 + 0.4 * XF), _F(np.zeros(sh)), _F(np.zeros(sh))]
    o = _grid_state(capi, oracle, x, y, z)
    a0 = _run(o, U, FU, 0.0, 0.02)[0]
    a1 = _run(o, U, FU, 1.0, 0.02)[0]
    o.close()
    d = np.abs(a0 - a1)[1:32, 1:8, 1:8]
    assert d.min() > 1e-6 * np.abs(a0).max()


def test_diffusion_closed_form_stretched(capi, geo, oracle):
    x = geo.tanh_axis(25, [0.4], 1.0 / 96.0, width=0.15)
    y = geo.tanh_axis(19, [0.6], 1.0 / 64.0, width=0.2)
    z = geo.tanh_axis(17, [0.5], 1.0 / 48.0, width=0.25)
    xc = [geo.metrics(a)[0] for a in (x, y, z)]
    X, Y, Z = np.meshgrid(*xc, indexing="ij")
    A = [(0.9, -0.4, 1.7), (-1.1, 0.3, 0.5), (0.2, 0.8, -0.6)]
    n0, nvec = 0.01, (0.004, -0.002, 0.003)
    visc = _F(n0 + nvec[0] * X + nvec[1] * Y + nvec[2] * Z)
    U = [_F(0.3 + a[0] * X + a[1] * Y + a[2] * Z) for a in A]
    sh = X.shape
    FU = [_F(np.zeros(sh)) for _ in range(3)]
    dt = 0.05
    o = _grid_state(capi, oracle, x, y, z)
    NL = [_F(np.full(sh, 7.0)) for _ in range(3)]
    _run(o, U, FU, 0.5, dt, visc=visc, NL=NL)
    o.close()
    for c in range(3):
        want = U[c] + 0.5 * dt * sum(A[c][d] * nvec[d] for d in range(3))
        got = NL[c]
        assert np.max(np.abs(got[1:-2, 1:-2, 1:-2] - want[1:-2, 1:-2, 1:-2])) <= 1e-12
        # entries outside the interior keep their values
        m = np.ones(sh, bool)
        m[1:-2, 1:-2, 1:-2] = False
        assert np.all(got[m] == 7.0)


def test_constant_viscosity_is_reinv(capi, geo, oracle):
    """visc = None uses the reinv of set_solve_ad (viscTot = reinv,
    MAIN.F90:926), bit for bit the same as a constant array"""
    w = geo.config_c2(24)
    sh = w.iblank.shape
    rng = np.random.default_rng(3)
    U = [_F(rng.standard_normal(sh)) for _ in range(3)]
    FU = [_F(rng.standard_normal(sh)) for _ in range(3)]
    o = _grid_state(capi, oracle, *w.grid, reinv=1.0 / 300.0, dt=0.01)
    a = _run(o, U, FU, 0.5, 0.01)
    b = _run(o, U, FU, 0.5, 0.01, visc=_F(np.full(sh, 1.0 / 300.0)))
    o.close()
    for c in range(3):
        assert np.array_equal(a[c], b[c])


# ---------------------------------------------------------------------------
# the reference itself

def _reference_case(geo, kind):
    if kind == "c2":
        w = geo.config_c2(40)
        return w.grid, w.iblank, w.ghost, w.dead
    if kind == "long":
        x = geo.tanh_axis(401, [0.3, 0.7], 1.0 / 1200.0, width=0.05)
        y = geo.uniform_axis(6)
        z = geo.tanh_axis(6, [0.5], 1.0 / 8.0, width=0.3)
        sh = (x.size, y.size, z.size)
        rng = np.random.default_rng(4)
        ghost = np.zeros(sh, np.int8, order="F")
        ghost[1:-2, 1:-2, 1:-2] = (rng.random((sh[0] - 3, sh[1] - 3, sh[2] - 3)) < 0.05).astype(np.int8)
        return (x, y, z), np.zeros(sh, np.int8, order="F"), ghost, np.zeros(sh, np.int8, order="F")
    raise ValueError(kind)


@pytest.mark.skipif(not _ref_mg.available(), reason="needs the reference tree")
@pytest.mark.parametrize("kind,alfa,visc_kind", [("c2", 0.0, "array"), ("c2", 0.6, "array"), ("c2", 1.0, "const"),
                                                 ("long", 0.8, "array")])
def test_rhs_advec_diff_matches_reference(capi, geo, oracle, tmp_path, kind, alfa, visc_kind):
    (x, y, z), ib, gh, de = _reference_case(geo, kind)
    sh = ib.shape
    rng = np.random.default_rng(17)
    U = [_F(rng.standard_normal(sh)) for _ in range(3)]
    FU = []
    for _ in range(3):
        f = rng.standard_normal(sh)
        f[rng.random(sh) < 0.1] = 0.0
        FU.append(_F(f))
    reinv, dt = 1.0 / 250.0, 0.013
    visc = _F(reinv * (1.0 + 0.5 * rng.random(sh))) if visc_kind == "array" else _F(np.full(sh, reinv))
    NL0 = [_F(rng.standard_normal(sh)) for _ in range(3)]
    prm = capi.default_params()
    ref, ns = _ref_mg.run_rhs_advec_diff(tmp_path, x, y, z, prm, ib, gh, de, dt, alfa, *U, *FU, visc, *NL0)
    o = _grid_state(capi, oracle, x, y, z, reinv=reinv, dt=dt)
    NL = [a.copy(order="F") for a in NL0]
    _run(o, U, FU, alfa, dt, visc=visc if visc_kind == "array" else None, NL=NL)
    o.close()
    for c in range(3):
        assert np.array_equal(NL[c], ref[c]), (c, np.max(np.abs(NL[c] - ref[c])))
    # ghostNScoeff(8:10) of the ghost cells are nlu, nlv, nlw there
    g = np.nonzero(np.abs(gh.ravel(order="F")) == 1)[0]
    assert ns.shape == (3, g.size)
    for c in range(3):
        assert np.array_equal(ns[c], NL[c].ravel(order="F")[g])
