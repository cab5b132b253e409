"""Pins of the oracle's solve_ad loop (CORE_SOLVE_AD.F90:11-127) and the
routines it calls besides itsolv_ad, each by a property the definition fixes:

* set_outer_velocity_bc / set_outer_ghost_velocity (CORE_SET_OUTER_BC.F90:
  18-509, 739-823): after both, the mean of a ghost and its interior
  neighbour is the face value of the type: the given value (Dirichlet), the
  value times sin(2 pi f t) (pulsatile), the channel profile
  6 (y - y0)(y1 - y)/(y1 - y0)^2 times the value (user specified, x faces),
  checked inside each face (the later passes rewrite the edge ghosts);
  a zero-gradient ghost equals its interior neighbour, a symmetry face's
  normal component vanishes there while the others copy, and a periodic
  ghost equals the interior cell at the other end (to rounding); shear and
  user-specified y/z faces are refused;
* GCM_ghostcell_velocity (GCM_GHOSTCELL_UPDATE.F90:1-359): the trilinear
  Dirichlet interpolation reproduces linear fields, so with the body-intercept
  velocities taken from three linear fields the converged ghost cells hold
  the fields' values at their centres and the dead cells (those that do not
  take the inverse-distance branch) their values too;
* solve_ad's loop: at its converged state every component's residual is
  below the tolerance, one more GCM_ghostcell_velocity step moves nothing
  beyond it, and refilling the outer ghosts leaves the interior unchanged
  and moves the ghosts only by the loop's own convergence."""
import numpy as np
import pytest

INC = np.array([[0, 0, 0], [0, 1, 0], [1, 1, 0], [1, 0, 0], [0, 0, 1], [0, 1, 1], [1, 1, 1], [1, 0, 1]])


def _box(geo, nx, ny, nz):
    w = geo.Workload()
    w.grid = (geo.uniform_axis(nx + 1), geo.tanh_axis(ny + 1, [0.5], 1.0 / 64.0, width=0.3), geo.uniform_axis(nz + 1))
    sh = tuple(a.size for a in w.grid)
    w.iblank = np.zeros(sh, np.int8, order="F")
    w.ghost = np.zeros(sh, np.int8, order="F")
    w.dead = np.zeros(sh, np.int8, order="F")
    w.gcm = geo.empty_gcm()
    return w


def _params(capi, types, **kw):
    return capi.default_params(bc_vel=list(types), **kw)


def _rand(sh, seed):
    return np.asfortranarray(np.random.default_rng(seed).standard_normal(sh))


VALS = np.arange(1, 19, dtype=float) * 0.1          # ux1, ux2, vx1, ..., wz2
FREQ = np.linspace(0.3, 2.0, 18)


def test_outer_velocity_types(capi, oracle, geo):
    w = _box(geo, 10, 12, 8)
    x, y, z = w.grid
    nx, ny, nz = (a.size - 2 for a in w.grid)
    # x1 Dirichlet, x2 user specified, y1 zero gradient, y2 symmetry,
    # z1 pulsatile, z2 Dirichlet
    types = (1, 6, 2, 4, 3, 1)
    o = oracle.Oracle(capi, *w.grid, _params(capi, types))
    o.set_workload(w)
    y0, y1, t = 0.1, 0.9, 0.37
    o.set_outer_velocity(VALS, FREQ, y0, y1)
    U = [_rand(w.iblank.shape, s) for s in (1, 2, 3)]
    assert o.set_outer_velocity_bc(*U, t) == 0
    o.set_outer_ghost_velocity(*U)
    o.close()
    val = lambda a, c, side: VALS[6 * a + 2 * c + side]      # noqa: E731
    tol = 1e-15
    I, J, K = slice(1, nx), slice(1, ny), slice(1, nz)      # inside each face
    for c in range(3):
        # x1 Dirichlet
        m = 0.5 * (U[c][0, J, K] + U[c][1, J, K])
        assert np.allclose(m, val(0, c, 0), rtol=tol, atol=tol)
        # x2 user specified: the channel profile of y for u
        m = 0.5 * (U[c][nx, J, K] + U[c][nx - 1, J, K])
        yy = y[1:ny]
        expect = val(0, c, 1) * 6.0 * (yy - y0) * (y1 - yy) / (y1 - y0) ** 2 if c == 0 else val(0, c, 1) + 0 * yy
        assert np.allclose(m, expect[:, None], rtol=1e-14, atol=1e-15)
        # y1 zero gradient
        assert np.array_equal(U[c][I, 0, K], U[c][I, 1, K])
        # y2 symmetry: v vanishes at the face, u and w copy
        if c == 1:
            assert np.array_equal(U[c][I, ny, K], -U[c][I, ny - 1, K])
        else:
            assert np.array_equal(U[c][I, ny, K], U[c][I, ny - 1, K])
        # z1 pulsatile, z2 Dirichlet
        m = 0.5 * (U[c][I, J, 0] + U[c][I, J, 1])
        assert np.allclose(m, val(2, c, 0) * np.sin(2.0 * np.pi * FREQ[12 + 2 * c] * t), rtol=1e-14, atol=1e-15)
        m = 0.5 * (U[c][I, J, nz] + U[c][I, J, nz - 1])
        assert np.allclose(m, val(2, c, 1), rtol=tol, atol=tol)


def test_outer_velocity_periodic_and_pulsatile_ramp(capi, oracle, geo):
    w = _box(geo, 10, 12, 8)
    nx, ny, nz = (a.size - 2 for a in w.grid)
    o = oracle.Oracle(capi, *w.grid, _params(capi, (3, 1, 5, 5, 2, 2)))
    o.set_workload(w)
    o.set_outer_velocity(VALS, FREQ)
    U = [_rand(w.iblank.shape, s) for s in (4, 5, 6)]
    t = 1.7
    assert o.set_outer_velocity_bc(*U, t) == 0
    o.set_outer_ghost_velocity(*U)
    o.close()
    for c in range(3):
        # x1 pulsatile: u ramps as t / 30, v and w oscillate
        m = 0.5 * (U[c][0, 1:ny, 1:nz] + U[c][1, 1:ny, 1:nz])
        expect = VALS[0] * t / 30.0 if c == 0 else VALS[2 * c] * np.sin(2.0 * np.pi * FREQ[2 * c] * t)
        assert np.allclose(m, expect, rtol=1e-14, atol=1e-15)
        # periodic y: the ghost is the interior cell at the other end
        s = np.max(np.abs(U[c]))
        assert np.max(np.abs(U[c][1:nx, 0, 1:nz] - U[c][1:nx, ny - 1, 1:nz])) <= 4e-16 * s
        assert np.max(np.abs(U[c][1:nx, ny, 1:nz] - U[c][1:nx, 1, 1:nz])) <= 4e-16 * s


@pytest.mark.parametrize("types", [(7, 1, 1, 1, 1, 1), (1, 1, 6, 1, 1, 1), (1, 1, 1, 1, 1, 6)])
def test_outer_velocity_refused_types(capi, oracle, geo, types):
    w = _box(geo, 8, 8, 8)
    o = oracle.Oracle(capi, *w.grid, _params(capi, types))
    o.set_workload(w)
    U = [_rand(w.iblank.shape, s) for s in (1, 2, 3)]
    assert o.set_outer_velocity_bc(*U, 0.0) == -2
    o.close()


def _linear_fields(xc, yc, zc):
    A = np.array([[0.3, 1.1, -0.7, 0.4], [-0.2, 0.5, 0.9, -1.3], [0.8, -0.6, 0.2, 0.7]])
    return A, [lambda X, Y, Z, a=a: a[0] + a[1] * X + a[2] * Y + a[3] * Z for a in A]


def test_ghostcell_velocity_reproduces_linear_fields(capi, oracle, geo):
    w = geo.config_c1(16)
    g = w.gcm
    x, y, z = w.grid
    xc, yc, zc = (geo.metrics(t)[0] for t in (x, y, z))
    _, F = _linear_fields(xc, yc, zc)
    Xc, Yc, Zc = np.meshgrid(xc, yc, zc, indexing="ij")
    U = [np.asfortranarray(f(Xc, Yc, Zc)) for f in F]
    # scramble the ghost and dead values the update must recover
    rng = np.random.default_rng(2)
    tg = (g["iGhost"], g["jGhost"], g["kGhost"])
    td = (g["iDead"], g["jDead"], g["kDead"])
    for a in U:
        a[tg] = rng.standard_normal(tg[0].size)
        a[td] = rng.standard_normal(td[0].size)
    bi, bid = g["xyzBI"], g["xyzBIDead"]
    ubi = np.stack([f(bi[:, 0], bi[:, 1], bi[:, 2]) for f in F], axis=1)
    ubid = np.stack([f(bid[:, 0], bid[:, 1], bid[:, 2]) for f in F], axis=1)
    o = oracle.Oracle(capi, *w.grid, capi.default_params(itermax_gc=500))
    o.set_workload(w)
    o.set_gcm_vel(g["coeffGCMD"], g["coeffGCMDeadD"], ubi, ubid, g["probeLengthNormalizedD"])
    rc, sweeps, rm = o.ghostcell_velocity(*U, restol_gcv=1e-14)
    o.close()
    assert rc == 0 and sweeps > 1
    ref = [f(Xc, Yc, Zc) for f in F]
    for c in range(3):
        assert np.max(np.abs(U[c][tg] - ref[c][tg])) <= 1e-11 * np.max(np.abs(ref[c]))
    # dead cells without a pure solid cell in their stencil (no inverse-distance branch)
    plain = []
    for n in range(td[0].size):
        i, j, k = td[0][n], td[1][n], td[2][n]
        if j < 1 or j > w.iblank.shape[1] - 3 or k < 1 or k > w.iblank.shape[2] - 3:
            continue
        cells = INC + np.array([g["iDeadCellIndex"][n], g["jDeadCellIndex"][n], g["kDeadCellIndex"][n]])
        solid = any(w.iblank[a, b, cc] == 1 and w.dead[a, b, cc] == 0 for a, b, cc in cells)
        if not solid:
            plain.append(n)
    assert len(plain) > 0
    plain = np.array(plain)
    for c in range(3):
        sel = (td[0][plain], td[1][plain], td[2][plain])
        assert np.max(np.abs(U[c][sel] - ref[c][sel])) <= 1e-11 * np.max(np.abs(ref[c]))


def test_solve_ad_fixed_point_invariants(capi, oracle, geo):
    w = geo.config_c1(16)
    g = w.gcm
    sh = w.iblank.shape
    nx, ny, nz = (a - 2 for a in sh)
    p = _params(capi, (1, 2, 4, 4, 1, 1), itermax_gc=200)
    o = oracle.Oracle(capi, *w.grid, p)
    o.set_workload(w)
    o.set_solve_ad(1.0 / 100.0, 0.01, 1, 1.0)
    o.set_outer_velocity(np.r_[1.0, 0.0, np.zeros(16)])
    o.set_gcm_vel(g["coeffGCMD"], g["coeffGCMDeadD"], np.zeros((g["iGhost"].size, 3)),
                  np.zeros((g["iDead"].size, 3)), g["probeLengthNormalizedD"])
    rng = np.random.default_rng(7)
    U = [np.asfortranarray(0.1 * rng.standard_normal(sh)) for _ in range(3)]
    NL = [np.asfortranarray(0.05 * rng.standard_normal(sh)) for _ in range(3)]
    FU = [np.asfortranarray(np.zeros(sh)) for _ in range(3)]
    FU[0][...] = 1.0
    rc, iters, res, loc = o.solve_ad(*U, *NL, *FU, 0.0, 300, 1e-11, 1e-13)
    assert rc == 0 and res <= 1e-11, (iters, res)
    for c in range(3):
        assert o.calc_residual_ad(U[c], NL[c], *FU)[0] <= 1e-11
    before = [a.copy() for a in U]
    rc, sweeps, rm = o.ghostcell_velocity(*U, *NL, *FU, restol_gcv=1e-13)
    assert rm <= 1e-13
    assert o.set_outer_velocity_bc(*U, 0.0) == 0
    o.set_outer_ghost_velocity(*U)
    o.close()
    # the outer ghosts lag the interior by one pass of the loop (the face
    # values read the ghosts of the other axes), so they agree to the
    # convergence of the loop
    for a, b in zip(U, before):
        assert np.max(np.abs(a - b)) <= 1e-9 * max(1.0, np.max(np.abs(b)))
        assert np.max(np.abs(a[1:nx, 1:ny, 1:nz] - b[1:nx, 1:ny, 1:nz])) <= 1e-12 * max(1.0, np.max(np.abs(b)))
