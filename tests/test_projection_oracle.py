"""Pins of the oracle's velocity projection (the calls after solve_poisson,
CORE_TIME_STEP.F90:211-222) by closed forms on a stretched, masked grid.

* correct_vel (CORE_FRAC_STEP_ROUTINES.F90:134-210) and
  correct_ghostcell_velocity (GCM_GHOSTCELL_UPDATE.F90:789-852): the face
  interpolation with the metrics weights fx(i) = (x(i) - xc(i-1)) /
  (xc(i) - xc(i-1)) reproduces a pressure linear in xc, yc, zc at the face
  nodes, so the gradient is the constant (a, b, c) on any stretching; each
  component has its own constant, so a swapped axis or a shifted weight
  index fails.
* face_vel (CORE_FRAC_STEP_ROUTINES.F90:6-130): the same weights interpolate
  a velocity linear in xc (times anything in j, k) to x(i) exactly.
* update_pressure (CORE_FRAC_STEP_ROUTINES.F90:212-299): invariants (p equals
  the new pPrime, fluid cells off the domain edges unchanged, pure-solid
  cells zero).  The reference's own compiled code pins it bit for bit
  (tests/test_reference_projection.py).
"""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def w(geo):
    return geo.config_c2(24)


def _centres(x):
    xc = 0.5 * (x[1:] + x[:-1])
    return np.concatenate([xc, x[-1:]])     # xc(0..n), xc(n+1) = x(n+1)


def _oracle(capi, oracle, w, mix=0):
    p = capi.default_params(mix_GC_form=mix)
    o = oracle.Oracle(capi, *w.grid, p)
    o.set_workload(w)
    return o


def _rand(shape, seed):
    return np.asfortranarray(np.random.default_rng(seed).standard_normal(shape))


def _linear(w, a, b, c, off=0.3):
    xc, yc, zc = (_centres(t) for t in w.grid)
    return np.asfortranarray(off + a * xc[:, None, None] + b * yc[None, :, None] + c * zc[None, None, :])


def _interior(shape):
    m = np.zeros(shape, bool)
    m[1:-2, 1:-2, 1:-2] = True
    return m


def test_correct_vel_linear_pressure(capi, oracle, w):
    o = _oracle(capi, oracle, w)
    a, b, c, dt = 1.7, -0.6, 2.3, 0.37
    P = _linear(w, a, b, c)
    u0, v0, w0 = (_rand(w.shape, s) for s in (1, 2, 3))
    u, v, ww = u0.copy(order="F"), v0.copy(order="F"), w0.copy(order="F")
    o.correct_vel(P, dt, u, v, ww)
    m = _interior(w.shape)
    flag = 1.0 - w.iblank.astype(np.float64)
    for new, old, g in ((u, u0, a), (v, v0, b), (ww, w0, c)):
        assert np.array_equal(new[~m], old[~m])                      # only interior cells change
        solid = m & (w.iblank == 1)
        assert np.array_equal(new[solid], old[solid])                # flag = 0 on solid cells
        exp = old - dt * g * flag
        assert np.max(np.abs(new[m] - exp[m])) < 1e-9 * abs(dt * g)


@pytest.mark.parametrize("mix", [0, 1])
def test_correct_ghostcell_velocity_linear_pressure(capi, oracle, w, mix):
    o = _oracle(capi, oracle, w, mix)
    a, b, c, dt = -0.9, 1.3, 0.45, 0.21
    P = _linear(w, a, b, c)
    u0, v0, w0 = (_rand(w.shape, s) for s in (4, 5, 6))
    u, v, ww = u0.copy(order="F"), v0.copy(order="F"), w0.copy(order="F")
    o.correct_ghostcell_velocity(P, dt, u, v, ww)
    g = w.gcm
    gm = np.zeros(w.shape, bool)
    gm[g["iGhost"], g["jGhost"], g["kGhost"]] = True
    assert gm.sum() == g["iGhost"].size > 0
    for new, old, gr in ((u, u0, a), (v, v0, b), (ww, w0, c)):
        assert np.array_equal(new[~gm], old[~gm])
        got = new[g["iGhost"], g["jGhost"], g["kGhost"]]
        ref = old[g["iGhost"], g["jGhost"], g["kGhost"]]
        if mix == 0:
            assert np.array_equal(got, ref)                           # flag 0: no change
        else:
            exp = ref - dt * gr * g["alphaGhost"]
            assert np.max(np.abs(got - exp)) < 1e-9 * abs(dt * gr)


def test_face_vel_linear_velocity(capi, oracle, w):
    o = _oracle(capi, oracle, w)
    x, y, z = w.grid
    nx, ny, nz = (t.size - 2 for t in w.grid)
    xc, yc, zc = (_centres(t) for t in w.grid)
    sh = w.shape
    G1, G2, G3 = _rand((1,) + sh[1:], 7), _rand((sh[0], 1, sh[2]), 8), _rand(sh[:2] + (1,), 9)
    U = np.asfortranarray(xc[:, None, None] * G1)
    V = np.asfortranarray(yc[None, :, None] * G2)
    W = np.asfortranarray(zc[None, None, :] * G3)
    bcxu, bcyv, bczw = _rand((2, ny + 2, nz + 2), 10), _rand((nx + 2, 2, nz + 2), 11), _rand((nx + 2, ny + 2, 2), 12)
    o.set_vel_bc(bcxu, bcyv, bczw)
    f0 = [_rand(sh, s) for s in (13, 14, 15)]
    fu, fv, fw = (a.copy(order="F") for a in f0)
    o.face_vel(U, V, W, fu, fv, fw)
    # written entries and their expected values
    wu = np.zeros(sh, bool); wu[1:nx + 1, 1:ny, 1:nz] = True
    wv = np.zeros(sh, bool); wv[1:nx, 1:ny + 1, 1:nz] = True
    ww_ = np.zeros(sh, bool); ww_[1:nx, 1:ny, 1:nz + 1] = True
    eu = np.broadcast_to(x[:, None, None] * G1, sh).copy()
    eu[1, :, :] = bcxu[0]
    eu[nx, :, :] = bcxu[1]
    ev = np.broadcast_to(y[None, :, None] * G2, sh).copy()
    ev[:, 1, :] = bcyv[:, 0, :]
    ev[:, ny, :] = bcyv[:, 1, :]
    ew = np.broadcast_to(z[None, None, :] * G3, sh).copy()
    ew[:, :, 1] = bczw[:, :, 0]
    ew[:, :, nz] = bczw[:, :, 1]
    for got, old, mask, exp in ((fu, f0[0], wu, eu), (fv, f0[1], wv, ev), (fw, f0[2], ww_, ew)):
        assert np.array_equal(got[~mask], old[~mask])
        assert np.max(np.abs(got[mask] - exp[mask])) < 1e-12 * np.max(np.abs(exp[mask]))


def test_update_pressure_invariants(capi, oracle, w):
    o = _oracle(capi, oracle, w)
    nx, ny, nz = (t.size - 2 for t in w.grid)
    P0 = _rand(w.shape, 16)
    P = P0.copy(order="F")
    p = np.zeros_like(P)
    o.update_pressure(P, p)
    assert np.array_equal(p, P)
    edge = np.zeros(w.shape, bool)
    for k in (0, nz, nz + 1):
        edge[0:nx + 1, [0, ny], k] = True
        edge[[0, nx], 0:ny + 2, k] = True
    edge[np.ix_([0, nx], [0, ny], np.arange(nz + 2))] = True
    solid = (w.iblank == 1) & (w.dead != -1)
    solid[nx + 1, :, :] = False
    assert np.all(P[edge] == 0.0) and np.all(P[solid] == 0.0)
    keep = ~(edge | solid)
    assert np.array_equal(P[keep], P0[keep])
    assert np.any((w.dead == -1) & keep)                             # dead-cell candidates are kept
