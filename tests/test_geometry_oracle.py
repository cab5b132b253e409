"""The oracle's restatement of the geometry refresh (oracle/fc_geom_oracle.c)
pinned three ways:

* against the reference's own routines, compiled from its sources by
  oracle/ref.mk and driven by oracle/ref_harness.f90 (mode 5), where the
  reference tree is present: masks, lists, body intercepts, normals, image
  points, probe lengths and alphaGhost bit for bit; the Vandermonde
  coefficients to 1e-8 (the reference inverts its 8x8 matrices with LAPACK's
  DGETRF/DGETRI, which round differently from the oracle's LU; the matrices
  are formed in absolute coordinates and are ill-conditioned);
* against golden fixtures written by the same compiled code
  (tests/golden/geometry_*.npz), everywhere;
* against what the mathematics fixes: the coefficients reproduce every
  trilinear function at the evaluation point, the element and vertex
  normals recomputed with numpy, iblank of a sphere against an exact
  inside test away from the surface, ghost and dead cells by brute force
  from iblank, alphaGhost from the dead cells' distances, the body
  intercept of a sphere at the sphere's distance."""
import os

import numpy as np
import pytest

import _ref_mg
from geometry_cases import CASES, case

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
INC = np.array([[0, 0, 0], [0, 1, 0], [1, 1, 0], [1, 0, 0], [0, 0, 1], [0, 1, 1], [1, 1, 1], [1, 0, 1]])
INTEGER_KEYS = ("iblank", "ghostCellMark", "dead_cell", "bodyNum", "iGhost", "jGhost", "kGhost", "iCellIndex",
                "jCellIndex", "kCellIndex", "closestElementGC", "iDead", "jDead", "kDead", "iDeadCellIndex",
                "jDeadCellIndex", "kDeadCellIndex", "closestElementDead")
EXACT_KEYS = ("xBodyIntercept", "yBodyIntercept", "zBodyIntercept", "xBodyInterceptNorm", "yBodyInterceptNorm",
              "zBodyInterceptNorm", "probeLength", "xImagePoint", "yImagePoint", "zImagePoint", "alphaGhost",
              "xBodyInterceptDead", "yBodyInterceptDead", "zBodyInterceptDead", "xNormInterceptDead",
              "yNormInterceptDead", "zNormInterceptDead", "probeLengthDead", "xyzDCIP")
COEF_KEYS = ("coeffGCMD", "coeffGCMN", "coeffGCMDeadD", "coeffGCMDeadN")


def _run(capi, geom, geom_oracle, name):
    x, y, z, bodies, tk = case(name)
    o = geom_oracle.GeomOracle(capi, geom, x, y, z, bodies, membrane_tkns=tk)
    st = o.refresh()
    r = o.results()
    o.close()
    return (x, y, z, bodies, tk), st, r


def _check_against(r, ref, coef_tol=1e-8):
    assert r["nGhost"] == ref["nGhost"] and r["nDead"] == ref["nDead"]
    for k in INTEGER_KEYS + EXACT_KEYS:
        assert np.array_equal(r[k], ref[k]), k
    for k in COEF_KEYS:
        np.testing.assert_allclose(r[k], ref[k], rtol=0, atol=coef_tol, err_msg=k)


@pytest.mark.parametrize("name", ["sphere", "two_bodies", "plate"])
def test_oracle_matches_reference(name, capi, geom, geom_oracle, tmp_path):
    if not _ref_mg.available():
        pytest.skip("reference tree absent (the golden fixtures pin the oracle here)")
    (x, y, z, bodies, tk), st, r = _run(capi, geom, geom_oracle, name)
    ref = _ref_mg.run_geometry(tmp_path, capi, x, y, z, bodies, membrane_tkns=tk)
    _check_against(r, ref)
    for sa, sb in zip(r["surfaces"], ref["surfaces"]):
        for k in sb:
            assert np.array_equal(sa[k], sb[k]), k


@pytest.mark.parametrize("name", ["two_bodies", "plate"])
def test_oracle_matches_golden(name, capi, geom, geom_oracle, surface):
    g = np.load(os.path.join(GOLD, "geometry_%s.npz" % name))
    bodies = [surface.Surface(g["b%d_x" % b], g["b%d_y" % b], g["b%d_z" % b], g["b%d_tri" % b],
                              kind=int(g["b%d_kind" % b])) for b in range(int(g["nbody"]))]
    o = geom_oracle.GeomOracle(capi, geom, g["x"], g["y"], g["z"], bodies, membrane_tkns=float(g["tkns"]))
    o.refresh()
    r = o.results()
    o.close()
    ref = {k: g[k] for k in g.files}
    ref["nGhost"], ref["nDead"] = g["iGhost"].size, g["iDead"].size
    _check_against(r, ref)
    for b in range(len(bodies)):
        for k in geom.SURF_ELEM + geom.SURF_VERT:
            assert np.array_equal(r["surfaces"][b][k], g["b%d_%s" % (b, k)]), k


def _metrics(a):
    n = a.size - 2
    c = np.zeros(n + 2)
    c[0:n + 1] = 0.5 * (a[1:n + 2] + a[0:n + 1])
    c[n + 1] = a[n + 1]
    return c


def _basis(X, Y, Z):
    return np.stack([X * Y * Z, X * Y, X * Z, Y * Z, X, Y, Z, np.ones_like(X)], axis=-1)


@pytest.mark.parametrize("name", CASES)
def test_coefficients_reproduce_trilinear(name, capi, geom, geom_oracle):
    """sum_r c_r f(Q_r) = f(eval) for f in {xyz, xy, xz, yz, x, y, z, 1}:
    Q_r the host-cell corners, the ghost/dead cell's own corner replaced by
    the body intercept (Dirichlet form) or by the normal derivative there
    (Neumann form).  Ghost rows are evaluated at the ghost cell, dead rows at
    the image point (GCM_SET_BOUNDARY.F90:404-406, 1469-1471)."""
    (x, y, z, bodies, tk), st, r = _run(capi, geom, geom_oracle, name)
    xc, yc, zc = _metrics(x), _metrics(y), _metrics(z)
    worst = 0.0
    for kind in ("ghost", "dead"):
        if kind == "ghost":
            I, J, K = r["iGhost"], r["jGhost"], r["kGhost"]
            C = np.stack([r["iCellIndex"], r["jCellIndex"], r["kCellIndex"]], 1)
            ev = np.stack([xc[I], yc[J], zc[K]], 1)
            cD, cN = r["coeffGCMD"], r["coeffGCMN"]
            nrm = np.stack([r["xBodyInterceptNorm"], r["yBodyInterceptNorm"], r["zBodyInterceptNorm"]], 1)
            # the relation uses the intercept moved by half the membrane thickness
            bi = np.stack([r["xBodyIntercept"], r["yBodyIntercept"], r["zBodyIntercept"]], 1)
            if tk > 0:
                bi = bi - 0.5 * tk * nrm
        else:
            I, J, K = r["iDead"], r["jDead"], r["kDead"]
            C = np.stack([r["iDeadCellIndex"], r["jDeadCellIndex"], r["kDeadCellIndex"]], 1)
            ev = r["xyzDCIP"][I, J, K, :]
            cD, cN = r["coeffGCMDeadD"], r["coeffGCMDeadN"]
            nrm = np.stack([r["xNormInterceptDead"], r["yNormInterceptDead"], r["zNormInterceptDead"]], 1)
            if tk > 0:
                continue      # the membrane's virtual intercept is not among the outputs
            bi = np.stack([r["xBodyInterceptDead"], r["yBodyInterceptDead"], r["zBodyInterceptDead"]], 1)
        for n in range(I.size):
            P = np.stack([xc[C[n, 0] + INC[:, 0]], yc[C[n, 1] + INC[:, 1]], zc[C[n, 2] + INC[:, 2]]], 1)
            self_row = np.all((C[n] + INC) == np.array([I[n], J[n], K[n]]), axis=1)
            VD = _basis(P[:, 0], P[:, 1], P[:, 2])
            VN = VD.copy()
            if self_row.any():
                b, m = bi[n], nrm[n]
                VD[self_row] = _basis(b[0], b[1], b[2])
                VN[self_row] = [m[0] * b[1] * b[2] + m[1] * b[0] * b[2] + m[2] * b[0] * b[1], m[0] * b[1] + m[1] * b[0],
                                m[0] * b[2] + m[2] * b[0], m[1] * b[2] + m[2] * b[1], m[0], m[1], m[2], 0.0]
            e = _basis(ev[n, 0], ev[n, 1], ev[n, 2])
            worst = max(worst, float(np.max(np.abs(cD[n] @ VD - e))), float(np.max(np.abs(cN[n] @ VN - e))))
    assert worst < 1e-8, worst


def test_surface_metrics_numpy(capi, geom, geom_oracle):
    (x, y, z, bodies, tk), st, r = _run(capi, geom, geom_oracle, "two_bodies")
    for s, sm in zip(bodies, r["surfaces"]):
        P = np.stack([s.x, s.y, s.z], 1)
        t = s.tri.T - 1
        n = np.cross(P[t[:, 1]] - P[t[:, 0]], P[t[:, 2]] - P[t[:, 1]])
        a = np.linalg.norm(n, axis=1)
        np.testing.assert_allclose(sm["triElemArea"], 0.5 * a, rtol=1e-14)
        en = n / a[:, None]
        np.testing.assert_allclose(np.stack([sm["triElemNormX"], sm["triElemNormY"], sm["triElemNormZ"]], 1), en,
                                   atol=1e-14)
        np.testing.assert_allclose(np.stack([sm["triElemCentX"], sm["triElemCentY"], sm["triElemCentZ"]], 1),
                                   P[t].mean(axis=1), atol=1e-15)
        vn = np.zeros_like(P)
        for q in range(3):
            np.add.at(vn, t[:, q], en)
        vn /= np.linalg.norm(vn, axis=1)[:, None]
        np.testing.assert_allclose(np.stack([sm["xNormBodyMarker"], sm["yNormBodyMarker"], sm["zNormBodyMarker"]], 1),
                                   vn, atol=1e-13)


def test_sphere_iblank_and_marks(capi, geom, geom_oracle):
    """iblank of a closed sphere against the exact inside test, for cell
    centres farther from the sphere than the facets' sag; ghost and dead
    cells by brute force from iblank; lists in array order."""
    (x, y, z, bodies, tk), st, r = _run(capi, geom, geom_oracle, "sphere")
    xc, yc, zc = _metrics(x), _metrics(y), _metrics(z)
    X, Y, Z = np.meshgrid(xc, yc, zc, indexing="ij")
    R = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2)
    ib = r["iblank"]
    inner = np.zeros(ib.shape, bool)
    inner[1:-2, 1:-2, 1:-2] = True
    # the facets' largest sag: the sphere radius less the smallest distance
    # from the centre to a facet's plane
    s0 = bodies[0]
    P = np.stack([s0.x, s0.y, s0.z], 1) - 0.5
    t = s0.tri.T - 1
    nf = np.cross(P[t[:, 1]] - P[t[:, 0]], P[t[:, 2]] - P[t[:, 1]])
    sag = 0.2 - np.min(np.abs(np.sum(nf * P[t[:, 0]], axis=1)) / np.linalg.norm(nf, axis=1)) + 1e-12
    far = inner & (np.abs(R - 0.2) > sag)
    assert np.array_equal(ib[far] == 1, R[far] < 0.2)
    # ghost cells: interior fluid cells with an iblank face neighbour
    gh = np.zeros(ib.shape, bool)
    solid = ib == 1
    for ax in range(3):
        for d in (-1, 1):
            gh[1:-2, 1:-2, 1:-2] |= np.roll(solid, -d, axis=ax)[1:-2, 1:-2, 1:-2]
    gh &= inner & (ib == 0)
    assert np.array_equal(r["ghostCellMark"] == 1, gh)
    order = np.nonzero(gh.ravel(order="F"))[0]
    gi, gj, gk = np.unravel_index(order, ib.shape, order="F")
    assert np.array_equal(r["iGhost"], gi) and np.array_equal(r["jGhost"], gj) and np.array_equal(r["kGhost"], gk)
    dead = np.zeros(ib.shape, bool)
    for ax in range(3):
        for d in (-1, 1):
            dead |= np.roll(gh, d, axis=ax)
    dead &= solid
    assert np.array_equal(r["dead_cell"] == -1, dead)
    # alphaGhost = 1 - sqrt(sum over dead face neighbours of (ds/dc)^2), with
    # ds the dead cell's intercept distance (probeLengthDead / probe)
    dsD = np.zeros(ib.shape)
    dsD[r["iDead"], r["jDead"], r["kDead"]] = r["probeLengthDead"] / 2.0
    dxc, dyc, dzc = np.diff(xc[:-1]), np.diff(yc[:-1]), np.diff(zc[:-1])
    alpha = np.zeros(r["nGhost"])
    for n, (i, j, k) in enumerate(zip(gi, gj, gk)):
        s = 0.0
        for (a, b, c, dc) in ((0, 0, -1, dzc[k - 1]), (0, -1, 0, dyc[j - 1]), (-1, 0, 0, dxc[i - 1]),
                              (1, 0, 0, dxc[i]), (0, 1, 0, dyc[j]), (0, 0, 1, dzc[k])):
            if dead[i + a, j + b, k + c]:
                s += (dsD[i + a, j + b, k + c] / dc) ** 2
        alpha[n] = 1.0 - np.sqrt(s)
    np.testing.assert_allclose(r["alphaGhost"], alpha, rtol=0, atol=1e-13)
    # body intercepts lie on the sphere up to the facets' sag
    bi = np.stack([r["xBodyIntercept"], r["yBodyIntercept"], r["zBodyIntercept"]], 1)
    rb = np.linalg.norm(bi - 0.5, axis=1)
    assert np.all(rb <= 0.2 + 1e-12) and np.all(rb >= 0.2 - sag)
