"""GPU parity of the immersed-boundary geometry refresh (fc_geom_*: surface
normals, set_iblank_body_fast, the smoothing passes, ghost and dead cells
with their image points and GCM coefficients) against the CPU oracle's
restatement of the same routines, on the same inputs.

Every result is compared with ==: the masks, lists and host cells are
integer work, and the floating-point results (intercepts, normals, image
points, probe lengths, alphaGhost, the Vandermonde coefficients) are computed
with the same operations in the same order on both sides (no fused
multiply-add on the device)."""
import numpy as np
import pytest

from geometry_cases import CASES, case

pytestmark = pytest.mark.gpu


def _compare(a, b):
    for k in ("iblank", "ghostCellMark", "dead_cell", "bodyNum", "ps1", "ps2"):
        assert np.array_equal(a[k], b[k]), k
    assert a["nGhost"] == b["nGhost"] and a["nDead"] == b["nDead"]
    for k, v in b.items():
        if k in ("surfaces", "nGhost", "nDead"):
            continue
        assert np.array_equal(a[k], v), k
    for sa, sb in zip(a["surfaces"], b["surfaces"]):
        for k in sb:
            assert np.array_equal(sa[k], sb[k]), k


@pytest.mark.parametrize("name", CASES)
def test_refresh_matches_oracle(name, capi, geom, geom_oracle):
    x, y, z, bodies, tk = case(name)
    print("case", name, "nodes", x.size - 2, y.size - 2, z.size - 2, flush=True)
    o = geom_oracle.GeomOracle(capi, geom, x, y, z, bodies, membrane_tkns=tk)
    so = o.refresh()
    ro = o.results()
    g = geom.GeometryRefresh(x, y, z, bodies, membrane_tkns=tk)
    sg = g.refresh()
    rg = g.results()
    g.close()
    o.close()
    assert ro["nGhost"] > 0 and ro["nDead"] > 0
    assert sg.n_undecided == so.n_undecided and sg.n_smoothed == so.n_smoothed
    assert sg.n_unselected_side == so.n_unselected_side
    _compare(rg, ro)


def test_handoff_matches_results(capi, geom):
    """GeometryRefresh.handoff (masks, lists, fc_geom_get_dead_ip) holds the
    same values as the full results (xyzDCIP gathered at the dead cells)."""
    x, y, z, bodies, tk = case("plate")
    g = geom.GeometryRefresh(x, y, z, bodies, membrane_tkns=tk)
    g.refresh()
    r = g.results()
    h = g.handoff()
    g.close()
    assert h["nDead"] > 0
    for k, v in h.items():
        if k == "xyzDCIP_dead":
            assert np.array_equal(v, r["xyzDCIP"][r["iDead"], r["jDead"], r["kDead"], :])
        else:
            assert np.array_equal(v, r[k]), k


def test_pitching_plate_sequence(capi, geom, geom_oracle, surface):
    """A membrane plate pitching as 0.3 sin(2 pi t): one handle refreshed at
    successive times (move_body, then refresh) beside one oracle state moved
    the same way.  Like the reference's xyzDCIP, the image-point array keeps
    the entries of cells that stopped being ghost or dead cells."""
    x, y, z, bodies, tk = case("plate")
    p0 = surface.plate((0.5, 0.5, 0.5), side=0.4, n=21, amplitude=0.01, seed=3)
    g = geom.GeometryRefresh(x, y, z, [p0], membrane_tkns=tk)
    o = geom_oracle.GeomOracle(capi, geom, x, y, z, [p0], membrane_tkns=tk)
    for t in (0.0, 0.1, 0.2, 0.35):
        s = surface.pitch(p0, (0.5, 0.5, 0.5), surface.pitch_angle(t))
        g.move_body(1, s)
        g.refresh()
        rg = g.results()
        o.move_body(1, s)
        o.refresh()
        ro = o.results()
        print("t", t, "nGhost", ro["nGhost"], "nDead", ro["nDead"], flush=True)
        _compare(rg, ro)
    o.close()
    g.close()


def test_refresh_argument_checks(capi, geom, surface):
    from feicar3d_amd.poisson import FcError
    import ctypes as C
    x, y, z, bodies, tk = case("sphere")
    with pytest.raises(ValueError):
        geom.GeometryRefresh(x, y, z, [])
    # the C entry points refuse what the Python mirror would not send
    L = geom.library()
    ga = [np.ascontiguousarray(a) for a in (x, y, z)]
    g = capi.FcGrid()
    g.nx, g.ny, g.nz = (a.size - 2 for a in ga)
    g.x, g.y, g.z = (a.ctypes.data_as(C.POINTER(C.c_double)) for a in ga)
    p = geom.make_params(bodies)
    h = C.c_void_p()
    for field, value in (("nBody", 0), ("nBody", 9)):
        q = geom.make_params(bodies)
        setattr(q, field, value)
        assert L.fc_geom_create(C.byref(g), C.byref(q), C.byref(h)) == -1
    q = geom.make_params(bodies)
    q.unstruc_surface_type[0] = 3
    assert L.fc_geom_create(C.byref(g), C.byref(q), C.byref(h)) == -1
    assert L.fc_geom_create(C.byref(g), C.byref(p), C.byref(h)) == 0
    st = geom.FcGeomStats()
    assert L.fc_geom_refresh(h, C.byref(st)) == -1          # no body set yet
    L.fc_geom_destroy(h)
    bad = surface.Surface(bodies[0].x, bodies[0].y, bodies[0].z, bodies[0].tri.copy())
    bad.tri[0, 0] = bad.npts + 1
    with pytest.raises(FcError):
        geom.GeometryRefresh(x, y, z, [bad])


def _stress(backend, inp, ntime=2):
    backend.refresh()
    for b, m in enumerate(inp["marker_vel"]):
        backend.set_marker_velocity(b + 1, *m)
    backend.body_intercept_values(inp["dt"], ntime)
    backend.marker_stress(inp["p"], inp["u"], inp["v"], inp["w"], inp["reinv"])
    return backend.stress_results()


@pytest.mark.parametrize("name", CASES)
def test_marker_stress_matches_oracle(name, capi, geom, geom_oracle):
    """GCM_SetBodyInterceptValues and compute_marker_stress on the device:
    the body-intercept velocities and dp/dn of every ghost and dead cell, the
    per-marker forces, shears, areas and stresses and the nine totals per
    body, all == the oracle (same operations in the same order; the totals
    are summed in element order on the device too)."""
    from geometry_cases import stress_inputs
    x, y, z, bodies, tk = case(name)
    inp = stress_inputs(name)
    print("case", name, "markers", [s.npts for s in bodies], flush=True)
    o = geom_oracle.GeomOracle(capi, geom, x, y, z, bodies, membrane_tkns=tk)
    ro = _stress(o, inp)
    g = geom.GeometryRefresh(x, y, z, bodies, membrane_tkns=tk)
    rg = _stress(g, inp)
    for k in geom.BI_NAMES:
        assert np.array_equal(rg[k], ro[k]), k
    for fg, fo in zip(rg["forces"], ro["forces"]):
        for k in fo:
            assert np.array_equal(fg[k], fo[k]), k
    # a second step on the same handle: first-step dp/dn and new fields
    inp2 = stress_inputs(name, seed=9)
    ro = _stress(o, inp2, ntime=1)
    rg = _stress(g, inp2, ntime=1)
    assert np.all(rg["dpdnBodyIntercept"] == 0.0)
    for k in geom.BI_NAMES:
        assert np.array_equal(rg[k], ro[k]), k
    for fg, fo in zip(rg["forces"], ro["forces"]):
        for k in fo:
            assert np.array_equal(fg[k], fo[k]), k
    g.close()
    o.close()


def test_marker_stress_call_order(capi, geom):
    from feicar3d_amd.poisson import FcError
    from geometry_cases import stress_inputs
    x, y, z, bodies, tk = case("sphere")
    inp = stress_inputs("sphere")
    g = geom.GeometryRefresh(x, y, z, bodies, membrane_tkns=tk)
    with pytest.raises(FcError):
        g.body_intercept_values(0.01, 2)           # before any refresh
    g.refresh()
    with pytest.raises(FcError):
        g.marker_stress(inp["p"], inp["u"], inp["v"], inp["w"], inp["reinv"])   # before the BI values
    with pytest.raises(FcError):
        g.body_intercept_values(0.0, 2)            # dt = 0 with ntime > 1
    g.close()
