"""Writes the golden fixtures of this directory from the reference's own
code: oracle/ref.mk compiles the reference's sources (under /root/reference)
into oracle/_ref/ref_harness, and this script runs it on small seeded cases
and stores inputs and outputs (numpy .npz, no pickles).

    python tests/golden/make_golden.py

The fixtures pin the oracle where the reference tree is absent
(tests/test_golden.py): mg_solve and the grid/level setup (CORE_MG.F90),
correct_vel / face_vel / update_pressure (CORE_FRAC_STEP_ROUTINES.F90),
set_solve_ad and rhs_advec_diff (CORE_SET_SOLVERS.F90), and the geometry
refresh of triangulated bodies (BOUNDARY_SET_ARCLENGTH_NORM.F90,
BOUNDARY_IBLANK_FAST.F90, GCM_SET_BOUNDARY.F90; python make_golden.py
--geometry writes only those)."""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import _pkg  # noqa: E402

_pkg.load_package()
import _ref_mg  # noqa: E402
import feicar3d_amd.capi as capi  # noqa: E402
import feicar3d_amd.geometry as geo  # noqa: E402


def small_grid():
    """12 x 10 x 9 cells, stretched in x, with a random solid region and
    dead-cell marks (the projection needs masks, not a consistent body)."""
    x = geo.tanh_axis(13, [0.4], 1.0 / 64.0, width=0.2)
    y = geo.uniform_axis(11)
    z = geo.tanh_axis(10, [0.5], 1.0 / 32.0, width=0.3)
    sh = (x.size, y.size, z.size)
    rng = np.random.default_rng(101)
    iblank = np.asfortranarray((rng.random(sh) < 0.2).astype(np.int8))
    dead = np.asfortranarray(np.where((iblank == 1) & (rng.random(sh) < 0.5), -1, 0).astype(np.int8))
    ghost = np.zeros(sh, np.int8, order="F")
    return x, y, z, iblank, ghost, dead


GEOMETRY_KEYS = ("iblank", "ghostCellMark", "dead_cell", "bodyNum", "iGhost", "jGhost", "kGhost", "iCellIndex",
                 "jCellIndex", "kCellIndex", "closestElementGC", "xBodyIntercept", "yBodyIntercept",
                 "zBodyIntercept", "xBodyInterceptNorm", "yBodyInterceptNorm", "zBodyInterceptNorm", "probeLength",
                 "xImagePoint", "yImagePoint", "zImagePoint", "coeffGCMD", "coeffGCMN", "alphaGhost", "iDead",
                 "jDead", "kDead", "iDeadCellIndex", "jDeadCellIndex", "kDeadCellIndex", "closestElementDead",
                 "xBodyInterceptDead", "yBodyInterceptDead", "zBodyInterceptDead", "xNormInterceptDead",
                 "yNormInterceptDead", "zNormInterceptDead", "probeLengthDead", "coeffGCMDeadD", "coeffGCMDeadN",
                 "xyzDCIP")


def write_geometry(tmp):
    """the reference's geometry refresh on two of the test cases: two closed
    bodies on a stretched grid, and a pitched membrane plate"""
    from geometry_cases import case
    for name in ("two_bodies", "plate"):
        x, y, z, bodies, tk = case(name)
        r = _ref_mg.run_geometry(tmp, capi, x, y, z, bodies, membrane_tkns=tk)
        d = dict(x=x, y=y, z=z, tkns=np.float64(tk), nbody=np.int32(len(bodies)))
        for b, s in enumerate(bodies):
            d.update({"b%d_x" % b: s.x, "b%d_y" % b: s.y, "b%d_z" % b: s.z, "b%d_tri" % b: s.tri,
                      "b%d_kind" % b: np.int32(s.kind)})
            for k, v in r["surfaces"][b].items():
                d["b%d_%s" % (b, k)] = v
        for k in GEOMETRY_KEYS:
            d[k] = r[k]
        np.savez_compressed(os.path.join(HERE, "geometry_%s.npz" % name), **d)


def write_stress(tmp):
    """GCM_SetBodyInterceptValues (ntime 2) and compute_marker_stress of the
    reference on the same two cases, with the seeded inputs of
    geometry_cases.stress_inputs (regenerated by the tests, not stored)"""
    from geometry_cases import case, stress_inputs
    import feicar3d_amd.geom as geom
    for name in ("two_bodies", "plate"):
        x, y, z, bodies, tk = case(name)
        r = _ref_mg.run_geometry(tmp, capi, x, y, z, bodies, membrane_tkns=tk, stress=stress_inputs(name))
        d = {k: r[k] for k in geom.BI_NAMES}
        for b, f in enumerate(r["forces"]):
            for k in ["totals"] + geom.FORCE_NAMES:
                d["b%d_%s" % (b, k)] = f[k]
        np.savez_compressed(os.path.join(HERE, "marker_stress_%s.npz" % name), **d)


def main():
    if not _ref_mg.available():
        raise SystemExit("the reference tree and its harness are needed to write the fixtures")
    tmp = tempfile.mkdtemp()
    write_geometry(tmp)
    write_stress(tmp)
    if "--geometry" in sys.argv:
        return
    prm = capi.default_params(iterInter=1, restol_pson=0.0)
    # mg_solve: two V-cycles on the C1 sphere at 16^3 from a random start
    w = geo.config_c1(16)
    rng = np.random.default_rng(7)
    var0 = np.asfortranarray(w.p_exact + 0.1 * rng.standard_normal(w.p_exact.shape))
    res = _ref_mg.run(tmp, *w.grid, prm, w.iblank, w.ghost, w.dead, var0, w.div)
    np.savez_compressed(os.path.join(HERE, "mg_solve_c1_16.npz"), x=w.grid[0], y=w.grid[1], z=w.grid[2],
                        iblank=w.iblank, ghost=w.ghost, dead=w.dead, var0=var0, rr=w.div,
                        iterInter=np.int32(prm.iterInter), var=res["var"], resm=np.float64(res["resm"]))
    # projection on the small stretched grid
    x, y, z, ib, gh, de = small_grid()
    sh = ib.shape
    rng = np.random.default_rng(11)
    f = lambda s: np.asfortranarray(rng.standard_normal(s))  # noqa: E731
    nx, ny, nz = sh[0] - 2, sh[1] - 2, sh[2] - 2
    st = dict(P=f(sh), u=f(sh), v=f(sh), w=f(sh), fu=f(sh), fv=f(sh), fw=f(sh),
              bcxu=f((2, ny + 2, nz + 2)), bcyv=f((nx + 2, 2, nz + 2)), bczw=f((nx + 2, ny + 2, 2)))
    dt = 0.0137
    (u, v, ww), (fu, fv, fw), (P, p) = _ref_mg.run_projection(
        tmp, x, y, z, capi.default_params(), ib, gh, de, st["P"], dt, st["u"], st["v"], st["w"], st["fu"],
        st["fv"], st["fw"], st["bcxu"], st["bcyv"], st["bczw"])
    np.savez_compressed(os.path.join(HERE, "projection_small.npz"), x=x, y=y, z=z, iblank=ib, ghost=gh, dead=de,
                        dt=np.float64(dt), **{"in_" + k: a for k, a in st.items()}, u=u, v=v, w=ww, fu=fu, fv=fv,
                        fw=fw, P=P, p=p)
    # set_solve_ad on the same grid
    reinv, dts = 1.0 / 150.0, 0.02
    c = _ref_mg.run_ad_coefficients(tmp, x, y, z, capi.default_params(), ib, gh, de, reinv, dts)
    np.savez_compressed(os.path.join(HERE, "set_solve_ad_small.npz"), x=x, y=y, z=z, reinv=np.float64(reinv),
                        dt=np.float64(dts), amx=c[0], apx=c[1], amy=c[2], apy=c[3], amz=c[4], apz=c[5])
    # rhs_advec_diff on the same grid: random velocities, face velocities with
    # zeros, a random viscTot, nlu/nlv/nlw with values outside the interior
    rng = np.random.default_rng(23)
    f = lambda s: np.asfortranarray(rng.standard_normal(s))  # noqa: E731
    ra = dict(uo=f(sh), vo=f(sh), wo=f(sh), fu=f(sh), fv=f(sh), fw=f(sh),
              visc=np.asfortranarray(0.004 * (1.0 + rng.random(sh))), nlu=f(sh), nlv=f(sh), nlw=f(sh))
    for k in ("fu", "fv", "fw"):
        ra[k][rng.random(sh) < 0.1] = 0.0
    dta, alfa = 0.011, 0.7
    (nlu, nlv, nlw), _ = _ref_mg.run_rhs_advec_diff(tmp, x, y, z, capi.default_params(), ib, gh, de, dta, alfa,
                                                    *(ra[k] for k in ("uo", "vo", "wo", "fu", "fv", "fw", "visc",
                                                                      "nlu", "nlv", "nlw")))
    np.savez_compressed(os.path.join(HERE, "rhs_advec_diff_small.npz"), x=x, y=y, z=z, dt=np.float64(dta),
                        alfa=np.float64(alfa), **{"in_" + k: a for k, a in ra.items()}, nlu=nlu, nlv=nlv, nlw=nlw)
    print("wrote", sorted(n for n in os.listdir(HERE) if n.endswith(".npz")))


if __name__ == "__main__":
    main()
