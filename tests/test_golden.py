"""The oracle against golden fixtures written by the reference's own compiled
code (tests/golden/make_golden.py runs oracle/_ref/ref_harness, built from the
reference's sources by oracle/ref.mk).  These run wherever the oracle builds,
with or without the reference tree, and must agree bit for bit:
- mg_solve, two V-cycles on the C1 sphere at 16^3 (CORE_MG.F90:1572-1848);
- correct_vel, face_vel, update_pressure on a small stretched masked grid
  (CORE_FRAC_STEP_ROUTINES.F90:6-299);
- set_solve_ad on the same grid (CORE_SET_SOLVERS.F90:591-676);
- rhs_advec_diff on the same grid, random velocities and viscTot
  (CORE_SET_SOLVERS.F90:10-526)."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLD, name))     # allow_pickle stays False


def _F(a):
    return np.asfortranarray(a)


def test_golden_mg_solve(capi, oracle):
    g = _load("mg_solve_c1_16.npz")
    prm = capi.default_params(iterInter=int(g["iterInter"]), restol_pson=0.0)
    o = oracle.Oracle(capi, g["x"], g["y"], g["z"], prm, ordering=oracle.ORDER_REFERENCE)
    o.set_masks(_F(g["iblank"]), _F(g["ghost"]), _F(g["dead"]))
    a = _F(g["var0"].copy())
    o.mg_solve(a, _F(g["rr"]))
    assert np.array_equal(a, g["var"])
    assert o.calc_residual(a, _F(g["rr"])) == float(g["resm"])
    o.close()


def test_golden_projection(capi, oracle):
    g = _load("projection_small.npz")
    o = oracle.Oracle(capi, g["x"], g["y"], g["z"], capi.default_params())
    o.set_masks(_F(g["iblank"]), _F(g["ghost"]), _F(g["dead"]))
    o.set_vel_bc(g["in_bcxu"], g["in_bcyv"], g["in_bczw"])
    dt = float(g["dt"])
    u, v, w = (_F(g["in_" + k].copy()) for k in ("u", "v", "w"))
    o.correct_vel(_F(g["in_P"]), dt, u, v, w)
    assert np.array_equal(u, g["u"]) and np.array_equal(v, g["v"]) and np.array_equal(w, g["w"])
    fu, fv, fw = (_F(g["in_" + k].copy()) for k in ("fu", "fv", "fw"))
    o.face_vel(u, v, w, fu, fv, fw)
    assert np.array_equal(fu, g["fu"]) and np.array_equal(fv, g["fv"]) and np.array_equal(fw, g["fw"])
    P = _F(g["in_P"].copy())
    p = np.zeros_like(P)
    o.update_pressure(P, p)
    assert np.array_equal(P, g["P"]) and np.array_equal(p, g["p"])
    assert np.count_nonzero(P == 0.0) > 0
    o.close()


def test_golden_set_solve_ad(capi, oracle):
    g = _load("set_solve_ad_small.npz")
    o = oracle.Oracle(capi, g["x"], g["y"], g["z"], capi.default_params())
    o.set_solve_ad(float(g["reinv"]), float(g["dt"]), 1, 1.0)
    for got, k in zip(o.get_solve_ad(), ("amx", "apx", "amy", "apy", "amz", "apz")):
        assert np.array_equal(got, g[k]), k
    o.close()


def test_golden_rhs_advec_diff(capi, oracle):
    g = _load("rhs_advec_diff_small.npz")
    o = oracle.Oracle(capi, g["x"], g["y"], g["z"], capi.default_params())
    o.set_solve_ad(0.004, float(g["dt"]), 1, 1.0)
    nl = [_F(g["in_" + k]) for k in ("nlu", "nlv", "nlw")]
    ins = [_F(g["in_" + k]) for k in ("uo", "vo", "wo", "fu", "fv", "fw")]
    assert o.rhs_advec_diff(*ins, float(g["alfa"]), float(g["dt"]), *nl, visc=_F(g["in_visc"])) == 0
    o.close()
    for got, k in zip(nl, ("nlu", "nlv", "nlw")):
        assert np.array_equal(got, g[k]), k
