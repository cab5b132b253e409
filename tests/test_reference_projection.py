"""The oracle's velocity projection against the reference's own code.

oracle/ref.mk compiles CORE_FRAC_STEP_ROUTINES.F90 unmodified with flang
beside the multigrid half, and oracle/ref_harness.f90 (mode 2) runs
metrics, correct_vel (:134-210), face_vel (:6-130) and update_pressure
(:212-299) on one state.  The oracle has to give the same bits after each
call.  correct_ghostcell_velocity lives in GCM_GHOSTCELL_UPDATE.F90, which
flang does not compile; the closed forms of tests/test_projection_oracle.py
pin it.

Skipped when neither the reference tree nor a built harness is present."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import _ref_mg  # noqa: E402

pytestmark = pytest.mark.skipif(not _ref_mg.available(), reason="reference tree and harness absent")


def _state(w, seed):
    rng = np.random.default_rng(seed)
    sh = w.shape
    nx, ny, nz = (t.size - 2 for t in w.grid)
    f = lambda s: np.asfortranarray(rng.standard_normal(s))  # noqa: E731
    return dict(P=f(sh), u=f(sh), v=f(sh), w=f(sh), fu=f(sh), fv=f(sh), fw=f(sh),
                bcxu=f((2, ny + 2, nz + 2)), bcyv=f((nx + 2, 2, nz + 2)), bczw=f((nx + 2, ny + 2, 2)))


@pytest.mark.parametrize("cfg", ["c1", "c2"])
def test_projection_matches_reference(capi, geo, oracle, tmp_path, cfg):
    w = geo.config_c1(16) if cfg == "c1" else geo.config_c2(24)
    prm = capi.default_params()
    s = _state(w, 21)
    dt = 0.0137
    (ru, rv, rw), (rfu, rfv, rfw), (rP, rp) = _ref_mg.run_projection(
        tmp_path, *w.grid, prm, w.iblank, w.ghost, w.dead, s["P"], dt, s["u"], s["v"], s["w"],
        s["fu"], s["fv"], s["fw"], s["bcxu"], s["bcyv"], s["bczw"])
    o = oracle.Oracle(capi, *w.grid, prm)
    o.set_workload(w)
    o.set_vel_bc(s["bcxu"], s["bcyv"], s["bczw"])
    u, v, ww = (s[k].copy(order="F") for k in ("u", "v", "w"))
    o.correct_vel(s["P"], dt, u, v, ww)
    assert np.array_equal(u, ru) and np.array_equal(v, rv) and np.array_equal(ww, rw)
    assert not np.array_equal(u, s["u"])
    fu, fv, fw = (s[k].copy(order="F") for k in ("fu", "fv", "fw"))
    o.face_vel(u, v, ww, fu, fv, fw)
    assert np.array_equal(fu, rfu) and np.array_equal(fv, rfv) and np.array_equal(fw, rfw)
    P = s["P"].copy(order="F")
    p = np.zeros_like(P)
    o.update_pressure(P, p)
    assert np.array_equal(P, rP) and np.array_equal(p, rp)
    assert np.count_nonzero(P == 0.0) > 0
