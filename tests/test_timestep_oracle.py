"""The time-step driver's host logic (feicar3d_amd.timestep) without a GPU,
on the oracle's backends (tests/timestep_backends.py): one step of the
pitching plate runs the reference's sequence to convergence, and the lists
the driver hands to the flow solver are the refresh's (the dead cells'
image points gathered from xyzDCIP at iDead, jDead, kDead)."""
import numpy as np

from geometry_cases import case


def test_pitching_plate_step_on_the_oracle(capi, geom, surface, geom_oracle, oracle):
    import feicar3d_amd.timestep as ts
    import timestep_backends as tb
    x, y, z, _, tk = case("plate")
    plate = surface.plate((0.5, 0.5, 0.5), side=0.4, n=21, amplitude=0.01, seed=3)
    c = (0.5, 0.5, 0.5)
    b = tb.make(capi, geom, ts, x, y, z, [plate], tk, gpu=False)
    s, mv = ts.pitching_plate_motion(plate, c, b.dt, b.dt)
    # the marker velocities are the rigid rotation's: omega x (r - c)
    om = 0.6 * np.pi * np.cos(2.0 * np.pi * b.dt)
    assert np.allclose(mv[0], -om * (s.y - c[1])) and np.allclose(mv[1], om * (s.x - c[0]))
    r = b.step([s], [mv])
    assert r["ad_rc"] == 0 and r["poisson_rc"] == 0
    assert r["nGhost"] > 0 and r["nDead"] > 0
    assert np.all(np.isfinite(r["totals"])) and np.all(np.isfinite(b.p))
    full = b.geom.results()
    h = b.geom.handoff()
    g = ts.gcm_lists(h, b.geom.bi_values())
    assert np.array_equal(g["xyzDCIP"], full["xyzDCIP"][full["iDead"], full["jDead"], full["kDead"], :])
    for k in ("iGhost", "kCellIndex", "coeffGCMN", "alphaGhost", "iDead", "probeLengthDead", "coeffGCMDeadN"):
        assert np.array_equal(g[k], full[k]), k
    b.geom.close()
