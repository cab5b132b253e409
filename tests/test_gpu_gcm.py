"""GCM_ghostcell_pressure (GCM_GHOSTCELL_UPDATE.F90:361-739) on the GPU
against the oracle in the cases the default workloads leave out: the hybrid
ghost update (mix_GC_form = 1, with its res_laplacian output), dead cells on
the inverse-distance branch with a nonzero previous pressure, and a whole
solve_poisson with mix_GC_form = 1.  The oracle's forms are pinned in
test_oracle_pins_gcm.py."""
import numpy as np
import pytest

from test_oracle_pins_gcm import idw_workload

pytestmark = pytest.mark.gpu


def _pair(capi, oracle, gpu, w, g, ordering, **kw):
    p = capi.default_params(**kw)
    o = oracle.Oracle(capi, *w.grid, p, ordering=ordering)
    o.set_masks(w.iblank, w.ghost, w.dead)
    o.set_gcm(g)
    s = gpu.PoissonSolver(*w.grid, p)
    s.set_masks(w.iblank, w.ghost, w.dead)
    s.set_gcm(g)
    return o, s


@pytest.mark.parametrize("sweeps", [1, 4])
def test_hybrid_ghost_update(capi, geo, oracle, gpu, sweeps):
    w = geo.config_c2(24)
    o, s = _pair(capi, oracle, gpu, w, w.gcm, oracle.ORDER_COLORED, mix_GC_form=1, itermax_gc=sweeps,
                 restol_gcp=0.0)
    rng = np.random.default_rng(12)
    q0 = np.asfortranarray(w.p_exact + 0.1 * rng.standard_normal(w.div.shape))
    a, b = q0.copy(order="F"), q0.copy(order="F")
    _, st_o, rl_o = o.ghostcell_pressure(a, w.div)
    _, st_g, rl_g = s.ghostcell_pressure(b, w.div)
    o.close()
    s.close()
    assert st_g.gcm_sweeps == st_o.gcm_sweeps == sweeps
    assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(a))
    assert np.max(np.abs(rl_o - rl_g)) <= 1e-10 * max(1.0, np.max(np.abs(rl_o)))
    assert np.count_nonzero(rl_o) > 0


def test_inverse_distance_dead_cells_with_previous_pressure(capi, geo, oracle, gpu):
    w = geo.config_c2(24)
    g, changed = idw_workload(geo, w)
    assert len(changed) >= 10
    rng = np.random.default_rng(13)
    pprev = np.asfortranarray(rng.standard_normal(w.div.shape))
    q0 = np.asfortranarray(w.p_exact + 0.1 * rng.standard_normal(w.div.shape))
    for ordering, sweeps in ((oracle.ORDER_COLORED, 1), (oracle.ORDER_COLORED, 3), (oracle.ORDER_REFERENCE, 400)):
        o, s = _pair(capi, oracle, gpu, w, g, ordering, itermax_gc=sweeps,
                     restol_gcp=0.0 if sweeps < 400 else 1e-14)
        a, b = q0.copy(order="F"), q0.copy(order="F")
        o.ghostcell_pressure(a, w.div, pprev)
        s.ghostcell_pressure(b, w.div, pprev)
        o.close()
        s.close()
        err = np.max(np.abs(a - b)) / np.max(np.abs(a))
        print("ordering %d sweeps %d: max rel diff %.2e" % (ordering, sweeps, err), flush=True)
        assert err <= (1e-12 if sweeps < 400 else 1e-10)
        idx = tuple(np.array([[g["iDead"][n], g["jDead"][n], g["kDead"][n]] for n in changed]).T)
        assert np.max(np.abs(a[idx] - b[idx])) <= 1e-12 * np.max(np.abs(a[idx]))


def test_solve_poisson_hybrid_ghosts(capi, geo, oracle, gpu):
    """solve_poisson with mix_GC_form = 1 (the iterative hybrid GCM sweeps
    inside the outer loop): 20 outer iterations of 5 GCM sweeps and one
    V-cycle each on the GPU against the oracle running the same iterations in
    the GPU's (colored) orderings.  (With the closed box's right-hand side
    the hybrid rows make the system incompatible, so the loop runs to its
    iteration cap in both, as the reference's does.)"""
    w = geo.config_c1(16)
    p = capi.default_params(mix_GC_form=1, restol_pson=0.0, iterInter=0, itermax_pson=20, restol_gcp=0.0,
                            itermax_gc=5)
    o = oracle.Oracle(capi, *w.grid, p, ordering=oracle.ORDER_COLORED)
    o.set_workload(w)
    a = np.zeros_like(w.div)
    rc_o, st_o, rl_o = o.solve_poisson(a, w.div)
    o.close()
    s = gpu.PoissonSolver(*w.grid, p)
    s.set_workload(w)
    b = np.zeros_like(w.div)
    rc, st, rl = s.solve_poisson(b, w.div)
    s.close()
    assert rc == 1 and rc_o == 1 and st.outer_iters == st_o.outer_iters == 20
    f = (w.iblank == 0)
    f[0] = f[-2:] = False
    f[:, 0] = f[:, -2:] = False
    f[:, :, 0] = f[:, :, -2:] = False
    err = np.max(np.abs(b[f] - a[f])) / np.max(np.abs(a[f]))
    print("mix_GC_form=1, 20 outer iterations: max rel diff %.2e" % err, flush=True)
    assert err < 1e-10
    assert np.max(np.abs(rl - rl_o)) <= 1e-9 * max(1.0, np.max(np.abs(rl_o)))
