"""GPU parity of solve_ad's loop (CORE_SOLVE_AD.F90:11-127) and the routines
it calls besides itsolv_ad, through the C-ABI against the oracle (pinned in
test_solve_ad_oracle.py):

* set_outer_velocity_bc / set_outer_ghost_velocity: every face type the
  build carries, the nine outer-face arrays and the ghost layers, to 1e-15
  relative (sin of the pulsatile faces may differ in the last bit);
* GCM_ghostcell_velocity: the GPU's Jacobi steps and the oracle's list-order
  sweeps converge to the same ghost and dead velocities (1e-11 relative),
  with mix_GC_form 0 and 1;
* solve_ad: the GPU loop (colored line order, Jacobi GCM steps) and the
  oracle's loop in the same line order reach the same velocities and face
  velocities (1e-9 relative) with Dirichlet, symmetry and zero-gradient
  faces, the latter through enforce_global_mass_consv."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

VALS = np.arange(1, 19, dtype=float) * 0.1
FREQ = np.linspace(0.3, 2.0, 18)


def _box(geo, nx, ny, nz):
    w = geo.Workload()
    w.grid = (geo.uniform_axis(nx + 1), geo.tanh_axis(ny + 1, [0.5], 1.0 / 64.0, width=0.3), geo.uniform_axis(nz + 1))
    sh = tuple(a.size for a in w.grid)
    w.iblank = np.zeros(sh, np.int8, order="F")
    w.ghost = np.zeros(sh, np.int8, order="F")
    w.dead = np.zeros(sh, np.int8, order="F")
    w.gcm = geo.empty_gcm()
    return w


def _rand(sh, seed, scale=1.0):
    return np.asfortranarray(scale * np.random.default_rng(seed).standard_normal(sh))


@pytest.mark.parametrize("types", [(1, 6, 2, 4, 3, 1), (3, 1, 5, 5, 2, 2), (5, 5, 1, 2, 4, 4)])
def test_outer_velocity_bc(capi, geo, oracle, gpu, types):
    w = _box(geo, 20, 24, 12)
    p = capi.default_params(bc_vel=list(types))
    sh = w.iblank.shape
    U = [_rand(sh, s) for s in (1, 2, 3)]
    t = 0.37
    o = oracle.Oracle(capi, *w.grid, p)
    o.set_workload(w)
    o.set_outer_velocity(VALS, FREQ, 0.1, 0.9)
    Uo = [a.copy(order="F") for a in U]
    assert o.set_outer_velocity_bc(*Uo, t) == 0
    o.set_outer_ghost_velocity(*Uo)
    bo = o.get_outer_velocity_bc()
    o.close()
    s = gpu.PoissonSolver(*w.grid, p)
    s.set_workload(w)
    s.vel_upload(*U)
    s.vel_set_outer(VALS, FREQ, 0.1, 0.9)
    s.outer_velocity_bc(t)
    s.outer_ghost_velocity()
    bg = s.vel_get_outer_all()
    Ug = s.vel_download()[:3]
    s.close()
    for a, b in zip(bg, bo):
        assert np.max(np.abs(a - b)) <= 1e-15 * max(1.0, np.max(np.abs(b)))
    for a, b in zip(Ug, Uo):
        assert np.max(np.abs(a - b)) <= 1e-15 * max(1.0, np.max(np.abs(b)))


def test_outer_velocity_refuses_shear(capi, geo, gpu):
    w = _box(geo, 8, 8, 8)
    s = gpu.PoissonSolver(*w.grid, capi.default_params(bc_vel=[7, 1, 1, 1, 1, 1]))
    s.set_workload(w)
    s.vel_upload(*[_rand(w.iblank.shape, k) for k in (1, 2, 3)])
    with pytest.raises(gpu.FcError, match="shear"):
        s.outer_velocity_bc(0.0)
    s.close()


def _gcm_vel_inputs(w, seed):
    g = w.gcm
    rng = np.random.default_rng(seed)
    return (g["coeffGCMD"], g["coeffGCMDeadD"], rng.standard_normal((g["iGhost"].size, 3)) * 0.2,
            rng.standard_normal((g["iDead"].size, 3)) * 0.2, g["probeLengthNormalizedD"])


@pytest.mark.parametrize("mix", [0, 1])
def test_ghostcell_velocity(capi, geo, oracle, gpu, mix):
    w = geo.config_c2(24)
    sh = w.iblank.shape
    p = capi.default_params(mix_GC_form=mix, itermax_gc=2000)
    gv = _gcm_vel_inputs(w, 3)
    U = [_rand(sh, s) for s in (4, 5, 6)]
    NL = [_rand(sh, s, 0.1) for s in (7, 8, 9)]
    FU = [_rand(sh, s, 0.5) for s in (10, 11, 12)]
    o = oracle.Oracle(capi, *w.grid, p, ordering=oracle.ORDER_COLORED)
    o.set_workload(w)
    o.set_solve_ad(1.0 / 100.0, 0.01, 1, 1.0)
    o.set_gcm_vel(*gv)
    Uo = [a.copy(order="F") for a in U]
    rc, sw, rm = o.ghostcell_velocity(*Uo, *NL, *FU, restol_gcv=1e-14)
    o.close()
    assert rc == 0
    s = gpu.PoissonSolver(*w.grid, p)
    s.set_workload(w)
    s.set_solve_ad(1.0 / 100.0, 0.01, 1, 1.0)
    s.vel_upload(*U, *FU)
    s.ad_upload(*NL)
    s.set_gcm_vel(*gv)
    rc, sw_g, rm_g = s.ghostcell_velocity(2, 1e-14)
    Ug = s.vel_download()[:3]
    s.close()
    print("mix %d: oracle %d sweeps, GPU %d steps" % (mix, sw, sw_g), flush=True)
    assert rc == 0
    for a, b in zip(Ug, Uo):
        assert np.max(np.abs(a - b)) <= 1e-11 * np.max(np.abs(b))


@pytest.mark.parametrize("types,mix", [((1, 2, 4, 4, 1, 1), 0), ((1, 2, 4, 4, 1, 1), 1), ((1, 1, 1, 1, 1, 1), 0)])
def test_solve_ad(capi, geo, oracle, gpu, types, mix):
    w = geo.config_c1(16)
    g = w.gcm
    sh = w.iblank.shape
    p = capi.default_params(bc_vel=list(types), mix_GC_form=mix, itermax_gc=500)
    vals = np.r_[1.0, 0.0, np.zeros(16)]
    gv = (g["coeffGCMD"], g["coeffGCMDeadD"], np.zeros((g["iGhost"].size, 3)), np.zeros((g["iDead"].size, 3)),
          g["probeLengthNormalizedD"])
    rng = np.random.default_rng(7)
    U = [np.asfortranarray(0.1 * rng.standard_normal(sh)) for _ in range(3)]
    NL = [np.asfortranarray(0.05 * rng.standard_normal(sh)) for _ in range(3)]
    FU = [np.asfortranarray(np.zeros(sh)) for _ in range(3)]
    FU[0][...] = 1.0
    o = oracle.Oracle(capi, *w.grid, p, ordering=oracle.ORDER_COLORED)
    o.set_workload(w)
    o.set_solve_ad(1.0 / 100.0, 0.01, 1, 1.0)
    o.set_outer_velocity(vals)
    o.set_gcm_vel(*gv)
    Uo = [a.copy(order="F") for a in U]
    Fo = [a.copy(order="F") for a in FU]
    rc, it_o, res_o, _ = o.solve_ad(*Uo, *NL, *Fo, 0.0, 400, 1e-12, 1e-14)
    o.close()
    assert rc == 0, (it_o, res_o)
    s = gpu.PoissonSolver(*w.grid, p)
    s.set_workload(w)
    s.set_solve_ad(1.0 / 100.0, 0.01, 1, 1.0)
    s.vel_set_outer(vals)
    s.set_gcm_vel(*gv)
    Ug = [a.copy(order="F") for a in U]
    Fg = [a.copy(order="F") for a in FU]
    rc, it_g, res_g, loc = s.solve_ad(*Ug, *NL, *Fg, 0.0, 400, 1e-12, 1e-14)
    s.close()
    print("types %s mix %d: oracle %d passes, GPU %d (res %.2e)" % (types, mix, it_o, it_g, res_g), flush=True)
    assert rc == 0 and res_g <= 1e-12
    for a, b in zip(Ug + Fg, Uo + Fo):
        assert np.max(np.abs(a - b)) <= 1e-9 * max(1.0, np.max(np.abs(b)))
