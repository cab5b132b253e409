"""GPU parity of whole calls through the C-ABI: mg_solve V-cycles against the
oracle in the same ordering, and solve_poisson fixed points against the
oracle's reference ordering and the manufactured pressure (tolerances:
1e-10 relative L2 after mean removal for FP64, 1e-8 for mixed precision)."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _flag_interior(w):
    f = (w.iblank == 0) & (w.ghost == 0)
    m = np.zeros_like(f)
    m[1:-2, 1:-2, 1:-2] = True
    return f & m


def _mean_free_err(a, b, mask):
    x = a[mask] - a[mask].mean()
    y = b[mask] - b[mask].mean()
    return np.linalg.norm(x - y) / np.linalg.norm(y)


@pytest.fixture(scope="module")
def c1(geo):
    return geo.config_c1(16)


@pytest.fixture(scope="module")
def c2(geo):
    return geo.config_c2(24)


def test_mg_solve_vcycles_match_oracle(capi, geo, oracle, gpu, c2):
    """Three V-cycles of mg_solve (CORE_MG.F90:1572-1848) in the GPU ordering,
    FP64, against the oracle running the same ordering."""
    p = capi.default_params(iterInter=2, restol_pson=0.0)
    o = oracle.Oracle(capi, *c2.grid, p, ordering=oracle.ORDER_COLORED)
    o.set_workload(c2)
    s = gpu.PoissonSolver(*c2.grid, p)
    s.set_workload(c2)
    rng = np.random.default_rng(7)
    v0 = np.asfortranarray(c2.p_exact + 0.1 * rng.standard_normal(c2.p_exact.shape))
    a = v0.copy(order="F")
    b = v0.copy(order="F")
    o.mg_solve(a, c2.div)
    rc, st = s.mg_solve(b, c2.div)
    assert st.vcycles_fp64 == 3
    m = np.zeros(a.shape, bool)
    m[1:-2, 1:-2, 1:-2] = True
    assert np.max(np.abs(a[m] - b[m])) <= 1e-11 * np.max(np.abs(a[m]))


def test_mg_solve_post_smoothed_cycle(capi, geo, oracle, gpu, c2):
    """The V-cycle with one finest-level post-smoothing pass (cycle = 2),
    FP64, against the oracle running the same cycle and ordering."""
    p = capi.default_params(iterInter=2, restol_pson=0.0, cycle=2)
    o = oracle.Oracle(capi, *c2.grid, p, ordering=oracle.ORDER_COLORED)
    o.set_workload(c2)
    s = gpu.PoissonSolver(*c2.grid, p)
    s.set_workload(c2)
    rng = np.random.default_rng(8)
    v0 = np.asfortranarray(c2.p_exact + 0.1 * rng.standard_normal(c2.p_exact.shape))
    a = v0.copy(order="F")
    b = v0.copy(order="F")
    o.mg_solve(a, c2.div)
    s.mg_solve(b, c2.div)
    m = np.zeros(a.shape, bool)
    m[1:-2, 1:-2, 1:-2] = True
    assert np.max(np.abs(a[m] - b[m])) <= 1e-11 * np.max(np.abs(a[m]))


@pytest.fixture(scope="module")
def c3s(geo):
    return geo.config_c3(64, 48, 48)


def test_c3_shape(capi, geo, oracle, gpu, c3s):
    """The C3 shape at 64 x 48 x 48 cells: x coarsens one level further than
    y and z, which are then held (semi-coarsening), around a thin tilted
    plate on a stretched grid.  The FP64 solve_poisson fixed point against
    the oracle's; the accelerated FP64 and mixed solves against the
    manufactured answer."""
    w = c3s
    p = capi.default_params(restol_pson=1e-12, iterInter=1, itermax_pson=800, restol_gcp=1e-11, itermax_gc=100)
    o = oracle.Oracle(capi, *w.grid, p, ordering=oracle.ORDER_REFERENCE)
    o.set_workload(w)
    assert o.level_nodes[-1] == (5, 7, 7) and o.level_nodes[-2] == (9, 7, 7)
    a = np.zeros_like(w.div)
    o.solve_poisson(a, w.div)
    s = gpu.PoissonSolver(*w.grid, p)
    s.set_workload(w)
    b = np.zeros_like(w.div)
    rc, st, _ = s.solve_poisson(b, w.div)
    s.close()
    assert rc == 0 and st.converged == 1, st.as_dict()
    f = _flag_interior(w)
    print("fp64 fixed point: outer %d, vs oracle %.2e" % (st.outer_iters, _mean_free_err(b, a, f)), flush=True)
    assert _mean_free_err(b, a, f) < 1e-10
    for prec in (0, 1):
        q = capi.default_params(restol_pson=0.0, iterInter=0, itermax_pson=400, restol_gcp=1e-11, itermax_gc=100,
                                precision=prec, outer_rule=capi.OUTER_RELRES, rel_tol=1e-11 if prec == 0 else 1e-10)
        s = gpu.PoissonSolver(*w.grid, q)
        s.set_workload(w)
        b = np.zeros_like(w.div)
        rc, st, _ = s.solve_poisson(b, w.div)
        s.close()
        err = _mean_free_err(b, w.p_exact, f)
        print("accelerated prec %d: outer %d, rel resid %.2e, vs exact %.2e" % (prec, st.outer_iters, st.rel_resid, err),
              flush=True)
        assert rc == 0 and st.rel_resid <= q.rel_tol, st.as_dict()
        assert err < (1e-9 if prec == 0 else 1e-8)


@pytest.mark.parametrize("cfg", ["c1", "c2"])
def test_solve_poisson_fp64_fixed_point(capi, geo, oracle, gpu, c1, c2, cfg):
    w = c1 if cfg == "c1" else c2
    p = capi.default_params(restol_pson=1e-12, iterInter=1, itermax_pson=800, restol_gcp=1e-11, itermax_gc=100)
    o = oracle.Oracle(capi, *w.grid, p, ordering=oracle.ORDER_REFERENCE)
    o.set_workload(w)
    a = np.zeros_like(w.div)
    rc_o, st_o, _ = o.solve_poisson(a, w.div)
    s = gpu.PoissonSolver(*w.grid, p)
    s.set_workload(w)
    b = np.zeros_like(w.div)
    rc, st, _ = s.solve_poisson(b, w.div)
    assert rc == 0 and st.converged == 1, st.as_dict()
    f = _flag_interior(w)
    assert _mean_free_err(b, a, f) < 1e-10
    assert _mean_free_err(b, w.p_exact, f) < 1e-9


@pytest.mark.parametrize("cfg", ["c1", "c2"])
def test_solve_poisson_mixed(capi, geo, oracle, gpu, c1, c2, cfg):
    w = c1 if cfg == "c1" else c2
    p = capi.default_params(restol_pson=1e-12, iterInter=1, itermax_pson=800, restol_gcp=1e-11, itermax_gc=100,
                            precision=capi.PREC_MIXED)
    s = gpu.PoissonSolver(*w.grid, p)
    s.set_workload(w)
    b = np.zeros_like(w.div)
    rc, st, _ = s.solve_poisson(b, w.div)
    assert rc == 0 and st.vcycles_fp32 > 0 and st.vcycles_fp64 == 0, st.as_dict()
    f = _flag_interior(w)
    assert _mean_free_err(b, w.p_exact, f) < 1e-8


@pytest.mark.parametrize("walls", [0, 1])
@pytest.mark.parametrize("prec", [0, 1])
def test_solve_poisson_accelerated(capi, geo, oracle, gpu, c2, prec, walls):
    """Anderson-accelerated outer loop: same fixed point (the oracle's, in the
    reference's ordering), far fewer outer iterations than the plain loop;
    with the wall correction cycles (walls = 1; FP64 then runs as defect
    correction too) fewer still."""
    w = c2
    p = capi.default_params(restol_pson=0.0, iterInter=0, itermax_pson=400, restol_gcp=1e-11, itermax_gc=100,
                            precision=prec, outer_rule=capi.OUTER_RELRES, rel_tol=1e-11 if prec == 0 else 1e-10,
                            vc_walls=walls)
    s = gpu.PoissonSolver(*w.grid, p)
    s.set_workload(w)
    b = np.zeros_like(w.div)
    rc, st, _ = s.solve_poisson(b, w.div)
    s.close()
    print("prec %d walls %d: outer %d, rel %.2e" % (prec, walls, st.outer_iters, st.rel_resid), flush=True)
    assert rc == 0 and st.rel_resid <= p.rel_tol, st.as_dict()
    assert st.outer_iters < 150
    f = _flag_interior(w)
    assert _mean_free_err(b, w.p_exact, f) < (1e-9 if prec == 0 else 1e-8)
    q = capi.default_params(restol_pson=1e-13, iterInter=1, itermax_pson=2000, restol_gcp=1e-12, itermax_gc=100)
    o = oracle.Oracle(capi, *w.grid, q, ordering=oracle.ORDER_REFERENCE)
    o.set_workload(w)
    a = np.zeros_like(w.div)
    o.solve_poisson(a, w.div)
    o.close()
    assert _mean_free_err(b, a, f) < (1e-9 if prec == 0 else 1e-8)


def test_rhs_poisson(capi, geo, oracle, gpu, c1):
    """rhs_poisson (CORE_POISSON.F90:35-50) on the GPU against the oracle."""
    w = c1
    nx, ny, nz = (a.size - 2 for a in w.grid)
    fu, fv, fw = geo.face_velocities(nx, ny, nz, seed=11)
    p = capi.default_params()
    o = oracle.Oracle(capi, *w.grid, p)
    o.set_workload(w)
    rc, d_o, s_o, r_o = o.rhs_poisson(fu, fv, fw, 2.0)
    s = gpu.PoissonSolver(*w.grid, p)
    s.set_workload(w)
    d_g, s_g, r_g = s.rhs_poisson(fu, fv, fw, 2.0)
    assert np.max(np.abs(d_g - d_o)) <= 1e-13 * np.max(np.abs(d_o))
    assert abs(s_g - s_o) <= 1e-10 * max(1.0, abs(s_o)) and abs(r_g - r_o) <= 1e-12 * r_o


def test_fortran_driver(capi, geo, oracle, gpu, c1, tmp_path):
    """The Fortran host program drives solve_poisson through MODULE
    fc_poisson (ISO_C_BINDING) and gets the oracle's answer."""
    import feicar3d_amd.build as bld
    drv = bld.DRIVER
    assert os.path.exists(drv), "fortran driver not built"
    w = c1
    p = capi.default_params(restol_pson=1e-12, iterInter=1, itermax_pson=800, restol_gcp=1e-11, itermax_gc=100)
    case = tmp_path / "case.bin"
    out = tmp_path / "out.bin"
    import feicar3d_amd.fortran_io as fio
    fio.write_case(str(case), w, p)
    print("running", drv, flush=True)
    r = subprocess.run([drv, str(case), str(out)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=150)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout
    pp, st = fio.read_result(str(out), w.div.shape)
    assert st.converged == 1
    o = oracle.Oracle(capi, *w.grid, p)
    o.set_workload(w)
    a = np.zeros_like(w.div)
    o.solve_poisson(a, w.div)
    assert _mean_free_err(pp, a, _flag_interior(w)) < 1e-10


@pytest.mark.parametrize("walls", [0, 1])
@pytest.mark.parametrize("prec", [0, 1])
def test_graph_replays_match_eager(capi, geo, gpu, c2_48, prec, walls):
    """The accelerated loop launched as one HIP graph per outer iteration
    (replayed once per iteration, with host synchronisation between the
    replays) gives bit-identical pressures and iteration counts to the same
    kernels launched from the host, over alternating solves on one handle.
    (48^3: its GCM relation graph is acyclic, so the exact pass and the graph
    are used.)"""
    w = c2_48
    p = capi.default_params(restol_pson=0.0, iterInter=0, itermax_pson=400, restol_gcp=1e-11, itermax_gc=100,
                            precision=prec, outer_rule=capi.OUTER_RELRES, rel_tol=1e-10, vc_walls=walls)
    s = gpu.PoissonSolver(*w.grid, p)
    s.set_workload(w)
    out = []
    for mode in (1, 0, 1, 0):
        p.launch_mode = mode
        s.set_params(p)
        b = np.zeros_like(w.div)
        rc, st, _ = s.solve_poisson(b, w.div)
        assert rc == 0, st.as_dict()
        out.append((b, st.outer_iters))
    eager, iters = out[0]
    for b, it in out[1:]:
        assert it == iters
        assert np.array_equal(b, eager)


@pytest.fixture(scope="module")
def c2_48(geo):
    return geo.config_c2(48)


@pytest.fixture(scope="module")
def c2_48_oracle(capi, oracle, c2_48):
    """the oracle's converged solve_poisson on the 48^3 workload (reference
    ordering, tight tolerances)"""
    w = c2_48
    q = capi.default_params(restol_pson=1e-13, iterInter=1, itermax_pson=3000, restol_gcp=1e-12, itermax_gc=100)
    o = oracle.Oracle(capi, *w.grid, q, ordering=oracle.ORDER_REFERENCE)
    o.set_workload(w)
    a = np.zeros_like(w.div)
    o.solve_poisson(a, w.div)
    o.close()
    return a


@pytest.mark.parametrize("prec", [0, 1])
def test_wall_cycles_defect_correction(capi, geo, oracle, gpu, c2_48, c2_48_oracle, prec):
    """The wall correction cycles (vc_walls = 1) run where the outer loop is a
    defect correction with the exact GCM pass (an acyclic relation graph, as
    at 48^3 and 256^3; the 24^3 grid's graph has cycles and runs the
    iterative GCM sweeps without them): one FP64 residual and one V-cycle
    per outer iteration in both precisions, the oracle's fixed point, and
    fewer outer iterations than with held masked cells."""
    w = c2_48
    f = _flag_interior(w)
    a = c2_48_oracle
    iters = {}
    for walls in (0, 1):
        p = capi.default_params(restol_pson=0.0, iterInter=0, itermax_pson=400, restol_gcp=1e-11, itermax_gc=100,
                                precision=prec, outer_rule=capi.OUTER_RELRES, rel_tol=1e-11, vc_walls=walls)
        s = gpu.PoissonSolver(*w.grid, p)
        s.set_workload(w)
        b = np.zeros_like(w.div)
        rc, st, _ = s.solve_poisson(b, w.div)
        s.close()
        err_o, err_x = _mean_free_err(b, a, f), _mean_free_err(b, w.p_exact, f)
        print("prec %d walls %d: outer %d, rel %.2e, vs oracle %.2e, vs exact %.2e"
              % (prec, walls, st.outer_iters, st.rel_resid, err_o, err_x), flush=True)
        assert rc == 0 and st.converged == 1 and st.rel_resid <= p.rel_tol, st.as_dict()
        assert st.gcm_sweeps == st.outer_iters          # the exact pass, once per iteration
        if walls:
            assert st.refine_steps == st.outer_iters
            assert (st.vcycles_fp64, st.vcycles_fp32) == ((st.outer_iters, 0) if prec == 0 else (0, st.outer_iters))
        assert err_o < 1e-9 and err_x < 1e-9
        iters[walls] = st.outer_iters
    assert iters[1] < iters[0]


@pytest.mark.parametrize("prec", [0, 1])
def test_coarse_direct_in_wall_cycles(capi, geo, oracle, gpu, c2_48, c2_48_oracle, prec):
    """coarse_solve = 1 (FP64 direct solve of the coarsest level of the wall
    correction cycles, automatic level count): the oracle's fixed point, in
    no more outer iterations than the single line-SOR pass there."""
    w = c2_48
    f = _flag_interior(w)
    iters = {}
    for cs in (0, 1):
        p = capi.default_params(restol_pson=0.0, iterInter=0, itermax_pson=400, restol_gcp=1e-11, itermax_gc=100,
                                precision=prec, outer_rule=capi.OUTER_RELRES, rel_tol=1e-11, vc_walls=1,
                                coarse_solve=cs)
        s = gpu.PoissonSolver(*w.grid, p)
        s.set_workload(w)
        b = np.zeros_like(w.div)
        rc, st, _ = s.solve_poisson(b, w.div)
        s.close()
        err_o = _mean_free_err(b, c2_48_oracle, f)
        print("prec %d coarse_solve %d: outer %d, rel %.2e, vs oracle %.2e" % (prec, cs, st.outer_iters,
                                                                            st.rel_resid, err_o), flush=True)
        assert rc == 0 and st.converged == 1 and st.rel_resid <= p.rel_tol, st.as_dict()
        assert err_o < 1e-9
        iters[cs] = st.outer_iters
    assert iters[1] <= iters[0]


@pytest.mark.parametrize("prec", [0, 1])
def test_accelerated_dirichlet_face(capi, geo, oracle, gpu, c2, prec):
    """The accelerated loop with a Dirichlet pressure face (x = 1, p = 0):
    the solution is unique, nothing is pinned, and the loop reaches the
    reference loop's fixed point (the oracle, no mean removed).  Pinning
    p(1,1,1) there would leave a fixed point of the iteration that is not a
    solution."""
    w = c2
    pbc = [capi.PBC_NEUMANN, capi.PBC_DIRICHLET] + [capi.PBC_NEUMANN] * 4
    p = capi.default_params(restol_pson=0.0, iterInter=0, itermax_pson=400, restol_gcp=1e-12, itermax_gc=100,
                            precision=prec, outer_rule=capi.OUTER_RELRES, rel_tol=1e-11, pbc=pbc)
    s = gpu.PoissonSolver(*w.grid, p)
    s.set_workload(w)
    b = np.zeros_like(w.div)
    rc, st, _ = s.solve_poisson(b, w.div)
    s.close()
    print("prec %d: outer %d, rel %.2e" % (prec, st.outer_iters, st.rel_resid), flush=True)
    assert rc == 0 and st.rel_resid <= p.rel_tol, st.as_dict()
    q = capi.default_params(restol_pson=1e-14, iterInter=1, itermax_pson=4000, restol_gcp=1e-13, itermax_gc=100,
                            pbc=pbc)
    o = oracle.Oracle(capi, *w.grid, q, ordering=oracle.ORDER_REFERENCE)
    o.set_workload(w)
    a = np.zeros_like(w.div)
    rc_o, _, _ = o.solve_poisson(a, w.div)
    o.close()
    assert rc_o == 0
    f = _flag_interior(w)
    assert np.max(np.abs(b - a)[f]) <= 1e-8 * np.max(np.abs(a[f]))
