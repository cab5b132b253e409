"""itsolv (CORE_POISSON.F90:205-456), the plain line-SOR solver of
it_solver_type = 1, on the GPU: one fc_pson_itsolv call against the oracle's
itsolv in the GPU's (colored) line order, and solve_poisson with
it_solver_type = 1 against the oracle's fixed point in the reference's
order.  The oracle's itsolv is bit-exact with the reference's own compiled
itsolv in the reference order (test_reference_mg.py)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2(geo):
    return geo.config_c2(24)


@pytest.fixture(scope="module")
def c1(geo):
    return geo.config_c1(16)


def _flag_interior(w):
    f = (w.iblank == 0) & (w.ghost == 0)
    f[0] = f[-2:] = False
    f[:, 0] = f[:, -2:] = False
    f[:, :, 0] = f[:, :, -2:] = False
    return f


def _mean_free_err(a, b, m):
    x = a[m] - a[m].mean()
    y = b[m] - b[m].mean()
    return float(np.linalg.norm(x - y) / np.linalg.norm(y))


@pytest.mark.parametrize("omega", [1.0, 1.3])
def test_itsolv_call_matches_oracle(capi, oracle, gpu, c2, omega):
    w = c2
    p = capi.default_params(it_solver_type=capi.IT_SOLVER_TYPE_LSOR, omega_pson=omega)
    rng = np.random.default_rng(7)
    var = np.asfortranarray(rng.standard_normal(w.div.shape))
    r = np.asfortranarray(rng.standard_normal(w.div.shape))
    s = gpu.PoissonSolver(*w.grid, p)
    s.set_workload(w)
    o = oracle.Oracle(capi, *w.grid, p, ordering=oracle.ORDER_COLORED)
    o.set_workload(w)
    g, a = var.copy(order="F"), var.copy(order="F")
    for _ in range(3):
        s.itsolv(g, r)
        o.itsolv(a, r)
    s.close()
    o.close()
    n = w.div.shape
    inner = (slice(1, n[0] - 2), slice(1, n[1] - 2), slice(1, n[2] - 2))
    err = np.max(np.abs(g[inner] - a[inner])) / np.max(np.abs(a[inner]))
    assert err < 1e-12, err
    # itsolv writes interior rows only: the ghost layers keep their values
    ghost = np.ones(n, bool)
    ghost[inner] = False
    assert np.array_equal(g[ghost], var[ghost])
    # masked cells (iblank or ghostCellMark) are held
    held = ((w.iblank == 1) | (w.ghost == 1))
    assert np.array_equal(g[held], var[held])


def test_solve_poisson_lsor_fixed_point(capi, oracle, gpu, c1):
    w = c1
    p = capi.default_params(it_solver_type=capi.IT_SOLVER_TYPE_LSOR, omega_pson=1.5, restol_pson=0.0,
                            iterInter=0, itermax_pson=20000, restol_gcp=1e-13, itermax_gc=100)
    o = oracle.Oracle(capi, *w.grid, p, ordering=oracle.ORDER_REFERENCE)
    o.set_workload(w)
    a = np.zeros_like(w.div)
    rc_o, st_o, _ = o.solve_poisson(a, w.div)
    o.close()
    s = gpu.PoissonSolver(*w.grid, p)
    s.set_workload(w)
    b = np.zeros_like(w.div)
    rc, st, _ = s.solve_poisson(b, w.div)
    s.close()
    f = _flag_interior(w)
    print("LSOR: GPU %d outer iterations, oracle %d; vs oracle %.2e, vs exact %.2e"
          % (st.outer_iters, st_o.outer_iters, _mean_free_err(b, a, f), _mean_free_err(b, w.p_exact, f)),
          flush=True)
    assert rc == 0 and st.converged == 1 and rc_o == 0
    assert st.vcycles_fp64 == 0 and st.vcycles_fp32 == 0
    assert _mean_free_err(b, a, f) < 1e-10
    assert _mean_free_err(b, w.p_exact, f) < 1e-10


def test_lsor_mixed_precision_is_refused(capi, gpu, c1):
    w = c1
    p = capi.default_params(it_solver_type=capi.IT_SOLVER_TYPE_LSOR, precision=capi.PREC_MIXED)
    s = gpu.PoissonSolver(*w.grid, p)
    try:
        s.set_workload(w)
        b = np.zeros_like(w.div)
        with pytest.raises(gpu.FcError, match="FP64 only"):
            s.solve_poisson(b, w.div)
    finally:
        s.close()
