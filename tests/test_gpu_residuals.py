"""calc_residual and calc_residual2 (CORE_POISSON.F90:460-551, 679-724) on
the GPU through the C-ABI, with their loc output: the norms against the
oracle's, the locations against a numpy search of the first maximum in the
reference's k-j-i loop order; and the Fortran host program that calls the
path under the reference's own names (MODULE fc_poisson_ref)."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2(geo):
    return geo.config_c2(24)


def _first_max_loc(v):
    """1-based (i, j, k) of the first max |v| over the interior cells in the
    reference's loop order (k outer, i inner)"""
    a = np.abs(v[1:-2, 1:-2, 1:-2])
    t = int(np.argmax(a.ravel(order="F")))
    i, j, k = np.unravel_index(t, a.shape, order="F")
    return (i + 1, j + 1, k + 1), float(a.ravel(order="F")[t])


def _flag_residual(geo, w, var, r):
    res = np.zeros_like(var)
    x, y, z = w.grid
    nx, ny, nz = x.size - 2, y.size - 2, z.size - 2
    res[1:nx, 1:ny, 1:nz] = r[1:nx, 1:ny, 1:nz] - geo.apply_operator(var, x, y, z)
    res[(w.iblank == 1) | (w.ghost == 1)] = 0.0
    return res


def test_calc_residual_with_loc(capi, geo, oracle, gpu, c2):
    w = c2
    rng = np.random.default_rng(31)
    var = np.asfortranarray(rng.standard_normal(w.div.shape))
    r = np.asfortranarray(rng.standard_normal(w.div.shape))
    p = capi.default_params()
    s = gpu.PoissonSolver(*w.grid, p)
    s.set_workload(w)
    o = oracle.Oracle(capi, *w.grid, p)
    o.set_workload(w)
    resm, loc = s.calc_residual(var, r, with_loc=True)
    ref = o.calc_residual(var, r)
    assert abs(resm - ref) <= 1e-14 * ref
    res = _flag_residual(geo, w, var, r)
    l_np, m_np = _first_max_loc(res)
    assert loc == l_np
    # the norm also from numpy: RMS over flag cells, node-count normalised
    nodes = np.prod([a.size - 2 for a in w.grid])
    assert abs(resm - np.sqrt(np.sum(res ** 2) / nodes)) <= 1e-12 * resm
    # calc_residual2: max |var - p_temp| over iblank = 0
    pt = np.asfortranarray(var + 1e-3 * rng.standard_normal(w.div.shape))
    r2, loc2 = s.calc_residual2(var, pt, with_loc=True)
    assert r2 == o.calc_residual2(var, pt)
    d = np.where(w.iblank == 1, 0.0, var - pt)
    l2_np, m2_np = _first_max_loc(d)
    assert loc2 == l2_np and r2 == m2_np
    # every value zero: resm 0 and loc (0, 0, 0)
    r0, l0 = s.calc_residual2(var, var, with_loc=True)
    assert r0 == 0.0 and l0 == (0, 0, 0)
    s.close()
    o.close()


def test_fortran_reference_names_driver(capi, geo, oracle, gpu, tmp_path):
    """solve_poisson(), calc_residual, calc_residual2 and itsolv called from
    Fortran under the reference's names, against the oracle."""
    import feicar3d_amd.build as bld
    import feicar3d_amd.fortran_io as fio
    assert os.path.exists(bld.REF_DRIVER), "fc_ref_driver not built"
    w = geo.config_c1(16)
    p = capi.default_params(restol_pson=1e-12, iterInter=1, itermax_pson=800, restol_gcp=1e-11, itermax_gc=100)
    case, out = tmp_path / "case.bin", tmp_path / "out.bin"
    fio.write_case(str(case), w, p)
    r = subprocess.run([bld.REF_DRIVER, str(case), str(out)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=150)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout
    raw = open(out, "rb").read()
    n = w.div.size
    pp, st = fio.read_result(str(out), w.div.shape)
    off = 8 * n + fio.C_sizeof_stats()
    resm = np.frombuffer(raw, "<f8", 1, off)[0]
    loc = tuple(np.frombuffer(raw, "<i4", 3, off + 8))
    resm2 = np.frombuffer(raw, "<f8", 1, off + 20)[0]
    loc2 = tuple(np.frombuffer(raw, "<i4", 3, off + 28))
    var = np.frombuffer(raw, "<f8", n, off + 40).reshape(w.div.shape, order="F")
    assert st.converged == 1
    o = oracle.Oracle(capi, *w.grid, p)
    o.set_workload(w)
    a = np.zeros_like(w.div)
    o.solve_poisson(a, w.div)
    f = (w.iblank == 0) & (w.ghost == 0)
    f[0] = f[-2:] = False
    f[:, 0] = f[:, -2:] = False
    f[:, :, 0] = f[:, :, -2:] = False
    x1, y1 = pp[f] - pp[f].mean(), a[f] - a[f].mean()
    assert np.linalg.norm(x1 - y1) / np.linalg.norm(y1) < 1e-10
    # at the converged pressure the residual is ~1e-10 of its terms, so the
    # summation order shows at 1e-6: the norm to that, and the location is
    # one where |res| is the maximum up to that rounding
    assert abs(resm - o.calc_residual(pp, w.div)) <= 1e-5 * resm
    res = _flag_residual(geo, w, pp, w.div)
    l_np, m_np = _first_max_loc(res)
    assert abs(abs(res[loc]) - m_np) <= 1e-5 * m_np, (loc, l_np)
    d = np.where(w.iblank == 1, 0.0, pp)              # p_temp = p = 0
    assert loc2 == _first_max_loc(d)[0] and resm2 == _first_max_loc(d)[1]
    o.close()
    oc = oracle.Oracle(capi, *w.grid, p, ordering=oracle.ORDER_COLORED)
    oc.set_workload(w)
    v = np.zeros_like(w.div)
    oc.itsolv(v, w.div)
    oc.close()
    assert np.max(np.abs(var - v)) <= 1e-12 * np.max(np.abs(v))
