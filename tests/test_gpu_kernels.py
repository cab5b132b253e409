"""GPU kernel parity: each HIP level operation against the oracle running the
same operation in the same (colored) ordering, on seeded inputs."""
import numpy as np
import pytest

from conftest import interior, no_edges

pytestmark = pytest.mark.gpu


def _pair(capi, geo, oracle, gpu, w, **kw):
    p = capi.default_params(**kw)
    o = oracle.Oracle(capi, *w.grid, p, ordering=oracle.ORDER_COLORED)
    o.set_workload(w)
    s = gpu.PoissonSolver(*w.grid, p)
    s.set_workload(w)
    return o, s


@pytest.fixture(scope="module")
def small(geo):
    return geo.config_c2(24)


@pytest.fixture(scope="module")
def c1(geo):
    return geo.config_c1(16)


def _rand(shape, seed):
    r = np.random.default_rng(seed)
    return np.asfortranarray(r.standard_normal(shape))


def test_tdma_batched(capi, oracle, gpu):
    rng = np.random.default_rng(3)
    for n in (6, 17, 64, 130, 258, 515):
        ns = 37
        a = rng.uniform(0.2, 1.0, (ns, n))
        c = rng.uniform(0.2, 1.0, (ns, n))
        b = -(a + c) - rng.uniform(0.1, 2.0, (ns, n))
        r = rng.standard_normal((ns, n))
        a[:, 0] = 0
        c[:, -1] = 0
        u = gpu.tdma_batched(a, b, c, r)
        ref = np.stack([oracle.tdma(capi, a[s], b[s], c[s], r[s], 0, n - 1) for s in range(ns)])
        assert np.max(np.abs(u - ref)) / np.max(np.abs(ref)) < 1e-13, n
        u32 = gpu.tdma_batched(a, b, c, r, prec=capi.PREC_MIXED)
        assert np.max(np.abs(u32 - ref)) / np.max(np.abs(ref)) < 1e-5, n


def test_masks_match(capi, geo, oracle, gpu, small):
    o, s = _pair(capi, geo, oracle, gpu, small)
    assert s.nlev == o.nlev and s.level_nodes == o.level_nodes
    for lev in range(1, o.nlev + 1):
        for which in (3, 4):
            g = s.lvl_read(lev, which)
            r = o.level(lev, which)
            if lev == 1:
                g, r = g[:-1], r[:-1]      # index nx+1 is not set at level 1
            assert np.array_equal(g, r), (lev, which)


@pytest.mark.parametrize("lev", [1, 2])
def test_bc_fill(capi, geo, oracle, gpu, small, lev):
    o, s = _pair(capi, geo, oracle, gpu, small)
    phi = _rand(s.level_shape(lev), 10 + lev)
    s.lvl_write(lev, 0, phi)
    o.level(lev, 0)[:] = phi
    s.bc_fill(lev)
    o.bc_fill(lev)
    g, r = s.lvl_read(lev, 0), o.level(lev, 0)
    m = no_edges(g)
    assert np.max(np.abs(g[m] - r[m])) < 1e-15


@pytest.mark.parametrize("lev", [1, 2, 3])
def test_residual(capi, geo, oracle, gpu, small, lev):
    o, s = _pair(capi, geo, oracle, gpu, small)
    phi = _rand(s.level_shape(lev), 20 + lev)
    rhs = _rand(s.level_shape(lev), 30 + lev)
    for which, a in ((0, phi), (1, rhs)):
        s.lvl_write(lev, which, a)
        o.level(lev, which)[:] = a
    s.lvl_write(lev, 2, np.zeros(s.level_shape(lev)))
    o.level(lev, 2)[:] = 0
    s.residual_mg(lev)
    o.residual_mg(lev)
    g, r = s.lvl_read(lev, 2), o.level(lev, 2)
    assert np.max(np.abs(g - r)) <= 1e-12 * np.max(np.abs(r))


@pytest.mark.parametrize("lev", [1, 2])
def test_restrict(capi, geo, oracle, gpu, small, lev):
    o, s = _pair(capi, geo, oracle, gpu, small)
    res = _rand(s.level_shape(lev), 40 + lev)
    s.lvl_write(lev, 2, res)
    o.level(lev, 2)[:] = res
    s.mg_inject(lev)
    o.mg_inject(lev)
    g, r = s.lvl_read(lev + 1, 1), o.level(lev + 1, 1)
    assert np.max(np.abs(g - r)) <= 1e-13 * np.max(np.abs(r))


@pytest.mark.parametrize("lev", [2, 3])
def test_prolong(capi, geo, oracle, gpu, small, lev):
    o, s = _pair(capi, geo, oracle, gpu, small)
    pc = _rand(s.level_shape(lev), 50 + lev)
    pf = _rand(s.level_shape(lev - 1), 60 + lev)
    s.lvl_write(lev, 0, pc)
    o.level(lev, 0)[:] = pc
    s.lvl_write(lev - 1, 0, pf)
    o.level(lev - 1, 0)[:] = pf
    s.mg_interp(lev)
    o.mg_interp(lev)
    g, r = s.lvl_read(lev - 1, 0), o.level(lev - 1, 0)
    assert np.max(np.abs(g - r)) <= 1e-13 * np.max(np.abs(r))


@pytest.mark.parametrize("lev", [1, 2, 3])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_sweep(capi, geo, oracle, gpu, small, lev, d):
    o, s = _pair(capi, geo, oracle, gpu, small, omega_pson=1.2)
    phi = _rand(s.level_shape(lev), 70 + lev)
    rhs = _rand(s.level_shape(lev), 80 + lev)
    for which, a in ((0, phi), (1, rhs)):
        s.lvl_write(lev, which, a)
        o.level(lev, which)[:] = a
    for color in (0, 1):
        s.sweep(lev, d, color)
        o.sweep(lev, d, color)
        g, r = s.lvl_read(lev, 0), o.level(lev, 0)
        m = no_edges(g)
        err = np.max(np.abs(g[m] - r[m])) / np.max(np.abs(r[m]))
        assert err < 1e-12, (lev, d, color, err)


@pytest.fixture(scope="module")
def c2_96(geo):
    return geo.config_c2(96)


@pytest.mark.parametrize("prec", [0, 1])
def test_sweep_large_levels(capi, geo, oracle, gpu, c2_96, prec):
    """Half sweeps on a 96^3 stretched, masked grid: the finest level has
    4608 lines per colour (16 lines per block) and level 2 has 1152 (4 lines
    per block); both colours of all three directions against the oracle, in
    FP64 and in FP32."""
    o, s = _pair(capi, geo, oracle, gpu, c2_96, omega_pson=1.2)
    for lev in (1, 2):
        phi = _rand(s.level_shape(lev), 100 + lev)
        rhs = _rand(s.level_shape(lev), 110 + lev)
        if prec == 0:
            s.lvl_write(lev, 0, phi)
            s.lvl_write(lev, 1, rhs)
            o.level(lev, 0)[:] = phi
            o.level(lev, 1)[:] = rhs
        else:
            s.lvl_write(lev, 10, phi)
            s.lvl_write(lev, 11, rhs)
            o.level(lev, 0)[:] = phi.astype(np.float32).astype(np.float64)
            o.level(lev, 1)[:] = rhs.astype(np.float32).astype(np.float64)
        for d in (0, 1, 2):
            for color in (0, 1):
                s.sweep(lev, d, color, prec=prec)
                o.sweep(lev, d, color)
            g = s.lvl_read(lev, 0 if prec == 0 else 10).astype(np.float64)
            r = o.level(lev, 0)
            m = no_edges(g)
            err = np.max(np.abs(g[m] - r[m])) / np.max(np.abs(r[m]))
            print("level %d direction %d prec %d: rel err %.2e" % (lev, d, prec, err), flush=True)
            assert err < (1e-12 if prec == 0 else 2e-5), (lev, d, err)


def test_sweep_fp32(capi, geo, oracle, gpu, small):
    o, s = _pair(capi, geo, oracle, gpu, small)
    lev = 1
    phi = _rand(s.level_shape(lev), 90)
    rhs = _rand(s.level_shape(lev), 91)
    s.lvl_write(lev, 10, phi)
    s.lvl_write(lev, 11, rhs)
    o.level(lev, 0)[:] = phi.astype(np.float32).astype(np.float64)
    o.level(lev, 1)[:] = rhs.astype(np.float32).astype(np.float64)
    s.itsolv_mg(lev, prec=capi.PREC_MIXED)
    o.itsolv_mg(lev)
    g, r = s.lvl_read(lev, 10).astype(np.float64), o.level(lev, 0)
    m = no_edges(g)
    assert np.max(np.abs(g[m] - r[m])) / np.max(np.abs(r[m])) < 2e-5


def test_gcm_update(capi, geo, oracle, gpu, small):
    o, s = _pair(capi, geo, oracle, gpu, small, itermax_gc=200, restol_gcp=1e-14)
    p0 = np.asfortranarray(small.p_exact + _rand(small.p_exact.shape, 5) * 0.1)
    a = p0.copy(order="F")
    b = p0.copy(order="F")
    rc_o, st_o, _ = o.ghostcell_pressure(a, small.div)
    rc_g, st_g, _ = s.ghostcell_pressure(b, small.div)
    assert np.max(np.abs(a - b)) < 1e-12 * np.max(np.abs(a))
