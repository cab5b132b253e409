"""Pins of the CPU oracle against closed forms, library routines, brute force
and invariants.  The reference ships no tests or golden vectors, so these
are the checks that make the oracle trustworthy before it judges the GPU."""
import math

import numpy as np
import pytest
import scipy.linalg as sla
from scipy.interpolate import RegularGridInterpolator


def test_tdma_vs_banded(capi, oracle):
    """tdma (CORE_TDMA.F90:10-34) against LAPACK's banded solver."""
    rng = np.random.default_rng(0)
    for n in (2, 3, 9, 64, 257):
        a = rng.uniform(0.1, 1, n)
        c = rng.uniform(0.1, 1, n)
        b = -(a + c) - rng.uniform(0.01, 1, n)
        r = rng.standard_normal(n)
        a[0] = 0.0
        c[-1] = 0.0
        u = oracle.tdma(capi, a, b, c, r, 0, n - 1)
        ab = np.zeros((3, n))
        ab[0, 1:] = c[:-1]
        ab[1, :] = b
        ab[2, :-1] = a[1:]
        ref = sla.solve_banded((1, 1), ab, r)
        assert np.max(np.abs(u - ref)) <= 1e-13 * np.max(np.abs(ref))
    # a sub-range n1..n2 leaves the rest untouched
    n = 20
    a = np.full(n, 1.0); b = np.full(n, -4.0); c = np.full(n, 1.0); r = np.arange(n, dtype=float)
    u = oracle.tdma(capi, a, b, c, r, 5, 12)
    ab = np.zeros((3, 8)); ab[0, 1:] = 1; ab[1] = -4; ab[2, :-1] = 1
    assert np.allclose(u[5:13], sla.solve_banded((1, 1), ab, r[5:13]), rtol=1e-14)


@pytest.mark.parametrize("n", [9, 17, 33, 34, 65, 100, 257, 385, 513])
def test_level_count_closed_form(capi, oracle, geo, n):
    """MG_initial level count and node counts (CORE_MG.F90:69-175):
    levels = floor(log2((n-1)/4)) + 1, coarse nodes = floor(n/2) + 1."""
    x = geo.uniform_axis(n)
    o = oracle.Oracle(capi, x, geo.uniform_axis(9), geo.uniform_axis(9), capi.default_params())
    expect_levels = int(math.floor(math.log2((n - 1) / 4.0))) + 1
    assert o.nlev == max(expect_levels, 2 if n > 9 else 1) or o.nlev == expect_levels
    sizes = [n]
    for _ in range(1, o.nlev):
        sizes.append(sizes[-1] if len(sizes) > expect_levels - 1 else sizes[-1] // 2 + 1)
    assert [ln[0] for ln in o.level_nodes] == sizes
    # injected coarse nodes: x_c(i) = x_f(2i-1), last node on the boundary
    for lev in range(2, o.nlev + 1):
        xf = o.level_metric(lev - 1, 0)
        xc_ = o.level_metric(lev, 0)
        nc, nf = o.level_nodes[lev - 1][0], o.level_nodes[lev - 2][0]
        for i in range(1, nc):
            assert xc_[i] == xf[2 * i - 1]
        assert xc_[nc] == xf[nf]


def test_mg_closed_form(capi, oracle, geo):
    """test_mg (CORE_POISSON.F90:565-674): with no mask on a uniform grid the
    manufactured right-hand side is a sum of discrete Neumann eigenvectors, so
    the converged MG answer is known exactly."""
    for n in (16, 32):
        w = geo.config_test_mg(n)
        for order in (oracle.ORDER_REFERENCE, oracle.ORDER_COLORED):
            p = capi.default_params(restol_pson=1e-13, iterInter=400)
            o = oracle.Oracle(capi, *w.grid, p, ordering=order)
            o.set_workload(w)
            pp = np.zeros_like(w.div)
            o.mg_solve(pp, w.div)
            num = pp[1:n + 1, 1:n + 1, 1:n + 1]
            ex = geo.test_mg_exact(n)
            err = np.max(np.abs((num - num.mean()) - (ex - ex.mean()))) / np.max(np.abs(ex))
            assert err < 1e-11, (n, order, err)


def _interior_flag(w):
    f = (w.iblank == 0) & (w.ghost == 0)
    m = np.zeros_like(f)
    m[1:-2, 1:-2, 1:-2] = True
    return f & m


@pytest.mark.parametrize("cfg", ["c1_16", "c1_24", "c2_24"])
def test_manufactured_fixed_point(capi, oracle, geo, cfg):
    """solve_poisson on a masked grid with GCM ghost/dead cells converges to
    the manufactured pressure: p_m satisfies the flag-cell Poisson rows
    (numpy operator) and the GCM relations (scipy sparse solve), both written
    independently of the oracle."""
    kind, n = cfg.split("_")
    w = geo.config_c1(int(n)) if kind == "c1" else geo.config_c2(int(n))
    for order in (oracle.ORDER_REFERENCE, oracle.ORDER_COLORED):
        p = capi.default_params(restol_pson=1e-12, iterInter=1, itermax_pson=600, restol_gcp=1e-11, itermax_gc=100)
        o = oracle.Oracle(capi, *w.grid, p, ordering=order)
        o.set_workload(w)
        pp = np.zeros_like(w.div)
        rc, st, _ = o.solve_poisson(pp, w.div)
        assert rc == 0 and st.converged == 1
        f = _interior_flag(w)
        e = pp[f] - pp[f].mean()
        x = w.p_exact[f] - w.p_exact[f].mean()
        assert np.linalg.norm(e - x) / np.linalg.norm(x) < 1e-8, (cfg, order)


def test_gcm_fixed_point(capi, oracle, geo):
    """GCM_ghostcell_pressure iterated to convergence equals the solution of
    the GCM relations assembled as a sparse system (dead rows, ghost rows,
    self row = dp/dn)."""
    w = geo.config_c2(24)
    p = capi.default_params(restol_gcp=1e-15, itermax_gc=500)
    o = oracle.Oracle(capi, *w.grid, p)
    o.set_workload(w)
    rng = np.random.default_rng(1)
    q = np.asfortranarray(w.p_exact + 0.05 * rng.standard_normal(w.p_exact.shape))
    flag = (w.iblank == 0) & (w.ghost == 0)
    ref = q.copy(order="F")
    ref[~flag] = 0
    q0 = ref.copy(order="F")
    o.ghostcell_pressure(q0, w.div)
    # sparse solve of the relations with the same flag values
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    targets, M, idw = geo.gcm_operator(w.iblank.shape, w.iblank, w.dead, w.gcm, w.iblank.shape[1] - 3,
                                       w.iblank.shape[2] - 3)
    flat = ref.ravel(order="F").copy()
    is_t = np.zeros(flat.size, bool)
    is_t[targets] = True
    A = sp.identity(targets.size, format="csr") - M[:, targets]
    vals = spla.spsolve(A.tocsc(), M[:, ~is_t] @ flat[~is_t])
    got = q0.ravel(order="F")[targets]
    assert np.max(np.abs(got - vals)) < 1e-10 * max(1.0, np.max(np.abs(vals)))


def test_residual_matches_operator(capi, oracle, geo):
    """residual_mg (CORE_MG.F90:994-1026) equals rhs - A phi of an
    independent numpy operator, masked by igk, on a stretched grid."""
    w = geo.config_c2(24)
    o = oracle.Oracle(capi, *w.grid, capi.default_params())
    o.set_workload(w)
    rng = np.random.default_rng(2)
    phi = np.asfortranarray(rng.standard_normal(w.div.shape))
    rhs = np.asfortranarray(rng.standard_normal(w.div.shape))
    o.level(1, 0)[:] = phi
    o.level(1, 1)[:] = rhs
    o.residual_mg(1)
    res = o.level(1, 2)
    ref = (rhs[1:-2, 1:-2, 1:-2] - geo.apply_operator(phi, *w.grid)) * (1 - o.level(1, 4)[1:-2, 1:-2, 1:-2])
    assert np.max(np.abs(res[1:-2, 1:-2, 1:-2] - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_bc_reflection(capi, oracle, geo):
    """set_outer_pressure_bc + set_outer_ghost_pressure at level 1: Neumann
    ghost = p(1) - pgrad*dxc(1) (and p(n-1) + pgrad*dxc(n)); Dirichlet ghost =
    2*ppp - p(1); coarse levels homogeneous."""
    n = 12
    x = geo.tanh_axis(n + 1, [0.5], 0.02)
    y = geo.uniform_axis(n + 1)
    z = geo.uniform_axis(n + 1)
    p = capi.default_params(pbc=[2, 2, 1, 2, 2, 1], ppp=[0, 0, 0.7, 0, 0, -0.3])
    o = oracle.Oracle(capi, x, y, z, p)
    sh = (n + 3, n + 3, n + 3)
    rng = np.random.default_rng(3)
    pg = [np.asfortranarray(rng.standard_normal((n + 3, n + 3))) for _ in range(6)]
    o.set_bc(*pg)
    phi = np.asfortranarray(rng.standard_normal(sh))
    o.level(1, 0)[:] = phi
    o.bc_fill(1)
    q = o.level(1, 0)
    _, _, _, dxc, _ = geo.metrics(x)
    # nx = n + 1 cells: interior 1..n, outer ghost layers at 0 and n + 1
    J = slice(1, n + 1)
    K = slice(1, n + 1)
    assert np.allclose(q[0, J, K], phi[1, J, K] - pg[0][J, K] * dxc[1], rtol=1e-13, atol=1e-13)
    assert np.allclose(q[n + 1, J, K], phi[n, J, K] + pg[1][J, K] * dxc[n + 1], rtol=1e-13, atol=1e-13)
    I = slice(1, n + 1)
    assert np.allclose(q[I, 0, K], 2 * 0.7 - phi[I, 1, K], rtol=1e-14)
    assert np.allclose(q[I, J, n + 1], 2 * -0.3 - phi[I, J, n], rtol=1e-14)
    # coarse level: homogeneous versions
    phi2 = np.asfortranarray(rng.standard_normal(o.level(2, 0).shape))
    o.level(2, 0)[:] = phi2
    o.bc_fill(2)
    q2 = o.level(2, 0)
    nc = o.level_nodes[1][0]
    assert np.array_equal(q2[0, 1:nc - 1, 1:nc - 1], phi2[1, 1:nc - 1, 1:nc - 1])
    assert np.array_equal(q2[1:nc - 1, 0, 1:nc - 1], -phi2[1:nc - 1, 1, 1:nc - 1])


def test_transfers_trilinear(capi, oracle, geo):
    """mg_inject / mg_interp without iblank cells are trilinear interpolation
    between cell centres (CORE_MG.F90:1127-1141, 1293-1307): checked against
    scipy's RegularGridInterpolator on a stretched grid."""
    n = 16
    x = geo.tanh_axis(n + 1, [0.4], 0.02)
    y = geo.tanh_axis(n + 1, [0.6], 0.03)
    z = geo.uniform_axis(n + 1)
    o = oracle.Oracle(capi, x, y, z, capi.default_params())
    sh = (n + 3,) * 3
    z8 = np.zeros(sh, dtype=np.int8, order="F")
    o.set_masks(z8, z8, z8)
    rng = np.random.default_rng(4)
    res = np.asfortranarray(rng.standard_normal(sh))
    o.level(1, 2)[:] = res
    o.mg_inject(1)
    rhs2 = o.level(2, 1)
    xf, yf, zf = (o.level_metric(1, 3), o.level_metric(1, 4), o.level_metric(1, 5))
    xc2, yc2, zc2 = (o.level_metric(2, 3), o.level_metric(2, 4), o.level_metric(2, 5))
    f = RegularGridInterpolator((xf[:n + 1], yf[:n + 1], zf[:n + 1]), res[:n + 1, :n + 1, :n + 1])
    nc = o.level_nodes[1][0]
    pts = np.stack(np.meshgrid(xc2[1:nc], yc2[1:nc], zc2[1:nc], indexing="ij"), -1)
    assert np.max(np.abs(rhs2[1:nc, 1:nc, 1:nc] - f(pts))) < 1e-12
    # prolongation
    phi2 = np.asfortranarray(rng.standard_normal(o.level(2, 0).shape))
    phi1 = np.asfortranarray(rng.standard_normal(sh))
    o.level(2, 0)[:] = phi2
    o.level(1, 0)[:] = phi1
    o.mg_interp(2)
    d = o.level(1, 0) - phi1
    g = RegularGridInterpolator((xc2[:nc + 1], yc2[:nc + 1], zc2[:nc + 1]), phi2[:nc + 1, :nc + 1, :nc + 1])
    # nx = n + 1 cells: fine interior 1..n
    pts = np.stack(np.meshgrid(xf[1:n + 1], yf[1:n + 1], zf[1:n + 1], indexing="ij"), -1)
    ref = g(pts)
    on = lambda v: (v == 1) | (v == n)
    I, J, K = np.meshgrid(np.arange(1, n + 1), np.arange(1, n + 1), np.arange(1, n + 1), indexing="ij")
    edge = (on(I).astype(int) + on(J).astype(int) + on(K).astype(int)) >= 2
    dd = d[1:n + 1, 1:n + 1, 1:n + 1]
    assert np.max(np.abs(dd[~edge] - ref[~edge])) < 1e-12
    assert np.all(dd[edge] == 0.0)       # edge cells are skipped (:1254-1273)


def test_restriction_volume_branch(capi, oracle, geo):
    """The volume-average branch of mg_inject keeps only the last non-iblank
    child: sum0 is assigned, not accumulated (CORE_MG.F90:1159), so the
    result is res(last)*vol(last)/sum(vol)."""
    n = 8
    x = geo.tanh_axis(n + 1, [0.5], 0.05)
    y = geo.uniform_axis(n + 1)
    z = geo.uniform_axis(n + 1)
    o = oracle.Oracle(capi, x, y, z, capi.default_params())
    sh = (n + 3,) * 3
    ib = np.zeros(sh, dtype=np.int8, order="F")
    ib[3, 3, 3] = 1                       # one iblank child of coarse cell (2,2,2)
    ib[4, 4, 4] = 1
    z8 = np.zeros(sh, dtype=np.int8, order="F")
    o.set_masks(ib, z8, z8)
    rng = np.random.default_rng(5)
    res = np.asfortranarray(rng.standard_normal(sh))
    o.level(1, 2)[:] = res
    o.mg_inject(1)
    xf = o.level_metric(1, 0); yf = o.level_metric(1, 1); zf = o.level_metric(1, 2)
    kids = [(i, j, k) for k in (3, 4) for j in (3, 4) for i in (3, 4) if ib[i, j, k] == 0]
    vol = lambda i, j, k: (xf[i + 1] - xf[i]) * (yf[j + 1] - yf[j]) * (zf[k + 1] - zf[k])
    last = kids[-1]
    expect = res[last] * vol(*last) / sum(vol(*c) for c in kids)
    assert abs(o.level(2, 1)[2, 2, 2] - expect) < 1e-14
    accumulated = sum(res[c] * vol(*c) for c in kids) / sum(vol(*c) for c in kids)
    assert abs(o.level(2, 1)[2, 2, 2] - accumulated) > 1e-6


def test_mg_prepare_brute_force(capi, oracle, geo):
    """Coarse masks: igk = any child igk, ibk = (iblank volume >= half)
    (CORE_MG.F90:1462-1503), brute force over children on a stretched grid."""
    w = geo.config_c2(24)
    o = oracle.Oracle(capi, *w.grid, capi.default_params())
    o.set_workload(w)
    for lev in range(2, o.nlev + 1):
        f_igk = o.level(lev - 1, 4)
        f_ibk = o.level(lev - 1, 3)
        xf = o.level_metric(lev - 1, 0); yf = o.level_metric(lev - 1, 1); zf = o.level_metric(lev - 1, 2)
        nc = o.level_nodes[lev - 1]
        igk = o.level(lev, 4)
        ibk = o.level(lev, 3)
        for k in range(1, nc[2]):
            for j in range(1, nc[1]):
                for i in range(1, nc[0]):
                    ch = [(ii, jj, kk) for kk in (2 * k - 1, 2 * k) for jj in (2 * j - 1, 2 * j) for ii in (2 * i - 1, 2 * i)]
                    v = [(xf[a + 1] - xf[a]) * (yf[b + 1] - yf[b]) * (zf[c + 1] - zf[c]) for a, b, c in ch]
                    assert igk[i, j, k] == int(any(f_igk[c] == 1 for c in ch))
                    frac = sum(vv for vv, c in zip(v, ch) if f_ibk[c] == 1) / sum(v)
                    assert ibk[i, j, k] == int(frac >= 0.5)


def test_outer_loop_stalls_on_incompatible_rhs(capi, oracle, geo):
    """A mean-removed divergence of a random velocity field is not in the
    range of the coupled operator (the GCM closure is not conservative); the
    reference's outer loop then stalls at a constant max|p - p_temp| and
    never meets restol_gcp.  This is why the synthetic workloads use the
    manufactured right-hand side."""
    w = geo.config_c1(16)
    div = geo.divergence_rhs(*w.grid, w.iblank, w.ghost, seed=1)
    p = capi.default_params(restol_pson=1e-13, iterInter=0, itermax_pson=300, restol_gcp=1e-12, itermax_gc=20)
    o = oracle.Oracle(capi, *w.grid, p)
    o.set_workload(w)
    pp = np.zeros_like(div)
    rc, st, _ = o.solve_poisson(pp, div)
    assert rc == 1 and st.maxres2 > 1e-10


def test_itsolv_orderings_share_the_fixed_point(capi, oracle, geo):
    """itsolv (CORE_POISSON.F90:205-456) in the reference's line order and in
    the GPU's colored order: repeated passes converge to the same solution of
    the flag-cell rows with the masked cells and the ghost layers held, which
    a sparse direct solve of those rows (numpy operator) gives too."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    w = geo.config_c2(12)
    rng = np.random.default_rng(3)
    r = np.asfortranarray(rng.standard_normal(w.div.shape))
    v0 = np.asfortranarray(rng.standard_normal(w.div.shape))
    out = []
    for order in (oracle.ORDER_REFERENCE, oracle.ORDER_COLORED):
        o = oracle.Oracle(capi, *w.grid, capi.default_params(it_solver_type=1, omega_pson=1.4), ordering=order)
        o.set_workload(w)
        v = v0.copy(order="F")
        for _ in range(400):
            o.itsolv(v, r)
        o.close()
        out.append(v)
    # direct solve: unknowns are the flag cells; every other cell keeps v0
    x, y, z = w.grid
    nx, ny, nz = x.size - 2, y.size - 2, z.size - 2
    flag = np.zeros(w.div.shape, bool)
    flag[1:nx, 1:ny, 1:nz] = ((w.iblank == 0) & (w.ghost == 0))[1:nx, 1:ny, 1:nz]
    idx = -np.ones(w.div.shape, np.int64)
    idx[flag] = np.arange(flag.sum())
    met = [geo.metrics(a) for a in (x, y, z)]
    rows, cols, vals = [], [], []
    b = np.zeros(flag.sum())
    for (i, j, k) in zip(*np.nonzero(flag)):
        q = idx[i, j, k]
        cs = [(met[0][4][i] * met[0][2][i], (i - 1, j, k)), (met[0][4][i + 1] * met[0][2][i], (i + 1, j, k)),
              (met[1][4][j] * met[1][2][j], (i, j - 1, k)), (met[1][4][j + 1] * met[1][2][j], (i, j + 1, k)),
              (met[2][4][k] * met[2][2][k], (i, j, k - 1)), (met[2][4][k + 1] * met[2][2][k], (i, j, k + 1))]
        b[q] = r[i, j, k]
        rows.append(q); cols.append(q); vals.append(-sum(c for c, _ in cs))
        for c, nb in cs:
            if flag[nb]:
                rows.append(q); cols.append(idx[nb]); vals.append(c)
            else:
                b[q] -= c * v0[nb]
    A = sp.csr_matrix((vals, (rows, cols)), shape=(b.size, b.size))
    xs = spla.spsolve(A.tocsc(), b)
    for v in out:
        assert np.max(np.abs(v[flag] - xs)) <= 1e-10 * np.max(np.abs(xs))
        assert np.array_equal(v[~flag], v0[~flag])


def _empty_box(geo, nx, ny, nz):
    """uniform box of nx x ny x nz cells without bodies"""
    w = geo.Workload()
    w.grid = (geo.uniform_axis(nx + 1), geo.uniform_axis(ny + 1), geo.uniform_axis(nz + 1))
    sh = tuple(a.size for a in w.grid)
    w.iblank = np.zeros(sh, np.int8, order="F")
    w.ghost = np.zeros(sh, np.int8, order="F")
    w.dead = np.zeros(sh, np.int8, order="F")
    w.gcm = geo.empty_gcm()
    return w


def _periodic_params(capi, axes, **kw):
    bcv = [capi.BC_TYPE_DIRICHLET] * 6
    for a in axes:
        bcv[2 * a] = bcv[2 * a + 1] = capi.BC_TYPE_PERIODIC
    return capi.default_params(bc_vel=bcv, **kw)


@pytest.mark.parametrize("lev", [1, 2])
def test_periodic_fill_wraps(capi, oracle, geo, lev):
    """Periodic x and y (set_outer_pressure_bc, CORE_SET_OUTER_BC.F90:
    1007-1026): the ghost layer takes the value of the interior cell at the
    other end (to rounding: 2 (a+b)/2 - a), on the finest and a coarse
    level; the z faces keep the Neumann reflection."""
    w = _empty_box(geo, 16, 12, 10)
    o = oracle.Oracle(capi, *w.grid, _periodic_params(capi, (0, 1)))
    o.set_workload(w)
    nx, ny, nz = o.level_nodes[lev - 1]
    phi = o.level(lev, 0)
    rng = np.random.default_rng(9)
    phi[...] = rng.standard_normal(phi.shape)
    o.bc_fill(lev)
    tol = 4e-16 * np.max(np.abs(phi))
    assert np.max(np.abs(phi[0, 1:ny, 1:nz] - phi[nx - 1, 1:ny, 1:nz])) <= tol
    assert np.max(np.abs(phi[nx, 1:ny, 1:nz] - phi[1, 1:ny, 1:nz])) <= tol
    assert np.max(np.abs(phi[1:nx, 0, 1:nz] - phi[1:nx, ny - 1, 1:nz])) <= tol
    assert np.max(np.abs(phi[1:nx, ny, 1:nz] - phi[1:nx, 1, 1:nz])) <= tol
    assert np.array_equal(phi[1:nx, 1:ny, 0], phi[1:nx, 1:ny, 1])
    assert np.array_equal(phi[1:nx, 1:ny, nz], phi[1:nx, 1:ny, nz - 1])
    o.close()


def test_periodic_solve_brute_force(capi, oracle, geo):
    """solve_poisson with periodic x faces on a box without bodies equals the
    least-squares solution of the 7-point system assembled here with x
    neighbours wrapping around (cell 1's left neighbour is cell nx-1) and
    zero-flux y and z faces (after mean removal)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    w = _empty_box(geo, 12, 10, 8)
    p = _periodic_params(capi, (0,), restol_pson=1e-14, iterInter=1, itermax_pson=3000, restol_gcp=1e-13)
    o = oracle.Oracle(capi, *w.grid, p)
    o.set_workload(w)
    nx, ny, nz = o.level_nodes[0]
    dxinv, dyinv, dzinv, dxcinv, dycinv, dzcinv = (o.level_metric(1, k) for k in range(6, 12))
    rng = np.random.default_rng(5)
    div = np.zeros((nx + 2, ny + 2, nz + 2), order="F")
    div[1:nx, 1:ny, 1:nz] = rng.standard_normal((nx - 1, ny - 1, nz - 1))
    vol = 1.0 / (dxinv[1:nx, None, None] * dyinv[None, 1:ny, None] * dzinv[None, None, 1:nz])
    div[1:nx, 1:ny, 1:nz] -= np.sum(vol * div[1:nx, 1:ny, 1:nz]) / np.sum(vol)
    pres = np.zeros_like(div)
    rc, st, _ = o.solve_poisson(pres, div)
    o.close()
    assert rc == 0, st.as_dict()
    m = (nx - 1) * (ny - 1) * (nz - 1)
    lin = lambda i, j, k: (i - 1) + (nx - 1) * ((j - 1) + (ny - 1) * (k - 1))   # noqa: E731
    rows, cols, vals = [], [], []
    for k in range(1, nz):
        for j in range(1, ny):
            for i in range(1, nx):
                t = lin(i, j, k)
                faces = ((i - 1 if i > 1 else nx - 1, j, k, dxcinv[i] * dxinv[i]),
                         (i + 1 if i < nx - 1 else 1, j, k, dxcinv[i + 1] * dxinv[i]),
                         (i, j - 1, k, dycinv[j] * dyinv[j] if j > 1 else 0.0),
                         (i, j + 1, k, dycinv[j + 1] * dyinv[j] if j < ny - 1 else 0.0),
                         (i, j, k - 1, dzcinv[k] * dzinv[k] if k > 1 else 0.0),
                         (i, j, k + 1, dzcinv[k + 1] * dzinv[k] if k < nz - 1 else 0.0))
                for (a, b, c, cf) in faces:
                    if cf == 0.0:
                        continue
                    rows += [t, t]
                    cols += [lin(a, b, c), t]
                    vals += [cf, -cf]
    rhs = div[1:nx, 1:ny, 1:nz].ravel(order="F")
    got = pres[1:nx, 1:ny, 1:nz]
    x = got - got.mean()

    def lsq(keep):
        A = sp.csr_matrix(([v for v, kp in zip(vals, keep) if kp], ([r for r, kp in zip(rows, keep) if kp],
                                                                     [c for c, kp in zip(cols, keep) if kp])),
                          shape=(m, m))
        y = spla.lsqr(A, rhs, atol=1e-15, btol=1e-15, iter_lim=20000)[0].reshape((nx - 1, ny - 1, nz - 1), order="F")
        return y - y.mean()

    y = lsq([True] * len(vals))
    assert np.linalg.norm(x - y) <= 1e-9 * np.linalg.norm(y)
    # without the wrapping x faces (zero-flux there) the answer is another one
    ci = np.array(cols) % (nx - 1) + 1
    ri = np.array(rows) % (nx - 1) + 1
    wrap = ((ri == 1) & (ci == nx - 1)) | ((ri == nx - 1) & (ci == 1))
    wrap_pairs = set(k for k in range(0, len(vals), 2) if wrap[k])
    keep = [not (k - (k % 2) in wrap_pairs) for k in range(len(vals))]
    y_neu = lsq(keep)
    assert np.linalg.norm(x - y_neu) > 1e-2 * np.linalg.norm(y_neu)
