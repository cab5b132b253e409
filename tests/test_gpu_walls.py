"""The correction V-cycles that treat immersed bodies as zero-flux walls
(fc_pson_params.vc_walls, new: the reference's V-cycles hold every masked
cell, CORE_MG.F90:663-731).  Each wall operation runs on the GPU through the
level API (FC_LVL_WALLS) and is checked against what it must compute, stated
independently in numpy:

* masks: a cell is free on level 1 if iblank = 0 (flag and ghost cells), on a
  coarse level if one of its restriction children is free; a free cell's
  face is a wall when the interior cell across it is held (brute force);
* residual: rhs - A_w phi with A_w the 7-point operator whose wall faces
  carry no flux (assembled face by face from the level metrics);
* half sweep: after one colour of one direction (omega = 1) every swept free
  cell away from the domain faces has zero A_w residual, and no other cell
  moved;
* prolongation: a coarse field equal to a constant on free cells (garbage on
  held ones) adds exactly that constant to every free fine cell, and where
  no held coarse cell is near it equals the reference's trilinear
  prolongation;
* coarsest-level direct solve (coarse_solve = 1): zero wall residual on the
  groups of free cells that touch a Dirichlet face, the group's
  volume-weighted right-hand-side mean on the others (scipy labels the
  groups).
The solve-level effect (same fixed point, fewer outer iterations) is in
test_gpu_solve.py."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WALLS = 16     # FC_LVL_WALLS


@pytest.fixture(scope="module")
def w(geo):
    return geo.config_c2(24)


@pytest.fixture(scope="module")
def pair(capi, oracle, gpu, w):
    p = capi.default_params()
    s = gpu.PoissonSolver(*w.grid, p)
    s.set_workload(w)
    o = oracle.Oracle(capi, *w.grid, p)
    o.set_workload(w)
    yield s, o
    s.close()
    o.close()


def _coarsened(nodes):
    """per level: (ic, jc, kc) = 1 where the axis was coarsened from the
    finer level"""
    out = [(0, 0, 0)]
    for a, b in zip(nodes[:-1], nodes[1:]):
        out.append(tuple(int(x != y) for x, y in zip(a, b)))
    return out


def _children(i, c):
    return ((2 * i - 1, 2 * i) if c else (i, i))


def np_wall_masks(w, nodes):
    """(mask, pflag) per level, brute force"""
    cz = _coarsened(nodes)
    free, masks, pflags = [], [], []
    for l, (nx, ny, nz) in enumerate(nodes):
        f = np.zeros((nx + 2, ny + 2, nz + 2), dtype=bool)
        ic, jc, kc = cz[l]
        for k in range(1, nz):
            for j in range(1, ny):
                for i in range(1, nx):
                    if l == 0:
                        f[i, j, k] = w.iblank[i, j, k] == 0
                    else:
                        i1, i2 = _children(i, ic)
                        j1, j2 = _children(j, jc)
                        k1, k2 = _children(k, kc)
                        f[i, j, k] = free[l - 1][i1:i2 + 1, j1:j2 + 1, k1:k2 + 1].any()
        free.append(f)
        held = np.zeros_like(f)
        held[1:nx, 1:ny, 1:nz] = ~f[1:nx, 1:ny, 1:nz]
        m = np.zeros(f.shape, dtype=np.int8)
        m[held] = 1
        for bit, (di, dj, dk) in zip((2, 4, 8, 16, 32, 64),
                                     ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1))):
            sh = np.zeros_like(held)
            src = held[max(0, di):held.shape[0] + min(0, di), max(0, dj):held.shape[1] + min(0, dj),
                       max(0, dk):held.shape[2] + min(0, dk)]
            sh[max(0, -di):held.shape[0] + min(0, -di), max(0, -dj):held.shape[1] + min(0, -dj),
               max(0, -dk):held.shape[2] + min(0, -dk)] = src
            fi = np.zeros_like(held)
            fi[1:nx, 1:ny, 1:nz] = f[1:nx, 1:ny, 1:nz]
            m[fi & sh] |= bit
        masks.append(m)
        if l > 0:
            pf = np.zeros(f.shape, dtype=np.int8)
            for k in range(0, nz + 1):
                for j in range(0, ny + 1):
                    for i in range(0, nx + 1):
                        blk = held[i:min(i + ic, nx + 1) + 1, j:min(j + jc, ny + 1) + 1, k:min(k + kc, nz + 1) + 1]
                        pf[i, j, k] = int(blk.any())
            pflags.append(pf)
        else:
            pflags.append(None)
    return masks, pflags


def np_wall_residual(o, lev, phi, rhs, mask):
    """rhs - A_w phi on interior cells (i = 1..nx-1), zero on held cells;
    A_w phi = sum over open faces of c_f (phi_f - phi_c)"""
    nx, ny, nz = (a - 2 for a in phi.shape)
    met = [o.level_metric(lev, k) for k in range(12)]
    dxinv, dyinv, dzinv, dxcinv, dycinv, dzcinv = met[6:12]
    I, J, K = np.meshgrid(np.arange(1, nx), np.arange(1, ny), np.arange(1, nz), indexing="ij")
    c = phi[1:nx, 1:ny, 1:nz]
    mk = mask[1:nx, 1:ny, 1:nz].astype(np.int64)
    faces = (((-1, 0, 0), dxcinv[I] * dxinv[I], 2), ((1, 0, 0), dxcinv[I + 1] * dxinv[I], 4),
             ((0, -1, 0), dycinv[J] * dyinv[J], 8), ((0, 1, 0), dycinv[J + 1] * dyinv[J], 16),
             ((0, 0, -1), dzcinv[K] * dzinv[K], 32), ((0, 0, 1), dzcinv[K + 1] * dzinv[K], 64))
    A = np.zeros_like(c)
    for (di, dj, dk), coef, bit in faces:
        nb = phi[1 + di:nx + di, 1 + dj:ny + dj, 1 + dk:nz + dk]
        A += np.where(mk & bit, 0.0, coef * (nb - c))
    r = np.where(mk & 1, 0.0, rhs[1:nx, 1:ny, 1:nz] - A)
    return r


def _rand(shape, seed):
    return np.asfortranarray(np.random.default_rng(seed).standard_normal(shape))


def test_wall_masks(pair, w):
    s, o = pair
    masks, pflags = np_wall_masks(w, s.level_nodes)
    for lev in range(1, s.nlev + 1):
        g = s.lvl_read(lev, 5)
        n = g.shape
        inner = (slice(1, n[0] - 2), slice(1, n[1] - 2), slice(1, n[2] - 2))
        assert np.array_equal(g[inner], masks[lev - 1][inner]), lev
        if lev > 1:
            gp = s.lvl_read(lev, 6)
            sub = (slice(0, n[0] - 1), slice(0, n[1] - 1), slice(0, n[2] - 1))
            assert np.array_equal(gp[sub], pflags[lev - 1][sub]), lev
    # the body is present as walls on level 1 and still on the coarsest level
    assert np.count_nonzero(s.lvl_read(1, 5) & 0x7e) > 0
    assert np.count_nonzero(s.lvl_read(s.nlev, 5) & 1) > 0


@pytest.mark.parametrize("lev", [1, 2, 3])
def test_wall_residual(capi, pair, lev):
    s, o = pair
    sh = s.level_shape(lev)
    phi, rhs = _rand(sh, 100 + lev), _rand(sh, 110 + lev)
    s.lvl_write(lev, 0, phi)
    s.lvl_write(lev, 1, rhs)
    s.lvl_write(lev, 2, np.zeros(sh))
    s.residual_mg(lev, prec=capi.PREC_FP64 | WALLS)
    g = s.lvl_read(lev, 2)
    ref = np_wall_residual(o, lev, phi, rhs, s.lvl_read(lev, 5))
    nx, ny, nz = (a - 2 for a in sh)
    err = np.max(np.abs(g[1:nx, 1:ny, 1:nz] - ref)) / np.max(np.abs(ref))
    assert err < 1e-13, err
    # the i = nx column keeps the res = rhs copy (CORE_MG.F90:1672-1678)
    assert np.array_equal(g[nx, 1:ny, 1:nz], rhs[nx, 1:ny, 1:nz])


@pytest.mark.parametrize("lev", [1, 2])
@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("prec", [0, 1])
def test_wall_sweep_zero_line_residual(capi, pair, lev, d, prec):
    s, o = pair
    sh = s.level_shape(lev)
    nx, ny, nz = (a - 2 for a in sh)
    mask = s.lvl_read(lev, 5)
    rhs = _rand(sh, 130 + lev)
    phi = _rand(sh, 120 + lev)
    wphi, wrhs = (0, 1) if prec == 0 else (10, 11)
    s.lvl_write(lev, wrhs, rhs)
    I, J, K = np.meshgrid(np.arange(sh[0]), np.arange(sh[1]), np.arange(sh[2]), indexing="ij")
    for color in (0, 1):
        s.lvl_write(lev, wphi, phi)
        s.sweep(lev, d, color, prec=prec | WALLS)
        new = s.lvl_read(lev, wphi).astype(np.float64)
        a, b = ((J, K), (I, K), (I, J))[d]
        swept = ((a + b) % 2 == color)
        inner = np.zeros(sh, bool)
        inner[2:nx - 1, 2:ny - 1, 2:nz - 1] = True        # away from the domain faces
        free = (mask & 1) == 0
        rhs_t = rhs.astype(np.float32).astype(np.float64) if prec else rhs
        r = np.zeros(sh)
        r[1:nx, 1:ny, 1:nz] = np_wall_residual(o, lev, new, rhs_t, mask)
        sel = swept & inner & free
        scale = np.max(np.abs(rhs_t[sel])) + np.max(np.abs(new[sel]))
        tol = 1e-12 if prec == 0 else 2e-4
        assert np.max(np.abs(r[sel])) <= tol * scale * np.max(np.abs(o.level_metric(lev, 6))) ** 2, (color,)
        # nothing else moved: the other colour, held cells
        phi_t = phi.astype(np.float32).astype(np.float64) if prec else phi
        keep = np.zeros(sh, bool)
        keep[1:nx, 1:ny, 1:nz] = True
        keep &= ~swept | ~free
        assert np.array_equal(new[keep], phi_t[keep])


@pytest.mark.parametrize("lev", [2, 3])
def test_wall_prolong(capi, pair, lev):
    s, o = pair
    shc, shf = s.level_shape(lev), s.level_shape(lev - 1)
    mc = s.lvl_read(lev, 5)
    mf = s.lvl_read(lev - 1, 5)
    held_c = (mc & 1) == 1
    # constant on the free coarse cells (ghost layers included), garbage on held ones
    C0 = 0.375
    pc = np.full(shc, C0, order="F")
    pc[held_c] = 1.0e6
    pf = _rand(shf, 140 + lev)
    s.lvl_write(lev, 0, pc)
    s.lvl_write(lev - 1, 0, pf)
    s.mg_interp(lev, prec=capi.PREC_FP64 | WALLS)
    g = s.lvl_read(lev - 1, 0)
    nx, ny, nz = (a - 2 for a in shf)
    I, J, K = np.meshgrid(np.arange(shf[0]), np.arange(shf[1]), np.arange(shf[2]), indexing="ij")
    onf = ((I == 1) | (I == nx - 1)).astype(int) + ((J == 1) | (J == ny - 1)).astype(int) + \
          ((K == 1) | (K == nz - 1)).astype(int)
    inner = (I >= 1) & (I <= nx - 1) & (J >= 1) & (J <= ny - 1) & (K >= 1) & (K <= nz - 1)
    free_f = (mf & 1) == 0
    upd = inner & free_f & (onf < 2)      # mg_interp skips cells on two domain faces
    assert np.max(np.abs(g[upd] - (pf[upd] + C0))) <= 1e-14 * (1.0 + np.max(np.abs(pf)))
    assert np.array_equal(g[~upd], pf[~upd])
    # a random coarse field: equal to the reference prolongation wherever no
    # held coarse cell is in the interpolation block and the reference's
    # volume branch does not apply
    pc = _rand(shc, 150 + lev)
    s.lvl_write(lev, 0, pc)
    s.lvl_write(lev - 1, 0, pf)
    s.mg_interp(lev, prec=capi.PREC_FP64 | WALLS)
    gw = s.lvl_read(lev - 1, 0)
    s.lvl_write(lev, 0, pc)
    s.lvl_write(lev - 1, 0, pf)
    s.mg_interp(lev, prec=capi.PREC_FP64)
    gr = s.lvl_read(lev - 1, 0)
    ibk_c = s.lvl_read(lev, 3)
    ibk_f = s.lvl_read(lev - 1, 3)
    cz = _coarsened(s.level_nodes)[lev - 1]
    gi = np.where(cz[0], I // 2, I)
    gj = np.where(cz[1], J // 2, J)
    gk = np.where(cz[2], K // 2, K)
    clean = np.zeros(shf, bool)
    ok = upd & (ibk_f != 1)
    for idx in zip(*np.nonzero(ok)):
        a, b, c = gi[idx], gj[idx], gk[idx]
        blk = (slice(a, a + cz[0] + 1), slice(b, b + cz[1] + 1), slice(c, c + cz[2] + 1))
        clean[idx] = not held_c[blk].any() and not (ibk_c[blk] == 1).any()
    assert clean.sum() > 0.5 * ok.sum()
    assert np.array_equal(gw[clean], gr[clean])


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("dirichlet", [False, True])
def test_coarse_direct_solve(capi, gpu, oracle, w, prec, dirichlet):
    """coarse_solve = 1: the coarsest level solved directly.  On each group
    of connected free cells the wall residual is zero where the group touches
    a Dirichlet face; elsewhere it equals the volume-weighted mean of the
    right-hand side over the group (what the projection onto solvable
    right-hand sides removes), and the group's first cell (Fortran order) is
    zero to rounding.  Groups from scipy's 6-connected labelling of the free cells."""
    import scipy.ndimage as ndi
    pbc = [capi.PBC_NEUMANN] * 6
    if dirichlet:
        pbc[1] = capi.PBC_DIRICHLET                  # x = 1
    p = capi.default_params(coarse_solve=1, vc_walls=1, pbc=pbc)
    s = gpu.PoissonSolver(*w.grid, p)
    o = oracle.Oracle(capi, *w.grid, p)
    try:
        s.set_workload(w)
        o.set_workload(w)
        lev = s.nlev
        sh = s.level_shape(lev)
        nx, ny, nz = (a - 2 for a in sh)
        mask = s.lvl_read(lev, 5)
        rhs = _rand(sh, 150)
        wphi, wrhs = (0, 1) if prec == 0 else (10, 11)
        s.lvl_write(lev, wrhs, rhs)
        s.lvl_write(lev, wphi, np.zeros(sh, order="F"))
        s.coarse_solve(prec=prec)
        e = s.lvl_read(lev, wphi).astype(np.float64)
        rhs_t = rhs.astype(np.float32).astype(np.float64) if prec else rhs
        r = np_wall_residual(o, lev, e, rhs_t, mask)
        dxinv, dyinv, dzinv = (o.level_metric(lev, k) for k in (6, 7, 8))
    finally:
        s.close()
        o.close()
    free = (mask[1:nx, 1:ny, 1:nz] & 1) == 0
    lab, ng = ndi.label(free)
    assert ng >= 1
    vol = 1.0 / (dxinv[1:nx, None, None] * dyinv[None, 1:ny, None] * dzinv[None, None, 1:nz])
    rin = rhs_t[1:nx, 1:ny, 1:nz]
    ein = e[1:nx, 1:ny, 1:nz]
    cmax = max(np.max(np.abs(o_)) for o_ in (dxinv, dyinv, dzinv)) ** 2
    tol = (1e-10 if prec == 0 else 1e-6) * (np.max(np.abs(rin)) + cmax * np.max(np.abs(ein)))
    print("coarsest level %s: %d free cells in %d groups" % ((nx, ny, nz), int(free.sum()), ng), flush=True)
    for g in range(1, ng + 1):
        sel = lab == g
        touches = dirichlet and sel[-1, :, :].any()
        expect = 0.0 if touches else np.sum(vol[sel] * rin[sel]) / np.sum(vol[sel])
        assert np.max(np.abs(r[sel] - expect)) <= tol, (g, touches)
        if not touches:
            cells = np.argwhere(sel)
            flat = np.ravel_multi_index(cells.T, sel.shape, order="F")
            first = tuple(cells[np.argmin(flat)])
            assert abs(ein[first]) <= 1e-10 * np.max(np.abs(ein))
    # held cells keep zero
    assert np.all(ein[~free] == 0.0)
