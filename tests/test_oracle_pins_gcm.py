"""Pins of the oracle's rhs_poisson, enforce_global_mass_consv,
GCM_enforce_p_compatibility, hybrid
(mix_GC_form = 1) ghost update and inverse-distance dead-cell values, each
by a property the mathematics fixes, not by restating the oracle's loops:

* rhs_poisson (CORE_POISSON.F90:9-79, correct_ghost_massflux
  GCM_GHOSTCELL_UPDATE.F90:742-787): face velocities linear in the face
  coordinate have a constant discrete divergence a + b + c on every cell of a
  stretched grid, so div = dtinv (a+b+c) flag, sumd = (a+b+c) x flag volume
  (plus the weighted ghost-cell fluxes with mix_GC_form = 1), resDiv =
  |a+b+c|;
* enforce_global_mass_consv (CORE_SET_OUTER_BC.F90:515-735): the same
  linear face velocities carry the net flux (a+b+c) V_flag (plus the
  weighted ghost volumes with mix_GC_form = 1); with outflow faces at x = 1
  and z = 0 of the unit cube the fluid area is 2, the correction is
  -flux / 2, only those faces' velocities (and bcxu/bczw) move, by +corr at
  x = 1 and -corr at z = 0, and the corrected faces carry no net flux;
* GCM_enforce_p_compatibility (GCM_GHOSTCELL_UPDATE.F90:865-1003): for
  p = a x^2 + b y^2 + c z^2 on a uniform grid each cell's discrete
  Laplacian flux is 2 (a+b+c) times its volume, so the correction on the
  zero-gradient faces is -2 (a+b+c) V_flag / outflow_area (V_flag + ghost
  volume, less the res_laplacian volume terms, with mix_GC_form = 1), only
  those faces change, and a Dirichlet face suppresses it;
* the hybrid ghost update (GCM_GHOSTCELL_UPDATE.F90:612-637): its fixed
  point p_G = (1-alpha) sum_r c_r p_r + alpha (sum_f w_f p_f + r_G/coeffG)
  with w_f = c_f / sum c (the Jacobi average of the cell's Poisson row),
  assembled as a sparse system and solved with scipy, equals the oracle's
  converged sweeps;
* the inverse-distance dead-cell value (GCM_GHOSTCELL_UPDATE.F90:459-506):
  on dead cells whose stencil holds a pure solid cell the value is the
  1/d^2-weighted average of the previous pressure over the stencil's fluid
  and dead cells (d from the image point to the cell centres): a constant
  previous pressure is reproduced, and a random one gives the average
  computed here from the geometry."""
import numpy as np
import pytest

INC = np.array([[0, 0, 0], [0, 1, 0], [1, 1, 0], [1, 0, 0], [0, 0, 1], [0, 1, 1], [1, 1, 1], [1, 0, 1]])


def _faces_linear(w, a, b, c):
    """face_u(i,j,k) = a x(i), face_v = b y(j), face_w = c z(k) on the whole
    field"""
    x, y, z = w.grid
    shp = w.div.shape
    fu = np.asfortranarray(np.broadcast_to(a * x[:, None, None], shp))
    fv = np.asfortranarray(np.broadcast_to(b * y[None, :, None], shp))
    fw = np.asfortranarray(np.broadcast_to(c * z[None, None, :], shp))
    return fu, fv, fw


def _vol(geo, w):
    x, y, z = w.grid
    dx, dy, dz = (geo.metrics(t)[1] for t in (x, y, z))
    return dx[:, None, None] * dy[None, :, None] * dz[None, None, :]


def _interior(shape):
    m = np.zeros(shape, bool)
    m[1:-2, 1:-2, 1:-2] = True
    return m


@pytest.mark.parametrize("mix", [0, 1])
def test_rhs_poisson_linear_faces(capi, oracle, geo, mix):
    w = geo.config_c2(24)
    a, b, c, dtinv = 0.7, -0.3, 1.9, 4.0
    s_abc = a + b + c
    o = oracle.Oracle(capi, *w.grid, capi.default_params(mix_GC_form=mix))
    o.set_workload(w)
    fu, fv, fw = _faces_linear(w, a, b, c)
    rc, div, sumd, resdiv = o.rhs_poisson(fu, fv, fw, dtinv)
    o.close()
    assert rc == 0
    flag = _interior(w.div.shape) & (w.iblank == 0) & (w.ghost == 0)
    vol = _vol(geo, w)
    assert np.allclose(div[flag], dtinv * s_abc, rtol=1e-12, atol=0)
    expect = s_abc * np.sum(vol[flag])
    gmask = np.zeros(w.div.shape, bool)
    g = w.gcm
    gmask[g["iGhost"], g["jGhost"], g["kGhost"]] = True
    if mix:
        wgt = np.cos(np.pi / 2.0 * (1.0 - g["alphaGhost"]))
        expect += s_abc * np.sum(wgt * vol[g["iGhost"], g["jGhost"], g["kGhost"]])
        assert np.allclose(div[g["iGhost"], g["jGhost"], g["kGhost"]], dtinv * s_abc * wgt, rtol=1e-12, atol=1e-14)
    else:
        assert np.all(div[gmask] == 0.0)
    assert np.all(div[~(flag | gmask)] == 0.0)
    assert abs(sumd - expect) <= 1e-12 * abs(expect)
    assert abs(resdiv - abs(s_abc)) <= 1e-12 * abs(s_abc)


@pytest.mark.parametrize("mix", [0, 1])
def test_mass_consv_linear_faces(capi, oracle, geo, mix):
    w = geo.config_c1(16)
    a, b, c, dtinv = 0.7, -0.3, 1.9, 4.0
    s_abc = a + b + c
    bcv = [capi.BC_TYPE_DIRICHLET] * 6
    bcv[1] = capi.BC_TYPE_ZERO_GRADIENT          # outflow at x = 1
    bcv[4] = capi.BC_TYPE_ZERO_GRADIENT          # and at z = 0
    o = oracle.Oracle(capi, *w.grid, capi.default_params(mix_GC_form=mix, bc_vel=bcv))
    o.set_workload(w)
    nx, ny, nz = (t.size - 2 for t in w.grid)
    vb0 = [np.zeros(sh, order="F") for sh in ((2, ny + 2, nz + 2), (nx + 2, 2, nz + 2), (nx + 2, ny + 2, 2))]
    o.set_vel_bc(*vb0)
    fu, fv, fw = _faces_linear(w, a, b, c)
    f0 = [fu.copy(), fv.copy(), fw.copy()]
    rc, div, sumd, resdiv = o.rhs_poisson(fu, fv, fw, dtinv)
    vb = o.get_vel_bc()
    area = o.outflow_area()
    o.close()
    assert rc == 0
    assert abs(area - 2.0) <= 1e-14
    flag = _interior(w.div.shape) & (w.iblank == 0) & (w.ghost == 0)
    vol = _vol(geo, w)
    flux = s_abc * np.sum(vol[flag])
    if mix:
        g = w.gcm
        wgt = np.cos(np.pi / 2.0 * (1.0 - g["alphaGhost"]))
        flux += s_abc * np.sum(wgt * vol[g["iGhost"], g["jGhost"], g["kGhost"]])
    corr = -flux / 2.0
    du, dv, dw = fu - f0[0], fv - f0[1], fw - f0[2]
    assert np.allclose(du[nx, 1:ny, 1:nz], corr, rtol=1e-12, atol=0)
    du[nx, 1:ny, 1:nz] = 0.0
    assert np.allclose(dw[1:nx, 1:ny, 1], -corr, rtol=1e-12, atol=0)
    dw[1:nx, 1:ny, 1] = 0.0
    assert not du.any() and not dv.any() and not dw.any()
    assert np.allclose(vb[0][1, 0:ny + 1, 0:nz + 1], corr, rtol=1e-12, atol=0)
    vb[0][1, 0:ny + 1, 0:nz + 1] = 0.0
    assert np.allclose(vb[2][0:nx + 1, 0:ny + 1, 0], -corr, rtol=1e-12, atol=0)
    vb[2][0:nx + 1, 0:ny + 1, 0] = 0.0
    assert not any(v.any() for v in vb)
    # the corrected faces carry no net flux: the correction's area is the
    # whole fluid part of the two faces, which the sphere does not reach
    assert abs(sumd) <= 1e-12 * abs(flux)


@pytest.mark.parametrize("mix", [0, 1])
def test_enforce_p_compatibility_quadratic(capi, oracle, geo, mix):
    w = geo.config_c1(16)
    a, b, c = 0.4, 1.3, -0.6
    x, y, z = w.grid
    xc, yc, zc = (geo.metrics(t)[0] for t in (x, y, z))
    pres = np.asfortranarray(a * xc[:, None, None] ** 2 + b * yc[None, :, None] ** 2 + c * zc[None, None, :] ** 2)
    area = 0.37
    bcv = [capi.BC_TYPE_DIRICHLET] * 6
    bcv[1] = capi.BC_TYPE_ZERO_GRADIENT          # outflow at x = 1
    bcv[4] = capi.BC_TYPE_ZERO_GRADIENT          # and at z = 0
    prm = capi.default_params(mix_GC_form=mix, bc_vel=bcv, itermax_gc=1)
    o = oracle.Oracle(capi, *w.grid, prm)
    o.set_workload(w)
    nx, ny, nz = (t.size - 2 for t in (x, y, z))
    planes = [np.full(s, 0.25, order="F") for s in [(ny + 2, nz + 2)] * 2 + [(nx + 2, nz + 2)] * 2
              + [(nx + 2, ny + 2)] * 2]
    o.set_bc(*planes, outflow_area=area)
    extra = 0.0
    vol = _vol(geo, w)
    if mix:
        # res_laplacian of the hybrid ghost update on this pressure
        q = pres.copy(order="F")
        _, _, reslap = o.ghostcell_pressure(q, w.div)
        g = w.gcm
        gv = vol[g["iGhost"], g["jGhost"], g["kGhost"]]
        extra = 2.0 * (a + b + c) * np.sum(gv) - np.sum(reslap * gv)
    o.enforce_p_compatibility(pres)
    after = o.get_bc()
    flag = _interior(w.div.shape) & (w.iblank == 0) & (w.ghost == 0)
    flux = 2.0 * (a + b + c) * np.sum(vol[flag]) + extra
    corr = -flux / area
    # the face planes are set on 0..n of their two indices
    sl = {0: (slice(0, ny + 1), slice(0, nz + 1)), 1: (slice(0, ny + 1), slice(0, nz + 1)),
          2: (slice(0, nx + 1), slice(0, nz + 1)), 3: (slice(0, nx + 1), slice(0, nz + 1)),
          4: (slice(0, nx + 1), slice(0, ny + 1)), 5: (slice(0, nx + 1), slice(0, ny + 1))}
    for f in range(6):
        d = after[f][sl[f]] - planes[f][sl[f]]
        if bcv[f] == capi.BC_TYPE_ZERO_GRADIENT:
            sign = -1.0 if f % 2 == 0 else 1.0
            assert np.allclose(d, sign * corr, rtol=1e-10, atol=1e-12 * abs(corr)), f
        else:
            assert np.all(after[f] == planes[f]), f
    o.close()
    # a Dirichlet pressure face: no correction at all
    pbc = [capi.PBC_NEUMANN] * 6
    pbc[3] = capi.PBC_DIRICHLET
    o = oracle.Oracle(capi, *w.grid, capi.default_params(mix_GC_form=mix, bc_vel=bcv, pbc=pbc))
    o.set_workload(w)
    o.set_bc(*planes, outflow_area=area)
    o.enforce_p_compatibility(pres)
    assert all(np.array_equal(u, v) for u, v in zip(o.get_bc(), planes))
    o.close()


def test_gcm_hybrid_fixed_point(capi, oracle, geo):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    w = geo.config_c1(16)
    x, y, z = w.grid
    mets = [geo.metrics(t) for t in (x, y, z)]
    shp = w.div.shape
    nx2, ny2 = shp[0], shp[1]
    lin = lambda i, j, k: i + nx2 * (j + ny2 * k)   # noqa: E731
    rng = np.random.default_rng(4)
    flag = (w.iblank == 0) & (w.ghost == 0)
    base = np.asfortranarray(rng.standard_normal(shp))
    base[~flag] = 0.0
    r = np.asfortranarray(rng.standard_normal(shp))
    o = oracle.Oracle(capi, *w.grid, capi.default_params(mix_GC_form=1, restol_gcp=1e-15, itermax_gc=3000))
    o.set_workload(w)
    q = base.copy(order="F")
    rc, st, _ = o.ghostcell_pressure(q, r)
    o.close()
    assert st.gcm_change < 1e-13
    # sparse system: dead rows as assembled by gcm_operator, ghost rows hybrid
    targets, M, idw = geo.gcm_operator(shp, w.iblank, w.dead, w.gcm, ny2 - 3, shp[2] - 3)
    assert idw.size == 0
    g = w.gcm
    nG = g["iGhost"].size
    gl = lin(g["iGhost"], g["jGhost"], g["kGhost"])
    M = M.tolil()
    const = np.zeros(targets.size)
    row_of = {int(t): n for n, t in enumerate(targets)}
    for n in range(nG):
        i, j, k = int(g["iGhost"][n]), int(g["jGhost"][n]), int(g["kGhost"][n])
        row = row_of[int(gl[n])]
        al = g["alphaGhost"][n]
        M[row, :] = M[row, :] * (1.0 - al)
        cf = [(mets[0][4][i] * mets[0][2][i], lin(i - 1, j, k)), (mets[0][4][i + 1] * mets[0][2][i], lin(i + 1, j, k)),
              (mets[1][4][j] * mets[1][2][j], lin(i, j - 1, k)), (mets[1][4][j + 1] * mets[1][2][j], lin(i, j + 1, k)),
              (mets[2][4][k] * mets[2][2][k], lin(i, j, k - 1)), (mets[2][4][k + 1] * mets[2][2][k], lin(i, j, k + 1))]
        csum = sum(cv for cv, _ in cf)
        for cv, col in cf:
            M[row, col] = M[row, col] + al * cv / csum
        const[row] = -al * r[i, j, k] / csum
    M = M.tocsr()
    flat = base.ravel(order="F").copy()
    is_t = np.zeros(flat.size, bool)
    is_t[targets] = True
    A = sp.identity(targets.size, format="csr") - M[:, targets]
    vals = spla.spsolve(A.tocsc(), M[:, ~is_t] @ flat[~is_t] + const)
    got = q.ravel(order="F")[targets]
    assert np.max(np.abs(got - vals)) < 1e-10 * max(1.0, np.max(np.abs(vals)))


def idw_workload(geo, w, limit=40):
    """w with the host boxes of some dead cells moved onto boxes (holding
    the dead cell) whose stencil has a pure solid cell and at least two fluid
    or dead cells besides the dead cell itself, and the image point inside
    that box: their values then take the inverse-distance branch."""
    g = {k: np.array(v, copy=True) for k, v in w.gcm.items()}
    xc, yc, zc = (geo.metrics(t)[0] for t in w.grid)
    nx, ny, nz = (t.size - 2 for t in w.grid)
    changed = []
    for n in range(g["iDead"].size):
        if len(changed) >= limit:
            break
        i, j, k = int(g["iDead"][n]), int(g["jDead"][n]), int(g["kDead"][n])
        if not (1 <= j <= ny - 1 and 1 <= k <= nz - 1):
            continue
        for di in (-1, 0):
            done = False
            for dj in (-1, 0):
                for dk in (-1, 0):
                    b = np.array([i + di, j + dj, k + dk])
                    if b.min() < 0 or b[0] > nx or b[1] > ny or b[2] > nz:
                        continue
                    st = b[None, :] + INC
                    ib = w.iblank[st[:, 0], st[:, 1], st[:, 2]]
                    de = w.dead[st[:, 0], st[:, 1], st[:, 2]]
                    self_ = (st[:, 0] == i) & (st[:, 1] == j) & (st[:, 2] == k)
                    pure = (ib == 1) & (de == 0)
                    valid = ((ib == 0) | (de == -1)) & ~self_
                    if pure.any() and valid.sum() >= 2:
                        g["iDeadCellIndex"][n], g["jDeadCellIndex"][n], g["kDeadCellIndex"][n] = b
                        t = np.array([0.37, 0.61, 0.44])
                        g["xyzDCIP"][n] = [xc[b[0]] + t[0] * (xc[b[0] + 1] - xc[b[0]]),
                                           yc[b[1]] + t[1] * (yc[b[1] + 1] - yc[b[1]]),
                                           zc[b[2]] + t[2] * (zc[b[2] + 1] - zc[b[2]])]
                        changed.append(n)
                        done = True
                        break
                if done:
                    break
            if done:
                break
    return g, changed


def test_dead_cell_inverse_distance(capi, oracle, geo):
    w = geo.config_c2(24)
    g, changed = idw_workload(geo, w)
    assert len(changed) >= 10
    xc, yc, zc = (geo.metrics(t)[0] for t in w.grid)
    p = capi.default_params(itermax_gc=1, restol_gcp=0.0)
    o = oracle.Oracle(capi, *w.grid, p)
    o.set_masks(w.iblank, w.ghost, w.dead)
    o.set_gcm(g)
    rng = np.random.default_rng(8)
    # a constant previous pressure is reproduced
    for pprev, check in ((np.full(w.div.shape, 2.5, order="F"), "const"),
                         (np.asfortranarray(rng.standard_normal(w.div.shape)), "avg")):
        q = np.asfortranarray(rng.standard_normal(w.div.shape))
        o.ghostcell_pressure(q, w.div, pprev)
        for n in changed:
            i, j, k = int(g["iDead"][n]), int(g["jDead"][n]), int(g["kDead"][n])
            b = np.array([g["iDeadCellIndex"][n], g["jDeadCellIndex"][n], g["kDeadCellIndex"][n]])
            st = b[None, :] + INC
            ib = w.iblank[st[:, 0], st[:, 1], st[:, 2]]
            de = w.dead[st[:, 0], st[:, 1], st[:, 2]]
            self_ = (st[:, 0] == i) & (st[:, 1] == j) & (st[:, 2] == k)
            valid = ((ib == 0) | (de == -1)) & ~self_
            ip = g["xyzDCIP"][n]
            d2 = (xc[st[:, 0]] - ip[0]) ** 2 + (yc[st[:, 1]] - ip[1]) ** 2 + (zc[st[:, 2]] - ip[2]) ** 2
            wgt = np.where(valid, 1.0 / d2, 0.0)
            expect = np.sum(wgt * pprev[st[:, 0], st[:, 1], st[:, 2]]) / np.sum(wgt)
            assert abs(q[i, j, k] - expect) <= 1e-13 * max(1.0, abs(expect)), (check, n)
    o.close()
