"""Synthetic workloads for the pressure-Poisson path: grids, immersed bodies,
iblank / ghostCellMark / dead_cell masks, GCM ghost- and dead-cell lists with
their interpolation coefficients, and seeded divergence right-hand sides.

This is the host-side producer of the hot path's inputs (the reference makes
them in BOUNDARY_*.F90 / GCM_SET_BOUNDARY.F90 from triangulated surfaces);
here bodies are analytic spheres and ellipsoids.  The list construction
follows the reference's rules:

* ghost cells: fluid cells with a face neighbour that is iblank, interior
  only, listed k-outer / j / i-inner (GCM_SET_BOUNDARY.F90:800-876, 236-276);
* dead cells: iblank face neighbours of interior ghost cells
  (GCM_SET_BOUNDARY.F90:946-1048);
* image points IP = BI + probeLengthNormalized*(GC - BI) for ghost cells and
  IP = DC + probeLengthNormalizedD*(BI - DC) for dead cells, host cells found
  with xc(i) <= xIP < xc(i+1) (ghost) and xc(i) <= xIP <= xc(i+1) (dead),
  last match wins (GCM_SET_BOUNDARY.F90:326-390, 1268-1359);
* coefficients from the 8x8 Neumann Vandermonde matrix of
  GCM_Calc_VanMatrixDN (GCM_SET_BOUNDARY.F90:607-797), ghost rows evaluated
  at the ghost-cell centre (the reference passes xGC as the evaluation point,
  :404-406), dead rows at the image point.  The inverse is taken in
  coordinates relative to the first stencil cell, which spans the same
  bilinear-trilinear space as the reference's absolute coordinates;
* alphaGhost = 1 - sqrt(sum (ds/dxc)^2) over adjacent dead cells
  (GCM_SET_BOUNDARY.F90:1155-1213, 1483-1489).

smooth_iblank (GCM_SET_BOUNDARY.F90:1618-1697) is not applied.
"""
import math

import numpy as np

# GCM_MEMORY_ALLOCATE.F90:593-623
INC_I = np.array([0, 0, 1, 1, 0, 0, 1, 1])
INC_J = np.array([0, 1, 1, 0, 0, 1, 1, 0])
INC_K = np.array([0, 0, 0, 0, 1, 1, 1, 1])


# ----------------------------------------------------------------------------
# deterministic generator (identical sequence in C and numpy)
def splitmix64_uniform(seed, n, start=0):
    """n values in [-1, 1): splitmix64 state seed + (i+1)*golden, output
    (z >> 11) * 2^-53 * 2 - 1."""
    with np.errstate(over="ignore"):
        i = np.arange(start + 1, start + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return 2.0 * u - 1.0


# ----------------------------------------------------------------------------
# grids
def _ghost_nodes(xi):
    """x(1..n) -> x(0..n+1) with make_grid's mirrored ghost nodes
    (CORE_GRID.F90:29-31)."""
    n = xi.size
    x = np.empty(n + 2)
    x[1:n + 1] = xi
    x[0] = xi[0] * 2.0 - xi[1]
    x[n + 1] = xi[-1] + xi[-1] - xi[-2]
    return x


def uniform_axis(nodes, lo=0.0, hi=1.0):
    """make_grid UNIFORM_GRID branch (CORE_GRID.F90:12-17)."""
    delta = (hi - lo) / float(nodes - 1)
    xi = lo + np.arange(nodes, dtype=np.float64) * delta
    return _ghost_nodes(xi)


def tanh_axis(nodes, centers, hmin, width=0.06, lo=0.0, hi=1.0):
    """Stretched nodes x(1..n) on [lo,hi] whose node density is
    1 + A*sum_c sech^2((x - c)/width); A is chosen so that the smallest
    spacing equals hmin.  The cumulative density is closed-form (tanh), so
    node i sits where the normalised cumulative equals (i-1)/(n-1)."""
    centers = np.asarray(centers, dtype=np.float64)

    def F(x, A):
        return (x - lo) + A * width * np.sum(np.tanh((x[..., None] - centers) / width)
                                           - np.tanh((lo - centers) / width), axis=-1)

    def nodes_for(A):
        xi = np.linspace(lo, hi, nodes)
        target = (xi - lo) / (hi - lo) * F(np.array(hi), A)
        a = np.full(nodes, lo)
        b = np.full(nodes, hi)
        for _ in range(80):
            m = 0.5 * (a + b)
            f = F(m, A) - target
            a = np.where(f < 0, m, a)
            b = np.where(f < 0, b, m)
        x = 0.5 * (a + b)
        x[0], x[-1] = lo, hi
        return x

    lo_a, hi_a = 0.0, 1.0e4
    if np.min(np.diff(nodes_for(0.0))) <= hmin:
        return _ghost_nodes(nodes_for(0.0))
    for _ in range(100):
        A = 0.5 * (lo_a + hi_a)
        h = np.min(np.diff(nodes_for(A)))
        if h > hmin:
            lo_a = A
        else:
            hi_a = A
    return _ghost_nodes(nodes_for(hi_a))


def metrics(x):
    """xc, dx, dxinv, dxc, dxcinv of CORE_GRID.F90:106-150 (index 0..n+1)."""
    n = x.size - 2
    xc = np.zeros(n + 2)
    dx = np.zeros(n + 2)
    dxc = np.zeros(n + 2)
    xc[0:n + 1] = 0.5 * (x[1:n + 2] + x[0:n + 1])
    dx[0:n + 1] = x[1:n + 2] - x[0:n + 1]
    xc[n + 1] = x[n + 1]
    dxc[1:n + 1] = xc[1:n + 1] - xc[0:n]
    dxinv = np.zeros(n + 2)
    dxinv[0:n + 1] = 1.0 / dx[0:n + 1]
    dxcinv = np.zeros(n + 2)
    dxcinv[1:n + 1] = 1.0 / dxc[1:n + 1]
    return xc, dx, dxinv, dxc, dxcinv


# ----------------------------------------------------------------------------
# bodies
class Sphere:
    def __init__(self, center, radius):
        self.c = np.asarray(center, dtype=np.float64)
        self.r = float(radius)

    def inside(self, X, Y, Z):
        return (X - self.c[0]) ** 2 + (Y - self.c[1]) ** 2 + (Z - self.c[2]) ** 2 < self.r ** 2

    def closest(self, P):
        d = P - self.c
        nrm = np.sqrt(np.sum(d * d, axis=1))
        nrm = np.where(nrm == 0.0, 1e-300, nrm)
        return self.c + d * (self.r / nrm)[:, None]

    def distance(self, P):
        return np.abs(np.sqrt(np.sum((P - self.c) ** 2, axis=1)) - self.r)


class Ellipsoid:
    """Semi-axes (a,b,c), rotated by angz degrees about z (the only rotation
    the reference's ellipsoid input supports, BOUNDARY_MARKER_INIT.F90:135)."""

    def __init__(self, center, axes, angz_deg=0.0):
        self.c = np.asarray(center, dtype=np.float64)
        self.ax = np.asarray(axes, dtype=np.float64)
        t = math.radians(angz_deg)
        self.R = np.array([[math.cos(t), -math.sin(t), 0.0], [math.sin(t), math.cos(t), 0.0], [0.0, 0.0, 1.0]])

    def _local(self, P):
        return (P - self.c) @ self.R          # row vector times R = R^T p

    def inside(self, X, Y, Z):
        P = np.stack([X - self.c[0], Y - self.c[1], Z - self.c[2]], axis=-1)
        L = P @ self.R
        return np.sum((L / self.ax) ** 2, axis=-1) < 1.0

    def closest(self, P):
        L = self._local(P)
        a2 = self.ax ** 2
        g = np.sum((L / self.ax) ** 2, axis=1)
        # root t of sum(a_i^2 u_i^2 / (a_i^2 + t)^2) = 1 on (-min a^2, inf)
        lo = np.where(g >= 1.0, 0.0, -np.min(a2) * (1.0 - 1e-15))
        hi = np.where(g >= 1.0, np.max(a2) * (np.sqrt(g) + 1.0), 0.0)
        for _ in range(200):
            t = 0.5 * (lo + hi)
            f = np.sum(a2 * L ** 2 / (a2 + t[:, None]) ** 2, axis=1) - 1.0
            lo = np.where(f > 0, t, lo)
            hi = np.where(f > 0, hi, t)
        t = 0.5 * (lo + hi)
        Q = a2 * L / (a2 + t[:, None])
        return Q @ self.R.T + self.c

    def distance(self, P):
        return np.sqrt(np.sum((self.closest(P) - P) ** 2, axis=1))


# ----------------------------------------------------------------------------
def _vandermonde_rows(X, Y, Z):
    """basis [xyz, xy, xz, yz, x, y, z, 1] (GCM_SET_BOUNDARY.F90:640-660)."""
    one = np.ones_like(X)
    return np.stack([X * Y * Z, X * Y, X * Z, Y * Z, X, Y, Z, one], axis=-1)


def _neumann_row(X, Y, Z, N1, N2, N3):
    """normal-derivative row at the body intercept (GCM_SET_BOUNDARY.F90:
    682-689)."""
    zero = np.zeros_like(X)
    return np.stack([N1 * Y * Z + N2 * X * Z + N3 * X * Y, N1 * Y + N2 * X, N1 * Z + N3 * X,
                     N2 * Z + N3 * Y, N1, N2, N3, zero], axis=-1)


def _coefficients(hostI, hostJ, hostK, self_ijk, BI, NRM, EVAL, xc, yc, zc, dirichlet=False):
    """coeffGCMN(1:8) of GCM_Calc_VanMatrixDN for a batch of points; with
    dirichlet, coeffGCMD(1:8): the self row is the basis at the body
    intercept instead of the normal derivative there."""
    m = hostI.size
    if m == 0:
        return np.zeros((0, 8))
    ii = hostI[:, None] + INC_I[None, :]
    jj = hostJ[:, None] + INC_J[None, :]
    kk = hostK[:, None] + INC_K[None, :]
    ox, oy, oz = xc[hostI], yc[hostJ], zc[hostK]
    X = xc[ii] - ox[:, None]
    Y = yc[jj] - oy[:, None]
    Z = zc[kk] - oz[:, None]
    V = _vandermonde_rows(X, Y, Z)                                       # (m,8,8)
    is_self = (ii == self_ijk[:, 0:1]) & (jj == self_ijk[:, 1:2]) & (kk == self_ijk[:, 2:3])
    bx = (BI[:, 0] - ox)[:, None] * np.ones((1, 8))
    by = (BI[:, 1] - oy)[:, None] * np.ones((1, 8))
    bz = (BI[:, 2] - oz)[:, None] * np.ones((1, 8))
    n1 = NRM[:, 0:1] * np.ones((1, 8))
    n2 = NRM[:, 1:2] * np.ones((1, 8))
    n3 = NRM[:, 2:3] * np.ones((1, 8))
    VN = _vandermonde_rows(bx, by, bz) if dirichlet else _neumann_row(bx, by, bz, n1, n2, n3)
    V = np.where(is_self[..., None], VN, V)
    Vinv = np.linalg.inv(V)
    e = _vandermonde_rows((EVAL[:, 0] - ox), (EVAL[:, 1] - oy), (EVAL[:, 2] - oz))  # (m,8)
    # coeff(iRow) = sum_m Vinv(m, iRow) * basis_m(eval)
    return np.einsum("nm,nmr->nr", e, Vinv)


def _host_index(xc, xip, lo, hi):
    """last i in [lo,hi] with xc(i) <= xIP < xc(i+1) (or <= for dead cells;
    for strictly increasing xc both give the same index)."""
    i = np.searchsorted(xc, xip, side="right") - 1
    ok = (i >= lo) & (i <= hi)
    return np.where(ok, i, -1)


class Workload:
    """Everything solve_poisson needs, in the reference's layouts."""

    def __init__(self):
        self.grid = None
        self.iblank = self.ghost = self.dead = None
        self.gcm = None
        self.div = None
        self.info = {}

    @property
    def shape(self):
        return self.iblank.shape

    def nodes(self):
        return tuple(a.size - 2 for a in self.grid)


def build_geometry(x, y, z, bodies, probe=2.0, probe_dead=2.0):
    """iblank / ghostCellMark / dead_cell masks and GCM lists for analytic
    bodies on the grid with node coordinates x(0:nx+1), y, z."""
    nx, ny, nz = x.size - 2, y.size - 2, z.size - 2
    xc, dx, dxinv, dxc, dxcinv = metrics(x)
    yc, dy, dyinv, dyc, dycinv = metrics(y)
    zc, dz, dzinv, dzc, dzcinv = metrics(z)
    shp = (nx + 2, ny + 2, nz + 2)
    iblank = np.zeros(shp, dtype=np.int8, order="F")
    body_of = np.zeros(shp, dtype=np.int16, order="F")
    X, Y, Z = np.meshgrid(xc[1:nx], yc[1:ny], zc[1:nz], indexing="ij")
    for b, body in enumerate(bodies):
        m = body.inside(X, Y, Z)
        iblank[1:nx, 1:ny, 1:nz][m] = 1
        body_of[1:nx, 1:ny, 1:nz][m] = b + 1
    del X, Y, Z
    # GCM_define_ghostcells (GCM_SET_BOUNDARY.F90:800-876)
    ghost = np.zeros(shp, dtype=np.int8, order="F")
    fl = (iblank == 0)
    ib = (iblank == 1)
    sub = (slice(1, nx), slice(1, ny), slice(1, nz))
    g = np.zeros((nx - 1, ny - 1, nz - 1), dtype=bool)
    for ax in range(3):
        for d in (-1, 1):
            sl = list(sub)
            s = sl[ax]
            sl[ax] = slice(s.start + d, s.stop + d)
            g |= ib[tuple(sl)]
    ghost[sub] = (g & fl[sub]).astype(np.int8)
    # GCM_set_deadcells dead marks (:946-961)
    dead = np.zeros(shp, dtype=np.int8, order="F")
    gi = ghost[sub] == 1
    for ax in range(3):
        for d in (-1, 1):
            sl = list(sub)
            s = sl[ax]
            sl[ax] = slice(s.start + d, s.stop + d)
            tgt = tuple(sl)
            mark = gi & ib[tgt]
            dview = dead[tgt]
            dview[mark] = -1
    # ghost list in Fortran order (k outer, j, i inner)
    gF = np.nonzero(ghost.ravel(order="F"))[0]
    iG, jG, kG = np.unravel_index(gF, shp, order="F")
    dF = np.nonzero((np.abs(dead) == 1).ravel(order="F"))[0]
    iD, jD, kD = np.unravel_index(dF, shp, order="F")
    # keep dead cells of i in 1..nx-1, j in 0..ny, k in 0..nz (:1035-1046)
    keep = (iD >= 1) & (iD <= nx - 1) & (jD <= ny) & (kD <= nz)
    iD, jD, kD = iD[keep], jD[keep], kD[keep]

    def nearest_surface(P):
        best = None
        bestd = None
        for body in bodies:
            q = body.closest(P)
            d = np.sqrt(np.sum((q - P) ** 2, axis=1))
            if best is None:
                best, bestd = q, d
            else:
                sel = d < bestd
                best = np.where(sel[:, None], q, best)
                bestd = np.where(sel, d, bestd)
        return best, bestd

    # ghost cells: body intercept, image point, host cell, coefficients
    GC = np.stack([xc[iG], yc[jG], zc[kG]], axis=1)
    BI, ds = nearest_surface(GC)
    ds = np.where(ds == 0.0, 1e-300, ds)
    slope = (GC - BI) / ds[:, None]
    nrmG = -slope
    IP = BI - (BI - GC) * probe
    maxd = np.maximum(np.maximum(dx[iG], dy[jG]), dz[kG])
    iR = 2 * np.rint(maxd / dx[iG] + 0.5).astype(int)
    jR = 2 * np.rint(maxd / dy[jG] + 0.5).astype(int)
    kR = 2 * np.rint(maxd / dz[kG] + 0.5).astype(int)
    iC = _host_index(xc, IP[:, 0], np.maximum(iG - iR, 0), np.minimum(iG + iR, nx))
    jC = _host_index(yc, IP[:, 1], np.maximum(jG - jR, -1), np.minimum(jG + jR, ny))
    kC = _host_index(zc, IP[:, 2], np.maximum(kG - kR, -1), np.minimum(kG + kR, nz))
    if np.any(iC < 0) or np.any(jC < 0) or np.any(kC < 0):
        raise ValueError("failed to find the hosting box for an image point of a ghost node")
    cG = _coefficients(iC, jC, kC, np.stack([iG, jG, kG], 1), BI, nrmG, GC, xc, yc, zc)
    cGD = _coefficients(iC, jC, kC, np.stack([iG, jG, kG], 1), BI, nrmG, GC, xc, yc, zc, dirichlet=True)
    # dead cells
    DC = np.stack([xc[iD], yc[jD], zc[kD]], axis=1)
    BId, dsd = nearest_surface(DC)
    dsd = np.where(dsd == 0.0, 1e-300, dsd)
    slopeD = (BId - DC) / dsd[:, None]
    nrmD = -slopeD
    IPd = DC + (BId - DC) * probe_dead
    plD = dsd * probe_dead
    maxd = np.maximum(np.maximum(dx[iD], dy[jD]), dz[kD])
    iR = 2 * np.rint(maxd / dx[iD] + 0.5).astype(int)
    jR = 2 * np.rint(maxd / dy[jD] + 0.5).astype(int)
    kR = 2 * np.rint(maxd / dz[kD] + 0.5).astype(int)
    iDC = _host_index(xc, IPd[:, 0], np.maximum(iD - iR, 0), np.minimum(iD + iR, nx))
    jDC = _host_index(yc, IPd[:, 1], np.maximum(jD - jR, -1), np.minimum(jD + jR, ny))
    kDC = _host_index(zc, IPd[:, 2], np.maximum(kD - kR, -1), np.minimum(kD + kR, nz))
    if np.any(iDC < 0) or np.any(jDC < 0) or np.any(kDC < 0):
        raise ValueError("failed to find the hosting box for the image point of a dead cell")
    cD = _coefficients(iDC, jDC, kDC, np.stack([iD, jD, kD], 1), BId, nrmD, IPd, xc, yc, zc)
    cDD = _coefficients(iDC, jDC, kDC, np.stack([iD, jD, kD], 1), BId, nrmD, IPd, xc, yc, zc, dirichlet=True)
    # alphaGhost (:1155-1213, 1483-1489)
    gidx = -np.ones(shp, dtype=np.int64, order="F")
    gidx[iG, jG, kG] = np.arange(iG.size)
    acc = np.zeros(iG.size)
    inner = (kD >= 1) & (kD <= nz - 1)
    for (di, dj, dk, dcv) in ((1, 0, 0, dxc[np.minimum(iD + 1, nx + 1)]), (-1, 0, 0, dxc[iD]),
                              (0, 1, 0, dyc[np.minimum(jD + 1, ny + 1)]), (0, -1, 0, dyc[jD]),
                              (0, 0, 1, dzc[np.minimum(kD + 1, nz + 1)]), (0, 0, -1, dzc[kD])):
        ii, jj, kk = iD + di, jD + dj, kD + dk
        ok = inner & (ii >= 0) & (ii <= nx + 1) & (jj >= 0) & (jj <= ny + 1) & (kk >= 0) & (kk <= nz + 1)
        ok &= (kk >= 1) & (kk <= nz - 1)
        n = np.where(ok, gidx[np.clip(ii, 0, nx + 1), np.clip(jj, 0, ny + 1), np.clip(kk, 0, nz + 1)], -1)
        sel = n >= 0
        np.add.at(acc, n[sel], (dsd[sel] / dcv[sel]) ** 2)
    alpha = 1.0 - np.sqrt(acc)
    gcm = dict(
        iGhost=(iG).astype(np.int32), jGhost=jG.astype(np.int32), kGhost=kG.astype(np.int32),
        iCellIndex=iC.astype(np.int32), jCellIndex=jC.astype(np.int32), kCellIndex=kC.astype(np.int32),
        coeffGCMN=np.ascontiguousarray(cG), alphaGhost=alpha, dpdnBodyIntercept=np.zeros(iG.size),
        iDead=iD.astype(np.int32), jDead=jD.astype(np.int32), kDead=kD.astype(np.int32),
        iDeadCellIndex=iDC.astype(np.int32), jDeadCellIndex=jDC.astype(np.int32),
        kDeadCellIndex=kDC.astype(np.int32), coeffGCMDeadN=np.ascontiguousarray(cD),
        probeLengthDead=plD, dpdnInterceptDead=np.zeros(iD.size),
        xyzDCIP=np.ascontiguousarray(IPd),
        # the velocity relations of GCM_ghostcell_velocity: Dirichlet
        # coefficients, body-intercept points, probeLengthNormalizedD
        coeffGCMD=np.ascontiguousarray(cGD), coeffGCMDeadD=np.ascontiguousarray(cDD),
        xyzBI=np.ascontiguousarray(BI), xyzBIDead=np.ascontiguousarray(BId), probeLengthNormalizedD=probe_dead)
    return iblank, ghost, dead, gcm


def divergence_rhs(x, y, z, iblank, ghost, seed, dtinv=1.0, remove_mean=True):
    """div of rhs_poisson (CORE_POISSON.F90:35-50) for face velocities drawn
    from splitmix64(seed): face_u(1:nx,1:ny-1,1:nz-1), then face_v
    (1:nx-1,1:ny,1:nz-1), then face_w(1:nx-1,1:ny-1,1:nz), Fortran order.
    With remove_mean the flag-cell volume-weighted mean is subtracted on flag
    cells."""
    nx, ny, nz = x.size - 2, y.size - 2, z.size - 2
    xc, dx, dxinv, _, _ = metrics(x)
    yc, dy, dyinv, _, _ = metrics(y)
    zc, dz, dzinv, _, _ = metrics(z)
    nu = nx * (ny - 1) * (nz - 1)
    nv = (nx - 1) * ny * (nz - 1)
    nw = (nx - 1) * (ny - 1) * nz
    r = splitmix64_uniform(seed, nu + nv + nw)
    fu = r[:nu].reshape((nx, ny - 1, nz - 1), order="F")
    fv = r[nu:nu + nv].reshape((nx - 1, ny, nz - 1), order="F")
    fw = r[nu + nv:].reshape((nx - 1, ny - 1, nz), order="F")
    d = dtinv * ((fu[1:nx, :, :] - fu[0:nx - 1, :, :]) * dxinv[1:nx, None, None]
                 + (fv[:, 1:ny, :] - fv[:, 0:ny - 1, :]) * dyinv[None, 1:ny, None]
                 + (fw[:, :, 1:nz] - fw[:, :, 0:nz - 1]) * dzinv[None, None, 1:nz])
    flag = (1 - iblank[1:nx, 1:ny, 1:nz].astype(np.float64)) * (1 - ghost[1:nx, 1:ny, 1:nz].astype(np.float64))
    d = d * flag
    if remove_mean:
        vol = dx[1:nx, None, None] * dy[None, 1:ny, None] * dz[None, None, 1:nz]
        mean = np.sum(d * vol * flag) / np.sum(vol * flag)
        d = (d - mean) * flag
    div = np.zeros((nx + 2, ny + 2, nz + 2), order="F")
    div[1:nx, 1:ny, 1:nz] = d
    return div


def _smooth_noise(shape, seed, passes=4):
    """seeded uniform[-1,1) noise (Fortran order over the interior) filtered
    by `passes` separable 1-2-1 passes; a smooth field with all scales."""
    n = int(np.prod(shape))
    f = splitmix64_uniform(seed, n).reshape(shape, order="F")
    for _ in range(passes):
        for ax in range(3):
            g = f.copy()
            sl_c = [slice(None)] * 3
            sl_m = [slice(None)] * 3
            sl_p = [slice(None)] * 3
            sl_c[ax] = slice(1, -1)
            sl_m[ax] = slice(0, -2)
            sl_p[ax] = slice(2, None)
            g[tuple(sl_c)] = 0.25 * f[tuple(sl_m)] + 0.5 * f[tuple(sl_c)] + 0.25 * f[tuple(sl_p)]
            f = g
    return f


def gcm_operator(shape, iblank, dead, gcm, ny_inner, nz_inner):
    """Sparse linear relations of the GCM update at its fixed point
    (GCM_GHOSTCELL_UPDATE.F90:425-643 with dpdn = 0 and mix_GC_form = 0):
    rows p(cell) - sum_r coeff_r p(stencil_r) for every updated dead cell and
    every ghost cell, the self row dropped (it multiplies dpdn).  Dead cells
    whose stencil holds a pure solid cell take the inverse-distance branch,
    which reads the previous-step pressure instead; they are returned in
    idw_dead so the caller can pin them."""
    import scipy.sparse as sp
    nx2, ny2, nz2 = shape
    lin = lambda i, j, k: i + nx2 * (j + ny2 * k)
    rows, cols, vals = [], [], []
    targets = []
    idw_dead = []
    g = gcm
    for kind in ("dead", "ghost"):
        if kind == "dead":
            I, J, K = g["iDead"], g["jDead"], g["kDead"]
            CI, CJ, CK = g["iDeadCellIndex"], g["jDeadCellIndex"], g["kDeadCellIndex"]
            CO = g["coeffGCMDeadN"]
            keep = (J >= 1) & (J <= ny_inner) & (K >= 1) & (K <= nz_inner)
        else:
            I, J, K = g["iGhost"], g["jGhost"], g["kGhost"]
            CI, CJ, CK = g["iCellIndex"], g["jCellIndex"], g["kCellIndex"]
            CO = g["coeffGCMN"]
            keep = np.ones(I.size, dtype=bool)
        for n in np.nonzero(keep)[0]:
            t = lin(I[n], J[n], K[n])
            si = CI[n] + INC_I
            sj = CJ[n] + INC_J
            sk = CK[n] + INC_K
            if kind == "dead":
                pure = (iblank[si, sj, sk] == 1) & (dead[si, sj, sk] == 0)
                if np.any(pure):
                    idw_dead.append(t)
                    continue
            targets.append(t)
            r = len(targets) - 1
            for q in range(8):
                c = lin(si[q], sj[q], sk[q])
                if c == t:
                    continue
                rows.append(r)
                cols.append(c)
                vals.append(CO[n, q])
    M = sp.csr_matrix((vals, (rows, cols)), shape=(len(targets), nx2 * ny2 * nz2))
    return np.array(targets, dtype=np.int64), M, np.array(idw_dead, dtype=np.int64)


def apply_operator(p, x, y, z):
    """A p on interior cells with the 7-point operator of residual_mg /
    calc_residual (CORE_POISSON.F90:490-533); ghost layers taken from p."""
    xc, dx, dxinv, dxc, dxcinv = metrics(x)
    yc, dy, dyinv, dyc, dycinv = metrics(y)
    zc, dz, dzinv, dzc, dzcinv = metrics(z)
    nx, ny, nz = x.size - 2, y.size - 2, z.size - 2
    bmx = (dxcinv[1:nx] * dxinv[1:nx])[:, None, None]
    bpx = (dxcinv[2:nx + 1] * dxinv[1:nx])[:, None, None]
    bmy = (dycinv[1:ny] * dyinv[1:ny])[None, :, None]
    bpy = (dycinv[2:ny + 1] * dyinv[1:ny])[None, :, None]
    bmz = (dzcinv[1:nz] * dzinv[1:nz])[None, None, :]
    bpz = (dzcinv[2:nz + 1] * dzinv[1:nz])[None, None, :]
    c = p[1:nx, 1:ny, 1:nz]
    out = (bmx * (p[0:nx - 1, 1:ny, 1:nz] - c) + bpx * (p[2:nx + 1, 1:ny, 1:nz] - c)
           + bmy * (p[1:nx, 0:ny - 1, 1:nz] - c) + bpy * (p[1:nx, 2:ny + 1, 1:nz] - c)
           + bmz * (p[1:nx, 1:ny, 0:nz - 1] - c) + bpz * (p[1:nx, 1:ny, 2:nz + 1] - c))
    return out


def manufactured_rhs(x, y, z, iblank, ghost, dead, gcm, seed, passes=4):
    """A right-hand side in the range of the coupled operator (flag-cell
    Poisson rows + GCM ghost/dead rows + homogeneous Neumann outer
    reflection), built from a seeded smooth pressure field p_m: p_m is drawn
    on flag cells, its ghost/dead values are solved from the GCM relations,
    outer ghost layers are reflected (p(0)=p(1), CORE_SET_OUTER_BC.F90:
    1088-1170 with pgrad=0), and div = (A p_m) on flag cells.  The converged
    solve then equals p_m up to a constant."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    nx, ny, nz = x.size - 2, y.size - 2, z.size - 2
    shp = (nx + 2, ny + 2, nz + 2)
    pm = np.zeros(shp, order="F")
    pm[1:nx, 1:ny, 1:nz] = _smooth_noise((nx - 1, ny - 1, nz - 1), seed, passes)
    flag = (iblank == 0) & (ghost == 0)
    pm[~flag] = 0.0
    pm[0, :, :] = 0.0
    targets, M, idw = gcm_operator(shp, iblank, dead, gcm, ny - 1, nz - 1)
    if targets.size:
        flat = pm.ravel(order="F")
        # stencils may reach the outer ghost layers, whose values are the
        # homogeneous-Neumann reflections p(0)=p(1), p(n)=p(n-1); those cells
        # join the unknowns with their reflection rows
        cols = np.unique(M.indices)
        ci, cj, ck = np.unravel_index(cols, shp, order="F")
        onp = ((ci == 0) | (ci == nx)).astype(int) + ((cj == 0) | (cj == ny)).astype(int) + ((ck == 0) | (ck == nz)).astype(int)
        if np.any(onp >= 2):
            raise ValueError("a GCM stencil reaches an edge ghost cell")
        gl = cols[onp == 1]
        gi, gj, gk = np.unravel_index(gl, shp, order="F")
        mi = np.where(gi == 0, 1, np.where(gi == nx, nx - 1, gi))
        mj = np.where(gj == 0, 1, np.where(gj == ny, ny - 1, gj))
        mk = np.where(gk == 0, 1, np.where(gk == nz, nz - 1, gk))
        mirror = mi + (nx + 2) * (mj + (ny + 2) * mk)
        R = sp.csr_matrix((np.ones(gl.size), (np.arange(gl.size), mirror)), shape=(gl.size, flat.size))
        unknown = np.concatenate([targets, gl])
        Mall = sp.vstack([M, R]).tocsr()
        is_u = np.zeros(flat.size, dtype=bool)
        is_u[unknown] = True
        A = sp.identity(unknown.size, format="csr") - Mall[:, unknown]
        rhs = Mall[:, ~is_u] @ flat[~is_u]
        import warnings
        with warnings.catch_warnings():
            warnings.simplefilter("error", spla.MatrixRankWarning)
            try:
                vals = spla.spsolve(A.tocsc(), rhs)
            except spla.MatrixRankWarning:
                raise ValueError("the GCM relations of this geometry are singular (no manufactured pressure)") from None
        if not np.all(np.isfinite(vals)):
            raise ValueError("the GCM relations of this geometry gave non-finite ghost values")
        flat[unknown] = vals
        pm = flat.reshape(shp, order="F")
    for (src, dst) in ((1, 0), (nx - 1, nx)):
        pm[dst, 1:ny, 1:nz] = pm[src, 1:ny, 1:nz]
    for (src, dst) in ((1, 0), (ny - 1, ny)):
        pm[0:nx + 1, dst, 1:nz] = pm[0:nx + 1, src, 1:nz]
    for (src, dst) in ((1, 0), (nz - 1, nz)):
        pm[0:nx + 1, 1:ny, dst] = pm[0:nx + 1, 1:ny, src]
    div = np.zeros(shp, order="F")
    div[1:nx, 1:ny, 1:nz] = apply_operator(pm, x, y, z) * flag[1:nx, 1:ny, 1:nz]
    return div, pm, idw


def face_velocities(nx, ny, nz, seed):
    """The face-velocity fields behind divergence_rhs, in full-field layout
    (0:nx+1, 0:ny+1, 0:nz+1) as the reference stores face_u/v/w."""
    nu = nx * (ny - 1) * (nz - 1)
    nv = (nx - 1) * ny * (nz - 1)
    nw = (nx - 1) * (ny - 1) * nz
    r = splitmix64_uniform(seed, nu + nv + nw)
    shp = (nx + 2, ny + 2, nz + 2)
    fu = np.zeros(shp, order="F")
    fv = np.zeros(shp, order="F")
    fw = np.zeros(shp, order="F")
    fu[1:nx + 1, 1:ny, 1:nz] = r[:nu].reshape((nx, ny - 1, nz - 1), order="F")
    fv[1:nx, 1:ny + 1, 1:nz] = r[nu:nu + nv].reshape((nx - 1, ny, nz - 1), order="F")
    fw[1:nx, 1:ny, 1:nz + 1] = r[nu + nv:].reshape((nx - 1, ny - 1, nz), order="F")
    return fu, fv, fw


# ----------------------------------------------------------------------------
# the named workloads
def config_c1(n=64, seed=1):
    """Unit cube, uniform (n+1)^3 nodes, sphere r=0.2 at the centre."""
    x = uniform_axis(n + 1)
    y = uniform_axis(n + 1)
    z = uniform_axis(n + 1)
    bodies = [Sphere((0.5, 0.5, 0.5), 0.2)]
    return _assemble(x, y, z, bodies, seed, "c1_uniform_sphere_%d" % n)


def config_c2(n=256, seed=2, hmin=1.0 / 1024.0, manufactured=True):
    """n^3 cells, tanh-stretched toward the bodies (clusters at x = 0.3, 0.7
    of width 0.12 and y = z = 0.5 of width 0.25, strongest stretching these
    widths allow; the smallest spacing is 1/540 at n = 256); sphere r=0.15 at
    (0.3,0.5,0.5) and ellipsoid 0.2/0.1/0.05 rotated 30 deg about z at
    (0.7,0.5,0.5).  Narrower clusters (needed for 1/1024) leave image-point
    host boxes on strongly anisotropic cells, where the GCM relations
    extrapolate with large weights and the ghost-cell iteration diverges."""
    x = tanh_axis(n + 1, [0.3, 0.7], hmin, width=0.12)
    y = tanh_axis(n + 1, [0.5], hmin, width=0.25)
    z = tanh_axis(n + 1, [0.5], hmin, width=0.25)
    bodies = [Sphere((0.3, 0.5, 0.5), 0.15), Ellipsoid((0.7, 0.5, 0.5), (0.2, 0.1, 0.05), 30.0)]
    return _assemble(x, y, z, bodies, seed, "c2_stretched_sphere_ellipsoid_%d" % n, manufactured)


def config_c3(nxc=512, nyc=384, nzc=384, seed=3, hmin=1.0 / 1024.0):
    """nxc x nyc x nzc cells on the unit box (the C3 shape, 512 x 384 x 384 at
    full size), tanh-stretched toward a thin plate at the centre: a 0.4-wide
    flat oblate ellipsoid (semi-axes 0.2 / 0.2 along x, z; half-thickness
    four of the smallest spacings along y) tilted 20 deg about z.  It stands
    in for the reference's triangulated membrane plate, whose iblank comes
    from the membrane-thickness rule of the mesh pipeline (out of scope
    here).  With x finer than y and z, x coarsens one level further than y
    and z, which are then held (semi-coarsening)."""
    x = tanh_axis(nxc + 1, [0.5], hmin, width=0.25)
    y = tanh_axis(nyc + 1, [0.5], hmin, width=0.25)
    z = tanh_axis(nzc + 1, [0.5], hmin, width=0.25)
    hy = float(np.min(np.diff(y[1:-1])))
    bodies = [Ellipsoid((0.5, 0.5, 0.5), (0.2, 4.0 * hy, 0.2), 20.0)]
    return _assemble(x, y, z, bodies, seed, "c3_stretched_plate_%dx%dx%d" % (nxc, nyc, nzc))


def config_request(nxc, nyc, nzc, seed):
    """Serving request: nxc x nyc x nzc cells on the unit box, one seeded
    random sphere or ellipsoid."""
    x = uniform_axis(nxc + 1)
    y = uniform_axis(nyc + 1)
    z = uniform_axis(nzc + 1)
    u = (splitmix64_uniform(seed + 7919, 8) + 1.0) * 0.5
    c = 0.4 + 0.2 * u[0:3]
    if u[3] < 0.5:
        bodies = [Sphere(c, 0.1 + 0.1 * u[4])]
    else:
        bodies = [Ellipsoid(c, (0.12 + 0.08 * u[4], 0.08 + 0.06 * u[5], 0.06 + 0.04 * u[6]), 90.0 * u[7])]
    return _assemble(x, y, z, bodies, seed, "request_%dx%dx%d_seed%d" % (nxc, nyc, nzc, seed))


def _assemble(x, y, z, bodies, seed, name, manufactured=True):
    """manufactured=False takes the seeded divergence right-hand side instead
    (no exact pressure; for geometries whose GCM relations are singular,
    where only the masks and lists are used)."""
    w = Workload()
    w.grid = (x, y, z)
    w.iblank, w.ghost, w.dead, w.gcm = build_geometry(x, y, z, bodies)
    if manufactured:
        w.div, w.p_exact, idw = manufactured_rhs(x, y, z, w.iblank, w.ghost, w.dead, w.gcm, seed)
    else:
        w.div = divergence_rhs(x, y, z, w.iblank, w.ghost, seed)
        w.p_exact, idw = None, np.zeros(0, dtype=np.int64)
    w.info = dict(name=name, seed=seed, nodes=[x.size - 2, y.size - 2, z.size - 2],
                  nGhost=int(w.gcm["iGhost"].size), nDead=int(w.gcm["iDead"].size),
                  nDeadIDW=int(idw.size), rhs="manufactured" if manufactured else "divergence",
                  hmin=[float(np.min(np.diff(a[1:-1]))) for a in (x, y, z)])
    return w


def config_test_mg(n=32, K=4):
    """test_mg's manufactured right-hand side (CORE_POISSON.F90:589-607) on
    an unmasked uniform unit cube with n cells per axis."""
    x = uniform_axis(n + 1)
    xc = metrics(x)[0]
    w = Workload()
    w.grid = (x, x.copy(), x.copy())
    shp = (n + 3, n + 3, n + 3)
    w.iblank = np.zeros(shp, dtype=np.int8, order="F")
    w.ghost = np.zeros(shp, dtype=np.int8, order="F")
    w.dead = np.zeros(shp, dtype=np.int8, order="F")
    w.gcm = empty_gcm()
    pi2 = 8.0 * math.atan(1.0)
    xx = (xc[1:n + 1] - 0.0) / 1.0 * pi2
    div = np.zeros(shp, order="F")
    acc = np.zeros((n, n, n))
    for k in range(1, K + 1):
        c = np.cos(xx * k)
        acc += 10.0 * c[:, None, None] * c[None, :, None] * c[None, None, :]
    div[1:n + 1, 1:n + 1, 1:n + 1] = acc
    w.div = div
    w.info = dict(name="test_mg_%d" % n, nodes=[n + 1] * 3, nGhost=0, nDead=0)
    return w


def test_mg_exact(n, K=4):
    """Exact discrete solution of test_mg's problem: each cos(2*pi*k*xc) is a
    discrete Neumann eigenvector of the 3-point Laplacian on the cell-centred
    uniform grid with reflected ghosts, eigenvalue -(4/h^2) sin^2(pi k/n)."""
    h = 1.0 / n
    xc = (np.arange(1, n + 1) - 0.5) * h
    p = np.zeros((n, n, n))
    for k in range(1, K + 1):
        c = np.cos(2.0 * math.pi * k * xc)
        lam = -(4.0 / h ** 2) * math.sin(math.pi * k / n) ** 2
        p += 10.0 * c[:, None, None] * c[None, :, None] * c[None, None, :] / (3.0 * lam)
    return p


def empty_gcm():
    z32 = np.zeros(0, dtype=np.int32)
    z64 = np.zeros(0)
    return dict(iGhost=z32, jGhost=z32, kGhost=z32, iCellIndex=z32, jCellIndex=z32, kCellIndex=z32,
                coeffGCMN=np.zeros((0, 8)), alphaGhost=z64, dpdnBodyIntercept=z64,
                iDead=z32, jDead=z32, kDead=z32, iDeadCellIndex=z32, jDeadCellIndex=z32,
                kDeadCellIndex=z32, coeffGCMDeadN=np.zeros((0, 8)), probeLengthDead=z64,
                dpdnInterceptDead=z64, xyzDCIP=np.zeros((0, 3)), coeffGCMD=np.zeros((0, 8)),
                coeffGCMDeadD=np.zeros((0, 8)), xyzBI=np.zeros((0, 3)), xyzBIDead=np.zeros((0, 3)),
                probeLengthNormalizedD=2.0)
