"""Triangulated body surfaces in the reference's unstructured-surface form:
marker coordinates xBodyMarker(1:nPts), yBodyMarker, zBodyMarker and the
connectivity triElemNeig(1:3, 1:nElem) with 1-based marker numbers
(AMODULES.F90:343-392), oriented as the reference leaves them after
determine_norm_dir_unstruc (BOUNDARY_SET_ARCLENGTH_NORM.F90:204-328): the
element normal (v2 - v1) x (v3 - v2) points into the body.

Workload generators for the geometry refresh: a sphere and a rotated
ellipsoid (closed solid bodies), a seeded, randomly perturbed square plate
(a membrane, the C3 / C5 body), and the plate's pitching motion
theta(t) = 0.3 sin(2 pi t) about z (the C5 prescribed motion)."""
import math

import numpy as np

from .geometry import splitmix64_uniform

SOLID_BODY = 1
MEMBRANE = 2


class Surface:
    """One body: markers (npts,) x3 and tri (3, nelem) int32, 1-based."""

    def __init__(self, x, y, z, tri, kind=SOLID_BODY, dim=3):
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.y = np.ascontiguousarray(y, dtype=np.float64)
        self.z = np.ascontiguousarray(z, dtype=np.float64)
        self.tri = np.asfortranarray(tri, dtype=np.int32)
        self.kind = kind
        self.dim = dim

    @property
    def npts(self):
        return self.x.size

    @property
    def nelem(self):
        return self.tri.shape[1]

    def moved(self, x, y, z):
        return Surface(x, y, z, self.tri, self.kind, self.dim)


def _icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v / np.linalg.norm(v, axis=1)[:, None], f


def unit_sphere(level):
    """Icosphere of 20*4^level triangles (level 5: 20480) on the unit sphere,
    0-based faces oriented outward."""
    v, f = _icosahedron()
    for _ in range(level):
        cache = {}
        verts = list(v)

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        v = np.array(verts)
        f = np.array(nf, dtype=np.int64)
    return v, f


def _inward(v, f, inside_point):
    """Orient faces so that (v2 - v1) x (v3 - v2) points toward inside_point."""
    p1, p2, p3 = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    n = np.cross(p2 - p1, p3 - p2)
    c = (p1 + p2 + p3) / 3.0
    flip = np.sum(n * (np.asarray(inside_point) - c), axis=1) < 0.0
    f = f.copy()
    f[flip, 1], f[flip, 2] = f[flip, 2].copy(), f[flip, 1].copy()
    return f


def sphere(center, radius, level=4):
    v, f = unit_sphere(level)
    p = np.asarray(center, dtype=np.float64) + radius * v
    f = _inward(p, f, center)
    return Surface(p[:, 0], p[:, 1], p[:, 2], (f + 1).T)


def ellipsoid(center, axes, angz_deg=0.0, level=4):
    """Semi-axes (a, b, c) rotated by angz about z (BOUNDARY_MARKER_INIT.F90:
    135-146)."""
    v, f = unit_sphere(level)
    t = math.radians(angz_deg)
    R = np.array([[math.cos(t), -math.sin(t), 0.0], [math.sin(t), math.cos(t), 0.0], [0.0, 0.0, 1.0]])
    p = (v * np.asarray(axes, dtype=np.float64)) @ R.T + np.asarray(center, dtype=np.float64)
    f = _inward(p, f, center)
    return Surface(p[:, 0], p[:, 1], p[:, 2], (f + 1).T)


def plate(center, side=0.4, n=101, amplitude=0.02, seed=3, angz_deg=0.0, smooth=False):
    """Square membrane of side `side` in its local x-z plane (normal along y),
    n x n markers, 2 (n-1)^2 triangles (n = 101: 20000), moved along the
    normal by amplitude * U(-1, 1) from splitmix64(seed): independently per
    marker, or (smooth) as the cosine-weighted interpolation of such values
    on a 6 x 6 lattice over the plate; then rotated by angz about z around
    the centre."""
    s = np.linspace(-0.5 * side, 0.5 * side, n)
    X, Z = np.meshgrid(s, s, indexing="ij")
    if smooth:
        L = 5
        v = splitmix64_uniform(seed, (L + 1) ** 2).reshape(L + 1, L + 1)
        u = (s - s[0]) / side * L
        i0 = np.minimum(u.astype(int), L - 1)
        w = 0.5 - 0.5 * np.cos(np.pi * (u - i0))
        a, b = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        ia, ib, wa, wb = i0[a], i0[b], w[a], w[b]
        Y = amplitude * ((1 - wa) * (1 - wb) * v[ia, ib] + wa * (1 - wb) * v[ia + 1, ib]
                         + (1 - wa) * wb * v[ia, ib + 1] + wa * wb * v[ia + 1, ib + 1])
    else:
        Y = amplitude * splitmix64_uniform(seed, n * n).reshape(n, n)
    p = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    f = []
    idx = np.arange(n * n).reshape(n, n)
    a, b = idx[:-1, :-1].ravel(), idx[1:, :-1].ravel()
    c, d = idx[1:, 1:].ravel(), idx[:-1, 1:].ravel()
    f = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)])
    return _placed(p, f, center, angz_deg)


def _placed(p, f, center, angz_deg):
    t = math.radians(angz_deg)
    R = np.array([[math.cos(t), -math.sin(t), 0.0], [math.sin(t), math.cos(t), 0.0], [0.0, 0.0, 1.0]])
    q = p @ R.T + np.asarray(center, dtype=np.float64)
    return Surface(q[:, 0], q[:, 1], q[:, 2], (f + 1).T, kind=MEMBRANE)


def pitch(surface, center, angle_rad):
    """Rigid rotation of the markers by angle about the z axis through
    center (the plate's prescribed pitching, BOUNDARY_MARKER_VEL.F90:5-37)."""
    c, s = math.cos(angle_rad), math.sin(angle_rad)
    x0, y0 = surface.x - center[0], surface.y - center[1]
    return surface.moved(center[0] + c * x0 - s * y0, center[1] + s * x0 + c * y0, surface.z.copy())


def pitch_angle(t):
    """theta(t) = 0.3 sin(2 pi t)."""
    return 0.3 * math.sin(2.0 * math.pi * t)


def membrane_thickness(x, y, z, factor=1.0):
    """membrane_tkns = factor * sqrt(dxMin^2 + dyMin^2 + dzMin^2) over the
    node spacings (BOUNDARY_MARKER_INIT.F90:428)."""
    d = [float(np.min(np.diff(a[1:-1]))) for a in (x, y, z)]
    return factor * math.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2)
