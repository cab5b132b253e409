"""Seeded geometry-refresh workloads shared by the oracle and GPU tests:
grids with triangulated bodies (closed solid bodies and a membrane plate)."""
import feicar3d_amd.geometry as geo
import feicar3d_amd.surface as surf

CASES = ["sphere", "two_bodies", "plate", "plate_stretched"]


def case(name):
    """x, y, z node arrays, the bodies and the membrane thickness."""
    if name == "sphere":
        n = 24
        x = y = z = geo.uniform_axis(n + 1)
        return x, y, z, [surf.sphere((0.5, 0.5, 0.5), 0.2, level=3)], 0.0
    if name == "two_bodies":
        n = 32
        x = geo.tanh_axis(n + 1, [0.3, 0.7], 1.0 / (4 * n), width=0.12)
        y = geo.tanh_axis(n + 1, [0.5], 1.0 / (4 * n), width=0.25)
        z = geo.uniform_axis(n + 1)
        bodies = [surf.sphere((0.3, 0.5, 0.5), 0.15, level=3),
                  surf.ellipsoid((0.7, 0.5, 0.5), (0.2, 0.1, 0.05), 30.0, level=3)]
        return x, y, z, bodies, 0.0
    if name == "plate":
        n = 32
        x = z = geo.uniform_axis(n + 1)
        y = geo.tanh_axis(n + 1, [0.5], 1.0 / (3 * n), width=0.3)
        p = surf.plate((0.5, 0.5, 0.5), side=0.4, n=21, amplitude=0.01, seed=3)
        return x, y, z, [surf.pitch(p, (0.5, 0.5, 0.5), 0.25)], surf.membrane_thickness(x, y, z, 1.0)
    if name == "plate_stretched":
        x = geo.tanh_axis(41, [0.5], 1.0 / 120, width=0.3)
        y = geo.tanh_axis(37, [0.5], 1.0 / 150, width=0.2)
        z = geo.tanh_axis(33, [0.5], 1.0 / 100, width=0.3)
        p = surf.plate((0.5, 0.5, 0.5), side=0.4, n=41, amplitude=0.02, seed=3)
        return x, y, z, [surf.pitch(p, (0.5, 0.5, 0.5), -0.3)], surf.membrane_thickness(x, y, z, 1.0)
    raise KeyError(name)


def stress_inputs(name, seed=5):
    """Marker velocities (current and previous step) and p, u, v, w fields
    for GCM_SetBodyInterceptValues and compute_marker_stress on a case:
    seeded splitmix64 values in (-1, 1) (0.1 of that for the markers)."""
    import numpy as np
    x, y, z, bodies, tk = case(name)
    sh = (x.size, y.size, z.size)
    n = sh[0] * sh[1] * sh[2]
    f = geo.splitmix64_uniform(seed, 4 * n)
    fields = [np.asfortranarray(f[q * n:(q + 1) * n].reshape(sh, order="F")) for q in range(4)]
    mv = []
    for b, s in enumerate(bodies):
        m = 0.1 * geo.splitmix64_uniform(seed + 100 + b, 6 * s.npts)
        mv.append(tuple(m[q * s.npts:(q + 1) * s.npts].copy() for q in range(6)))
    return dict(dt=0.01, reinv=1.0 / 200.0, marker_vel=mv, p=fields[0], u=fields[1], v=fields[2], w=fields[3])
