"""fc_pson_rhs_noalloc, the allocation-free variant of rhs_poisson
(fc_pson_rhs): called through the C-ABI on device arrays, it must give
fc_pson_rhs's divergence and sums to the bit, on two calls in a row (the
second on a smaller grid, which reuses the library's static scratch), and
refuse a grid with more interior rows than its scratch holds.  The test's
own device arrays come from the HIP runtime the library uses."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


class _Hip:
    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), nbytes) == 0
        self.bufs.append(p.value)
        return p.value

    def put(self, a):
        a = np.ascontiguousarray(a.ravel(order="F"))
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, _H2D) == 0
        return p

    def get(self, p, n):
        out = np.zeros(n)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, _D2H) == 0
        return out

    def sync(self):
        assert self.L.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _case(hip, capi, geo, gpu, w, seed, dtinv):
    nx, ny, nz = (a.size - 2 for a in w.grid)
    fu, fv, fw = geo.face_velocities(nx, ny, nz, seed=seed)
    s = gpu.PoissonSolver(*w.grid, capi.default_params())
    try:
        s.set_workload(w)
        div_h, sumd_h, rdiv_h = s.rhs_poisson(fu, fv, fw, dtinv)
        du, dv, dw = (hip.put(a) for a in (fu, fv, fw))
        div_d = hip.put(np.zeros(w.div.shape, order="F"))
        sums_d = hip.put(np.zeros(2))
        hip.sync()
        rc = s.rhs_poisson_noalloc(du, dv, dw, div_d, sums_d, dtinv)
        hip.sync()
        assert rc == 0
        div = hip.get(div_d, w.div.size).reshape(w.div.shape, order="F")
        sums = hip.get(sums_d, 2)
    finally:
        s.close()
        hip.free()
    assert np.array_equal(div[1:nx, 1:ny, 1:nz], div_h[1:nx, 1:ny, 1:nz])
    assert np.count_nonzero(div) == np.count_nonzero(div_h)
    assert sums[0] == sumd_h and sums[1] == rdiv_h
    return np.count_nonzero(div_h)


def test_rhs_noalloc_matches_rhs(capi, geo, gpu):
    gpu.library()
    hip = _Hip()
    # first call: the 40^3 stretched C2 grid; second: the smaller 16^3 C1 grid
    n1 = _case(hip, capi, geo, gpu, geo.config_c2(40), 21, 2.0)
    n2 = _case(hip, capi, geo, gpu, geo.config_c1(16), 22, 0.5)
    assert n1 > n2 > 0


def test_rhs_noalloc_refuses_oversized_grids(capi, geo, gpu):
    # 8 x 600 x 600 nodes: 599^2 = 358801 interior (j,k) rows > FC_NOALLOC_MAX_ROWS
    gpu.library()
    hip = _Hip()
    x, y = geo.uniform_axis(8), geo.uniform_axis(600)
    s = gpu.PoissonSolver(x, y, y.copy(), capi.default_params())
    try:
        p = hip.alloc(64)
        with pytest.raises(gpu.FcError, match="FC_NOALLOC_MAX_ROWS"):
            s.rhs_poisson_noalloc(p, p, p, p, p, 1.0)
    finally:
        s.close()
        hip.free()
