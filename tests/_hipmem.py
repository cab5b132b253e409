"""Device buffers for GPU tests through the HIP runtime the library itself
is linked against (libamdhip64.so.7, already loaded by the library, so the
same instance), without torch: hipMalloc, hipMemcpy, hipFree."""
import ctypes as C

import numpy as np

_H2D, _D2H = 1, 2
_rt = None


def runtime():
    global _rt
    if _rt is None:
        import feicar3d_amd.poisson as poisson
        poisson.library()
        rt = C.CDLL("libamdhip64.so.7")
        rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        rt.hipFree.argtypes = [C.c_void_p]
        _rt = rt
    return _rt


class DeviceBuffer:
    """nbytes of device memory; data_ptr() as a torch tensor has it."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        rc = runtime().hipMalloc(C.byref(p), max(self.nbytes, 1))
        if rc != 0:
            raise RuntimeError("hipMalloc failed: %d" % rc)
        self.ptr = p.value

    def data_ptr(self):
        return self.ptr

    def upload(self, a):
        a = np.asarray(a)
        assert a.nbytes == self.nbytes
        src = np.asfortranarray(a) if a.ndim > 1 else np.ascontiguousarray(a)
        assert runtime().hipMemcpy(self.ptr, src.ctypes.data, self.nbytes, _H2D) == 0

    def download(self, like):
        out = np.empty_like(like, order="F")
        assert runtime().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, _D2H) == 0
        return out

    def free(self):
        if self.ptr:
            runtime().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
