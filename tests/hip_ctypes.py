"""Device buffers and rocFFT through ctypes on the system ROCm runtime (the same
libamdhip64 the engine links), for the FFT cross-checks."""
import ctypes
import os

import numpy as np

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
_hip = None
_rocfft = None


def hip():
    global _hip
    if _hip is None:
        L = ctypes.CDLL(os.path.join(ROCM, "lib", "libamdhip64.so"))
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipFree.argtypes = [ctypes.c_void_p]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        L.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        _hip = L
    return _hip


class DeviceArray:
    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        rc = hip().hipMalloc(ctypes.byref(p), self.nbytes)
        if rc != 0:
            raise RuntimeError(f"hipMalloc rc={rc}")
        self.ptr = p.value
        hip().hipMemset(ctypes.c_void_p(self.ptr), 0, self.nbytes)

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        rc = hip().hipMemcpy(ctypes.c_void_p(d.ptr), a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy H2D rc={rc}")
        return d

    def to_numpy(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        rc = hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr), out.nbytes, 2)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy D2H rc={rc}")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        self.free()


def rocfft():
    global _rocfft
    if _rocfft is None:
        L = ctypes.CDLL(os.path.join(ROCM, "lib", "librocfft.so"))
        L.rocfft_setup()
        L.rocfft_plan_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), ctypes.c_size_t,
                                         ctypes.c_void_p]
        L.rocfft_execute.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                     ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p]
        L.rocfft_plan_destroy.argtypes = [ctypes.c_void_p]
        _rocfft = L
    return _rocfft


def rocfft_r2c(d_in, d_out, shape, batch):
    """Out-of-place single-precision real-forward transform; shape in (z, y, x) order."""
    L = rocfft()
    plan = ctypes.c_void_p()
    lengths = (ctypes.c_size_t * len(shape))(*shape[::-1])
    # placement notinplace = 1, transform real_forward = 2, precision single = 0
    rc = L.rocfft_plan_create(ctypes.byref(plan), 1, 2, 0, len(shape), lengths, batch, None)
    if rc != 0:
        raise RuntimeError(f"rocfft_plan_create rc={rc}")
    ib = (ctypes.c_void_p * 1)(d_in.ptr)
    ob = (ctypes.c_void_p * 1)(d_out.ptr)
    rc = L.rocfft_execute(plan, ib, ob, None)
    L.rocfft_plan_destroy(plan)
    if rc != 0:
        raise RuntimeError(f"rocfft_execute rc={rc}")
