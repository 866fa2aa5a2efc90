"""Device arrays for the GPU tests through the HIP runtime libygrid itself links (no torch, so
that one HIP runtime serves the process)."""
import ctypes
import os

import numpy as np

_HIP = None


def hip():
    global _HIP
    if _HIP is None:
        import yandasoft_amd as ya
        ya.lib()  # libygrid first: its HIP runtime is the one to share
        rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
        _HIP = ctypes.CDLL(os.path.join(rocm, "lib", "libamdhip64.so"))
    return _HIP


def _ok(rc):
    if rc != 0:
        raise RuntimeError(f"HIP error {rc}")


class DevArray:
    """A device copy of a numpy array (same shape and dtype); data_ptr() as torch tensors."""

    def __init__(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        self.shape, self.dtype, self.nbytes = a.shape, a.dtype, a.nbytes
        self.ptr = ctypes.c_void_p()
        _ok(hip().hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(max(self.nbytes, 1))))
        _ok(hip().hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(self.nbytes), 1))

    def data_ptr(self) -> int:
        return int(self.ptr.value)

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        _ok(hip().hipDeviceSynchronize())
        _ok(hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(self.nbytes), 2))
        return out

    def zero_(self):
        _ok(hip().hipDeviceSynchronize())
        _ok(hip().hipMemset(self.ptr, 0, ctypes.c_size_t(self.nbytes)))
        _ok(hip().hipDeviceSynchronize())

    def __del__(self):
        try:
            if self.ptr:
                hip().hipFree(self.ptr)
                self.ptr = ctypes.c_void_p()
        except Exception:
            pass
