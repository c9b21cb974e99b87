"""Device arrays for the GPU tests, through the HIP runtime's C API (ctypes): tests hand device pointers to the C ABI
(atom index arrays, device-resident parameters) the way a platform would."""
import ctypes

import numpy as np

_hip = None


def hip() -> ctypes.CDLL:
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


H2D, D2H = 1, 2


class DeviceArray:
    """A device copy of a numpy array (its dtype and shape kept for the download)."""

    def __init__(self, host: np.ndarray):
        host = np.ascontiguousarray(host)
        self.dtype, self.shape = host.dtype, host.shape
        self.nbytes = max(host.nbytes, 1)
        self.ptr = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.ptr), self.nbytes) == 0
        self.upload(host)

    def upload(self, host: np.ndarray) -> None:
        host = np.ascontiguousarray(host, dtype=self.dtype)
        assert host.nbytes <= self.nbytes
        if host.nbytes:
            assert hip().hipMemcpy(self.ptr, host.ctypes.data, host.nbytes, H2D) == 0

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        if out.nbytes:
            assert hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    @property
    def address(self) -> int:
        return self.ptr.value

    def free(self) -> None:
        if self.ptr:
            hip().hipDeviceSynchronize()
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()
