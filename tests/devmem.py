"""Minimal device buffers for tests, through the HIP runtime the wavefd_amd libraries link
(no torch in the test process, so only one HIP runtime is ever loaded)."""
import ctypes

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        from wavefd_amd import _lib
        _lib.load("iso")  # binds libamdhip64.so.7 the way the product libraries do
        _hip = ctypes.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DeviceArray:
    """float32 device buffer with the host array's shape; data_ptr() like a torch tensor."""

    def __init__(self, host):
        host = np.ascontiguousarray(host, np.float32)
        self.shape = host.shape
        self.nbytes = host.nbytes
        p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 4)) == 0
        self.ptr = p.value
        assert hip().hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, 1) == 0

    def data_ptr(self):
        return self.ptr

    def numpy(self):
        out = np.empty(self.shape, np.float32)
        hip().hipDeviceSynchronize()
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2) == 0
        return out

    def __del__(self):
        try:
            if self.ptr:
                hip().hipFree(self.ptr)
        except Exception:
            pass
