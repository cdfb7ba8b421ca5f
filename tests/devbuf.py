"""Device buffers for tests that call libgat_hip.so's C ABI directly."""
import ctypes as C

import numpy as np

from gat_amd import _native as N


class DevBuf:
    def __init__(self, nbytes: int):
        self.nbytes = max(int(nbytes), 1)
        p = C.c_void_p()
        N.check_hip(N.hip().gat_malloc(C.byref(p), self.nbytes))
        self.ptr = p

    @classmethod
    def of(cls, arr: np.ndarray) -> "DevBuf":
        a = np.ascontiguousarray(arr)
        b = cls(a.nbytes)
        if a.nbytes:
            N.check_hip(N.hip().gat_memcpy_h2d(b.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, None))
            N.check_hip(N.hip().gat_device_sync())
        return b

    def zero(self) -> "DevBuf":
        N.check_hip(N.hip().gat_memset(self.ptr, 0, self.nbytes, None))
        N.check_hip(N.hip().gat_device_sync())
        return self

    def get(self, dtype, count) -> np.ndarray:
        out = np.empty(count, dtype)
        if out.nbytes:
            N.check_hip(N.hip().gat_device_sync())
            N.check_hip(N.hip().gat_memcpy_d2h(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes, None))
            N.check_hip(N.hip().gat_device_sync())
        return out

    def __del__(self):
        if getattr(self, "ptr", None):
            N.hip().gat_free(self.ptr)
            self.ptr = None
