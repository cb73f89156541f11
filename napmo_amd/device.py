"""Minimal device buffers over the library's allocator (for callers without torch)."""
import ctypes

import numpy as np

from napmo_amd.system import cext


class DeviceArray(object):
    """float64 device buffer of ``shape``; ``ptr`` is the raw device address."""

    def __init__(self, shape):
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 8
        self.ptr = cext.napmo_device_malloc(max(self.nbytes, 8))
        if not self.ptr:
            raise cext.NapmoError(cext.lib.napmo_last_error().decode())

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = cls(a.shape)
        cext.check(cext.napmo_memcpy(d.ptr, a.ctypes.data, d.nbytes, 1))
        return d

    def to_host(self):
        out = np.empty(self.shape, dtype=np.float64)
        cext.check(cext.napmo_memcpy(out.ctypes.data, self.ptr, self.nbytes, 2))
        return out

    def __del__(self):
        if getattr(self, 'ptr', None):
            cext.napmo_device_free(self.ptr)
            self.ptr = None
