"""Device memory, streams and events over the C ABI (numpy on the host side)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import call, lib


class DeviceArray:
    """A device allocation with a numpy dtype and shape (C order)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = C.c_void_p()
        call("rts_malloc", C.byref(p), self.nbytes)
        self._ptr = p.value

    @property
    def ptr(self) -> int:
        return self._ptr

    @classmethod
    def from_host(cls, a: np.ndarray, stream=None) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a, stream)
        return d

    def upload(self, a: np.ndarray, stream=None) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        call("rts_memcpy_h2d", self._ptr, a.ctypes.data, self.nbytes, stream)
        call("rts_stream_sync", stream)

    def download(self, stream=None) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        call("rts_memcpy_d2h", out.ctypes.data, self._ptr, self.nbytes, stream)
        call("rts_stream_sync", stream)
        return out

    def fill(self, value: int = 0, stream=None) -> None:
        call("rts_fill_u8", self._ptr, value, self.nbytes, stream)

    def free(self) -> None:
        if self._ptr:
            lib.rts_free(self._ptr)
            self._ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = C.c_void_p()
        call("rts_stream_create", C.byref(p))
        self.handle = p.value

    def sync(self):
        call("rts_stream_sync", self.handle)

    def __del__(self):
        try:
            if self.handle:
                lib.rts_stream_destroy(self.handle)
        except Exception:
            pass


class Event:
    def __init__(self):
        p = C.c_void_p()
        call("rts_event_create", C.byref(p))
        self.handle = p.value

    def record(self, stream=None):
        call("rts_event_record", self.handle, stream)

    def sync(self):
        call("rts_event_sync", self.handle)

    def elapsed_ms(self, later: "Event") -> float:
        ms = C.c_float()
        call("rts_event_elapsed_ms", self.handle, later.handle, C.byref(ms))
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle:
                lib.rts_event_destroy(self.handle)
        except Exception:
            pass


def device_sync() -> None:
    call("rts_device_sync")


def gpu_count() -> int:
    n = C.c_int(0)
    st = lib.rts_device_count(C.byref(n))
    return n.value if st == 0 else 0
