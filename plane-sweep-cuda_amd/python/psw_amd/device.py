"""Device memory, streams and events through libpsw's C ABI (no torch)."""
import ctypes

import numpy as np

from ._lib import call, lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class DeviceArray:
    """A device allocation holding a dense numpy-shaped array."""

    def __init__(self, shape, dtype=np.float32):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        call("psw_malloc", ctypes.byref(p), self.nbytes)
        self._p = p.value

    @classmethod
    def from_host(cls, arr, dtype=None):
        arr = np.ascontiguousarray(arr, dtype=dtype if dtype is not None else arr.dtype)
        d = cls(arr.shape, arr.dtype)
        d.upload(arr)
        return d

    @property
    def ptr(self):
        return ctypes.c_void_p(self._p)

    def offset(self, elements):
        """Device pointer `elements` items into the array."""
        return ctypes.c_void_p(self._p + int(elements) * self.dtype.itemsize)

    def upload(self, arr, stream=None):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.nbytes == self.nbytes, (arr.shape, self.shape)
        call("psw_memcpy_htod", self.ptr, _ptr(arr), self.nbytes, stream)

    def to_host(self, stream=None):
        out = np.empty(self.shape, self.dtype)
        call("psw_memcpy_dtoh", _ptr(out), self.ptr, self.nbytes, stream)
        return out

    def fill(self, value, stream=None):
        if self.dtype == np.float32:
            call("psw_fill_f32", self.ptr, float(value), self.nbytes // 4, stream)
        else:
            assert self.nbytes % 4 == 0
            call("psw_fill_u32", self.ptr, int(value), self.nbytes // 4, stream)

    def view(self, i):
        """Non-owning view of item i along the first axis."""
        return DeviceView(self, i)

    def free(self):
        if getattr(self, "_p", None):
            lib.psw_free(ctypes.c_void_p(self._p))
            self._p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceView(DeviceArray):
    """Item i along the first axis of a DeviceArray (keeps the parent alive)."""

    def __init__(self, parent, i):
        assert 0 <= i < parent.shape[0]
        self.parent = parent
        self.shape = parent.shape[1:]
        self.dtype = parent.dtype
        self.nbytes = parent.nbytes // parent.shape[0]
        self._p = parent._p + int(i) * self.nbytes

    def free(self):
        self._p = None


class Stream:
    def __init__(self):
        s = ctypes.c_void_p()
        call("psw_stream_create", ctypes.byref(s))
        self.handle = s

    def synchronize(self):
        call("psw_stream_synchronize", self.handle)

    def __del__(self):
        try:
            lib.psw_stream_destroy(self.handle)
        except Exception:
            pass


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        call("psw_event_create", ctypes.byref(e))
        self.handle = e

    def record(self, stream=None):
        call("psw_event_record", self.handle, stream.handle if isinstance(stream, Stream) else stream)

    def synchronize(self):
        call("psw_event_synchronize", self.handle)

    def elapsed_ms(self, later):
        ms = ctypes.c_float()
        call("psw_event_elapsed_ms", ctypes.byref(ms), self.handle, later.handle)
        return ms.value

    def __del__(self):
        try:
            lib.psw_event_destroy(self.handle)
        except Exception:
            pass


def device_count():
    n = ctypes.c_int(0)
    st = lib.psw_device_count(ctypes.byref(n))
    return n.value if st == 0 else 0


def device_name():
    buf = ctypes.create_string_buffer(256)
    call("psw_device_name", buf, 256)
    return buf.value.decode()


def synchronize():
    call("psw_device_synchronize")
