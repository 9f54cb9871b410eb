"""Device memory, pinned host memory and streams over the libgls C ABI.

These replace the PyCUDA objects the reference uses (``gpuarray``,
``cuda.Stream``, ``cuda.aligned_zeros``; cuvarbase/lombscargle.py:129-191,
cuvarbase/core.py:27-40).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def _handle(stream):
    if stream is None:
        return None
    return stream.handle if isinstance(stream, Stream) else stream


class Stream:
    """A non-blocking HIP stream (replaces ``pycuda.driver.Stream``)."""

    def __init__(self):
        _lib.init(_current_device())
        h = C.c_void_p()
        _lib.call("gls_stream_create", C.byref(h))
        self.handle = h.value

    def synchronize(self):
        _lib.call("gls_stream_synchronize", self.handle)

    def is_done(self):
        d = C.c_int(0)
        _lib.call("gls_stream_query", self.handle, C.byref(d))
        return bool(d.value)

    def __del__(self):
        h, self.handle = getattr(self, "handle", None), None
        if h is not None and _lib._lib is not None:
            try:
                _lib.call("gls_stream_destroy", h)
            except Exception:  # pragma: no cover - interpreter shutdown
                pass


class Event:
    def __init__(self):
        _lib.init(_current_device())
        h = C.c_void_p()
        _lib.call("gls_event_create", C.byref(h))
        self.handle = h.value

    def record(self, stream=None):
        _lib.call("gls_event_record", self.handle, _handle(stream))
        return self

    def synchronize(self):
        _lib.call("gls_event_synchronize", self.handle)

    def query(self):
        d = C.c_int(0)
        _lib.call("gls_event_query", self.handle, C.byref(d))
        return bool(d.value)

    def time_since(self, start):
        ms = C.c_float(0)
        _lib.call("gls_event_elapsed_ms", start.handle, self.handle, C.byref(ms))
        return ms.value

    def __del__(self):
        h, self.handle = getattr(self, "handle", None), None
        if h is not None and _lib._lib is not None:
            try:
                _lib.call("gls_event_destroy", h)
            except Exception:  # pragma: no cover
                pass


_device = 0


def _current_device():
    return _device


def set_device(device):
    global _device
    _device = int(device)
    _lib.init(_device)


class DeviceArray:
    """A 1-D typed device buffer (the subset of ``gpuarray`` the path uses)."""

    def __init__(self, shape, dtype, zero=False, stream=None):
        _lib.init(_current_device())
        self.shape = (int(shape),) if np.isscalar(shape) else tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = self.size * self.dtype.itemsize
        p = C.c_void_p()
        _lib.call("gls_malloc", C.byref(p), C.c_size_t(self.nbytes))
        self.ptr = p.value or 0
        if zero:
            self.fill(0, stream=stream)

    @classmethod
    def zeros(cls, shape, dtype, stream=None):
        a = cls(shape, dtype, zero=True, stream=stream)
        if stream is None:
            _lib.call("gls_device_synchronize")
        return a

    @classmethod
    def to_device(cls, ary, stream=None):
        ary = np.ascontiguousarray(ary)
        a = cls(ary.shape, ary.dtype)
        a.set(ary) if stream is None else a.set_async(ary, stream=stream)
        return a

    def __len__(self):
        return self.shape[0] if self.shape else 1

    def fill(self, value=0, stream=None):
        if value != 0:
            raise ValueError("DeviceArray.fill supports only 0")
        _lib.call("gls_memset_async", self.ptr, 0, C.c_size_t(self.nbytes), _handle(stream))
        return self

    def set_async(self, ary, stream=None, offset=0):
        ary = np.ascontiguousarray(ary, dtype=self.dtype)
        nb = ary.nbytes
        if offset * self.dtype.itemsize + nb > self.nbytes:
            raise ValueError("host array larger than device buffer")
        _lib.call("gls_memcpy_h2d_async", self.ptr + offset * self.dtype.itemsize,
                  ary.ctypes.data, C.c_size_t(nb), _handle(stream))
        # keep the source alive until the caller synchronises
        self._pending_src = ary
        return self

    def set(self, ary):
        self.set_async(ary, stream=None)
        _lib.call("gls_device_synchronize")
        return self

    def get_async(self, ary=None, stream=None, count=None):
        count = self.size if count is None else int(count)
        if ary is None:
            ary = np.empty(count, dtype=self.dtype)
        if not ary.flags.c_contiguous or ary.dtype != self.dtype or ary.size < count:
            raise ValueError("destination must be a contiguous array of the same dtype")
        _lib.call("gls_memcpy_d2h_async", ary.ctypes.data, self.ptr,
                  C.c_size_t(count * self.dtype.itemsize), _handle(stream))
        return ary

    def get(self, ary=None, count=None):
        out = self.get_async(ary=ary, stream=None, count=count)
        _lib.call("gls_device_synchronize")
        return out.reshape(self.shape) if count is None and ary is None else out

    def free(self):
        p, self.ptr = getattr(self, "ptr", 0), 0
        if p and _lib._lib is not None:
            _lib.call("gls_free", C.c_void_p(p))

    def __del__(self):
        try:
            self.free()
        except Exception:  # pragma: no cover
            pass


class _HostBlock:
    def __init__(self, nbytes):
        p = C.c_void_p()
        _lib.call("gls_host_alloc", C.byref(p), C.c_size_t(max(int(nbytes), 1)))
        self.ptr = p.value

    def __del__(self):
        p, self.ptr = getattr(self, "ptr", None), None
        if p and _lib._lib is not None:
            try:
                _lib.call("gls_host_free", C.c_void_p(p))
            except Exception:  # pragma: no cover
                pass


def pinned_empty(shape, dtype):
    """Page-locked host array (the reference's ``cuda.aligned_zeros`` buffers
    are only aligned; these are pinned so D2H copies are truly async)."""
    _lib.init(_current_device())
    dtype = np.dtype(dtype)
    shape = (int(shape),) if np.isscalar(shape) else tuple(int(s) for s in shape)
    n = int(np.prod(shape)) if shape else 1
    blk = _HostBlock(n * dtype.itemsize)
    buf = (C.c_byte * max(n * dtype.itemsize, 1)).from_address(blk.ptr)
    buf._gls_owner = blk  # the array's base holds the pinned block alive
    # touch every page now: the process's first access to each page of a
    # fresh page-locked block faults it in (~1.5 us per 4 KiB page, 0.1 s for
    # 300 MB), which would otherwise land on the first copy into it
    np.frombuffer(buf, dtype=np.uint8)[::4096] = 0
    return np.frombuffer(buf, dtype=dtype, count=n).reshape(shape)


def pinned_zeros(shape, dtype):
    a = pinned_empty(shape, dtype)
    a[...] = 0
    return a


def synchronize():
    _lib.call("gls_device_synchronize")
