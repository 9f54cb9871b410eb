"""gls_memcpy_kernel_async: copies by a kernel between page-locked host memory
and device memory, at aligned and unaligned ends and sizes that leave a byte
tail; pageable host memory is refused."""
import ctypes as C

import numpy as np
import pytest

from cuvarbase_amd import _device, _lib

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nbytes,off", [(16, 0), (24 * 1001, 0), (1 << 20, 0), (4099, 0),
                                        (4099, 3), (777, 8)])
def test_kernel_copy_round_trip(nbytes, off):
    rng = np.random.default_rng(nbytes + off)
    src = _device.pinned_empty(nbytes + 64, np.uint8)
    src[:] = rng.integers(0, 256, size=src.size, dtype=np.uint8)
    back = _device.pinned_empty(nbytes + 64, np.uint8)
    back[:] = 0
    dev = _device.DeviceArray(nbytes + 64, np.uint8)
    s = _device.Stream()
    h = _device._handle(s)
    _lib.call("gls_memcpy_kernel_async", dev.ptr + off, src.ctypes.data + off,
              C.c_size_t(nbytes), h)
    _lib.call("gls_memcpy_kernel_async", back.ctypes.data + off, dev.ptr + off,
              C.c_size_t(nbytes), h)
    s.synchronize()
    np.testing.assert_array_equal(back[off:off + nbytes], src[off:off + nbytes])
    assert not back[:off].any() and not back[off + nbytes:].any()
    # device to device through the same path
    dev2 = _device.DeviceArray(nbytes + 64, np.uint8)
    _lib.call("gls_memcpy_kernel_async", dev2.ptr, dev.ptr + off, C.c_size_t(nbytes), h)
    s.synchronize()
    np.testing.assert_array_equal(dev2.get()[:nbytes], src[off:off + nbytes])


def test_kernel_copy_refuses_pageable_memory():
    dev = _device.DeviceArray(64, np.uint8)
    pageable = np.zeros(64, np.uint8)
    with pytest.raises(_lib.GLSError):
        _lib.call("gls_memcpy_kernel_async", dev.ptr, pageable.ctypes.data, C.c_size_t(64),
                  None)
