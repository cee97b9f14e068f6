"""Device buffers and a stream context over the pxd_* entry points (no GPU framework
needed).  Used by the tests and bench.py; every call goes to libparallax_amd.so."""
import ctypes

import numpy as np

from ._lib import lib, check
from .dtypes import NP_DTYPES, dtype_code


class DeviceBuffer:
    """A hipMalloc'd buffer of `nbytes` (freed with the object)."""

    def __init__(self, nbytes, zero=False):
        self.nbytes = int(nbytes)
        self.ptr = lib.pxd_malloc(max(self.nbytes, 1))
        if not self.ptr:
            raise MemoryError(f"pxd_malloc({self.nbytes}) failed")
        if zero:
            check(lib.pxd_memset(self.ptr, 0, max(self.nbytes, 1)), "memset")

    @classmethod
    def from_array(cls, a, pad=0):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes + pad)
        if a.nbytes:
            check(lib.pxd_memcpy_h2d(b.ptr, a.ctypes.data, a.nbytes), "h2d")
        return b

    def upload(self, a, offset=0):
        a = np.ascontiguousarray(a)
        if a.nbytes:
            check(lib.pxd_memcpy_h2d(self.ptr + offset, a.ctypes.data, a.nbytes), "h2d")

    def download(self, dtype, count, offset=0):
        out = np.empty(int(count), dtype=dtype)
        if out.nbytes:
            check(lib.pxd_memcpy_d2h(out.ctypes.data, self.ptr + offset, out.nbytes), "d2h")
        return out

    def addr(self, offset=0):
        return self.ptr + offset

    def free(self):
        if self.ptr:
            lib.pxd_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One HIP stream with its scratch (pxd_ctx_t)."""

    def __init__(self, stream=None):
        self.h = lib.pxd_ctx_create(stream)
        if not self.h:
            raise RuntimeError("pxd_ctx_create failed (no GPU?)")

    def sync(self):
        check(lib.pxd_sync(self.h), "sync")

    def timer_begin(self):
        check(lib.pxd_timer_begin(self.h), "timer_begin")

    def timer_end(self):
        ms = ctypes.c_double()
        check(lib.pxd_timer_end(self.h, ctypes.byref(ms)), "timer_end")
        return ms.value

    def close(self):
        if self.h:
            lib.pxd_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_count():
    return lib.pxr_device_count()


def stats():
    out = (ctypes.c_uint64 * 7)()
    check(lib.pxr_stats(out), "stats")
    keys = ["launches", "reduce", "scan", "sort", "compact", "transform", "seg"]
    return dict(zip(keys, list(out)))


def np_dtype(code):
    return NP_DTYPES[dtype_code(code)]
