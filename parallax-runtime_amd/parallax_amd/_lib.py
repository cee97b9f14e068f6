"""ctypes loader of libparallax_amd.so with the prototypes of parallax/runtime.h and
parallax/px_device.h.  Raises at import if the library has not been built."""
import ctypes
import pathlib

from .desc import PxDesc

LIB_PATH = pathlib.Path(__file__).resolve().parent.parent / "lib" / "libparallax_amd.so"

if not LIB_PATH.exists():
    raise ImportError(
        f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')"
    )

lib = ctypes.CDLL(str(LIB_PATH), mode=ctypes.RTLD_GLOBAL)

_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_dp = ctypes.POINTER(PxDesc)
_i = ctypes.c_int


def _proto(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


# ---- parallax/runtime.h
_proto("parallax_umalloc", _vp, _sz, ctypes.c_uint)
_proto("parallax_ufree", None, _vp)
_proto("parallax_arena_alloc", _vp, _sz, _sz)
_proto("parallax_arena_free", None, _vp)
_proto("parallax_arena_contains", _i, _vp)
_proto("parallax_heap_contains", _i, _vp)
_proto("parallax_kernel_load", _vp, _u32p, _sz)
_proto("parallax_kernel_launch_transform2", None, _vp, _vp, _vp, _sz, _sz, _sz)
_proto("parallax_kernel_launch_transform2_captures", None, _vp, _vp, _vp, _sz, _sz, _sz, _vp, _sz)
_proto("parallax_reduce", None, _vp, _vp, _sz, _sz, _vp)
_proto("parallax_transform_reduce", None, _vp, _vp, _vp, _sz, _sz, _sz, _vp)
_proto("parallax_argminmax", _sz, _vp, _vp, _sz, _sz, _i, _i, _i)
_proto("parallax_find", _sz, _vp, _vp, _sz, _sz, _i, _vp)
_proto("parallax_mismatch", _sz, _vp, _vp, _vp, _sz, _sz)
_proto("parallax_scan", None, _vp, _vp, _vp, _sz, _sz)
_proto("parallax_exclusive_scan", None, _vp, _vp, _vp, _vp, _vp, _sz, _sz, _vp)
_proto("parallax_sort", None, _vp, _vp, _sz, _sz)
_proto("parallax_copy_if", _sz, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _sz, _i)
_proto("parallax_kernel_register", None, ctypes.c_char_p, _u32p, _sz)
_proto("parallax_kernel_lookup", _vp, ctypes.c_char_p)
_proto("parallax_kernel_launch_with_captures", None, _vp, _vp, _sz, _vp, _sz, _sz)
# variadic entry points are called with explicit argument conversions
lib.parallax_kernel_launch.restype = None
lib.parallax_kernel_launch_transform.restype = None

# ---- parallax/px_device.h
_proto("pxr_device_count", _i)
_proto("pxr_device_info", _i, ctypes.c_char_p, _sz, ctypes.POINTER(_i), ctypes.POINTER(_sz), ctypes.POINTER(_sz))
_proto("pxr_last_error", _i, ctypes.c_char_p, _sz)
_proto("pxr_stats", _i, _u64p)
_proto("pxr_version", ctypes.c_char_p)
_proto("pxr_arena_info", _i, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_sz), ctypes.POINTER(_sz))
_proto("pxr_relocate_probe", _i, _vp, ctypes.c_uint)
_proto("pxr_builtin_descriptor", _i, ctypes.c_char_p, _u32p)
_proto("pxr_prof_enable", _i, _i)
_proto("pxr_prof_get", _i, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), _u64p)
_proto("pxd_malloc", _vp, _sz)
_proto("pxd_free", None, _vp)
_proto("pxd_memcpy_h2d", _i, _vp, _vp, _sz)
_proto("pxd_memcpy_d2h", _i, _vp, _vp, _sz)
_proto("pxd_memcpy_d2d", _i, _vp, _vp, _sz)
_proto("pxd_memset", _i, _vp, _i, _sz)
_proto("pxd_device_sync", _i)
_proto("pxd_ctx_create", _vp, _vp)
_proto("pxd_ctx_destroy", None, _vp)
_proto("pxd_ctx_stream", _vp, _vp)
_proto("pxd_sync", _i, _vp)
_proto("pxd_timer_begin", _i, _vp)
_proto("pxd_timer_end", _i, _vp, ctypes.POINTER(ctypes.c_double))
_proto("pxd_copy", _i, _vp, _vp, _vp, _sz)
_proto("pxd_reduce", _i, _vp, _i, _vp, _sz, _vp)
_proto("pxd_transform_reduce", _i, _vp, _dp, _vp, _sz, _vp)
_proto("pxd_inclusive_scan", _i, _vp, _i, _vp, _vp, _sz)
_proto("pxd_inclusive_scan_noalloc", _i, _i, _vp, _vp, _sz, _vp)
_proto("pxd_exclusive_scan", _i, _vp, _i, _vp, _vp, _sz, _vp)
_proto("pxd_sort", _i, _vp, _i, _vp, _sz)
_proto("pxd_copy_if", _i, _vp, _dp, _vp, _vp, _sz, _vp)
_proto("pxd_partition", _i, _vp, _dp, _vp, _vp, _sz, _vp)
_proto("pxd_transform", _i, _vp, _dp, _vp, _vp, _sz, _vp, _sz)
_proto("pxd_find", _i, _vp, _dp, _vp, _sz, _vp)
_proto("pxd_argminmax", _i, _vp, _i, _vp, _sz, _i, _i, _vp)
_proto("pxd_mismatch", _i, _vp, _i, _vp, _vp, _sz, _vp)
_proto("pxd_relocate_pointers", _i, _vp, _vp, _sz, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64)
_proto("pxb_seg_reduce", _i, _vp, _i, _vp, _u64p, _sz, _vp)
_proto("pxb_seg_inclusive_scan", _i, _vp, _i, _vp, _vp, _u64p, _sz)
_proto("pxb_seg_sort", _i, _vp, _i, _vp, _u64p, _sz)
_proto("pxb_seg_copy_if", _i, _vp, _dp, _vp, _vp, _u64p, _sz, _vp)
_proto("pxb_plan_create", _vp, _vp, _i, _i, _dp, _vp, _vp, _u64p, _sz, _vp)
_proto("pxb_plan_tile", _i, _vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double))
_proto("pxr_tile_override", _i, _i, _i, _i, _i)
_proto("pxb_plan_launch", _i, _vp)
_proto("pxb_plan_destroy", None, _vp)


# ---- parallax/px_serve.h
class PxsConfig(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_int), ("hbm_budget", ctypes.c_size_t),
                ("max_batch", ctypes.c_int), ("max_wait_us", ctypes.c_double), ("lanes", ctypes.c_int)]


class PxsRequest(ctypes.Structure):
    _fields_ = [("op", ctypes.c_int), ("dtype", ctypes.c_int), ("n", ctypes.c_size_t),
                ("d_in", ctypes.c_void_p), ("flags", PxDesc)]


class PxsStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("batches", ctypes.c_uint64), ("rejected", ctypes.c_uint64),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("mean_ms", ctypes.c_double), ("max_ms", ctypes.c_double),
                ("first_arrival_s", ctypes.c_double), ("last_completion_s", ctypes.c_double),
                ("req_per_s", ctypes.c_double), ("pool_bytes", ctypes.c_size_t),
                ("pool_peak_bytes", ctypes.c_size_t), ("failed", ctypes.c_uint64)]


_proto("pxs_config_default", None, ctypes.POINTER(PxsConfig))
_proto("pxs_create", _vp, ctypes.POINTER(PxsConfig))
_proto("pxs_destroy", None, _vp)
_proto("pxs_alloc", _vp, _vp, _sz)
_proto("pxs_free", None, _vp, _vp)
_proto("pxs_submit", ctypes.c_int64, _vp, ctypes.POINTER(PxsRequest))
_proto("pxs_submit_batch", _i, _vp, ctypes.POINTER(PxsRequest), _sz, ctypes.POINTER(ctypes.c_int64))
_proto("pxs_wait", _i, _vp, ctypes.c_int64, ctypes.c_double)
_proto("pxs_result", _i, _vp, ctypes.c_int64, ctypes.POINTER(_vp), _u64p, ctypes.POINTER(ctypes.c_double))
_proto("pxs_release", _i, _vp, ctypes.c_int64)
_proto("pxs_drain", _i, _vp)
_proto("pxs_get_stats", _i, _vp, ctypes.POINTER(PxsStats))
_proto("pxs_reset_stats", _i, _vp)


def last_error():
    buf = ctypes.create_string_buffer(512)
    code = lib.pxr_last_error(buf, 512)
    return code, buf.value.decode(errors="replace")


def check(rc, what="call"):
    """Raise RuntimeError for a negative return code."""
    if rc != 0:
        code, text = last_error()
        raise RuntimeError(f"{what} failed with {rc}: {text}")
    return rc
