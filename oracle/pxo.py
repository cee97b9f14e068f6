"""ctypes binding of the CPU oracle (oracle/libpx_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker (and the timed CPU baseline).  The product library never
uses it.
"""
import ctypes
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ORACLE_PATH = HERE / "libpx_oracle.so"

_NP = {0: np.int32, 1: np.uint32, 2: np.int64, 3: np.uint64, 4: np.float32, 5: np.float64}
_CODE = {np.dtype(v): k for k, v in _NP.items()}


def code(a):
    return _CODE[np.dtype(a.dtype if isinstance(a, np.ndarray) else a)]


class _Desc(ctypes.Structure):
    _fields_ = [("magic", ctypes.c_uint32), ("version", ctypes.c_uint32), ("op", ctypes.c_uint32),
                ("dtype", ctypes.c_uint32), ("aux", ctypes.c_uint32), ("fn", ctypes.c_uint32),
                ("flags", ctypes.c_uint32), ("fn2", ctypes.c_uint32), ("c0", ctypes.c_uint64),
                ("c1", ctypes.c_uint64)]


def _load():
    if not ORACLE_PATH.exists():
        raise ImportError(f"{ORACLE_PATH} missing: run `make -C oracle`")
    L = ctypes.CDLL(str(ORACLE_PATH))
    vp, sz, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
    dp = ctypes.POINTER(_Desc)
    dbl = ctypes.POINTER(ctypes.c_double)
    u64p = ctypes.POINTER(ctypes.c_uint64)
    for name, res, args in [
        ("pxo_reduce", None, [i, vp, sz, vp]),
        ("pxo_sum_exact", None, [i, vp, sz, dbl, dbl, dbl]),
        ("pxo_inclusive_scan", None, [i, vp, vp, sz]),
        ("pxo_exclusive_scan", None, [i, vp, vp, sz, vp]),
        ("pxo_scan_exact", None, [i, vp, sz, i, vp, vp, vp]),
        ("pxo_sort", None, [i, vp, sz]),
        ("pxo_pred_eval", i, [dp, vp]),
        ("pxo_copy_if", sz, [dp, vp, vp, sz]),
        ("pxo_remove_if", sz, [dp, vp, sz]),
        ("pxo_partition", sz, [dp, vp, sz]),
        ("pxo_stable_partition", sz, [dp, vp, sz]),
        ("pxo_unique", sz, [i, vp, sz]),
        ("pxo_count_if", ctypes.c_long, [dp, vp, sz]),
        ("pxo_find", sz, [dp, vp, sz]),
        ("pxo_mismatch", sz, [i, vp, vp, sz]),
        ("pxo_argminmax", sz, [i, vp, sz, i, i]),
        ("pxo_map", None, [dp, vp, sz, vp, sz]),
        ("pxo_transform", None, [dp, vp, vp, sz, vp, sz]),
        ("pxo_seg_reduce", None, [i, vp, u64p, sz, vp]),
        ("pxo_seg_sum_exact", None, [i, vp, u64p, sz, vp, vp]),
        ("pxo_seg_inclusive_scan", None, [i, vp, vp, u64p, sz]),
        ("pxo_seg_scan_exact", None, [i, vp, u64p, sz, vp, vp]),
        ("pxo_seg_sort", None, [i, vp, u64p, sz]),
        ("pxo_seg_unique", sz, [i, vp, vp, u64p, sz, u64p]),
    ]:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


lib = _load()


def desc(d):
    """Convert a parallax_amd PxDesc (or any object with the same fields) to the oracle's."""
    return _Desc(*(getattr(d, f) for f, _ in _Desc._fields_))


def _p(a):
    return a.ctypes.data if a.size else None


def reduce(a):
    out = np.zeros(1, dtype=a.dtype)
    lib.pxo_reduce(code(a), _p(a), a.size, out.ctypes.data)
    return out[0]


def sum_exact(a):
    hi, lo, ab = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib.pxo_sum_exact(code(a), _p(a), a.size, ctypes.byref(hi), ctypes.byref(lo), ctypes.byref(ab))
    return hi.value, lo.value, ab.value


def inclusive_scan(a):
    out = np.empty_like(a)
    lib.pxo_inclusive_scan(code(a), _p(a), _p(out), a.size)
    return out


def exclusive_scan(a, init=0):
    out = np.empty_like(a)
    iv = np.array([init], dtype=a.dtype)
    lib.pxo_exclusive_scan(code(a), _p(a), _p(out), a.size, iv.ctypes.data)
    return out


def scan_exact(a, exclusive=False, init=0):
    val = np.empty(a.size, dtype=np.float64)
    ab = np.empty(a.size, dtype=np.float64)
    iv = np.array([init], dtype=a.dtype)
    lib.pxo_scan_exact(code(a), _p(a), a.size, 1 if exclusive else 0, iv.ctypes.data, _p(val), _p(ab))
    return val, ab


def sort(a):
    out = a.copy()
    lib.pxo_sort(code(a), _p(out), out.size)
    return out


def copy_if(d, a):
    out = np.empty_like(a)
    k = lib.pxo_copy_if(ctypes.byref(desc(d)), _p(a), _p(out), a.size)
    return out[:k]


def remove_if(d, a):
    out = a.copy()
    k = lib.pxo_remove_if(ctypes.byref(desc(d)), _p(out), out.size)
    return out[:k]


def partition(d, a, stable=False):
    out = a.copy()
    f = lib.pxo_stable_partition if stable else lib.pxo_partition
    k = f(ctypes.byref(desc(d)), _p(out), out.size)
    return out, k


def unique(a):
    out = a.copy()
    k = lib.pxo_unique(code(a), _p(out), out.size)
    return out[:k]


def count_if(d, a):
    return lib.pxo_count_if(ctypes.byref(desc(d)), _p(a), a.size)


def find(d, a):
    """First index whose element passes descriptor d's predicate; a.size if none."""
    a = np.ascontiguousarray(a)
    return int(lib.pxo_find(desc(d), _p(a), a.size))


def mismatch(a, b):
    """First index with a[i] != b[i]; a.size if the ranges are equal."""
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return int(lib.pxo_mismatch(code(a), _p(a), _p(b), a.size))


def argminmax(a, want_max=False, want_last=False):
    """Index of the minimum (maximum) by <, ties to the first (last) index; 0 if empty."""
    a = np.ascontiguousarray(a)
    return int(lib.pxo_argminmax(code(a), _p(a), a.size, int(want_max), int(want_last)))


def pred_eval(d, x):
    v = np.array([x])
    return lib.pxo_pred_eval(ctypes.byref(desc(d)), v.ctypes.data)


def transform(d, a, out_dtype, captures=None):
    out = np.empty(a.size, dtype=out_dtype)
    cap = bytes(captures) if captures is not None else None
    cb = ctypes.create_string_buffer(cap, len(cap)) if cap else None
    lib.pxo_transform(ctypes.byref(desc(d)), _p(a), _p(out), a.size, cb, len(cap) if cap else 0)
    return out


def map_inplace(d, a, captures=None):
    out = a.copy()
    cap = bytes(captures) if captures is not None else None
    cb = ctypes.create_string_buffer(cap, len(cap)) if cap else None
    lib.pxo_map(ctypes.byref(desc(d)), _p(out), out.size, cb, len(cap) if cap else 0)
    return out


def _off(off):
    off = np.ascontiguousarray(off, dtype=np.uint64)
    return off, off.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def seg_reduce(a, off):
    off, op = _off(off)
    out = np.zeros(off.size - 1, dtype=a.dtype)
    lib.pxo_seg_reduce(code(a), _p(a), op, off.size - 1, _p(out))
    return out


def seg_sum_exact(a, off):
    off, op = _off(off)
    hi = np.zeros(off.size - 1)
    ab = np.zeros(off.size - 1)
    lib.pxo_seg_sum_exact(code(a), _p(a), op, off.size - 1, _p(hi), _p(ab))
    return hi, ab


def seg_inclusive_scan(a, off):
    off, op = _off(off)
    out = np.empty_like(a)
    lib.pxo_seg_inclusive_scan(code(a), _p(a), _p(out), op, off.size - 1)
    return out


def seg_scan_exact(a, off):
    """Per-segment inclusive-scan reference values and prefix sums of |x| (float64)."""
    off, op = _off(off)
    val = np.empty(a.size, dtype=np.float64)
    ab = np.empty(a.size, dtype=np.float64)
    lib.pxo_seg_scan_exact(code(a), _p(a), op, off.size - 1, _p(val), _p(ab))
    return val, ab


def seg_sort(a, off):
    off, op = _off(off)
    out = a.copy()
    lib.pxo_seg_sort(code(a), _p(out), op, off.size - 1)
    return out


def seg_unique(a, off):
    """Per-segment std::unique; returns (packed values, packed offsets)."""
    off, op = _off(off)
    out = np.empty_like(a)
    oo = np.zeros(off.size, dtype=np.uint64)
    k = lib.pxo_seg_unique(code(a), _p(a), _p(out), op, off.size - 1,
                           oo.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    return out[:k], oo
