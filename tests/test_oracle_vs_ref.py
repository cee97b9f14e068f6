"""The oracle restatement against the reference's OWN CPU fallback, compiled from its
include/parallax/stdpar.hpp into oracle/_ref (oracle/build_ref.sh).  Every reference
lookup misses (PARALLAX_FORCE_CPU=1 / no GPU), so each call runs the reference's host
loop; results must be bit-identical to the oracle's, including float folds, the
libstdc++ partition order and std::unique."""
import ctypes
import os

import numpy as np
import pytest

import pxo
from parallax_amd import make_desc, consts as C
from parallax_amd.dtypes import NP_DTYPES, PX_F32, PX_F64, PX_I32, PX_I64, PX_U32, PX_U64

REF = pxo.load_ref()
pytestmark = pytest.mark.skipif(REF is None, reason="oracle/_ref not built (no reference checkout)")


@pytest.fixture(autouse=True, scope="module")
def _reference_lookups_miss():
    """Every lookup of the reference's funnels misses while these tests run, so each call
    takes the reference's host loop (the setting is read at each lookup and restored
    afterwards, so it never leaks into GPU tests of the same session)."""
    with pytest.MonkeyPatch.context() as mp:
        mp.setenv("PARALLAX_FORCE_CPU", "1")
        yield

NAMES = {PX_I32: ("i32", ctypes.c_int32), PX_U32: ("u32", ctypes.c_uint32),
         PX_I64: ("i64", ctypes.c_int64), PX_U64: ("u64", ctypes.c_uint64),
         PX_F32: ("f32", ctypes.c_float), PX_F64: ("f64", ctypes.c_double)}
ALL = list(NAMES)


def fn(name, dt, res, *args):
    f = getattr(REF, f"{name}_{NAMES[dt][0]}")
    f.restype = res
    f.argtypes = list(args)
    return f


def data(dt, n, seed):
    rng = np.random.default_rng(seed)
    if dt in (PX_F32, PX_F64):
        return (rng.standard_normal(n) * 100).astype(NP_DTYPES[dt])
    lo = -500 if dt in (PX_I32, PX_I64) else 0
    return rng.integers(lo, 500, n).astype(NP_DTYPES[dt])


@pytest.mark.parametrize("dt", ALL)
def test_reduce_bitwise(dt):
    a = data(dt, 100003, dt)
    f = fn("ref_reduce", dt, NAMES[dt][1], ctypes.c_void_p, ctypes.c_size_t)
    got = np.array([f(a.ctypes.data, a.size)], dtype=a.dtype)
    assert got.tobytes() == np.array([pxo.reduce(a)], dtype=a.dtype).tobytes()


def test_reduce_config1_f32_uniform():
    # 2^20 float32 uniform [0,1) from seed 0: the reference fallback fold == oracle fold
    from parallax_amd import datagen
    a = datagen.uniform01_f32(0, 1 << 20)
    f = fn("ref_reduce", PX_F32, ctypes.c_float, ctypes.c_void_p, ctypes.c_size_t)
    assert np.float32(f(a.ctypes.data, a.size)) == pxo.reduce(a)


@pytest.mark.parametrize("dt", ALL)
def test_scans_bitwise(dt):
    a = data(dt, 50001, 20 + dt)
    out = np.empty_like(a)
    fn("ref_inclusive_scan", dt, None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t)(
        a.ctypes.data, out.ctypes.data, a.size)
    assert out.tobytes() == pxo.inclusive_scan(a).tobytes()
    init = NP_DTYPES[dt].type(3)
    fn("ref_exclusive_scan", dt, None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, NAMES[dt][1])(
        a.ctypes.data, out.ctypes.data, a.size, NAMES[dt][1](init))
    assert out.tobytes() == pxo.exclusive_scan(a, init).tobytes()


@pytest.mark.parametrize("dt", ALL)
def test_sort_unique_partition_bitwise(dt):
    a = data(dt, 40000, 40 + dt)
    s = a.copy()
    fn("ref_sort", dt, None, ctypes.c_void_p, ctypes.c_size_t)(s.ctypes.data, s.size)
    assert s.tobytes() == pxo.sort(a).tobytes()
    u = s.copy()
    k = fn("ref_unique", dt, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t)(u.ctypes.data, u.size)
    assert u[:k].tobytes() == pxo.unique(s).tobytes()
    c = NAMES[dt][1](NP_DTYPES[dt].type(100))
    d = make_desc(C.OP_FLAGS, dt, fn=C.PRED_GT, c0=100)
    out = np.empty_like(a)
    k = fn("ref_copy_if_gt", dt, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
           NAMES[dt][1])(a.ctypes.data, out.ctypes.data, a.size, c)
    assert out[:k].tobytes() == pxo.copy_if(d, a).tobytes()
    r = a.copy()
    k = fn("ref_remove_if_gt", dt, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t, NAMES[dt][1])(
        r.ctypes.data, r.size, c)
    assert r[:k].tobytes() == pxo.remove_if(d, a).tobytes()
    p = a.copy()
    k = fn("ref_partition_gt", dt, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t, NAMES[dt][1])(
        p.ctypes.data, p.size, c)
    po, pk = pxo.partition(d, a)
    assert k == pk and p.tobytes() == po.tobytes()


def test_copy_if_even_i32():
    from parallax_amd import datagen
    a = datagen.full_i32(1, 1 << 16, stream=6)
    out = np.empty_like(a)
    f = REF.ref_copy_if_even_i32
    f.restype = ctypes.c_size_t
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    k = f(a.ctypes.data, out.ctypes.data, a.size)
    d = make_desc(C.OP_FLAGS, PX_I32, fn=C.PRED_MOD_EQ, c0=2, c1=0)
    assert out[:k].tobytes() == pxo.copy_if(d, a).tobytes()
