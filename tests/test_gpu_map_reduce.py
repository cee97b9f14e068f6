"""Fused Σ map(x) -- the device pass behind the transform_reduce / count_if / all_of /
any_of / none_of funnels -- against the oracle's transform followed by the oracle's
reduce of the mapped array.  Integers bit-exact (wrap-around sums); floats within
2*ceil(log2 n)*eps*Σ|map(x)| of the double-double sum of the mapped values."""
import ctypes
import math

import numpy as np
import pytest

import pxo
from parallax_amd import lib, check, make_desc, descriptor_words, consts as C
from parallax_amd.device import DeviceBuffer
from parallax_amd.dtypes import NP_DTYPES, PX_F32, PX_F64, PX_I32, PX_I64, PX_U32, PX_U64

pytestmark = pytest.mark.gpu

ALL = [PX_I32, PX_U32, PX_I64, PX_U64, PX_F32, PX_F64]
FLOATS = (PX_F32, PX_F64)
MAPS = [(C.MAP_IDENTITY, 0, 0), (C.MAP_FILL, 9, 0), (C.MAP_ADD, 3, 0), (C.MAP_MUL, -2, 0),
        (C.MAP_AXPB, 3, -1), (C.MAP_NEG, 0, 0), (C.MAP_ABS, 0, 0), (C.MAP_SQUARE, 0, 0)]
WIDENING = [(PX_I32, PX_I64), (PX_U32, PX_U64), (PX_F32, PX_F64)]


def gen(dt, n, seed):
    rng = np.random.default_rng(seed)
    if dt in FLOATS:
        return (rng.standard_normal(n) * 3.0).astype(NP_DTYPES[dt])
    if dt in (PX_I32, PX_I64):
        return rng.integers(-1000, 1000, n).astype(NP_DTYPES[dt])
    return rng.integers(0, 1000, n).astype(NP_DTYPES[dt])


def bound(dt, n):
    return 2 * max(1, math.ceil(math.log2(max(n, 2)))) * float(np.finfo(NP_DTYPES[dt]).eps)


def run_fused(ctx, d, a, tout, shift):
    b = DeviceBuffer(a.nbytes + 64)
    b.upload(a, shift * a.itemsize)
    r = DeviceBuffer(16)
    check(lib.pxd_transform_reduce(ctx.h, ctypes.byref(d), b.addr(shift * a.itemsize), a.size, r.ptr),
          "transform_reduce")
    ctx.sync()
    return r.download(NP_DTYPES[tout], 1)[0]


def check_sum(got, y, tout):
    if tout in FLOATS:
        hi, lo, ab = pxo.sum_exact(y)
        assert abs(float(got) - (hi + lo)) <= bound(tout, y.size) * ab + 1e-300, (got, hi + lo)
    else:
        assert got == pxo.reduce(y), (got, pxo.reduce(y))


@pytest.mark.parametrize("dt", ALL)
@pytest.mark.parametrize("mp", MAPS)
@pytest.mark.parametrize("n", [0, 1, 7, 4099, 1_000_003])
def test_same_type_maps(gpu_ctx, dt, mp, n):
    fn, c0, c1 = mp
    if dt in (PX_U32, PX_U64):
        c0, c1 = abs(c0), abs(c1)
    a = gen(dt, n, 1000 + dt + 10 * fn)
    d = make_desc(C.OP_TRANSFORM, dt, aux=dt, fn=fn, c0=c0, c1=c1)
    got = run_fused(gpu_ctx, d, a, dt, n % 3)
    check_sum(got, pxo.transform(d, a, a.dtype), dt)


@pytest.mark.parametrize("pair", WIDENING)
@pytest.mark.parametrize("mp", MAPS)
def test_widening_maps(gpu_ctx, pair, mp):
    tin, tout = pair
    fn, c0, c1 = mp
    if tin == PX_U32:
        c0, c1 = abs(c0), abs(c1)
    a = gen(tin, 300_007, 1100 + fn)
    d = make_desc(C.OP_TRANSFORM, tin, aux=tout, fn=fn, c0=c0, c1=c1, const_dtype=tout)
    got = run_fused(gpu_ctx, d, a, tout, 1)
    check_sum(got, pxo.transform(d, a, NP_DTYPES[tout]), tout)


@pytest.mark.parametrize("tin", ALL)
@pytest.mark.parametrize("tout", ALL)
def test_converting_identity(gpu_ctx, tin, tout):
    a = gen(tin, 100_003, 1200 + tin)
    if tin in FLOATS and tout in (PX_U32, PX_U64):
        a = np.abs(a)  # float -> unsigned of a negative value is undefined in C++: not compared
    d = make_desc(C.OP_TRANSFORM, tin, aux=tout, fn=C.MAP_IDENTITY)
    got = run_fused(gpu_ctx, d, a, tout, 2)
    check_sum(got, pxo.transform(d, a, NP_DTYPES[tout]), tout)


PREDS = [(C.PRED_MOD_EQ, 2, 0, 0), (C.PRED_MOD_EQ, 8, -3, 0), (C.PRED_MOD_EQ, 3, 1, 0),
         (C.PRED_GT, 0, 0, 0), (C.PRED_IN_RANGE, -5, 100, 0), (C.PRED_LT, 10, 0, C.FLAG_NEGATE)]


@pytest.mark.parametrize("dt", ALL)
@pytest.mark.parametrize("pr", PREDS)
@pytest.mark.parametrize("tout", [PX_I32, PX_U32, PX_I64, PX_U64])
def test_count_if(gpu_ctx, dt, pr, tout):
    fn2, c0, c1, flags = pr
    if dt in (PX_U32, PX_U64):
        c0, c1 = abs(c0), abs(c1)
    a = gen(dt, 500_009, 1300 + dt)
    d = make_desc(C.OP_TRANSFORM, dt, aux=tout, fn=C.MAP_PRED, fn2=fn2, flags=flags, c0=c0, c1=c1)
    got = run_fused(gpu_ctx, d, a, tout, 1)
    fl = make_desc(C.OP_FLAGS, dt, fn=fn2, flags=flags, c0=c0, c1=c1)
    assert int(got) == pxo.count_if(fl, a)


def test_c_abi_fused_and_two_pass(gpu_ctx):
    """parallax_transform_reduce with host operands: a fused map (i32 -> i64 square) and
    one the fused kernel does not cover (i64 -> f64 axpb: transform + reduce on the
    device); unsupported maps are refused by the device entry point."""
    def load(d):
        w, n = descriptor_words(d)
        k = lib.parallax_kernel_load(w, n)
        assert k
        return k
    a = gen(PX_I32, 200_003, 1400)
    sq = make_desc(C.OP_TRANSFORM, PX_I32, aux=PX_I64, fn=C.MAP_SQUARE)
    r = np.zeros(1, dtype=np.int64)
    lib.parallax_transform_reduce(load(sq), load(make_desc(C.OP_REDUCE, PX_I64)), a.ctypes.data, a.size, 4, 8,
                                  r.ctypes.data)
    assert r[0] == int(np.sum(a.astype(np.int64) ** 2))
    b = gen(PX_I64, 200_003, 1401)
    ax = make_desc(C.OP_TRANSFORM, PX_I64, aux=PX_F64, fn=C.MAP_AXPB, c0=2.0, c1=1.0, const_dtype=PX_F64)
    rf = np.zeros(1, dtype=np.float64)
    lib.parallax_transform_reduce(load(ax), load(make_desc(C.OP_REDUCE, PX_F64)), b.ctypes.data, b.size, 8, 8,
                                  rf.ctypes.data)
    y = pxo.transform(ax, b, np.float64)
    hi, lo, ab = pxo.sum_exact(y)
    assert abs(rf[0] - (hi + lo)) <= bound(PX_F64, y.size) * ab
    d = DeviceBuffer.from_array(b)
    o = DeviceBuffer(16)
    assert lib.pxd_transform_reduce(gpu_ctx.h, ctypes.byref(ax), d.ptr, b.size, o.ptr) != 0
