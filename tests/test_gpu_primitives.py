"""Parity of the HIP primitives (device-resident entry points) with the oracle, for all
six element types, tile-boundary sizes, misaligned pointers and in-place calls.

Integer, sort, compaction and unique results must be bit-exact.  Float reduce and scan
must satisfy |err| <= 2*ceil(log2 n)*eps*sum|x| against the double-double oracle (eps =
machine epsilon of the element type; for scans sum|x| is taken over each prefix)."""
import ctypes
import math
import os

import numpy as np
import pytest

import pxo
from parallax_amd import lib, check, make_desc, descriptor_words, consts as C
from parallax_amd.device import DeviceBuffer
from parallax_amd.dtypes import NP_DTYPES, PX_F32, PX_F64, PX_I32, PX_I64, PX_U32, PX_U64, dtype_code

pytestmark = pytest.mark.gpu

ALL = [PX_I32, PX_U32, PX_I64, PX_U64, PX_F32, PX_F64]
SIZES = [1, 2, 3, 7, 64, 255, 256, 257, 1000, 2047, 2048, 2049, 4095, 4096, 4097, 65537, 1_000_003]
FLOATS = (PX_F32, PX_F64)


def gen(dt, n, seed, lo=None, hi=None):
    rng = np.random.default_rng(seed)
    if dt in FLOATS:
        return (rng.standard_normal(n) * 3.0).astype(NP_DTYPES[dt])
    if lo is None:
        lo = -(1 << 30) if dt in (PX_I32, PX_I64) else 0
        hi = 1 << 30
    return rng.integers(lo, hi, n).astype(NP_DTYPES[dt])


def bound(dt, n):
    eps = np.finfo(NP_DTYPES[dt]).eps
    return 2 * max(1, math.ceil(math.log2(max(n, 2)))) * eps


def upload(a, shift=0):
    """Device copy of `a` starting `shift` elements past a 256-byte boundary."""
    b = DeviceBuffer(a.nbytes + shift * a.itemsize + 64)
    b.upload(a, shift * a.itemsize)
    return b, b.addr(shift * a.itemsize)


@pytest.mark.parametrize("dt", ALL)
@pytest.mark.parametrize("n", SIZES)
def test_reduce(gpu_ctx, dt, n):
    a = gen(dt, n, 100 + n + dt)
    for shift in (0, 1):
        buf, p = upload(a, shift)
        out = DeviceBuffer(16)
        check(lib.pxd_reduce(gpu_ctx.h, dt, p, n, out.ptr), "reduce")
        gpu_ctx.sync()
        got = out.download(a.dtype, 1)[0]
        if dt in FLOATS:
            hi, lo, ab = pxo.sum_exact(a)
            err = abs(float(got) - (hi + lo))
            assert err <= bound(dt, n) * ab + 1e-300, (n, err, ab)
        else:
            assert got == pxo.reduce(a), (n, shift)


def test_reduce_empty_writes_zero(gpu_ctx):
    out = DeviceBuffer.from_array(np.array([123], dtype=np.int64))
    check(lib.pxd_reduce(gpu_ctx.h, PX_I64, out.ptr, 0, out.ptr), "reduce")
    gpu_ctx.sync()
    assert out.download(np.int64, 1)[0] == 0


def check_scan(dt, a, got, exclusive, init=0):
    if dt in FLOATS:
        val, ab = pxo.scan_exact(a, exclusive, init)
        err = np.abs(got.astype(np.float64) - val)
        assert np.all(err <= bound(dt, a.size) * ab + 1e-300), float((err / np.maximum(ab, 1e-300)).max())
    else:
        ref = pxo.exclusive_scan(a, init) if exclusive else pxo.inclusive_scan(a)
        assert np.array_equal(got, ref)


@pytest.mark.parametrize("dt", ALL)
@pytest.mark.parametrize("n", SIZES)
def test_inclusive_scan(gpu_ctx, dt, n):
    a = gen(dt, n, 200 + n + dt)
    ib, ip = upload(a, 0)
    ob, op = upload(np.zeros_like(a), 0)
    check(lib.pxd_inclusive_scan(gpu_ctx.h, dt, ip, op, n), "scan")
    gpu_ctx.sync()
    check_scan(dt, a, ob.download(a.dtype, n), False)


@pytest.mark.parametrize("dt", ALL)
@pytest.mark.parametrize("n", [1, 5, 4097, 100_003])
@pytest.mark.parametrize("shifts", [(1, 1), (1, 2), (3, 0)])
def test_scan_misaligned(gpu_ctx, dt, n, shifts):
    a = gen(dt, n, 300 + n + dt)
    ib, ip = upload(a, shifts[0])
    ob, op = upload(np.zeros_like(a), shifts[1])
    check(lib.pxd_inclusive_scan(gpu_ctx.h, dt, ip, op, n), "scan")
    gpu_ctx.sync()
    check_scan(dt, a, ob.download(a.dtype, n, shifts[1] * a.itemsize), False)


@pytest.mark.parametrize("dt", ALL)
@pytest.mark.parametrize("n", [1, 4096, 70001])
def test_exclusive_scan_with_init_and_in_place(gpu_ctx, dt, n):
    a = gen(dt, n, 400 + n + dt)
    init = np.array([5], dtype=a.dtype)
    b, p = upload(a, 0)
    check(lib.pxd_exclusive_scan(gpu_ctx.h, dt, p, p, n, init.ctypes.data), "xscan")
    gpu_ctx.sync()
    check_scan(dt, a, b.download(a.dtype, n), True, init[0])


def test_scan_u32_many_tiles_lookback_window(gpu_ctx):
    # more than 64 predecessors per look-back window, small values, wraps checked exactly
    a = gen(PX_U32, (1 << 22) + 13, 7, 0, 16)
    b, p = upload(a)
    o, q = upload(np.zeros_like(a))
    for _ in range(3):  # repeated launches on one context exercise the epoch / ticket reset
        check(lib.pxd_inclusive_scan(gpu_ctx.h, PX_U32, p, q, a.size), "scan")
    gpu_ctx.sync()
    assert np.array_equal(o.download(np.uint32, a.size), np.cumsum(a, dtype=np.uint32))


SORT_SIZES = [0, 1, 2, 3, 17, 1000, 1024, 4096, 8191, 8192, 8193, 16383, 16384, 16385, 70000, (1 << 20) + 3]


@pytest.mark.parametrize("dt", ALL)
@pytest.mark.parametrize("n", SORT_SIZES)
def test_sort(gpu_ctx, dt, n):
    a = gen(dt, n, 500 + n + dt)
    if n > 10:
        a[::7] = a[0]  # duplicates
    b, p = upload(a, 1 if n % 2 else 0)
    check(lib.pxd_sort(gpu_ctx.h, dt, p, n), "sort")
    gpu_ctx.sync()
    got = b.download(a.dtype, n, (1 if n % 2 else 0) * a.itemsize)
    assert got.tobytes() == pxo.sort(a).tobytes()


def test_sort_extremes(gpu_ctx):
    a = np.array([np.inf, -np.inf, 0.0, 1e300, -1e-300, np.inf, 3.0] * 3000, dtype=np.float64)
    b, p = upload(a)
    check(lib.pxd_sort(gpu_ctx.h, PX_F64, p, a.size), "sort")
    gpu_ctx.sync()
    assert b.download(np.float64, a.size).tobytes() == np.sort(a).tobytes()
    u = np.array([0xFFFFFFFF, 0, 0xFFFFFFFF, 5] * 5001, dtype=np.uint32)
    b, p = upload(u)
    check(lib.pxd_sort(gpu_ctx.h, PX_U32, p, u.size), "sort")
    gpu_ctx.sync()
    assert np.array_equal(b.download(np.uint32, u.size), np.sort(u))



# whole 16K-key (8-byte) and 20K-key (4-byte) tiles, partial last tiles
RADIX_SIZES = [1 << 20, 64 * 20480, (1 << 20) + 3, 3 * (1 << 20) + 7, (1 << 22) + 1]


@pytest.mark.parametrize("dt", ALL)
@pytest.mark.parametrize("n", RADIX_SIZES)
def test_sort_radix(gpu_ctx, dt, n):
    """Arrays at or above the radix threshold (full range of the type's values)."""
    rng = np.random.default_rng(900 + n % 97 + dt)
    if dt in FLOATS:
        a = (rng.standard_normal(n) * np.exp(rng.uniform(-30, 30, n))).astype(NP_DTYPES[dt])
    else:
        info = np.iinfo(NP_DTYPES[dt])
        a = rng.integers(info.min, info.max, n, dtype=NP_DTYPES[dt], endpoint=True)
    a[::5] = a[3]  # duplicates
    shift = n % 2
    b, p = upload(a, shift)
    check(lib.pxd_sort(gpu_ctx.h, dt, p, n), "sort")
    gpu_ctx.sync()
    got = b.download(a.dtype, n, shift * a.itemsize)
    assert got.tobytes() == pxo.sort(a).tobytes()


@pytest.mark.parametrize("dt", [PX_I32, PX_U32, PX_F32])
@pytest.mark.parametrize("n", [1 << 17, (1 << 17) + 3, 1 << 18, (1 << 18) + 5, 700_001, (1 << 20) - 1])
def test_sort_radix_4byte_below_2p20(gpu_ctx, dt, n):
    """4-byte keys take the radix path from 2^17 keys (256-thread digit-pass blocks below
    2^19, 512 below 2^20); 8-byte ones from 2^20."""
    a = gen(dt, n, 950 + n % 89 + dt)
    a[::3] = a[1]  # duplicates
    b, p = upload(a, n % 2)
    check(lib.pxd_sort(gpu_ctx.h, dt, p, n), "sort")
    gpu_ctx.sync()
    assert b.download(a.dtype, n, (n % 2) * a.itemsize).tobytes() == pxo.sort(a).tobytes()


@pytest.mark.parametrize("dt", ALL)
def test_sort_radix_patterns(gpu_ctx, dt):
    """All-equal, sorted, reversed and extreme keys through the radix path."""
    n = (1 << 20) + 5
    np_t = NP_DTYPES[dt]
    if dt in FLOATS:
        fi = np.finfo(np_t)
        ext = np.array([np.inf, -np.inf, fi.max, -fi.max, fi.tiny, -fi.tiny, fi.tiny / 4, 0.0, 1.0, -1.0], np_t)
    else:
        ii = np.iinfo(np_t)
        ext = np.array([ii.min, ii.max, 0, 1, ii.max - 1, ii.min + 1] + ([-1] if ii.min < 0 else []), np_t)
    base = gen(dt, n, 77 + dt)
    cases = [np.full(n, base[0]), np.sort(base), np.sort(base)[::-1].copy(), np.resize(ext, n)]
    for a in cases:
        a = np.ascontiguousarray(a, dtype=np_t)
        b, p = upload(a)
        check(lib.pxd_sort(gpu_ctx.h, dt, p, n), "sort")
        gpu_ctx.sync()
        got = b.download(np_t, n)
        assert got.tobytes() == pxo.sort(a).tobytes()


def _shared_digit_keys(dt, kind, n, rng):
    """Keys whose encoded radix digits are constant at some positions (those digit passes
    are skipped; with an odd number left the first skipped one copies the keys)."""
    if kind == "byte1":  # only digit 1 varies: one pass, copy before it
        return (rng.integers(0, 256, n, dtype=np.uint64) << 8).astype(NP_DTYPES[dt])
    if kind == "low24":  # top digit constant: three passes, copy after them
        return rng.integers(0, 1 << 20, n).astype(NP_DTYPES[dt])
    if kind == "mid":  # digit 2 constant: copy between passes
        x = rng.integers(0, 1 << 32, n, dtype=np.uint64)
        return ((x & 0xFF00FFFF) | 0x5A0000).astype(NP_DTYPES[dt])
    if kind == "low32":  # 8-byte keys below 2^32: four of eight passes
        return rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(NP_DTYPES[dt])
    if kind == "low40":  # five of eight passes
        return rng.integers(0, 1 << 40, n, dtype=np.uint64).astype(NP_DTYPES[dt])
    if kind == "top":  # only the top digit varies
        return (rng.integers(0, 128, n, dtype=np.uint64) << 56).astype(NP_DTYPES[dt])
    if kind == "unit":  # floats in [1, 2): sign and exponent constant
        return (1.0 + rng.random(n)).astype(NP_DTYPES[dt])
    if kind == "ints":  # small non-negative integers as floats: low mantissa digits zero
        return rng.integers(0, 256, n).astype(NP_DTYPES[dt])
    raise ValueError(kind)


SHARED_DIGIT_CASES = [(PX_U32, "byte1"), (PX_U32, "low24"), (PX_U32, "mid"), (PX_I32, "low24"),
                      (PX_U64, "low32"), (PX_I64, "low40"), (PX_U64, "top"), (PX_U64, "byte1"),
                      (PX_F32, "unit"), (PX_F64, "unit"), (PX_F64, "ints"), (PX_F32, "ints")]


@pytest.mark.parametrize("dt,kind", SHARED_DIGIT_CASES)
def test_sort_radix_shared_digits(gpu_ctx, dt, kind):
    """Radix sorts that skip the digit positions every key shares, on an odd and an
    even count of remaining passes, and twice in a row on one stream."""
    n = (1 << 20) + 7
    rng = np.random.default_rng(31 + dt)
    for rep in range(2):
        a = _shared_digit_keys(dt, kind, n + rep, rng)
        b, p = upload(a, rep)
        check(lib.pxd_sort(gpu_ctx.h, dt, p, n + rep), "sort")
        gpu_ctx.sync()
        got = b.download(a.dtype, n + rep, rep * a.itemsize)
        assert got.tobytes() == pxo.sort(a).tobytes(), (kind, rep)


def test_sort_radix_signed_zero_nan(gpu_ctx):
    """-0.0 / +0.0 are equal under operator<, NaNs unordered: the radix order puts -0.0
    before +0.0 and NaNs by sign at the ends; everything else matches std::sort."""
    if "sort_radix_min" in os.environ.get("PARALLAX_TUNE", ""):
        pytest.skip("the radix threshold is overridden (this order is the radix path's)")
    n = (1 << 20) + 1
    rng = np.random.default_rng(5)
    a = rng.standard_normal(n).astype(np.float32)
    a[::3] = 0.0
    a[1::3] = -0.0
    a[7::1001] = np.nan
    a[9::1003] = -np.nan
    b, p = upload(a)
    check(lib.pxd_sort(gpu_ctx.h, PX_F32, p, n), "sort")
    gpu_ctx.sync()
    got = b.download(np.float32, n)
    bits = got.view(np.uint32)
    neg_nan = np.isnan(got) & (bits >> 31 == 1)
    pos_nan = np.isnan(got) & (bits >> 31 == 0)
    k_neg, k_pos = int(neg_nan.sum()), int(pos_nan.sum())
    assert neg_nan[:k_neg].all() and pos_nan[n - k_pos:].all()
    mid = got[k_neg:n - k_pos]
    ref = np.sort(a[~np.isnan(a)])
    assert np.array_equal(mid, ref)  # equal values (-0.0 == +0.0)
    z = mid[mid == 0].view(np.uint32)
    assert (np.diff((z >> 31) ^ 1) >= 0).all()  # every -0.0 before every +0.0
    assert sorted(got.view(np.uint32).tolist()) == sorted(a.view(np.uint32).tolist())  # permutation

PREDS = [
    (C.PRED_GT, 0, 0, 0), (C.PRED_LT, 100, 0, C.FLAG_NEGATE), (C.PRED_MOD_EQ, 2, 0, 0),
    (C.PRED_IN_RANGE, -1000, 1000, 0), (C.PRED_NONZERO, 0, 0, 0), (C.PRED_TRUE, 0, 0, 0),
    (C.PRED_EQ, 12345, 0, 0),
]


@pytest.mark.parametrize("dt", ALL)
@pytest.mark.parametrize("pred", PREDS)
@pytest.mark.parametrize("n", [1, 4097, 300_001])
def test_copy_if(gpu_ctx, dt, pred, n):
    fn, c0, c1, fl = pred
    if dt in (PX_U32, PX_U64) and c0 < 0:
        c0 = 0
    a = gen(dt, n, 600 + n + dt, *((None, None) if dt in FLOATS else ((-5000, 5000) if dt in (PX_I32, PX_I64) else (0, 5000))))
    d = make_desc(C.OP_FLAGS, dt, fn=fn, flags=fl, c0=c0, c1=c1)
    ib, ip = upload(a, 1)
    sentinel = np.full(n, 7, dtype=a.dtype)
    ob, op = upload(sentinel)
    cnt = DeviceBuffer(8)
    check(lib.pxd_copy_if(gpu_ctx.h, ctypes.byref(d), ip, op, n, cnt.ptr), "copy_if")
    gpu_ctx.sync()
    k = int(cnt.download(np.uint64, 1)[0])
    ref = pxo.copy_if(d, a)
    assert k == ref.size
    got = ob.download(a.dtype, n)
    assert got[:k].tobytes() == ref.tobytes()
    assert np.all(got[k:] == 7)  # nothing written past the kept count


@pytest.mark.parametrize("dt", ALL)
@pytest.mark.parametrize("n", [1, 2, 4097, 200_003])
def test_unique(gpu_ctx, dt, n):
    a = np.sort(gen(dt, n, 700 + dt, *((None, None) if dt in FLOATS else (0, 50))))
    if dt in FLOATS:
        a = np.round(a, 1).astype(a.dtype)
    d = make_desc(C.OP_FLAGS, dt, fn=C.PRED_UNIQUE)
    ib, ip = upload(a, 3)
    ob, op = upload(np.zeros_like(a))
    cnt = DeviceBuffer(8)
    check(lib.pxd_copy_if(gpu_ctx.h, ctypes.byref(d), ip, op, n, cnt.ptr), "unique")
    gpu_ctx.sync()
    k = int(cnt.download(np.uint64, 1)[0])
    ref = pxo.unique(a)
    assert k == ref.size and ob.download(a.dtype, k).tobytes() == ref.tobytes()


@pytest.mark.parametrize("dt", ALL)
@pytest.mark.parametrize("kind", ["all", "half", "unique_most", "unique_few"])
def test_compaction_in_place(gpu_ctx, dt, kind):
    """input == output (remove_if / unique funnels): every tile writes below its own end,
    after the tiles before it have read theirs; many tiles, misaligned start."""
    n = 1_000_003
    if kind.startswith("unique"):
        hi = 1 << 40 if kind == "unique_most" else 300
        a = np.sort(gen(dt, n, 950 + dt, *((None, None) if dt in FLOATS else (0, min(hi, 1 << 30)))))
        if dt in FLOATS and kind == "unique_few":
            a = np.round(a, 1).astype(a.dtype)
        d = make_desc(C.OP_FLAGS, dt, fn=C.PRED_UNIQUE)
        ref = pxo.unique(a)
    else:
        a = gen(dt, n, 960 + dt)
        d = make_desc(C.OP_FLAGS, dt, fn=C.PRED_TRUE) if kind == "all" else \
            make_desc(C.OP_FLAGS, dt, fn=C.PRED_GT, c0=0 if dt not in (PX_U32, PX_U64) else 1 << 29)
        ref = pxo.copy_if(d, a)
    b, p = upload(a, 3)
    cnt = DeviceBuffer(8)
    check(lib.pxd_copy_if(gpu_ctx.h, ctypes.byref(d), p, p, n, cnt.ptr), "copy_if in place")
    gpu_ctx.sync()
    k = int(cnt.download(np.uint64, 1)[0])
    got = b.download(a.dtype, n, 3 * a.itemsize)
    assert k == ref.size and got[:k].tobytes() == ref.tobytes()
    assert got[k:].tobytes() == a[k:].tobytes()  # nothing written past the kept count


PART_SIZES = [1, 2, 3, 64, 4097, 65_536, 300_001, 1_000_003]


@pytest.mark.parametrize("dt", ALL)
@pytest.mark.parametrize("n", PART_SIZES)
@pytest.mark.parametrize("in_place", [False, True])
def test_partition_std_order(gpu_ctx, dt, n, in_place):
    """Bit-exact with std::partition (the partition funnel's fallback, reference
    stdpar.hpp:369), out of place and in place, misaligned input."""
    a = gen(dt, n, 800 + n + dt)
    d = make_desc(C.OP_FLAGS, dt, fn=C.PRED_GT, c0=0 if dt not in (PX_U32, PX_U64) else 1 << 29)
    ib, ip = upload(a, 1)
    if in_place:
        ob, op = ib, ip
    else:
        ob, op = upload(np.zeros_like(a), 2)
    nt = DeviceBuffer(8)
    check(lib.pxd_partition(gpu_ctx.h, ctypes.byref(d), ip, op, n, nt.ptr), "partition")
    gpu_ctx.sync()
    k = int(nt.download(np.uint64, 1)[0])
    ref, rk = pxo.partition(d, a)
    got = ob.download(a.dtype, n, (1 if in_place else 2) * a.itemsize)
    assert k == rk
    assert got.tobytes() == ref.tobytes()


@pytest.mark.parametrize("dt", [PX_I32, PX_U64, PX_F64])
@pytest.mark.parametrize("frac", [0.0, 0.01, 0.5, 0.99, 1.0])
def test_partition_std_order_split_fractions(gpu_ctx, dt, frac):
    # every element kept, none kept and lopsided splits (m = 0 and tiny m)
    n = 200_003
    rng = np.random.default_rng(int(frac * 100) + dt)
    flags = rng.random(n) < frac
    a = np.where(flags, 1, -1).astype(NP_DTYPES[dt]) if dt != PX_U64 else np.where(flags, 2, 1).astype(np.uint64)
    a = (a * (np.arange(n) % 1000 + 1)).astype(a.dtype)
    d = make_desc(C.OP_FLAGS, dt, fn=C.PRED_GT, c0=0 if dt != PX_U64 else 1000)
    b, p = upload(a)
    nt = DeviceBuffer(8)
    check(lib.pxd_partition(gpu_ctx.h, ctypes.byref(d), p, p, n, nt.ptr), "partition")
    gpu_ctx.sync()
    ref, rk = pxo.partition(d, a)
    assert int(nt.download(np.uint64, 1)[0]) == rk
    assert b.download(a.dtype, n).tobytes() == ref.tobytes()


@pytest.mark.parametrize("dt", ALL)
@pytest.mark.parametrize("n", [1, 4097, 300_001])
@pytest.mark.parametrize("in_place", [False, True])
def test_stable_partition_mode(gpu_ctx, dt, n, in_place):
    """Scatter mode PX_SCATTER_STABLE_PARTITION of the C ABI: std::stable_partition."""
    a = gen(dt, n, 900 + n + dt)
    d = make_desc(C.OP_FLAGS, dt, fn=C.PRED_GT, c0=0 if dt not in (PX_U32, PX_U64) else 1 << 29)
    kf = lib.parallax_kernel_load(*descriptor_words(d))
    ks = lib.parallax_kernel_load(*descriptor_words(make_desc(C.OP_SCAN, dt)))
    ka = lib.parallax_kernel_load(*descriptor_words(make_desc(C.OP_SCAN_ADD, dt)))
    kx = lib.parallax_kernel_load(*descriptor_words(make_desc(C.OP_SCATTER, dt, aux=C.SCATTER_STABLE_PARTITION)))
    assert kf and ks and ka and kx
    ib, ip = upload(a)
    ob, op = (ib, ip) if in_place else upload(np.zeros_like(a))
    k = lib.parallax_copy_if(kf, ks, ka, kx, ip, op, n, a.itemsize, int(dt in FLOATS))
    sref, sk = pxo.partition(d, a, stable=True)
    assert k == sk
    assert ob.download(a.dtype, n).tobytes() == sref.tobytes()


MAPS = [(C.MAP_IDENTITY, 0, 0), (C.MAP_FILL, 9, 0), (C.MAP_ADD, 3, 0), (C.MAP_MUL, -2, 0),
        (C.MAP_AXPB, 3, -1), (C.MAP_NEG, 0, 0), (C.MAP_ABS, 0, 0), (C.MAP_SQUARE, 0, 0)]


@pytest.mark.parametrize("dt", ALL)
@pytest.mark.parametrize("mp", MAPS)
def test_transform_same_type(gpu_ctx, dt, mp):
    fn, c0, c1 = mp
    if dt in (PX_U32, PX_U64):
        c0, c1 = abs(c0), abs(c1)
    a = gen(dt, 10007, 900 + dt, *((None, None) if dt in FLOATS else (-1000, 1000) if dt in (PX_I32, PX_I64) else (0, 1000)))
    d = make_desc(C.OP_TRANSFORM, dt, aux=dt, fn=fn, c0=c0, c1=c1)
    ib, ip = upload(a)
    ob, op = upload(np.zeros_like(a))
    check(lib.pxd_transform(gpu_ctx.h, ctypes.byref(d), ip, op, a.size, None, 0), "transform")
    gpu_ctx.sync()
    assert ob.download(a.dtype, a.size).tobytes() == pxo.transform(d, a, a.dtype).tobytes()


@pytest.mark.parametrize("pair", [(PX_F32, PX_F64), (PX_I32, PX_I64), (PX_I64, PX_F64), (PX_U32, PX_F32)])
def test_transform_converting_with_captures(gpu_ctx, pair):
    tin, tout = pair
    a = gen(tin, 5000, 950, *((None, None) if tin in FLOATS else (-100, 100) if tin in (PX_I32, PX_I64) else (0, 100)))
    d = make_desc(C.OP_TRANSFORM, tin, aux=tout, fn=C.MAP_AXPB)
    cap = np.array([2, 1], dtype=NP_DTYPES[tout]).tobytes()
    ib, ip = upload(a)
    ob, op = upload(np.zeros(a.size, dtype=NP_DTYPES[tout]))
    check(lib.pxd_transform(gpu_ctx.h, ctypes.byref(d), ip, op, a.size, cap, len(cap)), "transform")
    gpu_ctx.sync()
    ref = pxo.transform(d, a, NP_DTYPES[tout], cap)
    assert ob.download(NP_DTYPES[tout], a.size).tobytes() == ref.tobytes()


def test_search_entry_points(gpu_ctx):
    a = gen(PX_I32, 100_000, 960, -1000, 1000)
    a[77_777] = 5000
    a[88_888] = 5000
    a[1234] = -5000
    ib, ip = upload(a)
    idx = DeviceBuffer(8)
    d = make_desc(C.OP_FIND, PX_I32, fn=C.PRED_GT, c0=4000)
    check(lib.pxd_find(gpu_ctx.h, ctypes.byref(d), ip, a.size, idx.ptr), "find")
    gpu_ctx.sync()
    assert idx.download(np.uint64, 1)[0] == 77_777
    for want_max, want_last, expect in [(1, 0, 77_777), (1, 1, 88_888), (0, 0, 1234)]:
        check(lib.pxd_argminmax(gpu_ctx.h, PX_I32, ip, a.size, want_max, want_last, idx.ptr), "argminmax")
        gpu_ctx.sync()
        assert idx.download(np.uint64, 1)[0] == expect
    b = a.copy()
    b[99_000] += 1
    bb, bp = upload(b)
    check(lib.pxd_mismatch(gpu_ctx.h, PX_I32, ip, bp, a.size, idx.ptr), "mismatch")
    gpu_ctx.sync()
    assert idx.download(np.uint64, 1)[0] == 99_000


SEARCH_SIZES = [1, 3, 17, 1000, 4099, 65_537, 1_000_003, 3_000_017]


def _search(gpu_ctx, fn, *args):
    idx = DeviceBuffer(8)
    check(fn(gpu_ctx.h, *args, idx.ptr), "search")
    gpu_ctx.sync()
    return int(idx.download(np.uint64, 1)[0])


@pytest.mark.parametrize("dt", ALL)
@pytest.mark.parametrize("n", SEARCH_SIZES)
def test_search_against_oracle(gpu_ctx, dt, n):
    """find / argminmax / mismatch against the oracle: hits at the start, the end, inside
    and nowhere; ties both ways; pointers off a 16-byte boundary (mismatch with equal
    and with different offsets)."""
    rng = np.random.default_rng(990 + n % 61 + dt)
    t = NP_DTYPES[dt]
    a = rng.integers(0, 1000, n).astype(t)  # ties for argminmax
    big = make_desc(C.OP_FIND, dt, fn=C.PRED_GT, c0=5000)
    for shift in (0, 1, 3):
        for pos in sorted({0, n - 1, n // 2, (7 * n) // 9}):
            x = a.copy()
            x[pos] = t.type(7000)
            if pos + 5 < n:
                x[pos + 5] = t.type(7000)  # a later hit and a tie for the maximum
            b, p = upload(x, shift)
            got = _search(gpu_ctx, lib.pxd_find, ctypes.byref(big), p, n)
            assert got == pxo.find(big, x) == pos, (shift, pos)
        b, p = upload(a, shift)
        assert _search(gpu_ctx, lib.pxd_find, ctypes.byref(big), p, n) == pxo.find(big, a) == n
        for want_max in (0, 1):
            for want_last in (0, 1):
                got = _search(gpu_ctx, lib.pxd_argminmax, dt, p, n, want_max, want_last)
                assert got == pxo.argminmax(a, want_max, want_last), (shift, want_max, want_last)
        for shift_b in (shift, (shift + 1) % 4):
            y = a.copy()
            bb, pb = upload(y, shift_b)
            assert _search(gpu_ctx, lib.pxd_mismatch, dt, p, pb, n) == n
            k = int(rng.integers(0, n))
            y[k] += y.dtype.type(1)
            if k + 9 < n:
                y[k + 9] += y.dtype.type(1)
            bb, pb = upload(y, shift_b)
            assert _search(gpu_ctx, lib.pxd_mismatch, dt, p, pb, n) == pxo.mismatch(a, y) == k, (shift, shift_b)


def test_search_empty(gpu_ctx):
    b, p = upload(np.zeros(4, np.float32))
    d = make_desc(C.OP_FIND, PX_F32, fn=C.PRED_GT, c0=0.0)
    assert _search(gpu_ctx, lib.pxd_find, ctypes.byref(d), p, 0) == 0
    assert _search(gpu_ctx, lib.pxd_mismatch, PX_F32, p, p, 0) == 0
    assert _search(gpu_ctx, lib.pxd_argminmax, PX_F32, p, 0, 0, 0) == 0

def test_beyond_2p32_elements(gpu_ctx):
    """n = 2^32 + 4099 int32 elements (17 GB, filled on the device): reduce, inclusive
    scan and copy_if across the 32-bit index boundary, where compaction counts are 64-bit
    (two status words per tile).  Closed forms: every element is 0x01010101 except 20
    zeros straddling index 2^32."""
    n = (1 << 32) + 4099
    v = 0x01010101
    src = lib.pxd_malloc(n * 4 + 64)
    dst = lib.pxd_malloc(n * 4 + 64)
    word = lib.pxd_malloc(64)
    assert src and dst and word
    try:
        check(lib.pxd_memset(src, 1, n * 4), "memset")
        z0 = (1 << 32) - 10
        check(lib.pxd_memset(src + 4 * z0, 0, 20 * 4), "memset")

        def at(p, i, t=np.uint32):
            out = np.zeros(1, t)
            check(lib.pxd_memcpy_d2h(out.ctypes.data, p + i * out.itemsize, out.itemsize), "d2h")
            return out[0]

        check(lib.pxd_reduce(gpu_ctx.h, PX_U32, src, n, word), "reduce")
        gpu_ctx.sync()
        assert at(word, 0) == ((n - 20) * v) % (1 << 32)

        check(lib.pxd_inclusive_scan(gpu_ctx.h, PX_U32, src, dst, n), "scan")
        gpu_ctx.sync()
        for i in (0, 1, 12_345, z0 - 1, z0, z0 + 19, z0 + 20, (1 << 32), n - 1):
            kept_before = i + 1 - max(0, min(i + 1, z0 + 20) - z0)  # nonzero elements in [0, i]
            assert at(dst, i) == (kept_before * v) % (1 << 32), i

        nz = make_desc(C.OP_FLAGS, PX_U32, fn=C.PRED_NONZERO)
        check(lib.pxd_copy_if(gpu_ctx.h, ctypes.byref(nz), src, dst, n, word), "copy_if")
        gpu_ctx.sync()
        assert at(word, 0, np.uint64) == n - 20
        for i in (0, z0 - 1, z0, (1 << 32) - 1, n - 21):
            assert at(dst, i) == v, i
    finally:
        lib.pxd_free(src)
        lib.pxd_free(dst)
        lib.pxd_free(word)


@pytest.mark.parametrize("nbytes", [0, 1, 15, 16, 17, 4096, 1_000_003, 1 << 26])
@pytest.mark.parametrize("offs", [(0, 0), (3, 3), (1, 2), (4, 8), (8, 0)])
def test_copy_kernel(gpu_ctx, nbytes, offs):
    """pxd_copy / pxd_memcpy_d2d (the streaming-copy kernel): exact bytes for every
    alignment pair, nothing written outside the destination range."""
    rng = np.random.default_rng(nbytes + 7 * offs[0] + offs[1])
    src = rng.integers(0, 256, nbytes + 64, dtype=np.uint8)
    sb = DeviceBuffer.from_array(src)
    db = DeviceBuffer.from_array(np.full(nbytes + 64, 0xAB, dtype=np.uint8))
    check(lib.pxd_copy(gpu_ctx.h, db.addr(offs[1]), sb.addr(offs[0]), nbytes), "copy")
    gpu_ctx.sync()
    got = db.download(np.uint8, nbytes + 64)
    assert got[offs[1]:offs[1] + nbytes].tobytes() == src[offs[0]:offs[0] + nbytes].tobytes()
    assert np.all(got[:offs[1]] == 0xAB) and np.all(got[offs[1] + nbytes:] == 0xAB)
    db.upload(np.full(nbytes + 64, 0xCD, dtype=np.uint8))
    check(lib.pxd_memcpy_d2d(db.addr(offs[1]), sb.addr(offs[0]), nbytes), "memcpy d2d")
    got = db.download(np.uint8, nbytes + 64)
    assert got[offs[1]:offs[1] + nbytes].tobytes() == src[offs[0]:offs[0] + nbytes].tobytes()
    assert np.all(got[:offs[1]] == 0xCD) and np.all(got[offs[1] + nbytes:] == 0xCD)
