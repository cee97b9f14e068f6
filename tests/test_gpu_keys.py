"""Key transforms of the HIP engine, bit-exact against the oracle.

* Power-of-two remainder predicates, `x % m == c`, evaluate as one AND and one compare
  ((x & am) == ac, with the mask and value chosen from the sign of c); every case of C++
  truncated remainder is checked here: c = 0, 0 < c < m, -m < c < 0, |c| >= m, m = 1,
  for signed and unsigned types and values at the ends of the type.
* Signed integer radix sorts key on x - min(x): narrow ranges of either sign (whose
  upper digits are then shared and skipped), ranges at the ends of the type, the full
  range and all-equal arrays, plain (radix path) and batched (segmented radix path).
* Digits whose keys share some bits, plain radix path: every byte (random & mask) |
  fixed, and uniform [0, 1) doubles (whose top digit takes two values).
"""
import ctypes

import numpy as np
import pytest

import pxo
from parallax_amd import lib, check, make_desc, consts as C
from parallax_amd.device import DeviceBuffer
from parallax_amd.dtypes import NP_DTYPES, PX_F64, PX_I32, PX_I64, PX_U32, PX_U64

pytestmark = pytest.mark.gpu

INTS = [PX_I32, PX_I64, PX_U32, PX_U64]


def _edge_values(dt, n, seed):
    np_t = NP_DTYPES[dt]
    info = np.iinfo(np_t)
    rng = np.random.default_rng(seed)
    a = rng.integers(info.min, info.max, n, dtype=np_t, endpoint=True)
    a[: n // 4] = rng.integers(-70, 70, n // 4) if info.min < 0 else rng.integers(0, 70, n // 4)
    ends = np.array([info.min, info.min + 1, info.max, info.max - 1, 0, 1], dtype=np_t)
    a[n // 4 : n // 4 + ends.size] = ends
    if info.min < 0:
        a[n // 4 + ends.size : n // 4 + ends.size + 3] = np.array([-1, -2, -8], dtype=np_t)
    return a


@pytest.mark.parametrize("dt", INTS)
@pytest.mark.parametrize("m", [1, 2, 8, 1 << 20])
def test_copy_if_power_of_two_remainders(gpu_ctx, dt, m):
    signed = dt in (PX_I32, PX_I64)
    cs = sorted({0, 1, m - 1, m, m + 3, 5} | ({-1, -(m - 1), -m, -m - 2} if signed else set()))
    n = 70_001
    a = _edge_values(dt, n, 11 + m + dt)
    ib = DeviceBuffer.from_array(a)
    ob = DeviceBuffer(a.nbytes)
    cnt = DeviceBuffer(8)
    for c in cs:
        d = make_desc(C.OP_FLAGS, dt, fn=C.PRED_MOD_EQ, c0=m, c1=c)
        check(lib.pxd_copy_if(gpu_ctx.h, ctypes.byref(d), ib.ptr, ob.ptr, n, cnt.ptr), "copy_if")
        gpu_ctx.sync()
        k = int(cnt.download(np.uint64, 1)[0])
        ref = pxo.copy_if(d, a)
        assert k == ref.size, (m, c)
        assert ob.download(a.dtype, k).tobytes() == ref.tobytes(), (m, c)


def _signed_cases(dt, n, rng):
    np_t = NP_DTYPES[dt]
    info = np.iinfo(np_t)
    return {
        "narrow [-1000, 1000]": rng.integers(-1000, 1001, n).astype(np_t),
        "negative [-5000, -4000)": rng.integers(-5000, -4000, n).astype(np_t),
        "[-1, 0]": rng.integers(-1, 1, n, endpoint=True).astype(np_t),
        "low end": rng.integers(info.min, info.min + 300, n, dtype=np_t),
        "high end": rng.integers(info.max - 300, info.max, n, dtype=np_t, endpoint=True),
        "full range": rng.integers(info.min, info.max, n, dtype=np_t, endpoint=True),
        "both ends": np.where(rng.random(n) < 0.5, info.min, info.max).astype(np_t),
        "all equal": np.full(n, -7, dtype=np_t),
    }


@pytest.mark.parametrize("dt", [PX_I32, PX_I64])
def test_signed_radix_key_base(gpu_ctx, dt):
    n = (1 << 20) + 3  # the radix path for both widths, a partial last tile
    rng = np.random.default_rng(21 + dt)
    for name, a in _signed_cases(dt, n, rng).items():
        b = DeviceBuffer.from_array(a)
        check(lib.pxd_sort(gpu_ctx.h, dt, b.ptr, n), "sort")
        gpu_ctx.sync()
        assert b.download(a.dtype, n).tobytes() == pxo.sort(a).tobytes(), name


@pytest.mark.parametrize("dt", [PX_I32, PX_I64])
def test_signed_segmented_radix_key_base(gpu_ctx, dt):
    """One key base for the batch: segments of different ranges (narrow, negative, at
    the ends of the type) sorted together, and each range batched alone."""
    rng = np.random.default_rng(31 + dt)
    parts = list(_signed_cases(dt, 60_000, rng).values())
    lens = [3_000, 60_000, 1_025, 20_000, 513, 45_000, 9_000, 2_048]
    segs = [p[:l] for p, l in zip(parts, lens)]
    for batch in (segs, segs[:1], segs[2:3], segs[4:5]):
        a = np.concatenate(batch)
        off = np.concatenate([[0], np.cumsum([s.size for s in batch])]).astype(np.uint64)
        b = DeviceBuffer.from_array(a)
        check(lib.pxb_seg_sort(gpu_ctx.h, dt, b.ptr, off.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), len(batch)),
              "seg_sort")
        gpu_ctx.sync()
        assert b.download(a.dtype, a.size).tobytes() == pxo.seg_sort(a, off).tobytes()


@pytest.mark.parametrize("dt", [PX_U32, PX_U64, PX_I32])
@pytest.mark.parametrize("mask_set", [(0x0F, 0x30), (0x81, 0x00), (0x01, 0xFE), (0x7F, 0x80)])
def test_radix_partially_varying_digits(gpu_ctx, dt, mask_set):
    """Digits whose keys share some bits (a digit all keys share is skipped): each byte
    of every key is (random & mask) | fixed; full tiles and a partial last tile."""
    m, f = mask_set
    np_t = NP_DTYPES[dt]
    size = np.dtype(np_t).itemsize
    n = (1 << 20) + 5  # the radix path for both widths, a partial last tile
    rng = np.random.default_rng(41 + dt + m)
    b = (rng.integers(0, 256, (n, size), dtype=np.uint8) & m) | f
    a = b.view(np_t).reshape(n).copy()
    d = DeviceBuffer.from_array(a)
    check(lib.pxd_sort(gpu_ctx.h, dt, d.ptr, n), "sort")
    gpu_ctx.sync()
    assert d.download(a.dtype, n).tobytes() == pxo.sort(a).tobytes()


def test_radix_f64_unit_interval_top_digit(gpu_ctx):
    """Uniform [0, 1) doubles (the top digit takes two values), with 0, a tiny value and
    the largest double below 1."""
    n = (1 << 21) + 77
    a = np.random.default_rng(43).random(n)
    a[:3] = [0.0, 2.0 ** -40, 1.0 - 2.0 ** -53]
    d = DeviceBuffer.from_array(a)
    check(lib.pxd_sort(gpu_ctx.h, PX_F64, d.ptr, n), "sort")
    gpu_ctx.sync()
    assert d.download(a.dtype, n).tobytes() == pxo.sort(a).tobytes()
