"""pxd_inclusive_scan_noalloc (the inclusive scan that allocates nothing) against
pxd_inclusive_scan on the same inputs: bit-identical outputs (same kernel and tiling),
on two calls in a row, the second with a smaller input, for every element type; sizes
beyond the stated maximum are refused."""
import ctypes

import numpy as np
import pytest

import pxo
from parallax_amd import lib, check, consts as C
from parallax_amd.device import DeviceBuffer
from parallax_amd.dtypes import NP_DTYPES, PX_F32, PX_F64, PX_I32, PX_I64, PX_U32, PX_U64

pytestmark = pytest.mark.gpu

ALL = [PX_I32, PX_U32, PX_I64, PX_U64, PX_F32, PX_F64]


def gen(dt, n, seed):
    rng = np.random.default_rng(seed)
    if dt in (PX_F32, PX_F64):
        return (rng.standard_normal(n) * 3.0).astype(NP_DTYPES[dt])
    return rng.integers(0, 1 << 20, n).astype(NP_DTYPES[dt])


def scan_both(gpu_ctx, dt, a, shift):
    es = a.itemsize
    ib = DeviceBuffer(a.nbytes + 64)
    ib.upload(a, shift * es)
    ref_out = DeviceBuffer(a.nbytes + 64)
    na_out = DeviceBuffer(a.nbytes + 64)
    check(lib.pxd_inclusive_scan(gpu_ctx.h, dt, ib.addr(shift * es), ref_out.addr(shift * es), a.size), "scan")
    gpu_ctx.sync()
    check(lib.pxd_inclusive_scan_noalloc(dt, ib.addr(shift * es), na_out.addr(shift * es), a.size, None),
          "scan_noalloc")
    check(lib.pxd_device_sync(), "sync")
    return ref_out.download(a.dtype, a.size, shift * es), na_out.download(a.dtype, a.size, shift * es)


@pytest.mark.parametrize("dt", ALL)
def test_noalloc_scan_matches_the_allocating_scan(gpu_ctx, dt):
    # first call large (many tiles, look-back across them), second smaller and unaligned
    for n, shift, seed in ((3_000_017, 0, 1), (70_001, 1, 2)):
        a = gen(dt, n, seed + 10 * dt)
        ref, got = scan_both(gpu_ctx, dt, a, shift)
        assert got.tobytes() == ref.tobytes(), (dt, n)
        if dt not in (PX_F32, PX_F64):
            assert np.array_equal(got, pxo.inclusive_scan(a))


def test_noalloc_scan_refuses_more_than_the_limit(gpu_ctx):
    b = DeviceBuffer(64)
    assert lib.pxd_inclusive_scan_noalloc(PX_U32, b.ptr, b.ptr, (1 << 31) + 1, None) == C.E_ARG
    assert lib.pxd_inclusive_scan_noalloc(PX_U32, b.ptr, b.ptr, 0, None) == 0
