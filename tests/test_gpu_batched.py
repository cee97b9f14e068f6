"""Batched (segmented) forms, graph plans and the request server, against the oracle
applied per segment; plus "alone vs in the mix" identity."""
import ctypes
import threading

import numpy as np
import pytest

import pxo
from parallax_amd import lib, check, make_desc, consts as C, datagen
from parallax_amd._lib import PxsConfig, PxsRequest, PxsStats
from parallax_amd.device import Context, DeviceBuffer
from parallax_amd.dtypes import NP_DTYPES, PX_F32, PX_F64, PX_I32, PX_I64, PX_U32, PX_U64

pytestmark = pytest.mark.gpu

ALL = [PX_I32, PX_U32, PX_I64, PX_U64, PX_F32, PX_F64]
FLOATS = (PX_F32, PX_F64)


def seg_input(dt, seed, nseg=300, lo=1, hi=70_000):
    lens = datagen.log_uniform_lengths(seed, nseg, lo, hi)
    lens[::37] = 0  # empty segments
    lens[5] = 1
    off = datagen.offsets_from_lengths(lens)
    rng = np.random.default_rng(seed)
    n = int(off[-1])
    if dt in FLOATS:
        a = rng.standard_normal(n).astype(NP_DTYPES[dt])
    else:
        a = rng.integers(0, 1000, n).astype(NP_DTYPES[dt])
    return a, off


def u64p(off):
    return off.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def bound(dt, n):
    import math
    return 2 * max(1, math.ceil(math.log2(max(n, 2)))) * float(np.finfo(NP_DTYPES[dt]).eps)


@pytest.mark.parametrize("dt", ALL)
def test_seg_reduce(gpu_ctx, dt):
    a, off = seg_input(dt, 10 + dt)
    d = DeviceBuffer.from_array(a)
    out = DeviceBuffer(8 * (off.size - 1))
    check(lib.pxb_seg_reduce(gpu_ctx.h, dt, d.ptr, u64p(off), off.size - 1, out.ptr), "seg_reduce")
    gpu_ctx.sync()
    got = out.download(a.dtype, off.size - 1)
    if dt in FLOATS:
        hi, ab = pxo.seg_sum_exact(a, off)
        lens = np.diff(off.astype(np.int64))
        for s in range(off.size - 1):
            assert abs(float(got[s]) - hi[s]) <= bound(dt, lens[s]) * ab[s] + 1e-300, s
    else:
        assert np.array_equal(got, pxo.seg_reduce(a, off))


@pytest.mark.parametrize("dt", [PX_F32, PX_U64])
def test_seg_reduce_input_above_bit31(gpu_ctx, dt):
    """Segmented reduce of an input whose addresses have bit 31 of the low word set (the
    tile loads build a buffer base address from 32-bit halves; a sign-extended low half
    would corrupt it).  A 2 GiB + 128 MiB allocation always holds 64 MiB of such
    addresses in a row: at its start, or from the next low word 0x80000000."""
    a, off = seg_input(dt, 40 + dt)
    assert a.nbytes <= 48 << 20
    big = DeviceBuffer((1 << 31) + (128 << 20))
    low = big.ptr & 0xFFFFFFFF
    if low & 0x80000000 and (1 << 32) - low >= (64 << 20):
        shift = 0
    else:
        shift = ((1 << 31) - low) % (1 << 32)
    shift += a.itemsize  # not 16-byte aligned: the head and tail vectors are partial
    first, last = big.addr(shift), big.addr(shift) + a.nbytes - 1
    assert shift + a.nbytes <= big.nbytes
    assert first & 0x80000000 and last & 0x80000000 and first >> 32 == last >> 32
    big.upload(a, shift)
    out = DeviceBuffer(8 * (off.size - 1))
    check(lib.pxb_seg_reduce(gpu_ctx.h, dt, big.addr(shift), u64p(off), off.size - 1, out.ptr), "seg_reduce")
    gpu_ctx.sync()
    got = out.download(a.dtype, off.size - 1)
    if dt in FLOATS:
        hi, ab = pxo.seg_sum_exact(a, off)
        lens = np.diff(off.astype(np.int64))
        for s in range(off.size - 1):
            assert abs(float(got[s]) - hi[s]) <= bound(dt, lens[s]) * ab[s] + 1e-300, s
    else:
        assert np.array_equal(got, pxo.seg_reduce(a, off))
    big.free()


@pytest.mark.parametrize("dt", ALL)
def test_seg_scan(gpu_ctx, dt):
    a, off = seg_input(dt, 20 + dt)
    d = DeviceBuffer.from_array(a)
    o = DeviceBuffer(a.nbytes + 16)
    check(lib.pxb_seg_inclusive_scan(gpu_ctx.h, dt, d.ptr, o.ptr, u64p(off), off.size - 1), "seg_scan")
    gpu_ctx.sync()
    got = o.download(a.dtype, a.size)
    if dt in FLOATS:
        for s in range(off.size - 1):
            seg = a[off[s]:off[s + 1]]
            if seg.size == 0:
                continue
            val, ab = pxo.scan_exact(seg)
            err = np.abs(got[off[s]:off[s + 1]].astype(np.float64) - val)
            assert np.all(err <= bound(dt, seg.size) * ab + 1e-300), s
    else:
        assert np.array_equal(got, pxo.seg_inclusive_scan(a, off))


@pytest.mark.parametrize("dt", ALL)
def test_seg_sort(gpu_ctx, dt):
    a, off = seg_input(dt, 30 + dt, nseg=200, lo=1, hi=300_000)
    d = DeviceBuffer.from_array(a)
    check(lib.pxb_seg_sort(gpu_ctx.h, dt, d.ptr, u64p(off), off.size - 1), "seg_sort")
    gpu_ctx.sync()
    assert d.download(a.dtype, a.size).tobytes() == pxo.seg_sort(a, off).tobytes()



MIXED_LENS = [5, 20_000, 0, 8_192, 8_193, 16_384, 16_385, 3, 150_001, 40_000, 1, 70_000, 9_000, 33_000]
# Long segments that fill 16K-key tiles (one ragged): 1024-thread digit-pass blocks; the
# mixed lengths above hold 14 % more slots in 16K tiles than in 8K ones: 512-thread blocks.
LARGE_LENS = [5, 65_536, 131_072, 16_384, 0, 49_152, 300, 32_769, 98_304]


def radix_seg_input(dt, seed, kind, lens=MIXED_LENS):
    """Short and long segments mixed (long ones take the segmented radix path)."""
    rng = np.random.default_rng(seed)
    lens = np.array(lens, dtype=np.int64)
    off = datagen.offsets_from_lengths(lens)
    n = int(off[-1])
    t = NP_DTYPES[dt]
    if kind == "full":
        if dt in FLOATS:
            a = (rng.standard_normal(n) * np.exp(rng.uniform(-20, 20, n))).astype(t)
        else:
            info = np.iinfo(t)
            a = rng.integers(info.min, info.max, n, dtype=t, endpoint=True)
    elif kind == "narrow":  # 20 significant bits: three digits run (integers)
        a = rng.integers(0, 1 << 20, n).astype(t)
    elif kind == "byte":  # one digit varies: one pass plus one copy (integers: pass first)
        if dt == PX_F32:  # 1 + j/128: only digit 2 varies
            a = (1.0 + rng.integers(0, 128, n) / 128.0).astype(t)
        elif dt == PX_F64:  # 1 + j/16: only digit 6 varies
            a = (1.0 + rng.integers(0, 16, n) / 16.0).astype(t)
        else:
            a = rng.integers(0, 256, n).astype(t)
    else:  # "equal": each segment holds one value (values below 1000 across segments)
        a = np.repeat(rng.integers(0, 1000, lens.size).astype(t), lens)
    return a, off


@pytest.mark.parametrize("lens", ["mixed", "large"])
@pytest.mark.parametrize("kind", ["full", "narrow", "byte", "equal"])
@pytest.mark.parametrize("dt", ALL)
def test_seg_sort_radix(gpu_ctx, dt, kind, lens):
    a, off = radix_seg_input(dt, 60 + dt, kind, MIXED_LENS if lens == "mixed" else LARGE_LENS)
    d = DeviceBuffer.from_array(a)
    check(lib.pxb_seg_sort(gpu_ctx.h, dt, d.ptr, u64p(off), off.size - 1), "seg_sort")
    gpu_ctx.sync()
    assert d.download(a.dtype, a.size).tobytes() == pxo.seg_sort(a, off).tobytes()
    # again on the same context (scratch, counters and skip mask reused)
    b = a[::-1].copy()
    d2 = DeviceBuffer.from_array(b)
    check(lib.pxb_seg_sort(gpu_ctx.h, dt, d2.ptr, u64p(off), off.size - 1), "seg_sort")
    gpu_ctx.sync()
    assert d2.download(b.dtype, b.size).tobytes() == pxo.seg_sort(b, off).tobytes()


@pytest.mark.parametrize("dt", ALL)
def test_seg_sort_radix_unaligned(gpu_ctx, dt):
    """Segmented radix with the array one element past a 16-byte boundary (the counting
    pass reads each tile's unaligned head and tail one key at a time)."""
    a, off = radix_seg_input(dt, 80 + dt, "full")
    item = a.itemsize
    d = DeviceBuffer(a.nbytes + 64)
    d.upload(a, item)
    check(lib.pxb_seg_sort(gpu_ctx.h, dt, d.addr(item), u64p(off), off.size - 1), "seg_sort")
    gpu_ctx.sync()
    assert d.download(a.dtype, a.size, item).tobytes() == pxo.seg_sort(a, off).tobytes()


@pytest.mark.parametrize("dt", [PX_U64, PX_F32])
def test_seg_sort_radix_plan_replays(gpu_ctx, dt):
    """A captured plan with long segments: every replay sorts fresh data, with eager
    work on the same context between replays."""
    a, off = radix_seg_input(dt, 70 + dt, "narrow" if dt == PX_U64 else "full")
    d = DeviceBuffer.from_array(a)
    plan = lib.pxb_plan_create(gpu_ctx.h, C.OP_SORT, dt, None, d.ptr, d.ptr, u64p(off), off.size - 1, None)
    assert plan
    try:
        rng = np.random.default_rng(5)
        for r in range(3):
            x = rng.permutation(a)
            d.upload(x)
            check(lib.pxb_plan_launch(plan), "replay")
            gpu_ctx.sync()
            assert d.download(a.dtype, a.size).tobytes() == pxo.seg_sort(x, off).tobytes(), r
            y = rng.permutation(a)
            e = DeviceBuffer.from_array(y)
            check(lib.pxb_seg_sort(gpu_ctx.h, dt, e.ptr, u64p(off), off.size - 1), "eager")
            gpu_ctx.sync()
            assert e.download(a.dtype, a.size).tobytes() == pxo.seg_sort(y, off).tobytes(), r
    finally:
        lib.pxb_plan_destroy(plan)

@pytest.mark.parametrize("dt", ALL)
def test_seg_copy_if_and_unique(gpu_ctx, dt):
    a, off = seg_input(dt, 40 + dt)
    nseg = off.size - 1
    d = DeviceBuffer.from_array(a)
    o = DeviceBuffer(a.nbytes + 16)
    cnt = DeviceBuffer(8 * nseg)
    c0 = 0.0 if dt in FLOATS else 500
    fl = make_desc(C.OP_FLAGS, dt, fn=C.PRED_GT, c0=c0)
    check(lib.pxb_seg_copy_if(gpu_ctx.h, ctypes.byref(fl), d.ptr, o.ptr, u64p(off), nseg, cnt.ptr), "seg_copy_if")
    gpu_ctx.sync()
    got = o.download(a.dtype, a.size)
    counts = cnt.download(np.uint64, nseg)
    for s in range(nseg):
        ref = pxo.copy_if(fl, a[off[s]:off[s + 1]])
        assert counts[s] == ref.size
        assert got[off[s]:off[s] + ref.size].tobytes() == ref.tobytes()
    # unique after a segmented sort (the B2 shape)
    srt = pxo.seg_sort(a, off)
    d.upload(srt)
    ud = make_desc(C.OP_FLAGS, dt, fn=C.PRED_UNIQUE)
    check(lib.pxb_seg_copy_if(gpu_ctx.h, ctypes.byref(ud), d.ptr, o.ptr, u64p(off), nseg, cnt.ptr), "seg_unique")
    gpu_ctx.sync()
    got = o.download(a.dtype, a.size)
    counts = cnt.download(np.uint64, nseg)
    uq, uo = pxo.seg_unique(srt, off)
    for s in range(nseg):
        ref = uq[uo[s]:uo[s + 1]]
        assert counts[s] == ref.size and got[off[s]:off[s] + ref.size].tobytes() == ref.tobytes()


@pytest.mark.parametrize("dt", [PX_F32, PX_F64, PX_I64])
def test_alone_equals_in_the_mix(gpu_ctx, dt):
    a, off = seg_input(dt, 50 + dt, nseg=64)
    d = DeviceBuffer.from_array(a)
    out = DeviceBuffer(8 * (off.size - 1))
    check(lib.pxb_seg_reduce(gpu_ctx.h, dt, d.ptr, u64p(off), off.size - 1, out.ptr), "seg_reduce")
    o = DeviceBuffer(a.nbytes + 16)
    check(lib.pxb_seg_inclusive_scan(gpu_ctx.h, dt, d.ptr, o.ptr, u64p(off), off.size - 1), "seg_scan")
    gpu_ctx.sync()
    mix = out.download(a.dtype, off.size - 1)
    mixs = o.download(a.dtype, a.size)
    one = DeviceBuffer(8)
    for s in range(0, off.size - 1, 7):
        so = np.array([off[s], off[s + 1]], dtype=np.uint64)
        check(lib.pxb_seg_reduce(gpu_ctx.h, dt, d.ptr, u64p(so), 1, one.ptr), "seg_reduce alone")
        gpu_ctx.sync()
        assert one.download(a.dtype, 1).tobytes() == mix[s:s + 1].tobytes(), s
        check(lib.pxb_seg_inclusive_scan(gpu_ctx.h, dt, d.ptr, o.ptr, u64p(so), 1), "seg_scan alone")
        gpu_ctx.sync()
        alone = o.download(a.dtype, a.size)[off[s]:off[s + 1]]
        ref = mixs[off[s]:off[s + 1]]
        if dt in FLOATS:  # identical tiling; the look-back may sum its window in another order
            val, ab = pxo.scan_exact(a[off[s]:off[s + 1]])
            assert np.all(np.abs(alone.astype(np.float64) - val) <= bound(dt, alone.size) * ab + 1e-300)
            assert np.all(np.abs(ref.astype(np.float64) - val) <= bound(dt, ref.size) * ab + 1e-300)
        else:
            assert alone.tobytes() == ref.tobytes()


SCAN_TILES = [(256, 4), (256, 8), (512, 8), (1024, 4), (1024, 8)]
REDUCE_TILES = [(256, 4), (256, 8), (256, 16)]


@pytest.fixture
def no_tile_override():
    yield
    for op in (C.OP_REDUCE, C.OP_SCAN, C.OP_FLAGS):
        for dt in ALL:
            lib.pxr_tile_override(op, dt, 0, 0)


def check_seg_scan(a, off, got, dt):
    if dt in FLOATS:
        for s in range(off.size - 1):
            seg = a[off[s]:off[s + 1]]
            if seg.size:
                val, ab = pxo.scan_exact(seg)
                err = np.abs(got[off[s]:off[s + 1]].astype(np.float64) - val)
                assert np.all(err <= bound(dt, seg.size) * ab + 1e-300), s
    else:
        assert np.array_equal(got, pxo.seg_inclusive_scan(a, off))


@pytest.mark.parametrize("tile", SCAN_TILES)
@pytest.mark.parametrize("dt", [PX_U32, PX_I64, PX_F32, PX_F64])
def test_every_tile_shape(gpu_ctx, no_tile_override, dt, tile):
    """Every compiled look-back tile shape, forced through pxr_tile_override, gives the
    oracle's results: batched scan and copy_if, plain scan and copy_if on a pointer one
    element past a 16-byte boundary, and every batched reduce shape."""
    check(lib.pxr_tile_override(C.OP_SCAN, dt, *tile), "override scan")
    check(lib.pxr_tile_override(C.OP_FLAGS, dt, *tile), "override flags")
    a, off = seg_input(dt, 200 + dt + tile[0] + tile[1], nseg=90, hi=150_000)
    nseg = off.size - 1
    d = DeviceBuffer.from_array(a)
    o = DeviceBuffer(a.nbytes + 16)
    cnt = DeviceBuffer(8 * nseg)
    check(lib.pxb_seg_inclusive_scan(gpu_ctx.h, dt, d.ptr, o.ptr, u64p(off), nseg), "seg_scan")
    gpu_ctx.sync()
    check_seg_scan(a, off, o.download(a.dtype, a.size), dt)
    fl = make_desc(C.OP_FLAGS, dt, fn=C.PRED_GT, c0=0.0 if dt in FLOATS else 500)
    check(lib.pxb_seg_copy_if(gpu_ctx.h, ctypes.byref(fl), d.ptr, o.ptr, u64p(off), nseg, cnt.ptr), "seg_copy_if")
    gpu_ctx.sync()
    got = o.download(a.dtype, a.size)
    counts = cnt.download(np.uint64, nseg)
    for s in range(nseg):
        ref = pxo.copy_if(fl, a[off[s]:off[s + 1]])
        assert counts[s] == ref.size and got[off[s]:off[s] + ref.size].tobytes() == ref.tobytes()
    # plain calls, misaligned by one element, several tiles long
    n = 300_001
    es = a.itemsize
    b = a[:n] if a.size >= n else np.resize(a, n)
    pd = DeviceBuffer(b.nbytes + 64)
    pd.upload(b, offset=es)
    po = DeviceBuffer(b.nbytes + 64)
    check(lib.pxd_inclusive_scan(gpu_ctx.h, dt, pd.ptr + es, po.ptr + es, n), "scan")
    cn = DeviceBuffer(8)
    gpu_ctx.sync()
    got = po.download(b.dtype, n, offset=es)
    one = np.array([0, n], dtype=np.uint64)
    check_seg_scan(b, one, got, dt)
    check(lib.pxd_copy_if(gpu_ctx.h, ctypes.byref(fl), pd.ptr + es, po.ptr, n, cn.ptr), "copy_if")
    gpu_ctx.sync()
    ref = pxo.copy_if(fl, b)
    k = int(cn.download(np.uint64, 1)[0])
    assert k == ref.size and po.download(b.dtype, k).tobytes() == ref.tobytes()
    # batched reduce, every shape
    r = DeviceBuffer(8 * nseg)
    for rt in REDUCE_TILES:
        check(lib.pxr_tile_override(C.OP_REDUCE, dt, *rt), "override reduce")
        check(lib.pxb_seg_reduce(gpu_ctx.h, dt, d.ptr, u64p(off), nseg, r.ptr), "seg_reduce")
        gpu_ctx.sync()
        got = r.download(a.dtype, nseg)
        if dt in FLOATS:
            hi, ab = pxo.seg_sum_exact(a, off)
            lens = np.diff(off.astype(np.int64))
            for s in range(nseg):
                assert abs(float(got[s]) - hi[s]) <= bound(dt, lens[s]) * ab[s] + 1e-300, (rt, s)
        else:
            assert np.array_equal(got, pxo.seg_reduce(a, off)), rt


def test_tile_override_rejects_unsupported(no_tile_override):
    assert lib.pxr_tile_override(C.OP_SCAN, PX_U32, 128, 8) != 0
    assert lib.pxr_tile_override(C.OP_REDUCE, PX_U32, 512, 4) != 0
    assert lib.pxr_tile_override(C.OP_SORT, PX_U32, 256, 8) != 0
    assert lib.pxr_tile_override(C.OP_SCAN, 99, 256, 8) != 0


@pytest.mark.parametrize("op,dt", [(C.OP_SCAN, PX_U32), (C.OP_SCAN, PX_F64), (C.OP_FLAGS, PX_I32),
                                   (C.OP_REDUCE, PX_F32)])
def test_plan_autotune(gpu_ctx, no_tile_override, op, dt):
    """The plan's autotuner picks one of the candidate shapes, reports a timing, and the
    captured call gives the oracle's result.  An in-place scan plan leaves its input
    intact through tuning."""
    a, off = seg_input(dt, 300 + op + dt, nseg=200, hi=120_000)
    nseg = off.size - 1
    d = DeviceBuffer.from_array(a)
    o = DeviceBuffer(max(a.nbytes, 8 * nseg) + 16)
    cnt = DeviceBuffer(8 * nseg)
    fl = make_desc(C.OP_FLAGS, dt, fn=C.PRED_LT, c0=300)
    inplace = op == C.OP_SCAN
    ctx = Context()
    plan = lib.pxb_plan_create(ctx.h, op, dt, ctypes.byref(fl), d.ptr, d.ptr if inplace else o.ptr,
                               u64p(off), nseg, cnt.ptr)
    assert plan, "plan creation failed"
    blk, k, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    check(lib.pxb_plan_tile(plan, ctypes.byref(blk), ctypes.byref(k), ctypes.byref(ms)), "plan_tile")
    shapes = REDUCE_TILES if op == C.OP_REDUCE else SCAN_TILES
    assert (blk.value, k.value) in shapes
    assert ms.value > 0.0
    if inplace:
        ctx.sync()
        assert d.download(a.dtype, a.size).tobytes() == a.tobytes(), "input changed by tuning"
    check(lib.pxb_plan_launch(plan), "launch")
    ctx.sync()
    if op == C.OP_SCAN:
        check_seg_scan(a, off, d.download(a.dtype, a.size), dt)
    elif op == C.OP_FLAGS:
        counts = cnt.download(np.uint64, nseg)
        got = o.download(a.dtype, a.size)
        for s in range(nseg):
            ref = pxo.copy_if(fl, a[off[s]:off[s + 1]])
            assert counts[s] == ref.size and np.array_equal(got[off[s]:off[s] + ref.size], ref)
    else:
        got = o.download(a.dtype, nseg)
        hi, ab = pxo.seg_sum_exact(a, off)
        lens = np.diff(off.astype(np.int64))
        assert all(abs(float(got[s]) - hi[s]) <= bound(dt, lens[s]) * ab[s] + 1e-300 for s in range(nseg))
    lib.pxb_plan_destroy(plan)
    ctx.close()


def test_graph_plan_replays(gpu_ctx):
    """A captured plan replayed several times, with other eager work between replays,
    gives the eager result every time (no memset nodes; self-resetting scratch)."""
    other = Context()
    for op, dt in [(C.OP_REDUCE, PX_F64), (C.OP_SCAN, PX_U32), (C.OP_FLAGS, PX_I32), (C.OP_SORT, PX_U64)]:
        a, off = seg_input(dt, 60 + op, nseg=120, hi=100_000)
        nseg = off.size - 1
        d = DeviceBuffer.from_array(a)
        o = DeviceBuffer(max(a.nbytes, 8 * nseg) + 16)
        cnt = DeviceBuffer(8 * nseg)
        fl = make_desc(C.OP_FLAGS, dt, fn=C.PRED_LT, c0=300)
        ctx = Context()
        plan = lib.pxb_plan_create(ctx.h, op, dt, ctypes.byref(fl), d.ptr,
                                   d.ptr if op == C.OP_SORT else o.ptr, u64p(off), nseg, cnt.ptr)
        assert plan, "plan creation failed"
        for rep in range(3):
            if op == C.OP_SORT:
                d.upload(a)
            check(lib.pxb_plan_launch(plan), "plan launch")
            ctx.sync()
            # eager work on another context between replays
            tmp = DeviceBuffer.from_array(np.arange(100_000, dtype=np.uint32))
            tout = DeviceBuffer(8)
            check(lib.pxd_reduce(other.h, PX_U32, tmp.ptr, 100_000, tout.ptr), "eager")
            other.sync()
            if op == C.OP_REDUCE:
                got = o.download(a.dtype, nseg)
                hi, ab = pxo.seg_sum_exact(a, off)
                lens = np.diff(off.astype(np.int64))
                assert all(abs(float(got[s]) - hi[s]) <= bound(dt, lens[s]) * ab[s] + 1e-300 for s in range(nseg))
            elif op == C.OP_SCAN:
                assert np.array_equal(o.download(a.dtype, a.size), pxo.seg_inclusive_scan(a, off))
            elif op == C.OP_FLAGS:
                counts = cnt.download(np.uint64, nseg)
                got = o.download(a.dtype, a.size)
                for s in range(nseg):
                    ref = pxo.copy_if(fl, a[off[s]:off[s + 1]])
                    assert counts[s] == ref.size and np.array_equal(got[off[s]:off[s] + ref.size], ref)
            else:
                assert np.array_equal(d.download(a.dtype, a.size), pxo.seg_sort(a, off))
        lib.pxb_plan_destroy(plan)
        ctx.close()
    other.close()


def make_server(budget, slots=4, max_batch=4096, lanes=None):
    cfg = PxsConfig()
    lib.pxs_config_default(ctypes.byref(cfg))
    cfg.hbm_budget = budget
    cfg.slots = slots
    cfg.max_batch = max_batch
    if lanes is not None:
        cfg.lanes = lanes
    s = lib.pxs_create(ctypes.byref(cfg))
    assert s, "pxs_create failed"
    return s


def put(server, a):
    p = lib.pxs_alloc(server, max(a.nbytes, 1))
    assert p
    if a.nbytes:
        check(lib.pxd_memcpy_h2d(p, a.ctypes.data, a.nbytes), "h2d")
    return p


def check_served(server, rid, op, dt, a, fl=None):
    lib.pxs_wait(server, rid, -1.0)
    out = ctypes.c_void_p()
    cnt = ctypes.c_uint64()
    lat = ctypes.c_double()
    check(lib.pxs_result(server, rid, ctypes.byref(out), ctypes.byref(cnt), ctypes.byref(lat)), "result")
    npd = NP_DTYPES[dt]
    if op == C.OP_REDUCE:
        got = np.empty(1, dtype=npd)
        lib.pxd_memcpy_d2h(got.ctypes.data, out.value, got.nbytes)
        if dt in FLOATS:
            hi, lo, ab = pxo.sum_exact(a)
            assert abs(float(got[0]) - (hi + lo)) <= bound(dt, a.size) * ab + 1e-300
        else:
            assert got[0] == pxo.reduce(a)
    elif op == C.OP_SCAN:
        got = np.empty(a.size, dtype=npd)
        lib.pxd_memcpy_d2h(got.ctypes.data, out.value, got.nbytes)
        if dt in FLOATS:
            val, ab = pxo.scan_exact(a)
            assert np.all(np.abs(got.astype(np.float64) - val) <= bound(dt, a.size) * ab + 1e-300)
        else:
            assert np.array_equal(got, pxo.inclusive_scan(a))
    elif op == C.OP_SORT:
        got = np.empty(a.size, dtype=npd)
        lib.pxd_memcpy_d2h(got.ctypes.data, out.value, got.nbytes)
        assert got.tobytes() == pxo.sort(a).tobytes()
    else:
        ref = pxo.copy_if(fl, a)
        got = np.empty(ref.size, dtype=npd)
        if ref.size:
            lib.pxd_memcpy_d2h(got.ctypes.data, out.value, got.nbytes)
        assert cnt.value == ref.size and got.tobytes() == ref.tobytes()
    return lat.value


def mixed_requests(seed, count, max_log=18):
    rng = np.random.default_rng(seed)
    ops = [C.OP_REDUCE, C.OP_SCAN, C.OP_SORT, C.OP_FLAGS]
    dts = [PX_I32, PX_U64, PX_F32, PX_F64]
    reqs = []
    for i in range(count):
        op = ops[rng.choice(4, p=[0.3, 0.3, 0.25, 0.15])]
        dt = dts[rng.integers(0, 4)]
        n = int(2 ** rng.uniform(4, max_log))
        if dt in FLOATS:
            a = rng.uniform(-1000, 1000, n).astype(NP_DTYPES[dt])
        elif dt == PX_U64:
            a = rng.integers(0, 1001, n).astype(np.uint64)
        else:
            a = rng.integers(-1000, 1001, n).astype(np.int32)
        reqs.append((op, dt, a))
    return reqs


def submit(server, op, dt, a, p):
    fl = make_desc(C.OP_FLAGS, dt, fn=C.PRED_LT, c0=500 if dt == PX_U64 else 0)
    r = PxsRequest(op, dt, a.size, p, fl)
    rid = lib.pxs_submit(server, ctypes.byref(r))
    assert rid >= 0
    return rid, fl


@pytest.mark.parametrize("lanes", [1, 2])
def test_served_mix_matches_oracle_and_alone(lanes):
    """One stream, or two (reduces and the sorts' counting kernels on the second, the
    sorts' digit passes joining the first by an event)."""
    server = make_server(1 << 30, lanes=lanes)
    reqs = mixed_requests(4, 120)
    ptrs = [put(server, a) for _, _, a in reqs]
    # all at once (mixed, batched, overlapping)
    ids = [submit(server, op, dt, a, p) for (op, dt, a), p in zip(reqs, ptrs)]
    lib.pxs_drain(server)
    for (op, dt, a), (rid, fl) in zip(reqs, ids):
        check_served(server, rid, op, dt, a, fl)
        lib.pxs_release(server, rid)
    st = PxsStats()
    lib.pxs_get_stats(server, ctypes.byref(st))
    assert st.completed == 120 and st.batches <= 120
    # the same requests one at a time give the same results
    for (op, dt, a), p in list(zip(reqs, ptrs))[::10]:
        if op == C.OP_SORT:
            lib.pxd_memcpy_h2d(p, a.ctypes.data, a.nbytes)
        rid, fl = submit(server, op, dt, a, p)
        check_served(server, rid, op, dt, a, fl)
        lib.pxs_release(server, rid)
    lib.pxs_destroy(server)


def test_served_admission_waits_for_budget():
    # budget: the inputs plus room for about three outputs at a time
    n = 1 << 20
    arrays = [np.arange(n, dtype=np.uint32) + i for i in range(12)]
    server = make_server(12 * n * 4 + 3 * n * 4 + (1 << 20), slots=2)
    ptrs = [put(server, a) for a in arrays]
    ids = [submit(server, C.OP_SCAN, PX_U32, a, p)[0] for a, p in zip(arrays, ptrs)]
    for a, rid in zip(arrays, ids):  # consume in order, freeing outputs as we go
        check_served(server, rid, C.OP_SCAN, PX_U32, a)
        lib.pxs_release(server, rid)
    st = PxsStats()
    lib.pxs_get_stats(server, ctypes.byref(st))
    assert st.completed == 12 and st.pool_peak_bytes <= 12 * n * 4 + 3 * n * 4 + (1 << 20)
    lib.pxs_destroy(server)


def test_served_concurrent_submitters():
    server = make_server(1 << 30)
    reqs = mixed_requests(9, 80, max_log=16)
    ptrs = [put(server, a) for _, _, a in reqs]
    errors = []

    def client(k):
        try:
            for i in range(k, len(reqs), 4):
                op, dt, a = reqs[i]
                rid, fl = submit(server, op, dt, a, ptrs[i])
                check_served(server, rid, op, dt, a, fl)
                lib.pxs_release(server, rid)
        except Exception as e:  # pragma: no cover - reported below
            errors.append(e)

    ts = [threading.Thread(target=client, args=(k,)) for k in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
    lib.pxs_destroy(server)


def test_served_fuses_queued_requests():
    """Requests of one (op, dtype) that queue while every slot is busy run as one batch:
    four sorts of 2^24 keys hold the four slots for milliseconds, and 64 small reduces
    submitted meanwhile wait behind them."""
    server = make_server(1 << 32)
    rng = np.random.default_rng(12)
    big = [rng.random(1 << 24) for _ in range(4)]
    small = [rng.integers(-1000, 1000, 1000).astype(np.int32) for _ in range(64)]
    pb = [put(server, a) for a in big]
    ps = [put(server, a) for a in small]
    ids_b = [submit(server, C.OP_SORT, PX_F64, a, p) for a, p in zip(big, pb)]
    ids_s = [submit(server, C.OP_REDUCE, PX_I32, a, p) for a, p in zip(small, ps)]
    lib.pxs_drain(server)
    for a, (rid, fl) in zip(small, ids_s):
        check_served(server, rid, C.OP_REDUCE, PX_I32, a, fl)
    for a, p in zip(big, pb):  # sortedness only (sort parity is tested elsewhere)
        got = np.empty(a.size, np.float64)
        check(lib.pxd_memcpy_d2h(got.ctypes.data, p, got.nbytes), "d2h")
        assert np.all(got[1:] >= got[:-1])
    st = PxsStats()
    lib.pxs_get_stats(server, ctypes.byref(st))
    assert st.completed == 68 and st.batches < 68, (st.completed, st.batches)
    for rid, _ in ids_b + ids_s:
        lib.pxs_release(server, rid)
    lib.pxs_destroy(server)


def test_served_output_that_never_fits_is_refused():
    # the pool holds the input and little else: a scan output of the same size can never
    # fit beside it, so pxs_submit refuses it at once (nothing blocks the queue)
    n = 1 << 20
    a = np.arange(n, dtype=np.uint32)
    server = make_server(n * 4 + (1 << 16))
    p = put(server, a)
    fl = make_desc(C.OP_FLAGS, PX_U32, fn=C.PRED_LT, c0=10)
    rid = lib.pxs_submit(server, ctypes.byref(PxsRequest(C.OP_SCAN, PX_U32, n, p, fl)))
    assert rid == C.E_NOMEM
    # a reduce (one element of output) still runs
    rid = lib.pxs_submit(server, ctypes.byref(PxsRequest(C.OP_REDUCE, PX_U32, n, p, fl)))
    assert rid >= 0
    check_served(server, rid, C.OP_REDUCE, PX_U32, a)
    st = PxsStats()
    lib.pxs_get_stats(server, ctypes.byref(st))
    assert st.rejected == 1 and st.failed == 0
    lib.pxs_destroy(server)


def test_served_drain_before_release_does_not_block_forever():
    # room for the inputs plus about three outputs; a client that drains before releasing
    # gets PX_ENOMEM back (not a hang), releases what completed, and drains again
    n = 1 << 20
    arrays = [np.arange(n, dtype=np.uint32) * 3 + i for i in range(12)]
    server = make_server(12 * n * 4 + 3 * n * 4 + (1 << 16), slots=2)
    ptrs = [put(server, a) for a in arrays]
    ids = [submit(server, C.OP_SCAN, PX_U32, a, p)[0] for a, p in zip(arrays, ptrs)]
    done = set()
    stalls = 0
    for _ in range(100):
        rc = lib.pxs_drain(server)
        assert rc in (0, C.E_NOMEM)
        for a, rid in zip(arrays, ids):
            if rid in done:
                continue
            w = lib.pxs_wait(server, rid, 0.0)
            if w == 0:
                check_served(server, rid, C.OP_SCAN, PX_U32, a)
                check(lib.pxs_release(server, rid), "release")
                done.add(rid)
        if rc == 0 and len(done) == len(ids):
            break
        stalls += rc != 0
    assert len(done) == len(ids) and stalls >= 1
    lib.pxs_destroy(server)


def test_served_batch_scratch_wait_reports_enomem_then_runs():
    # inputs and the output fill the pool exactly: the copy_if batch's count words do not
    # fit, nothing runs, so pxs_wait returns PX_ENOMEM; freeing a client buffer lets it run
    n = 1 << 16
    a = np.random.default_rng(3).integers(-100, 100, n).astype(np.int32)
    server = make_server(2 * n * 4 + 256)
    p = put(server, a)
    spare = lib.pxs_alloc(server, 256)
    assert spare
    fl = make_desc(C.OP_FLAGS, PX_I32, fn=C.PRED_LT, c0=0)
    rid = lib.pxs_submit(server, ctypes.byref(PxsRequest(C.OP_FLAGS, PX_I32, n, p, fl)))
    assert rid >= 0
    assert lib.pxs_wait(server, rid, 5000.0) == C.E_NOMEM
    lib.pxs_free(server, spare)
    assert lib.pxs_wait(server, rid, 5000.0) == 0
    check_served(server, rid, C.OP_FLAGS, PX_I32, a, fl)
    lib.pxs_destroy(server)


def test_served_rejects_misaligned_and_overlapping_inputs():
    server = make_server(1 << 30)
    big = [np.random.default_rng(i).random(1 << 23) for i in range(4)]
    pb = [put(server, a) for a in big]
    x = np.arange(1 << 16, dtype=np.uint32)
    px = put(server, x)
    fl = make_desc(C.OP_FLAGS, PX_U32, fn=C.PRED_LT, c0=10)
    assert lib.pxs_submit(server, ctypes.byref(PxsRequest(C.OP_REDUCE, PX_U32, 100, px + 2, fl))) == C.E_ARG
    # four long sorts hold the slots; a sort of x queues behind them, so a reduce of x
    # submitted now would share its bytes with a pending in-place sort
    ids = [submit(server, C.OP_SORT, PX_F64, a, p)[0] for a, p in zip(big, pb)]
    rs = submit(server, C.OP_SORT, PX_U32, x, px)[0]
    assert lib.pxs_submit(server, ctypes.byref(PxsRequest(C.OP_REDUCE, PX_U32, x.size, px, fl))) == C.E_ARG
    check(lib.pxs_drain(server), "drain")
    check_served(server, rs, C.OP_SORT, PX_U32, x)
    for rid in ids + [rs]:
        lib.pxs_release(server, rid)
    # once the sort is done, the same reduce is accepted
    rid = lib.pxs_submit(server, ctypes.byref(PxsRequest(C.OP_REDUCE, PX_U32, x.size, px, fl)))
    assert rid >= 0
    check_served(server, rid, C.OP_REDUCE, PX_U32, np.sort(x))
    lib.pxs_destroy(server)


def test_served_sort_deep_in_the_pool_uses_little_scratch():
    # a sort of 2^22 u64 keys 6 GiB into an 8 GiB pool: its radix scratch (about twice
    # its own bytes) comes from the pool and does not depend on where the keys sit
    server = make_server(8 << 30)
    filler = lib.pxs_alloc(server, 6 << 30)
    assert filler
    a = np.random.default_rng(5).integers(0, 1 << 63, 1 << 22, dtype=np.uint64)
    p = put(server, a)
    rid = submit(server, C.OP_SORT, PX_U64, a, p)[0]
    check_served(server, rid, C.OP_SORT, PX_U64, a)
    st = PxsStats()
    lib.pxs_get_stats(server, ctypes.byref(st))
    assert st.pool_peak_bytes <= (6 << 30) + 3 * a.nbytes + (1 << 22), st.pool_peak_bytes
    lib.pxs_release(server, rid)
    lib.pxs_destroy(server)
