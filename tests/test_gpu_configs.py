"""Every headline configuration at its stated size, through the C ABI, against the oracle.

  P1 reduce 2^28 float32 normal(0,1)          P2 reduce 2^28 int64 uniform [-2^20, 2^20)
  P3 inclusive + exclusive scan 2^28 uint32 in [0, 16)
  P4 sort 2^26 uint32 (1% duplicates)         P5 sort 2^25 float64 (1% duplicates)
  P6 copy_if 2^28 int32, x % 2 == 0           (all from seed 1, parallax/runtime.h)
  B1 16,384 float32 segments, lengths log-uniform [1, 2^18], seed 2: segmented reduce and
     inclusive scan from hipGraph plans, every segment checked
  B2 4,096 uint64 segments, lengths log-uniform [16, 2^18], keys [0, 2^20), seed 3:
     segmented sort then unique from hipGraph plans, every segment checked
  S  the served mixed stream (seed 4, 400 requests of 2^10..2^24 elements, 30% reduce,
     30% scan, 25% sort, 15% copy_if over {int32, uint64, float32, float64}): every
     result checked, and a sample re-run alone gives the same bytes.

Bars: integer, sort, compaction and unique results bit-exact; float reduce and scan
within |err| <= 2*ceil(log2 n)*eps*sum|x| of the double-double oracle (scans: per prefix,
n = the segment's length)."""
import concurrent.futures
import ctypes
import math

import numpy as np
import pytest

import pxo
from parallax_amd import lib, check, make_desc, descriptor_words, consts as C, datagen
from parallax_amd._lib import PxsConfig, PxsRequest, PxsStats
from parallax_amd.device import Context, DeviceBuffer, stats
from parallax_amd.dtypes import NP_DTYPES, PX_F32, PX_F64, PX_I32, PX_I64, PX_U32, PX_U64

pytestmark = pytest.mark.gpu

FLOATS = (PX_F32, PX_F64)


def bound(dt, n):
    eps = float(np.finfo(NP_DTYPES[dt]).eps)
    return 2 * max(1, math.ceil(math.log2(max(int(n), 2)))) * eps


def load(d):
    k = lib.parallax_kernel_load(*descriptor_words(d))
    assert k, "parallax_kernel_load"
    return k


@pytest.fixture(scope="module")
def plain():
    return datagen.plain_suite_inputs(seed=1, scale=1.0)


def _launches():
    return stats()["launches"]


def test_p1_reduce_f32_2p28(gpu_ctx, plain):
    a = plain["reduce_f32"]
    assert a.size == 1 << 28
    d = DeviceBuffer.from_array(a)
    r = DeviceBuffer(8)
    n0 = _launches()
    lib.parallax_reduce(load(make_desc(C.OP_REDUCE, PX_F32)), d.ptr, a.size, 4, r.ptr)
    assert _launches() > n0
    got = float(r.download(np.float32, 1)[0])
    hi, lo, ab = pxo.sum_exact(a)
    err = abs(got - (hi + lo))
    print(f"P1 err/bound {err / (bound(PX_F32, a.size) * ab):.3e}")
    assert err <= bound(PX_F32, a.size) * ab


def test_p2_reduce_i64_2p28(gpu_ctx, plain):
    a = plain["reduce_i64"]
    d = DeviceBuffer.from_array(a)
    r = DeviceBuffer(8)
    lib.parallax_reduce(load(make_desc(C.OP_REDUCE, PX_I64)), d.ptr, a.size, 8, r.ptr)
    assert r.download(np.int64, 1)[0] == pxo.reduce(a)


def test_p3_scans_u32_2p28(gpu_ctx, plain):
    a = plain["scan_u32"]
    n = a.size
    ks, ka, kh = (load(make_desc(op, PX_U32)) for op in (C.OP_SCAN, C.OP_SCAN_ADD, C.OP_SCAN_SHIFT))
    d = DeviceBuffer.from_array(a)
    lib.parallax_scan(ks, ka, d.ptr, n, 4)  # in place, as the ABI scans
    ref = pxo.inclusive_scan(a)
    assert np.array_equal(d.download(np.uint32, n), ref)
    d.upload(a)
    o = DeviceBuffer(a.nbytes)
    init = np.zeros(1, dtype=np.uint32)
    lib.parallax_exclusive_scan(ks, ka, kh, d.ptr, o.ptr, n, 4, init.ctypes.data)
    got = o.download(np.uint32, n)
    assert np.array_equal(got, pxo.exclusive_scan(a, 0))
    assert np.array_equal(d.download(np.uint32, n), a)  # input untouched


@pytest.mark.parametrize("key,dt", [("sort_u32", PX_U32), ("sort_f64", PX_F64)])
def test_p4_p5_sorts(gpu_ctx, plain, key, dt):
    a = plain[key]
    assert a.size == (1 << 26 if dt == PX_U32 else 1 << 25)
    d = DeviceBuffer.from_array(a)
    lib.parallax_sort(load(make_desc(C.OP_SORT, dt)), d.ptr, a.size, a.itemsize)
    assert d.download(a.dtype, a.size).tobytes() == pxo.sort(a).tobytes()


def test_p6_copy_if_even_i32_2p28(gpu_ctx, plain):
    a = plain["copy_if_i32"]
    n = a.size
    fl = make_desc(C.OP_FLAGS, PX_I32, fn=C.PRED_MOD_EQ, c0=2, c1=0)
    d = DeviceBuffer.from_array(a)
    o = DeviceBuffer(a.nbytes)
    k = lib.parallax_copy_if(load(fl), load(make_desc(C.OP_SCAN, PX_I32)), load(make_desc(C.OP_SCAN_ADD, PX_I32)),
                             load(make_desc(C.OP_SCATTER, PX_I32, aux=C.SCATTER_COMPACT)), d.ptr, o.ptr, n, 4, 0)
    ref = pxo.copy_if(fl, a)
    assert k == ref.size
    assert o.download(np.int32, k).tobytes() == ref.tobytes()


# ------------------------------------------------------------------ batched
def _u64p(off):
    return off.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def test_b1_seg_reduce_and_scan_f32(gpu_ctx):
    x, off = datagen.batched_b1(seed=2, nseg=16384)
    nseg = off.size - 1
    lens = np.diff(off).astype(np.int64)
    ctx = Context()
    d = DeviceBuffer.from_array(x)
    r = DeviceBuffer(4 * nseg)
    s = DeviceBuffer(x.nbytes + 16)
    pr = lib.pxb_plan_create(ctx.h, C.OP_REDUCE, PX_F32, None, d.ptr, r.ptr, _u64p(off), nseg, None)
    ps = lib.pxb_plan_create(ctx.h, C.OP_SCAN, PX_F32, None, d.ptr, s.ptr, _u64p(off), nseg, None)
    assert pr and ps
    r.upload(np.full(nseg, np.nan, dtype=np.float32))
    check(lib.pxb_plan_launch(pr), "reduce replay")
    check(lib.pxb_plan_launch(ps), "scan replay")
    ctx.sync()
    # reduce: every segment within its bound
    got = r.download(np.float32, nseg).astype(np.float64)
    hi, ab = pxo.seg_sum_exact(x, off)
    fac = np.array([bound(PX_F32, l) for l in lens])
    ratio = np.abs(got - hi) / np.maximum(fac * ab, 1e-300)
    print(f"B1 reduce max err/bound {ratio.max():.3e}")
    assert np.all(np.abs(got - hi) <= fac * ab)
    # scan: every prefix of every segment within its segment's bound
    gs = s.download(np.float32, x.size).astype(np.float64)
    val, pab = pxo.seg_scan_exact(x, off)
    per = np.repeat(fac, lens)
    bad = np.abs(gs - val) > per * pab
    print(f"B1 scan max err/bound {float(np.max(np.abs(gs - val) / np.maximum(per * pab, 1e-300))):.3e}")
    assert not bad.any(), int(np.argmax(bad))
    lib.pxb_plan_destroy(pr)
    lib.pxb_plan_destroy(ps)
    ctx.close()


def test_b2_seg_sort_then_unique_u64(gpu_ctx):
    x, off = datagen.batched_b2(seed=3, nseg=4096)
    nseg = off.size - 1
    ctx = Context()
    d = DeviceBuffer.from_array(x)
    u = DeviceBuffer(x.nbytes + 16)
    c = DeviceBuffer(8 * nseg)
    ud = make_desc(C.OP_FLAGS, PX_U64, fn=C.PRED_UNIQUE)
    ps = lib.pxb_plan_create(ctx.h, C.OP_SORT, PX_U64, None, d.ptr, d.ptr, _u64p(off), nseg, None)
    assert ps
    assert np.array_equal(d.download(np.uint64, x.size), x)  # plan creation leaves the keys alone
    pu = lib.pxb_plan_create(ctx.h, C.OP_FLAGS, PX_U64, ctypes.byref(ud), d.ptr, u.ptr, _u64p(off), nseg, c.ptr)
    assert pu
    check(lib.pxb_plan_launch(ps), "sort replay")
    check(lib.pxb_plan_launch(pu), "unique replay")
    ctx.sync()
    srt = pxo.seg_sort(x, off)
    assert d.download(np.uint64, x.size).tobytes() == srt.tobytes()
    uq_ref, uo = pxo.seg_unique(srt, off)
    cnt = c.download(np.uint64, nseg)
    assert np.array_equal(cnt, np.diff(uo))
    got = u.download(np.uint64, x.size)
    # kept elements of segment s sit at off[s] .. off[s] + cnt[s]
    starts = np.repeat(off[:-1].astype(np.int64), cnt.astype(np.int64))
    within = np.arange(int(cnt.sum()), dtype=np.int64) - np.repeat(uo[:-1].astype(np.int64), cnt.astype(np.int64))
    assert got[starts + within].tobytes() == uq_ref.tobytes()
    lib.pxb_plan_destroy(ps)
    lib.pxb_plan_destroy(pu)
    ctx.close()


# ------------------------------------------------------------------ served
def _check_result(op, dt, a, c0, out_bytes, count):
    """None when the served result matches the oracle, else a description."""
    npd = NP_DTYPES[dt]
    if op == C.OP_REDUCE:
        got = np.frombuffer(out_bytes, dtype=npd)[0]
        if dt in FLOATS:
            hi, lo, ab = pxo.sum_exact(a)
            if abs(float(got) - (hi + lo)) > bound(dt, a.size) * ab + 1e-300:
                return f"reduce err {abs(float(got) - (hi + lo))}"
        elif got != pxo.reduce(a):
            return "reduce"
    elif op == C.OP_SCAN:
        got = np.frombuffer(out_bytes, dtype=npd)
        if dt in FLOATS:
            val, ab = pxo.scan_exact(a)
            if np.any(np.abs(got.astype(np.float64) - val) > bound(dt, a.size) * ab + 1e-300):
                return "scan bound"
        elif not np.array_equal(got, pxo.inclusive_scan(a)):
            return "scan"
    elif op == C.OP_SORT:
        if out_bytes != pxo.sort(a).tobytes():
            return "sort"
    else:
        ref = pxo.copy_if(make_desc(C.OP_FLAGS, dt, fn=C.PRED_LT, c0=c0), a)
        if count != ref.size or out_bytes[:ref.nbytes] != ref.tobytes():
            return f"copy_if count {count} vs {ref.size}"
    return None


def _fetch(server, rid, op, a):
    out = ctypes.c_void_p()
    cnt = ctypes.c_uint64()
    lat = ctypes.c_double()
    check(lib.pxs_result(server, rid, ctypes.byref(out), ctypes.byref(cnt), ctypes.byref(lat)), "result")
    nb = a.itemsize if op == C.OP_REDUCE else (cnt.value * a.itemsize if op == C.OP_FLAGS else a.nbytes)
    buf = np.empty(max(nb, 1), dtype=np.uint8)
    if nb:
        check(lib.pxd_memcpy_d2h(buf.ctypes.data, out.value, nb), "d2h")
    return buf[:nb].tobytes(), cnt.value


def test_served_mixed_stream_every_result(gpu_ctx):
    reqs = datagen.served_requests(seed=4, nreq=400)
    cfg = PxsConfig()
    lib.pxs_config_default(ctypes.byref(cfg))  # 64 GB pool, 8 slots on one stream
    server = lib.pxs_create(ctypes.byref(cfg))
    assert server
    ptrs = []
    for _, _, a, _ in reqs:
        p = lib.pxs_alloc(server, a.nbytes)
        assert p
        check(lib.pxd_memcpy_h2d(p, a.ctypes.data, a.nbytes), "h2d")
        ptrs.append(p)
    ids = []
    for (op, dt, a, c0), p in zip(reqs, ptrs):
        rid = lib.pxs_submit(server, ctypes.byref(PxsRequest(op, dt, a.size, p, make_desc(C.OP_FLAGS, dt, fn=C.PRED_LT, c0=c0))))
        assert rid >= 0
        ids.append(rid)
    check(lib.pxs_drain(server), "drain")
    st = PxsStats()
    lib.pxs_get_stats(server, ctypes.byref(st))
    assert st.completed == len(reqs) and st.failed == 0, (st.completed, st.failed)
    assert st.batches < len(reqs)  # requests were fused
    mixed = [_fetch(server, rid, op, a) for rid, (op, _, a, _) in zip(ids, reqs)]
    with concurrent.futures.ThreadPoolExecutor(max_workers=8) as ex:
        errs = list(ex.map(lambda i: _check_result(*reqs[i], *mixed[i]), range(len(reqs))))
    bad = [(i, e) for i, e in enumerate(errs) if e]
    assert not bad, bad[:5]
    for rid in ids:
        check(lib.pxs_release(server, rid), "release")
    # every 8th request alone: the same bytes as in the mix
    for i in range(0, len(reqs), 8):
        op, dt, a, c0 = reqs[i]
        if op == C.OP_SORT:
            check(lib.pxd_memcpy_h2d(ptrs[i], a.ctypes.data, a.nbytes), "restore")
        rid = lib.pxs_submit(server, ctypes.byref(PxsRequest(op, dt, a.size, ptrs[i], make_desc(C.OP_FLAGS, dt, fn=C.PRED_LT, c0=c0))))
        assert rid >= 0
        assert lib.pxs_wait(server, rid, -1.0) == 0
        assert _fetch(server, rid, op, a) == mixed[i], i
        check(lib.pxs_release(server, rid), "release")
    lib.pxs_destroy(server)
