"""Benchmark of the Parallax primitive hot path on one MI355X.

Default workload (the plain per-call configuration): one step calls, through the C ABI
of parallax/runtime.h with device-resident (HBM) operands, exactly as the reference's
funnels call it:
    reduce  2^28 float32 normal(0,1)          reduce  2^28 int64 uniform [-2^20, 2^20)
    inclusive scan 2^28 uint32 in [0,16)       exclusive scan 2^28 uint32 in [0,16)
    sort    2^26 uint32 (1% duplicates)        sort    2^25 float64 (1% duplicates)
    copy_if 2^28 int32, predicate x % 2 == 0   (all from seed 1)
`value` = algorithmic bytes of the step / wall time of the step (GB/s); algorithmic
bytes: reduce n*s, scans 2*n*s, sort 2*n*s, copy_if n*s + kept*s.  Each call's device
time is also measured with HIP events on the launcher's stream (pxr_prof_*), giving
per-primitive GB/s and the roofline of the dominant primitive.  Results are checked
against the oracle before timing.

Other workloads (--workload batched | served) run the batched and served forms.
"""
import argparse
import ctypes
import json
import os
import pathlib
import subprocess
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "parallax-runtime_amd"))
sys.path.insert(0, str(ROOT / "oracle"))

import numpy as np  # noqa: E402

HBM_PEAK_GBPS = 8000.0
METRIC = "GB/s and % HBM peak per primitive; batched/served req/s + p99 latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ------------------------------------------------------------------ plain suite
def plain_suite(args):
    from parallax_amd import lib, make_desc, descriptor_words, consts as C, datagen
    from parallax_amd.device import DeviceBuffer, stats
    from parallax_amd.dtypes import PX_F32, PX_F64, PX_I32, PX_I64, PX_U32

    scale = args.scale
    t0 = time.time()
    X = datagen.plain_suite_inputs(seed=1, scale=scale)
    log(f"[bench] inputs generated in {time.time() - t0:.1f}s")

    def load(d):
        w, n = descriptor_words(d)
        k = lib.parallax_kernel_load(w, n)
        if not k:
            raise RuntimeError("kernel load failed (no GPU?)")
        return k

    K = {
        "red_f32": load(make_desc(C.OP_REDUCE, PX_F32)),
        "red_i64": load(make_desc(C.OP_REDUCE, PX_I64)),
        "scan": load(make_desc(C.OP_SCAN, PX_U32)),
        "add": load(make_desc(C.OP_SCAN_ADD, PX_U32)),
        "shift": load(make_desc(C.OP_SCAN_SHIFT, PX_U32)),
        "sort_u32": load(make_desc(C.OP_SORT, PX_U32)),
        "sort_f64": load(make_desc(C.OP_SORT, PX_F64)),
        "flags": load(make_desc(C.OP_FLAGS, PX_I32, fn=C.PRED_MOD_EQ, c0=2, c1=0)),
        "scan_i32": load(make_desc(C.OP_SCAN, PX_I32)),
        "add_i32": load(make_desc(C.OP_SCAN_ADD, PX_I32)),
        "scatter": load(make_desc(C.OP_SCATTER, PX_I32, aux=C.SCATTER_COMPACT)),
    }
    nsteps = args.warmup + args.steps + 1  # +1: the verification call
    t0 = time.time()
    D = {
        "red_f32": DeviceBuffer.from_array(X["reduce_f32"]),
        "red_i64": DeviceBuffer.from_array(X["reduce_i64"]),
        "scan_in": DeviceBuffer.from_array(X["scan_u32"]),
        "scan_inplace": DeviceBuffer.from_array(X["scan_u32"]),
        "scan_out": DeviceBuffer(X["scan_u32"].nbytes),
        "cif_in": DeviceBuffer.from_array(X["copy_if_i32"]),
        "cif_out": DeviceBuffer(X["copy_if_i32"].nbytes),
        "res": DeviceBuffer(64),
    }
    sort_u32 = [DeviceBuffer.from_array(X["sort_u32"]) for _ in range(nsteps)]
    sort_f64 = [DeviceBuffer.from_array(X["sort_f64"]) for _ in range(nsteps)]
    lib.pxd_device_sync()
    log(f"[bench] inputs resident in HBM in {time.time() - t0:.1f}s")

    n28 = X["reduce_f32"].size
    n26 = X["sort_u32"].size
    n25 = X["sort_f64"].size
    init = np.zeros(1, dtype=np.uint32)
    kept_box = [0]

    def step(i):
        r = D["res"].ptr
        lib.parallax_reduce(K["red_f32"], D["red_f32"].ptr, n28, 4, r)
        lib.parallax_reduce(K["red_i64"], D["red_i64"].ptr, n28, 8, r + 8)
        lib.parallax_scan(K["scan"], K["add"], D["scan_inplace"].ptr, n28, 4)
        lib.parallax_exclusive_scan(K["scan"], K["add"], K["shift"], D["scan_in"].ptr,
                                    D["scan_out"].ptr, n28, 4, init.ctypes.data)
        lib.parallax_sort(K["sort_u32"], sort_u32[i].ptr, n26, 4)
        lib.parallax_sort(K["sort_f64"], sort_f64[i].ptr, n25, 8)
        kept_box[0] = lib.parallax_copy_if(K["flags"], K["scan_i32"], K["add_i32"], K["scatter"],
                                           D["cif_in"].ptr, D["cif_out"].ptr, n28, 4, 0)

    # ---- verification call (full size) against the oracle / exact closed forms
    t0 = time.time()
    step(nsteps - 1)
    parity = verify(X, D, sort_u32[nsteps - 1], sort_f64[nsteps - 1], kept_box[0])
    parity["c1_reduce_2p20_f32"] = verify_c1(lib, load, C, PX_F32)
    log(f"[bench] parity verified in {time.time() - t0:.1f}s: {parity}")

    for i in range(args.warmup):
        step(i)
    lib.pxd_device_sync()
    lib.pxr_prof_enable(1)
    launches0 = stats()["launches"]
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(args.warmup + i)
    lib.pxd_device_sync()
    wall = time.perf_counter() - t0
    assert stats()["launches"] - launches0 >= 7 * args.steps, "HIP path did not run"

    kept = kept_box[0]
    prims = [
        ("reduce:f32", "reduce_f32", n28 * 4),
        ("reduce:i64", "reduce_i64", n28 * 8),
        ("scan:u32", "inclusive_scan_u32", 2 * n28 * 4),
        ("exclusive_scan:u32", "exclusive_scan_u32", 2 * n28 * 4),
        ("sort:u32", "sort_u32", 2 * n26 * 4),
        ("sort:f64", "sort_f64", 2 * n25 * 8),
        ("copy_if:i32", "copy_if_i32", n28 * 4 + kept * 4),
    ]
    def prof(key):
        tot = ctypes.c_double()
        calls = ctypes.c_uint64()
        if lib.pxr_prof_get(key.encode(), ctypes.byref(tot), ctypes.byref(calls)) != 0:
            return 0.0, 0
        return tot.value, calls.value

    per = {}
    step_bytes = 0
    for key, name, bytes_ in prims:
        tot, calls = prof(key)
        ms = tot / max(calls, 1)
        gbps = bytes_ / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        per[name] = {"ms": round(ms, 4), "alg_bytes": bytes_, "gbps": round(gbps, 1),
                     "pct_hbm_peak": round(100 * gbps / HBM_PEAK_GBPS, 1)}
        step_bytes += bytes_
    # Kernels, one launch at a time (live HIP events on the launch stream): the
    # single-launch primitives, and the kernels a sort is made of: the radix digit
    # histograms (one read, n*s algorithmic bytes) and digit passes (one read + write,
    # 2*n*s), or for the bitonic path its tile sort, register passes and LDS merges
    # (2*n*s each).
    kern = {}
    for key, name, bytes_ in prims:
        if not key.startswith("sort"):
            tot, calls = prof(key)
            kern[name] = (tot, calls, bytes_)
    for dt, n, s in (("u32", n26, 4), ("f64", n25, 8)):
        for k, nb in (("sort_hist", n * s), ("sort_radix", 2 * n * s), ("sort_tile", 2 * n * s),
                      ("sort_global", 2 * n * s), ("sort_merge", 2 * n * s)):
            tot, calls = prof(f"{k}:{dt}")
            kern[f"{k}_{dt}"] = (tot, calls, nb)
    lib.pxr_prof_enable(0)
    kernels = {}
    for name, (tot, calls, bytes_) in kern.items():
        if calls == 0:
            continue
        avg = tot / calls
        kernels[name] = {"launches_per_step": round(calls / args.steps, 2), "avg_ms": round(avg, 4),
                         "ms_per_step": round(tot / args.steps, 4), "alg_bytes_per_launch": bytes_,
                         "gbps": round(bytes_ / (avg * 1e-3) / 1e9, 1)}
    ms_step = wall * 1e3 / args.steps
    value = step_bytes / (ms_step * 1e-3) / 1e9
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    # The drop-in path with host operands (the reference's interface): every call stages
    # its operand over PCIe and copies the result back.  Reported beside, never `value`.
    host = {}
    hx = {"red_f32": X["reduce_f32"], "red_i64": X["reduce_i64"], "scan": X["scan_u32"].copy(),
          "sort_u32": X["sort_u32"].copy(), "sort_f64": X["sort_f64"].copy(), "cif": X["copy_if_i32"]}
    hres = np.zeros(2, dtype=np.int64)
    hout = np.empty_like(X["copy_if_i32"])
    host_calls = [
        ("reduce_f32", n28 * 4, lambda: lib.parallax_reduce(K["red_f32"], hx["red_f32"].ctypes.data, n28, 4, hres.ctypes.data)),
        ("reduce_i64", n28 * 8, lambda: lib.parallax_reduce(K["red_i64"], hx["red_i64"].ctypes.data, n28, 8, hres.ctypes.data)),
        ("inclusive_scan_u32", 2 * n28 * 4, lambda: lib.parallax_scan(K["scan"], K["add"], hx["scan"].ctypes.data, n28, 4)),
        ("sort_u32", 2 * n26 * 4, lambda: lib.parallax_sort(K["sort_u32"], hx["sort_u32"].ctypes.data, n26, 4)),
        ("sort_f64", 2 * n25 * 8, lambda: lib.parallax_sort(K["sort_f64"], hx["sort_f64"].ctypes.data, n25, 8)),
        ("copy_if_i32", n28 * 4 + kept * 4, lambda: lib.parallax_copy_if(K["flags"], K["scan_i32"], K["add_i32"], K["scatter"],
                                                                        hx["cif"].ctypes.data, hout.ctypes.data, n28, 4, 0)),
    ]
    for name, bytes_, fn in host_calls:
        t0 = time.perf_counter()
        fn()
        dt = time.perf_counter() - t0
        host[name] = {"ms": round(dt * 1e3, 3), "gbps": round(bytes_ / dt / 1e9, 1)}
    log(f"[bench] C ABI with host operands (PCIe-inclusive): {host}")
    # Fused map-reduce (the transform_reduce / count_if funnels' device pass), one read of
    # the input: squared sum of the 2^28 f32 input in f64, count of even values of the
    # 2^28 i32 input.  Device time per call from the launcher's events.
    fused = {}
    sq = load(make_desc(C.OP_TRANSFORM, PX_F32, aux=PX_F64, fn=C.MAP_SQUARE))
    red64 = load(make_desc(C.OP_REDUCE, PX_F64))
    ev = load(make_desc(C.OP_TRANSFORM, PX_I32, aux=PX_I32, fn=C.MAP_PRED, fn2=C.PRED_MOD_EQ, c0=2, c1=0))
    red32 = load(make_desc(C.OP_REDUCE, PX_I32))
    fres = np.zeros(2, dtype=np.float64)
    cres = np.zeros(2, dtype=np.int32)
    lib.pxr_prof_enable(1)
    for _ in range(args.warmup + args.steps):
        lib.parallax_transform_reduce(sq, red64, D["red_f32"].ptr, n28, 4, 8, fres.ctypes.data)
        lib.parallax_transform_reduce(ev, red32, D["cif_in"].ptr, n28, 4, 4, cres.ctypes.data)
    for key, name, dt_in, nb in (("transform_reduce:f32", "square_sum_f32_to_f64", "f32", n28 * 4),
                                 ("transform_reduce:i32", "count_if_even_i32", "i32", n28 * 4)):
        tot, calls = prof(key)
        ms = tot / max(calls, 1)
        fused[name] = {"ms": round(ms, 4), "alg_bytes": nb, "gbps": round(nb / (ms * 1e-3) / 1e9, 1),
                       "pct_hbm_peak": round(100 * nb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 1)}
    lib.pxr_prof_enable(0)
    import math
    # reference: x*x is exact in f64; sum it in 80-bit extended precision, in chunks
    exact_ld = np.longdouble(0)
    for c0 in range(0, n28, 1 << 24):
        y = X["reduce_f32"][c0:c0 + (1 << 24)].astype(np.float64)
        exact_ld += np.sum(np.square(y).astype(np.longdouble))
    exact = float(exact_ld)
    bnd = 2 * math.ceil(math.log2(n28)) * float(np.finfo(np.float64).eps) * exact
    assert abs(fres[0] - exact) <= bnd, ("square sum", fres[0], exact)
    even = int(np.count_nonzero((X["copy_if_i32"] & 1) == 0))
    assert int(cres[0]) == even, ("count_if", int(cres[0]), even)
    fused["parity"] = {"square_sum_err_over_bound": abs(fres[0] - exact) / bnd, "count_if": "exact"}
    log(f"[bench] fused map-reduce: {fused}")
    search = search_suite(args, X, D, n28)
    log(f"[bench] search: {search}")
    ach = kernels[dom]["gbps"]
    roof = {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBPS,
            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBPS, 4),
            "traffic": measured_traffic(dom),
            "units_per_launch": f"{kernels[dom]['alg_bytes_per_launch']} algorithmic bytes per launch",
            "avg_launch_ms": kernels[dom]["avg_ms"]}
    return {
        "value": round(value, 1), "unit": "GB/s", "ms_per_step": round(ms_step, 3),
        "dtype": "mixed(f32,i64,u32,f64,i32)",
        "config": {"workload": "plain per-call primitives via the C ABI, HBM-resident operands: "
                               "reduce 2^28 f32 + 2^28 i64, inclusive + exclusive scan 2^28 u32, "
                               "sort 2^26 u32 + 2^25 f64, copy_if 2^28 i32 (x%2==0), seed 1",
                   "scale": scale},
        "roofline": roof, "per_primitive": per, "kernels": kernels, "parity": parity,
        "c_abi_host_operands": host, "fused_map_reduce": fused, "search": search,
    }


def search_suite(args, X, D, n):
    """find / argminmax / mismatch over the suite's 2^28 inputs (device-resident, one
    stream; device time per call from events, median of --steps calls): find with a
    predicate no element passes (the whole array is read), argmin of the f32 input,
    argmax of the i32 input, mismatch of the i32 input with a copy differing 5 elements
    from the end.  Algorithmic bytes: n*s (mismatch 2*n*s)."""
    import ctypes
    import statistics
    from parallax_amd import lib, check, make_desc, consts as C
    from parallax_amd.device import Context, DeviceBuffer
    from parallax_amd.dtypes import PX_F32, PX_I32
    ctx = Context()
    idx = DeviceBuffer(8)
    other = DeviceBuffer.from_array(X["copy_if_i32"])
    k = n - 5
    bumped = np.array([X["copy_if_i32"][k] ^ 1], dtype=np.int32)
    check(lib.pxd_memcpy_h2d(other.addr(4 * k), bumped.ctypes.data, 4), "upload")
    never = make_desc(C.OP_FIND, PX_I32, fn=C.PRED_GT, c0=(1 << 31) - 1)
    cases = [
        ("find_i32_no_hit", n * 4, n, lambda: lib.pxd_find(ctx.h, ctypes.byref(never), D["cif_in"].ptr, n, idx.ptr)),
        ("argmin_f32", n * 4, int(np.argmin(X["reduce_f32"])),
         lambda: lib.pxd_argminmax(ctx.h, PX_F32, D["red_f32"].ptr, n, 0, 0, idx.ptr)),
        ("argmax_i32", n * 4, int(np.argmax(X["copy_if_i32"])),
         lambda: lib.pxd_argminmax(ctx.h, PX_I32, D["cif_in"].ptr, n, 1, 0, idx.ptr)),
        ("mismatch_i32", 2 * n * 4, k, lambda: lib.pxd_mismatch(ctx.h, PX_I32, D["cif_in"].ptr, other.ptr, n, idx.ptr)),
    ]
    out = {}
    for name, nb, expect, fn in cases:
        ms = []
        for i in range(args.warmup + args.steps):
            ctx.timer_begin()
            check(fn(), name)
            t = ctx.timer_end()
            if i >= args.warmup:
                ms.append(t)
        ctx.sync()
        got = int(idx.download(np.uint64, 1)[0])
        assert got == expect, (name, got, expect)
        m = statistics.median(ms)
        out[name] = {"ms": round(m, 4), "alg_bytes": nb, "gbps": round(nb / (m * 1e-3) / 1e9, 1),
                     "pct_hbm_peak": round(100 * nb / (m * 1e-3) / 1e9 / HBM_PEAK_GBPS, 1), "index": got}
    return out


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (profiles/traffic.txt: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE per dispatch, scaled
    by the same counters on a copy of known size), or None when not measured."""
    f = ROOT / "profiles" / "traffic.txt"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        parts = line.split()
        if len(parts) >= 4 and not line.startswith("#") and parts[0] == kernel:
            try:
                return int(float(parts[1]) + float(parts[2]))
            except ValueError:
                return None
    return None


def verify(X, D, su, sf, kept):
    """Full-size checks: integers, sort and compaction bit-exact; floats within the bound."""
    import math
    res = D["res"].download(np.uint8, 16)
    out = {}
    f32 = res[:4].view(np.float32)[0]
    a = X["reduce_f32"]
    exact = float(np.sum(a, dtype=np.float64))
    ab = float(np.sum(np.abs(a), dtype=np.float64))
    bound = 2 * math.ceil(math.log2(a.size)) * float(np.finfo(np.float32).eps) * ab
    err = abs(float(f32) - exact)
    assert err <= bound, ("reduce f32", err, bound)
    out["reduce_f32_err_over_bound"] = err / bound
    i64 = res[8:16].view(np.int64)[0]
    assert i64 == np.sum(X["reduce_i64"], dtype=np.int64), "reduce i64"
    s = X["scan_u32"]
    cs = np.cumsum(s, dtype=np.uint32)
    assert np.array_equal(D["scan_inplace"].download(np.uint32, s.size), cs), "inclusive scan"
    ex = D["scan_out"].download(np.uint32, s.size)
    assert ex[0] == 0 and np.array_equal(ex[1:], cs[:-1]), "exclusive scan"
    assert np.array_equal(su.download(np.uint32, X["sort_u32"].size), np.sort(X["sort_u32"])), "sort u32"
    assert sf.download(np.float64, X["sort_f64"].size).tobytes() == np.sort(X["sort_f64"]).tobytes(), "sort f64"
    c = X["copy_if_i32"]
    ref = c[np.fmod(c, 2) == 0]
    assert kept == ref.size and np.array_equal(D["cif_out"].download(np.int32, kept), ref), "copy_if"
    out["bit_exact"] = ["reduce_i64", "scans_u32", "sort_u32", "sort_f64", "copy_if_i32"]
    return out


def verify_c1(lib, load, C, PX_F32):
    """C1: sum of 2^20 float32 uniform [0,1) (seed 0) on the GPU through the C ABI, with
    its error against the FP64 sum and the stated bound."""
    import math
    from parallax_amd import datagen, make_desc
    a = datagen.uniform01_f32(0, 1 << 20)
    r = ctypes.c_float()
    lib.parallax_reduce(load(make_desc(C.OP_REDUCE, PX_F32)), a.ctypes.data, a.size, 4, ctypes.byref(r))
    exact = math.fsum(a.astype(np.float64))
    bnd = 2 * math.ceil(math.log2(a.size)) * float(np.finfo(np.float32).eps) * exact
    err = abs(float(r.value) - exact)
    assert err <= bnd, ("C1", err, bnd)
    return {"gpu_sum": float(r.value), "fp64_sum": exact, "abs_err": err, "bound": bnd, "err_over_bound": err / bnd}


# ------------------------------------------------------------------ batched forms
def batched(args):
    """B1: 16,384 f32 segments, lengths log-uniform [1, 2^18], seed 2: segmented reduce and
    inclusive scan, each one persistent launch replayed from a hipGraph.
    B2: 4,096 u64 segments, lengths log-uniform [16, 2^18], keys uniform [0, 2^20), seed 3:
    segmented sort then segmented unique, replayed from hipGraphs.  value = algorithmic
    bytes of B1 + B2 / device time of one replay of the four graphs."""
    import math
    from parallax_amd import lib, check, make_desc, consts as C, datagen
    from parallax_amd.device import Context, DeviceBuffer
    from parallax_amd.dtypes import PX_F32, PX_U64

    s = args.scale
    nseg1 = max(1, int(16384 * s))
    lens1 = datagen.log_uniform_lengths(2, nseg1, 1, 1 << 18, stream=1)
    off1 = datagen.offsets_from_lengths(lens1)
    x1 = datagen.normal_f32(2, int(off1[-1]), stream=2)
    nseg2 = max(1, int(4096 * s))
    lens2 = datagen.log_uniform_lengths(3, nseg2, 16, 1 << 18, stream=1)
    off2 = datagen.offsets_from_lengths(lens2)
    x2 = datagen.int_range(3, int(off2[-1]), 0, 1 << 20, np.uint64, stream=2)
    log(f"[bench] batched inputs: B1 {off1[-1]} f32 in {nseg1} segments, B2 {off2[-1]} u64 in {nseg2}")
    ctx = Context()
    d1 = DeviceBuffer.from_array(x1)
    r1 = DeviceBuffer(8 * nseg1)
    sc1 = DeviceBuffer(x1.nbytes + 16)
    d2 = DeviceBuffer.from_array(x2)
    u2 = DeviceBuffer(x2.nbytes + 16)
    c2 = DeviceBuffer(8 * nseg2)
    o1 = off1.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
    o2 = off2.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
    ud = make_desc(C.OP_FLAGS, PX_U64, fn=C.PRED_UNIQUE)
    plans = {
        "seg_reduce_f32": lib.pxb_plan_create(ctx.h, C.OP_REDUCE, PX_F32, None, d1.ptr, r1.ptr, o1, nseg1, None),
        "seg_scan_f32": lib.pxb_plan_create(ctx.h, C.OP_SCAN, PX_F32, None, d1.ptr, sc1.ptr, o1, nseg1, None),
        "seg_sort_u64": lib.pxb_plan_create(ctx.h, C.OP_SORT, PX_U64, None, d2.ptr, d2.ptr, o2, nseg2, None),
        "seg_unique_u64": lib.pxb_plan_create(ctx.h, C.OP_FLAGS, PX_U64, ctypes.byref(ud), d2.ptr, u2.ptr, o2, nseg2, c2.ptr),
    }
    assert all(plans.values()), plans
    tiles = {}
    for k, p in plans.items():
        blk, it, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        check(lib.pxb_plan_tile(p, ctypes.byref(blk), ctypes.byref(it), ctypes.byref(ms)), "plan_tile")
        if blk.value:
            tiles[k] = {"threads": blk.value, "vectors_per_thread": it.value, "tuned_ms": round(ms.value, 4)}
    # parity (after the plan warm-up runs): per segment vs float64 reference / numpy
    ctx.sync()
    got = r1.download(np.float32, nseg1)
    idx = off1[:-1].astype(np.int64)
    exact = np.add.reduceat(x1.astype(np.float64), idx) if x1.size else np.zeros(nseg1)
    absum = np.add.reduceat(np.abs(x1.astype(np.float64)), idx)
    exact[lens1 == 0] = 0
    absum[lens1 == 0] = 0
    eps = float(np.finfo(np.float32).eps)
    bnd = np.array([2 * max(1, math.ceil(math.log2(max(int(l), 2)))) * eps for l in lens1]) * absum
    ratio = np.abs(got.astype(np.float64) - exact) / np.maximum(bnd, 1e-300)
    assert np.all(ratio <= 1.0), float(ratio.max())
    # sort + unique (the unique plan ran on the pre-sort data during warm-up: re-run in order)
    d2.upload(x2)
    check(lib.pxb_plan_launch(plans["seg_sort_u64"]), "sort replay")
    check(lib.pxb_plan_launch(plans["seg_unique_u64"]), "unique replay")
    ctx.sync()
    srt = d2.download(np.uint64, x2.size)
    cnt = c2.download(np.uint64, nseg2)
    uq = u2.download(np.uint64, x2.size)
    kept_total = int(cnt.sum())
    for i in range(0, nseg2, max(1, nseg2 // 64)):
        seg = np.sort(x2[off2[i]:off2[i + 1]])
        assert np.array_equal(srt[off2[i]:off2[i + 1]], seg)
        u = np.unique(seg)
        assert cnt[i] == u.size and np.array_equal(uq[off2[i]:off2[i] + u.size], u)
    parity = {"seg_reduce_f32_max_err_over_bound": float(ratio.max()),
              "sort_unique_u64": "bit-exact (sampled 64 segments)"}
    # timed replays
    times = {k: [] for k in plans}
    for it in range(args.warmup + args.steps):
        d2.upload(x2)  # fresh unsorted keys for the sort replay (outside the timed graphs)
        for k in ["seg_reduce_f32", "seg_scan_f32", "seg_sort_u64", "seg_unique_u64"]:
            ctx.timer_begin()
            check(lib.pxb_plan_launch(plans[k]), k)
            ms = ctx.timer_end()
            if it >= args.warmup:
                times[k].append(ms)
    n1, n2 = int(off1[-1]), int(off2[-1])
    alg = {"seg_reduce_f32": n1 * 4 + (nseg1 + 1) * 8, "seg_scan_f32": 2 * n1 * 4 + (nseg1 + 1) * 8,
           "seg_sort_u64": 2 * n2 * 8 + (nseg2 + 1) * 8, "seg_unique_u64": n2 * 8 + kept_total * 8 + (nseg2 + 1) * 8}
    per = {}
    total_ms = 0.0
    for k in plans:
        ms = float(np.median(times[k]))
        total_ms += ms
        gbps = alg[k] / (ms * 1e-3) / 1e9
        per[k] = {"ms": round(ms, 4), "alg_bytes": alg[k], "gbps": round(gbps, 1),
                  "pct_hbm_peak": round(100 * gbps / HBM_PEAK_GBPS, 1)}
    for p in plans.values():
        lib.pxb_plan_destroy(p)
    value = sum(alg.values()) / (total_ms * 1e-3) / 1e9
    dom = max(per, key=lambda k: per[k]["ms"])
    return {"value": round(value, 1), "unit": "GB/s", "ms_per_step": round(total_ms, 3),
            "dtype": "mixed(f32,u64)",
            "config": {"workload": f"batched: B1 {nseg1} f32 segments log-uniform [1,2^18] (seg reduce + "
                                   f"inclusive scan), B2 {nseg2} u64 segments log-uniform [16,2^18] "
                                   "(seg sort + unique); hipGraph replays", "scale": s},
            "roofline": {"bound": "hbm", "kernel": dom, "achieved": per[dom]["gbps"], "peak": HBM_PEAK_GBPS,
                         "unit": "GB/s", "frac": round(per[dom]["gbps"] / HBM_PEAK_GBPS, 4), "traffic": None},
            "per_primitive": per, "autotuned_tiles": tiles, "parity": parity}


# ------------------------------------------------------------------ served form
def served(args):
    """Served mixed stream (seed 4): requests of 30% reduce, 30% scan, 25% sort, 15%
    copy_if; sizes log-uniform [2^10, 2^24]; dtypes uniform over {i32, u64, f32, f64};
    values uniform in [-1000, 1000] ([0, 1000] unsigned); 64 GB budget; inputs resident in
    the server's pool.  Reported: saturated req/s (all requests queued at once; median of
    --steps passes after --warmup), p50/p99 latency under Poisson arrivals at half that
    rate, and the same requests issued as plain per-call synchronous calls
    (device-resident operands; C ABI with host operands) for the speed-up."""
    import threading
    from parallax_amd import lib, check, make_desc, descriptor_words, consts as C, datagen
    from parallax_amd._lib import PxsConfig, PxsRequest, PxsStats
    from parallax_amd.device import Context
    from parallax_amd.dtypes import NP_DTYPES, PX_F32, PX_F64, PX_I32, PX_I64, PX_U64

    nreq = max(8, int(400 * args.scale))
    rng_u = datagen.uniform01_f64(4, 4 * nreq, stream=1)
    ops = np.array([C.OP_REDUCE, C.OP_SCAN, C.OP_SORT, C.OP_FLAGS])
    op_of = ops[np.searchsorted(np.cumsum([0.3, 0.3, 0.25, 0.15]), rng_u[:nreq], side="right").clip(0, 3)]
    dts = [PX_I32, PX_U64, PX_F32, PX_F64]
    dt_of = [dts[int(u * 4) % 4] for u in rng_u[nreq:2 * nreq]]
    sizes = np.floor(2.0 ** (10 + 14 * rng_u[2 * nreq:3 * nreq])).astype(np.int64)
    cfg = PxsConfig()
    lib.pxs_config_default(ctypes.byref(cfg))
    server = lib.pxs_create(ctypes.byref(cfg))
    assert server, "pxs_create failed"
    reqs = []
    for i in range(nreq):
        dt, n = dt_of[i], int(sizes[i])
        if dt == PX_U64:
            a = datagen.int_range(4, n, 0, 1001, np.uint64, stream=100 + i)
        elif dt == PX_I32:
            a = datagen.int_range(4, n, -1000, 1001, np.int32, stream=100 + i)
        else:
            a = (datagen.uniform01_f64(4, n, stream=100 + i) * 2000.0 - 1000.0).astype(NP_DTYPES[dt])
        p = lib.pxs_alloc(server, a.nbytes)
        assert p
        check(lib.pxd_memcpy_h2d(p, a.ctypes.data, a.nbytes), "upload")
        fl = make_desc(C.OP_FLAGS, dt, fn=C.PRED_LT, c0=500 if dt == PX_U64 else 0)
        reqs.append((int(op_of[i]), dt, a, p, fl))
    total_bytes = sum(a.nbytes * (1 if op == C.OP_REDUCE else 2) for op, _, a, _, _ in reqs)
    log(f"[bench] served: {nreq} requests, {total_bytes / 1e9:.2f} GB algorithmic")

    def restore_sorts():
        for op, dt, a, p, _ in reqs:
            if op == C.OP_SORT:
                check(lib.pxd_memcpy_h2d(p, a.ctypes.data, a.nbytes), "restore")

    def run_served(arrivals):
        lib.pxs_reset_stats(server)
        ids = []
        t0 = time.perf_counter()
        for (op, dt, a, p, fl), t_arr in zip(reqs, arrivals):
            d = t0 + t_arr - time.perf_counter()
            if d > 0:
                time.sleep(d)
            rid = lib.pxs_submit(server, ctypes.byref(PxsRequest(op, dt, a.size, p, fl)))
            assert rid >= 0
            ids.append(rid)
        lib.pxs_drain(server)
        st = PxsStats()
        lib.pxs_get_stats(server, ctypes.byref(st))
        for rid in ids:
            lib.pxs_release(server, rid)
        return st

    # warm-up passes, then saturated throughput: the median of `steps` passes (one pass is
    # only a few ms of device work, so single passes vary by several percent)
    for _ in range(max(1, args.warmup)):
        restore_sorts()
        run_served([0.0] * nreq)
    passes = []
    for _ in range(max(1, args.steps)):
        restore_sorts()
        passes.append(run_served([0.0] * nreq))
    passes.sort(key=lambda st: st.req_per_s)
    sat = passes[len(passes) // 2]
    # Poisson arrivals at half the saturated rate
    lam = 0.5 * sat.req_per_s
    gaps = -np.log1p(-datagen.uniform01_f64(4, nreq, stream=9)) / lam
    restore_sorts()
    poi = run_served(np.cumsum(gaps) - gaps[0])
    # plain per-call: the same requests, one synchronous call after another
    ctx = Context()
    out = lib.pxd_malloc(max(a.nbytes for _, _, a, _, _ in reqs) + 64)
    cnt = lib.pxd_malloc(64)
    restore_sorts()

    def plain_all():
        for op, dt, a, p, fl in reqs:
            if op == C.OP_REDUCE:
                check(lib.pxd_reduce(ctx.h, dt, p, a.size, out), "reduce")
            elif op == C.OP_SCAN:
                check(lib.pxd_inclusive_scan(ctx.h, dt, p, out, a.size), "scan")
            elif op == C.OP_SORT:
                check(lib.pxd_sort(ctx.h, dt, p, a.size), "sort")
            else:
                check(lib.pxd_copy_if(ctx.h, ctypes.byref(fl), p, out, a.size, cnt), "copy_if")
            ctx.sync()
    plain_all()
    plain_t = []
    for _ in range(5):
        restore_sorts()
        t0 = time.perf_counter()
        plain_all()
        plain_t.append(time.perf_counter() - t0)
    plain_rps = nreq / sorted(plain_t)[2]
    lib.pxd_free(out)
    lib.pxd_free(cnt)
    lib.pxs_destroy(server)

    # the reference's own per-call path: the C ABI of parallax/runtime.h with host
    # operands (each call stages its operand to the device and its result back)
    def load(d):
        w, nw = descriptor_words(d)
        k = lib.parallax_kernel_load(w, nw)
        assert k, "kernel load"
        return k
    K = {}
    for dt in set(dt_of):
        K[dt] = {"red": load(make_desc(C.OP_REDUCE, dt)), "scan": load(make_desc(C.OP_SCAN, dt)),
                 "add": load(make_desc(C.OP_SCAN_ADD, dt)), "sort": load(make_desc(C.OP_SORT, dt)),
                 "iscan": load(make_desc(C.OP_SCAN, PX_I64)), "iadd": load(make_desc(C.OP_SCAN_ADD, PX_I64)),
                 "scatter": load(make_desc(C.OP_SCATTER, dt, aux=C.SCATTER_COMPACT))}
    res = np.zeros(16, dtype=np.uint8)

    def abi_all(work):
        for (op, dt, a, p, fl), h in zip(reqs, work):
            k, es = K[dt], a.itemsize
            if op == C.OP_REDUCE:
                lib.parallax_reduce(k["red"], h.ctypes.data, h.size, es, res.ctypes.data)
            elif op == C.OP_SCAN:
                lib.parallax_scan(k["scan"], k["add"], h.ctypes.data, h.size, es)
            elif op == C.OP_SORT:
                lib.parallax_sort(k["sort"], h.ctypes.data, h.size, es)
            else:
                o = np.empty_like(h)
                lib.parallax_copy_if(load_cached(fl), k["iscan"], k["iadd"], k["scatter"],
                                     h.ctypes.data, o.ctypes.data, h.size, es, 0)
    flk = {}

    def load_cached(fl):
        key = id(fl)
        if key not in flk:
            flk[key] = load(fl)
        return flk[key]
    abi_all([a.copy() for _, _, a, _, _ in reqs])
    abi_t = []
    for _ in range(3):
        work = [a.copy() for _, _, a, _, _ in reqs]
        t0 = time.perf_counter()
        abi_all(work)
        abi_t.append(time.perf_counter() - t0)
    abi_rps = nreq / sorted(abi_t)[1]
    return {"value": round(sat.req_per_s, 1), "unit": "req/s", "ms_per_step": round(1e3 / sat.req_per_s, 4),
            "dtype": "mixed(i32,u64,f32,f64)",
            "config": {"workload": f"served mixed stream: {nreq} requests (30% reduce, 30% scan, 25% sort, "
                                   "15% copy_if), sizes log-uniform [2^10,2^24], dtypes {i32,u64,f32,f64}, "
                                   "seed 4, 64 GB pool, 4 slots", "scale": args.scale},
            "req_per_s": {"served_saturated": round(sat.req_per_s, 1), "served_poisson": round(poi.req_per_s, 1),
                          "saturated_passes": len(passes), "saturated_min": round(passes[0].req_per_s, 1),
                          "saturated_max": round(passes[-1].req_per_s, 1),
                          "plain_per_call_c_abi_host_operands": round(abi_rps, 1),
                          "plain_per_call_device_resident": round(plain_rps, 1)},
            "latency_ms": {"poisson_rate_req_s": round(lam, 1), "p50": round(poi.p50_ms, 3), "p99": round(poi.p99_ms, 3),
                           "saturated_p50": round(sat.p50_ms, 3), "saturated_p99": round(sat.p99_ms, 3)},
            "speedup_vs_plain": {"c_abi_host_operands": round(sat.req_per_s / abi_rps, 2),
                                 "device_resident": round(sat.req_per_s / plain_rps, 2)},
            "roofline": None}


# ------------------------------------------------------------------ counter reference
def copy_probe():
    """Three device-to-device copies of exactly 1 GiB (1 GiB read + 1 GiB written per
    dispatch): the reference that converts FETCH_SIZE / WRITE_SIZE readings to bytes."""
    from parallax_amd import lib, check
    nb = 1 << 30
    a, b = lib.pxd_malloc(nb), lib.pxd_malloc(nb)
    assert a and b, "pxd_malloc"
    check(lib.pxd_memset(a, 1, nb), "memset")
    for _ in range(3):
        check(lib.pxd_memcpy_d2d(b, a, nb), "d2d")
    lib.pxd_device_sync()
    lib.pxd_free(a)
    lib.pxd_free(b)
    log("[bench] copy probe: 3 x 1 GiB device-to-device copies")


# ------------------------------------------------------------------ CPU baseline
CPU_SAMPLE = {"n_stream": 1 << 28, "n_sort_u32": 1 << 26, "n_sort_f64": 1 << 25}


def cpu_baseline_child():
    """Runs in a child process pinned to one core with PARALLAX_FORCE_CPU=1: the
    reference's own fallback (oracle/_ref, built from its stdpar.hpp) on a bounded
    sample of the same suite; prints one JSON line."""
    import pxo
    from parallax_amd import datagen

    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except Exception:
        pass
    R = pxo.load_ref()
    kind = "reference"
    ns, nu, nf = CPU_SAMPLE["n_stream"], CPU_SAMPLE["n_sort_u32"], CPU_SAMPLE["n_sort_f64"]
    f32 = datagen.normal_f32(1, ns, stream=1)
    i64 = datagen.int_range(1, ns, -(1 << 20), 1 << 20, np.int64, stream=2)
    u32 = datagen.int_range(1, ns, 0, 16, np.uint32, stream=3)
    su = datagen.with_duplicates(datagen.full_u32(1, nu, stream=4), 1, 1, stream=40)
    sf = datagen.with_duplicates(datagen.uniform01_f64(1, nf, stream=5), 1, 1, stream=50)
    ci = datagen.full_i32(1, ns, stream=6)
    out_u = np.empty_like(u32)
    out_i = np.empty_like(ci)
    vp, sz = ctypes.c_void_p, ctypes.c_size_t

    def f(name, res, *a):
        fn = getattr(R, name)
        fn.restype = res
        fn.argtypes = list(a)
        return fn

    if R is None:
        raise SystemExit("oracle/_ref not built")
    calls = [
        ("reduce_f32", ns * 4, lambda: f("ref_reduce_f32", ctypes.c_float, vp, sz)(f32.ctypes.data, ns)),
        ("reduce_i64", ns * 8, lambda: f("ref_reduce_i64", ctypes.c_int64, vp, sz)(i64.ctypes.data, ns)),
        ("inclusive_scan_u32", 2 * ns * 4, lambda: f("ref_inclusive_scan_u32", None, vp, vp, sz)(u32.ctypes.data, out_u.ctypes.data, ns)),
        ("exclusive_scan_u32", 2 * ns * 4, lambda: f("ref_exclusive_scan_u32", None, vp, vp, sz, ctypes.c_uint32)(u32.ctypes.data, out_u.ctypes.data, ns, 0)),
        ("sort_u32", 2 * nu * 4, None),
        ("sort_f64", 2 * nf * 8, None),
        ("copy_if_i32", None, lambda: f("ref_copy_if_even_i32", sz, vp, vp, sz)(ci.ctypes.data, out_i.ctypes.data, ns)),
    ]
    times = {}
    total_bytes = 0
    for name, bytes_, fn in calls:
        reps = []
        for _ in range(1 if name.startswith("sort") else 3):
            if name == "sort_u32":
                w = su.copy()
                t = time.perf_counter()
                f("ref_sort_u32", None, vp, sz)(w.ctypes.data, nu)
            elif name == "sort_f64":
                w = sf.copy()
                t = time.perf_counter()
                f("ref_sort_f64", None, vp, sz)(w.ctypes.data, nf)
            else:
                t = time.perf_counter()
                r = fn()
            reps.append(time.perf_counter() - t)
            if name == "copy_if_i32":
                bytes_ = ns * 4 + int(r) * 4
        times[name] = sorted(reps)[len(reps) // 2]
        total_bytes += bytes_
    total = sum(times.values())
    # C1: std::reduce over 2^20 f32 uniform [0,1) (seed 0) through the fallback, against FP64
    import math
    c1 = datagen.uniform01_f32(0, 1 << 20)
    fb = float(f("ref_reduce_f32", ctypes.c_float, vp, sz)(c1.ctypes.data, c1.size))
    exact = math.fsum(c1.astype(np.float64))
    bnd = 2 * math.ceil(math.log2(c1.size)) * float(np.finfo(np.float32).eps) * exact
    c1r = {"fallback_sum": fb, "fp64_sum": exact, "abs_err": abs(fb - exact), "bound": bnd,
           "err_over_bound": abs(fb - exact) / bnd}
    cpu = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except Exception:
        pass
    print(json.dumps({"value": round(total_bytes / total / 1e9, 3), "unit": "GB/s", "cores": 1,
                      "kind": kind,
                      "sample": (f"reference CPU fallback (stdpar.hpp host loops, g++ -O3, 1 core; median of 3 "
                                 f"for the streaming primitives, one run per sort) "
                                 f"on the full suite: streams 2^28, sort u32 2^26, sort f64 2^25, same "
                                 f"generators; host {cpu}"),
                      "seconds": {k: round(v, 4) for k, v in times.items()}, "c1_reduce_2p20_f32": c1r}))


def cpu_baseline():
    env = dict(os.environ, PARALLAX_FORCE_CPU="1")
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--cpu-baseline-child"], env=env,
                       capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        log("[bench] cpu baseline failed:", r.stderr[-2000:])
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])


# ------------------------------------------------------------------ main
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="plain", choices=["plain", "batched", "served"])
    ap.add_argument("--scale", type=float, default=1.0, help="size multiplier (1.0 = the configuration)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-baseline-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--copy-probe", action="store_true",
                    help="copy 1 GiB device to device 3 times (the known-size reference that scales "
                         "FETCH_SIZE / WRITE_SIZE counter readings)")
    args = ap.parse_args()
    if args.cpu_baseline_child:
        cpu_baseline_child()
        return
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only (the primitives run on one device); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "plain":
        r = plain_suite(args)
    elif args.workload == "batched":
        r = batched(args)
    else:
        r = served(args)
    cpu = None if (args.no_cpu_baseline or args.workload != "plain") else cpu_baseline()
    line = {"metric": METRIC, "value": r["value"], "unit": r["unit"], "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["ms_per_step"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": r["dtype"], "data": "synthetic", "config": r["config"],
            "roofline": r.get("roofline"), "cpu_baseline": cpu}
    for k in ("per_primitive", "kernels", "autotuned_tiles", "parity", "c_abi_host_operands",
              "fused_map_reduce", "search", "latency_ms", "req_per_s", "speedup_vs_plain"):
        if k in r:
            line[k] = r[k]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
