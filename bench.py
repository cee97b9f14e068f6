"""Benchmark of the Parallax primitive hot path on one MI355X.

Default run = the plain per-call configuration (the headline `value`), then the batched
and the served forms (the "batched" / "served" keys of the same JSON line):

plain   one step calls, through the C ABI of parallax/runtime.h with HBM-resident
        operands, exactly as the reference's funnels call it:
          reduce 2^28 float32 normal(0,1)         reduce 2^28 int64 uniform [-2^20, 2^20)
          inclusive + exclusive scan 2^28 uint32 in [0,16)
          sort 2^26 uint32 (1% duplicates)        sort 2^25 float64 (1% duplicates)
          copy_if 2^28 int32, x % 2 == 0          (all from seed 1)
        value = algorithmic bytes of one step / wall time of the step (GB/s); algorithmic
        bytes: reduce n*s, scans 2*n*s, sort 2*n*s, copy_if n*s + kept*s.  Kernel launch
        times come from HIP events on the launcher's stream over the timed steps
        (pxr_prof_*): they give `roofline` for the kernel with the most device time per
        step.  Per-call device times are also taken one call at a time after a 1 GiB write
        that flushes the L2 and the 256 MB Infinity Cache (median of 20 after 3).
batched B1: 16,384 float32 segments (lengths log-uniform [1, 2^18], seed 2), segmented
        reduce and inclusive scan; B2: 4,096 uint64 segments (lengths log-uniform
        [16, 2^18], keys [0, 2^20), seed 3), segmented sort then unique.  Each is one
        hipGraph replay (autotuned tiles); the per-call leg runs the same 40,960 segments
        as synchronous device-resident calls, one per segment.
served  400 requests (seed 4; 30% reduce, 30% scan, 25% sort, 15% copy_if; 2^10..2^24
        elements; int32 / uint64 / float32 / float64): saturated req/s, p50/p99 latency
        under Poisson arrivals at half that rate, every result checked against the
        oracle, and the same requests as plain synchronous calls (device-resident, and the
        C ABI with host operands).
Results are checked against the oracle before anything is timed.
"""
import argparse
import concurrent.futures
import ctypes
import json
import math
import os
import pathlib
import statistics
import subprocess
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "parallax-runtime_amd"))
sys.path.insert(0, str(ROOT / "oracle"))

import numpy as np  # noqa: E402

HBM_PEAK_GBPS = 8000.0
METRIC = "GB/s and % HBM peak per primitive; batched/served req/s + p99 latency"
FLUSH_BYTES = 1 << 30  # > the 256 MB Infinity Cache + every L2
RUNS, WARM = 20, 3     # flushed per-call medians


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gbps(nbytes, ms):
    return nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0


def rate(nbytes, ms):
    g = gbps(nbytes, ms)
    return {"ms": round(ms, 4), "alg_bytes": int(nbytes), "gbps": round(g, 1),
            "pct_hbm_peak": round(100 * g / HBM_PEAK_GBPS, 1)}


class Flush:
    """A 1 GiB device buffer whose synchronous overwrite evicts the L2s and the MALL."""

    def __init__(self, lib):
        from parallax_amd.device import DeviceBuffer
        self.lib = lib
        self.buf = DeviceBuffer(FLUSH_BYTES)
        self.k = 0

    def __call__(self):
        from parallax_amd import check
        self.k += 1
        check(self.lib.pxd_memset(self.buf.ptr, self.k & 0xff, FLUSH_BYTES), "flush")


def load_kernel(lib, d):
    from parallax_amd import descriptor_words
    k = lib.parallax_kernel_load(*descriptor_words(d))
    if not k:
        raise RuntimeError("kernel load failed (no GPU?)")
    return k


def prof_get(lib, key):
    tot = ctypes.c_double()
    calls = ctypes.c_uint64()
    if lib.pxr_prof_get(key.encode(), ctypes.byref(tot), ctypes.byref(calls)) != 0:
        return 0.0, 0
    return tot.value, calls.value


# ------------------------------------------------------------------ copy ceiling
def copy_ceiling(lib, flush):
    """The streaming-copy kernel (pxd_copy), 1 GiB -> 1 GiB, flushed, median of 20 after
    3: the read-and-write bandwidth every look-back kernel is compared with."""
    from parallax_amd import check
    from parallax_amd.device import Context, DeviceBuffer
    ctx = Context()
    nb = 1 << 30
    a, b = DeviceBuffer(nb), DeviceBuffer(nb)
    check(lib.pxd_memset(a.ptr, 1, nb), "memset")
    ms = []
    for i in range(WARM + RUNS):
        flush()
        ctx.timer_begin()
        check(lib.pxd_copy(ctx.h, b.ptr, a.ptr, nb), "copy")
        t = ctx.timer_end()
        if i >= WARM:
            ms.append(t)
    m = statistics.median(ms)
    return {"kernel": "k_copy16", "bytes": nb, "ms": round(m, 4), "gbps": round(gbps(2 * nb, m), 1),
            "pct_hbm_peak": round(100 * gbps(2 * nb, m) / HBM_PEAK_GBPS, 1)}


# ------------------------------------------------------------------ plain suite
def plain_suite(args, flush, copy):
    from parallax_amd import lib, make_desc, consts as C, datagen
    from parallax_amd.device import DeviceBuffer, stats
    from parallax_amd.dtypes import PX_F32, PX_F64, PX_I32, PX_I64, PX_U32

    t0 = time.time()
    X = datagen.plain_suite_inputs(seed=1, scale=args.scale)
    log(f"[bench] plain inputs generated in {time.time() - t0:.1f}s")
    L = lambda d: load_kernel(lib, d)  # noqa: E731
    K = {
        "red_f32": L(make_desc(C.OP_REDUCE, PX_F32)), "red_i64": L(make_desc(C.OP_REDUCE, PX_I64)),
        "scan": L(make_desc(C.OP_SCAN, PX_U32)), "add": L(make_desc(C.OP_SCAN_ADD, PX_U32)),
        "shift": L(make_desc(C.OP_SCAN_SHIFT, PX_U32)),
        "sort_u32": L(make_desc(C.OP_SORT, PX_U32)), "sort_f64": L(make_desc(C.OP_SORT, PX_F64)),
        "flags": L(make_desc(C.OP_FLAGS, PX_I32, fn=C.PRED_MOD_EQ, c0=2, c1=0)),
        "scan_i32": L(make_desc(C.OP_SCAN, PX_I32)), "add_i32": L(make_desc(C.OP_SCAN_ADD, PX_I32)),
        "scatter": L(make_desc(C.OP_SCATTER, PX_I32, aux=C.SCATTER_COMPACT)),
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
        "u32_src": DeviceBuffer.from_array(X["sort_u32"]),
        "f64_src": DeviceBuffer.from_array(X["sort_f64"]),
    }
    sort_u32 = [DeviceBuffer(X["sort_u32"].nbytes) for _ in range(nsteps)]
    sort_f64 = [DeviceBuffer(X["sort_f64"].nbytes) for _ in range(nsteps)]
    for b in sort_u32:
        lib.pxd_memcpy_d2d(b.ptr, D["u32_src"].ptr, X["sort_u32"].nbytes)
    for b in sort_f64:
        lib.pxd_memcpy_d2d(b.ptr, D["f64_src"].ptr, X["sort_f64"].nbytes)
    lib.pxd_device_sync()
    log(f"[bench] plain inputs resident in HBM in {time.time() - t0:.1f}s")

    n28, n26, n25 = X["reduce_f32"].size, X["sort_u32"].size, X["sort_f64"].size
    init = np.zeros(1, dtype=np.uint32)
    kept_box = [0]
    calls = {
        "reduce_f32": lambda: lib.parallax_reduce(K["red_f32"], D["red_f32"].ptr, n28, 4, D["res"].ptr),
        "reduce_i64": lambda: lib.parallax_reduce(K["red_i64"], D["red_i64"].ptr, n28, 8, D["res"].ptr + 8),
        "inclusive_scan_u32": lambda: lib.parallax_scan(K["scan"], K["add"], D["scan_inplace"].ptr, n28, 4),
        "exclusive_scan_u32": lambda: lib.parallax_exclusive_scan(K["scan"], K["add"], K["shift"], D["scan_in"].ptr,
                                                                  D["scan_out"].ptr, n28, 4, init.ctypes.data),
        "copy_if_i32": lambda: kept_box.__setitem__(0, lib.parallax_copy_if(
            K["flags"], K["scan_i32"], K["add_i32"], K["scatter"], D["cif_in"].ptr, D["cif_out"].ptr, n28, 4, 0)),
    }

    def step(i):
        calls["reduce_f32"]()
        calls["reduce_i64"]()
        calls["inclusive_scan_u32"]()
        calls["exclusive_scan_u32"]()
        lib.parallax_sort(K["sort_u32"], sort_u32[i].ptr, n26, 4)
        lib.parallax_sort(K["sort_f64"], sort_f64[i].ptr, n25, 8)
        calls["copy_if_i32"]()

    # ---- verification call (full size): integers, sort and compaction bit-exact,
    # floats within the bound
    t0 = time.time()
    step(nsteps - 1)
    parity = verify(X, D, sort_u32[nsteps - 1], sort_f64[nsteps - 1], kept_box[0])
    parity["c1_reduce_2p20_f32"] = verify_c1(lib, C, PX_F32)
    log(f"[bench] parity verified in {time.time() - t0:.1f}s: {parity}")

    # ---- timed steps (wall clock, the launcher's profiling events off)
    for i in range(args.warmup):
        step(i)
    lib.pxd_device_sync()
    lib.pxr_prof_enable(0)
    launches0 = stats()["launches"]
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(args.warmup + i)
    lib.pxd_device_sync()
    wall = time.perf_counter() - t0
    assert stats()["launches"] - launches0 >= 7 * args.steps, "HIP path did not run"
    # ---- the same steps again with the launcher's events on: device time per call and
    # per kernel launch (the sort buffers get a fresh unsorted copy before each step,
    # outside every event pair)
    lib.pxr_prof_enable(1)
    for i in range(args.steps):
        lib.pxd_memcpy_d2d(sort_u32[args.warmup + i].ptr, D["u32_src"].ptr, X["sort_u32"].nbytes)
        lib.pxd_memcpy_d2d(sort_f64[args.warmup + i].ptr, D["f64_src"].ptr, X["sort_f64"].nbytes)
        step(args.warmup + i)
    lib.pxd_device_sync()
    kept = kept_box[0]
    prims = [("reduce:f32", "reduce_f32", n28 * 4), ("reduce:i64", "reduce_i64", n28 * 8),
             ("scan:u32", "inclusive_scan_u32", 2 * n28 * 4), ("exclusive_scan:u32", "exclusive_scan_u32", 2 * n28 * 4),
             ("sort:u32", "sort_u32", 2 * n26 * 4), ("sort:f64", "sort_f64", 2 * n25 * 8),
             ("copy_if:i32", "copy_if_i32", n28 * 4 + kept * 4)]
    timed = {}
    step_bytes = 0
    for key, name, nb in prims:
        tot, c = prof_get(lib, key)
        timed[name] = rate(nb, tot / max(c, 1))
        step_bytes += nb
    # kernels, one launch at a time: the single-launch primitives, and the sort's radix
    # histogram (n*s) and digit passes (2*n*s each), or its bitonic passes (2*n*s each)
    kern = {}
    for key, name, nb in prims:
        if not key.startswith("sort"):
            tot, c = prof_get(lib, key)
            kern[name] = (tot, c, nb)
    for dt, n, s in (("u32", n26, 4), ("f64", n25, 8)):
        for k, nb in (("sort_hist", n * s), ("sort_radix", 2 * n * s), ("sort_tile", 2 * n * s),
                      ("sort_global", 2 * n * s), ("sort_merge", 2 * n * s)):
            tot, c = prof_get(lib, f"{k}:{dt}")
            kern[f"{k}_{dt}"] = (tot, c, nb)
    lib.pxr_prof_enable(0)
    kernels = {}
    for name, (tot, c, nb) in kern.items():
        if c == 0:
            continue
        avg = tot / c
        kernels[name] = {"launches_per_step": round(c / args.steps, 2), "avg_ms": round(avg, 4),
                         "ms_per_step": round(tot / args.steps, 4), "alg_bytes_per_launch": nb,
                         "gbps": round(gbps(nb, avg), 1)}
    ms_step = wall * 1e3 / args.steps
    value = gbps(step_bytes, ms_step)

    # ---- per-call device time after a cache flush (median of RUNS after WARM)
    def flushed(name, fn, prep=None):
        ms = []
        for i in range(WARM + RUNS):
            if prep:
                prep()
            flush()
            lib.pxr_prof_enable(1)
            fn()
            tot, c = prof_get(lib, name)
            if i >= WARM:
                ms.append(tot)
        lib.pxr_prof_enable(0)
        return statistics.median(ms)

    su, sf = sort_u32[0], sort_f64[0]
    per = {}
    for key, name, nb in prims:
        if name == "sort_u32":
            m = flushed(key, lambda: lib.parallax_sort(K["sort_u32"], su.ptr, n26, 4),
                        lambda: lib.pxd_memcpy_d2d(su.ptr, D["u32_src"].ptr, n26 * 4))
        elif name == "sort_f64":
            m = flushed(key, lambda: lib.parallax_sort(K["sort_f64"], sf.ptr, n25, 8),
                        lambda: lib.pxd_memcpy_d2d(sf.ptr, D["f64_src"].ptr, n25 * 8))
        else:
            m = flushed(key, calls[name])
        per[name] = rate(nb, m)
        if name not in ("sort_u32", "sort_f64", "reduce_f32", "reduce_i64"):
            per[name]["frac_vs_copy"] = round(per[name]["gbps"] / copy["gbps"], 4)
    log(f"[bench] per-call (flushed medians): {per}")

    host = host_operand_leg(lib, K, X, n28, n26, n25, kept)
    fused = fused_suite(args, lib, D, X, n28)
    search = search_suite(args, X, D, n28)
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    ach = kernels[dom]["gbps"]
    roof = {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBPS, 4), "traffic": measured_traffic(dom),
            "units_per_launch": f"{kernels[dom]['alg_bytes_per_launch']} algorithmic bytes per launch",
            "avg_launch_ms": kernels[dom]["avg_ms"], "copy_ceiling_gbps": copy["gbps"],
            "frac_vs_copy": round(ach / copy["gbps"], 4)}
    return {
        "value": round(value, 1), "unit": "GB/s", "ms_per_step": round(ms_step, 3),
        "dtype": "mixed(f32,i64,u32,f64,i32)",
        "config": {"workload": "plain per-call primitives via the C ABI, HBM-resident operands: "
                               "reduce 2^28 f32 + 2^28 i64, inclusive + exclusive scan 2^28 u32, "
                               "sort 2^26 u32 + 2^25 f64, copy_if 2^28 i32 (x%2==0), seed 1",
                   "scale": args.scale},
        "roofline": roof, "per_primitive": per, "per_primitive_timed_steps": timed, "kernels": kernels,
        "copy_ceiling": copy, "parity": parity, "c_abi_host_operands": host, "fused_map_reduce": fused,
        "search": search,
    }


def host_operand_leg(lib, K, X, n28, n26, n25, kept):
    """The drop-in path with host operands (the reference's own interface): each call
    stages its operand over PCIe and its result back.  Reported beside `value`, never as
    it; best of 3."""
    hx = {"red_f32": X["reduce_f32"], "red_i64": X["reduce_i64"], "scan": X["scan_u32"].copy(),
          "sort_u32": X["sort_u32"].copy(), "sort_f64": X["sort_f64"].copy(), "cif": X["copy_if_i32"]}
    hres = np.zeros(2, dtype=np.int64)
    hout = np.empty_like(X["copy_if_i32"])
    hout[:] = 0
    calls = [
        ("reduce_f32", n28 * 4, lambda: lib.parallax_reduce(K["red_f32"], hx["red_f32"].ctypes.data, n28, 4, hres.ctypes.data)),
        ("reduce_i64", n28 * 8, lambda: lib.parallax_reduce(K["red_i64"], hx["red_i64"].ctypes.data, n28, 8, hres.ctypes.data)),
        ("inclusive_scan_u32", 2 * n28 * 4, lambda: lib.parallax_scan(K["scan"], K["add"], hx["scan"].ctypes.data, n28, 4)),
        ("sort_u32", 2 * n26 * 4, lambda: lib.parallax_sort(K["sort_u32"], hx["sort_u32"].ctypes.data, n26, 4)),
        ("sort_f64", 2 * n25 * 8, lambda: lib.parallax_sort(K["sort_f64"], hx["sort_f64"].ctypes.data, n25, 8)),
        ("copy_if_i32", n28 * 4 + kept * 4, lambda: lib.parallax_copy_if(K["flags"], K["scan_i32"], K["add_i32"], K["scatter"],
                                                                        hx["cif"].ctypes.data, hout.ctypes.data, n28, 4, 0)),
    ]
    host = {}
    for name, nb, fn in calls:
        best = float("inf")
        for _ in range(3):
            t0 = time.perf_counter()
            fn()
            best = min(best, time.perf_counter() - t0)
        host[name] = {"ms": round(best * 1e3, 3), "gbps": round(nb / best / 1e9, 1)}
    assert np.array_equal(hx["sort_u32"], np.sort(X["sort_u32"])), "host-operand sort"
    log(f"[bench] C ABI with host operands (PCIe-inclusive): {host}")
    return host


def fused_suite(args, lib, D, X, n28):
    """Fused map-reduce (the transform_reduce / count_if funnels' device pass), one read of
    the input: squared sum of the 2^28 f32 input in f64, count of even values of the
    2^28 i32 input."""
    from parallax_amd import make_desc, consts as C
    from parallax_amd.dtypes import PX_F32, PX_F64, PX_I32
    sq = load_kernel(lib, make_desc(C.OP_TRANSFORM, PX_F32, aux=PX_F64, fn=C.MAP_SQUARE))
    red64 = load_kernel(lib, make_desc(C.OP_REDUCE, PX_F64))
    ev = load_kernel(lib, make_desc(C.OP_TRANSFORM, PX_I32, aux=PX_I32, fn=C.MAP_PRED, fn2=C.PRED_MOD_EQ, c0=2, c1=0))
    red32 = load_kernel(lib, make_desc(C.OP_REDUCE, PX_I32))
    fres = np.zeros(2, dtype=np.float64)
    cres = np.zeros(2, dtype=np.int32)
    lib.pxr_prof_enable(1)
    for _ in range(args.warmup + args.steps):
        lib.parallax_transform_reduce(sq, red64, D["red_f32"].ptr, n28, 4, 8, fres.ctypes.data)
        lib.parallax_transform_reduce(ev, red32, D["cif_in"].ptr, n28, 4, 4, cres.ctypes.data)
    fused = {}
    for key, name, nb in (("transform_reduce:f32", "square_sum_f32_to_f64", n28 * 4),
                          ("transform_reduce:i32", "count_if_even_i32", n28 * 4)):
        tot, c = prof_get(lib, key)
        fused[name] = rate(nb, tot / max(c, 1))
    lib.pxr_prof_enable(0)
    # reference: x*x is exact in f64; summed in 80-bit extended precision, in chunks
    exact_ld = np.longdouble(0)
    for c0 in range(0, n28, 1 << 24):
        y = X["reduce_f32"][c0:c0 + (1 << 24)].astype(np.float64)
        exact_ld += np.sum(np.square(y).astype(np.longdouble))
    exact = float(exact_ld)
    bnd = 2 * math.ceil(math.log2(n28)) * float(np.finfo(np.float64).eps) * exact
    assert abs(fres[0] - exact) <= bnd, ("square sum", fres[0], exact)
    even = int(np.count_nonzero((X["copy_if_i32"] & 1) == 0))
    assert int(cres[0]) == even, ("count_if", int(cres[0]), even)
    fused["parity"] = {"square_sum_err_over_bound": float(abs(fres[0] - exact) / bnd), "count_if": "exact"}
    log(f"[bench] fused map-reduce: {fused}")
    return fused


def search_suite(args, X, D, n):
    """find / argminmax / mismatch over the suite's 2^28 inputs (device-resident, one
    stream; device time per call from events, median of --steps calls): find with a
    predicate no element passes (the whole array is read), argmin of the f32 input,
    argmax of the i32 input, mismatch of the i32 input with a copy differing 5 elements
    from the end.  Algorithmic bytes: n*s (mismatch 2*n*s)."""
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
        out[name] = dict(rate(nb, statistics.median(ms)), index=got)
    log(f"[bench] search: {out}")
    return out


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (profiles/traffic.txt: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE per dispatch, scaled
    by the same counters on the streaming copy of known size), or None when not measured."""
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
    import pxo
    from parallax_amd import make_desc, consts as C
    from parallax_amd.dtypes import PX_I32
    res = D["res"].download(np.uint8, 16)
    out = {}
    f32 = res[:4].view(np.float32)[0]
    a = X["reduce_f32"]
    hi, lo, ab = pxo.sum_exact(a)
    bound = 2 * math.ceil(math.log2(a.size)) * float(np.finfo(np.float32).eps) * ab
    err = abs(float(f32) - (hi + lo))
    assert err <= bound, ("reduce f32", err, bound)
    out["reduce_f32_err_over_bound"] = err / bound
    assert res[8:16].view(np.int64)[0] == pxo.reduce(X["reduce_i64"]), "reduce i64"
    s = X["scan_u32"]
    assert np.array_equal(D["scan_inplace"].download(np.uint32, s.size), pxo.inclusive_scan(s)), "inclusive scan"
    assert np.array_equal(D["scan_out"].download(np.uint32, s.size), pxo.exclusive_scan(s, 0)), "exclusive scan"
    assert su.download(np.uint32, X["sort_u32"].size).tobytes() == pxo.sort(X["sort_u32"]).tobytes(), "sort u32"
    assert sf.download(np.float64, X["sort_f64"].size).tobytes() == pxo.sort(X["sort_f64"]).tobytes(), "sort f64"
    ref = pxo.copy_if(make_desc(C.OP_FLAGS, PX_I32, fn=C.PRED_MOD_EQ, c0=2, c1=0), X["copy_if_i32"])
    assert kept == ref.size and D["cif_out"].download(np.int32, kept).tobytes() == ref.tobytes(), "copy_if"
    out["bit_exact"] = ["reduce_i64", "inclusive_scan_u32", "exclusive_scan_u32", "sort_u32", "sort_f64", "copy_if_i32"]
    return out


def verify_c1(lib, C, PX_F32):
    """C1: sum of 2^20 float32 uniform [0,1) (seed 0) on the GPU through the C ABI, with
    its error against the FP64 sum and the stated bound."""
    from parallax_amd import datagen, make_desc
    a = datagen.uniform01_f32(0, 1 << 20)
    r = ctypes.c_float()
    lib.parallax_reduce(load_kernel(lib, make_desc(C.OP_REDUCE, PX_F32)), a.ctypes.data, a.size, 4, ctypes.byref(r))
    exact = math.fsum(a.astype(np.float64))
    bnd = 2 * math.ceil(math.log2(a.size)) * float(np.finfo(np.float32).eps) * exact
    err = abs(float(r.value) - exact)
    assert err <= bnd, ("C1", err, bnd)
    return {"gpu_sum": float(r.value), "fp64_sum": exact, "abs_err": err, "bound": bnd, "err_over_bound": err / bnd}


# ------------------------------------------------------------------ batched forms
def batched(args, flush):
    """B1 (seg reduce + inclusive scan, 16,384 f32 segments) and B2 (seg sort + unique,
    4,096 u64 segments) as hipGraph replays of autotuned plans, every segment checked;
    the per-call leg runs the same segments as synchronous device-resident calls."""
    import pxo
    from parallax_amd import lib, check, make_desc, consts as C, datagen
    from parallax_amd.device import Context, DeviceBuffer
    from parallax_amd.dtypes import PX_F32, PX_U64

    nseg1 = max(1, int(16384 * args.scale))
    nseg2 = max(1, int(4096 * args.scale))
    x1, off1 = datagen.batched_b1(seed=2, nseg=nseg1)
    x2, off2 = datagen.batched_b2(seed=3, nseg=nseg2)
    lens1 = np.diff(off1).astype(np.int64)
    log(f"[bench] batched inputs: B1 {off1[-1]} f32 in {nseg1} segments, B2 {off2[-1]} u64 in {nseg2}")
    ctx = Context()
    d1 = DeviceBuffer.from_array(x1)
    r1 = DeviceBuffer(8 * nseg1)
    sc1 = DeviceBuffer(x1.nbytes + 16)
    d2 = DeviceBuffer.from_array(x2)
    p2 = DeviceBuffer.from_array(x2)  # pristine keys: restored before each sort replay
    u2 = DeviceBuffer(x2.nbytes + 16)
    c2 = DeviceBuffer(8 * nseg2)
    o1 = off1.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
    o2 = off2.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
    ud = make_desc(C.OP_FLAGS, PX_U64, fn=C.PRED_UNIQUE)
    order = ["seg_reduce_f32", "seg_scan_f32", "seg_sort_u64", "seg_unique_u64"]
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

    # ---- timed replays (flushed, median), then parity of the last replay of each
    times = {k: [] for k in plans}
    for i in range(WARM + RUNS):
        for k in order:
            if k == "seg_sort_u64":
                check(lib.pxd_copy(ctx.h, d2.ptr, p2.ptr, x2.nbytes), "restore keys")
                ctx.sync()
            flush()
            ctx.timer_begin()
            check(lib.pxb_plan_launch(plans[k]), k)
            t = ctx.timer_end()
            if i >= WARM:
                times[k].append(t)
    ctx.sync()
    t0 = time.time()
    got = r1.download(np.float32, nseg1).astype(np.float64)
    hi, ab = pxo.seg_sum_exact(x1, off1)
    eps = float(np.finfo(np.float32).eps)
    fac = np.array([2 * max(1, math.ceil(math.log2(max(int(n), 2)))) * eps for n in lens1])
    red_ratio = float(np.max(np.abs(got - hi) / np.maximum(fac * ab, 1e-300)))
    assert red_ratio <= 1.0, red_ratio
    gs = sc1.download(np.float32, x1.size).astype(np.float64)
    val, pab = pxo.seg_scan_exact(x1, off1)
    per_elem = np.repeat(fac, lens1) * pab
    scan_ratio = float(np.max(np.abs(gs - val) / np.maximum(per_elem, 1e-300)))
    assert scan_ratio <= 1.0, scan_ratio
    del gs, val, pab, per_elem
    srt = pxo.seg_sort(x2, off2)
    assert d2.download(np.uint64, x2.size).tobytes() == srt.tobytes(), "seg sort"
    uq, uo = pxo.seg_unique(srt, off2)
    cnt = c2.download(np.uint64, nseg2)
    assert np.array_equal(cnt, np.diff(uo)), "seg unique counts"
    ug = u2.download(np.uint64, x2.size)
    starts = np.repeat(off2[:-1].astype(np.int64), cnt.astype(np.int64))
    within = np.arange(int(cnt.sum()), dtype=np.int64) - np.repeat(uo[:-1].astype(np.int64), cnt.astype(np.int64))
    assert ug[starts + within].tobytes() == uq.tobytes(), "seg unique"
    kept_total = int(cnt.sum())
    parity = {"seg_reduce_f32_max_err_over_bound": red_ratio, "seg_scan_f32_max_err_over_bound": scan_ratio,
              "seg_sort_u64": f"bit-exact, all {nseg2} segments", "seg_unique_u64": f"bit-exact, all {nseg2} segments"}
    log(f"[bench] batched parity in {time.time() - t0:.1f}s: {parity}")

    # ---- per-call leg: the same segments, one synchronous device-resident call each
    # Each call is bracketed by events and waited for (a synchronous call, as the
    # reference's C ABI makes them): the loop's wall time includes the launch and wait
    # of every call; the sum of the event times is its device time alone.
    t_beg, t_end = lib.pxd_timer_begin, lib.pxd_timer_end
    dev_ms = ctypes.c_double()

    def per_call(k):
        acc = 0.0
        t0 = time.perf_counter()
        if k == "seg_reduce_f32":
            for s in range(nseg1):
                t_beg(ctx.h)
                lib.pxd_reduce(ctx.h, PX_F32, d1.addr(4 * int(off1[s])), int(lens1[s]), r1.addr(4 * s))
                t_end(ctx.h, ctypes.byref(dev_ms))
                acc += dev_ms.value
        elif k == "seg_scan_f32":
            for s in range(nseg1):
                t_beg(ctx.h)
                lib.pxd_inclusive_scan(ctx.h, PX_F32, d1.addr(4 * int(off1[s])), sc1.addr(4 * int(off1[s])), int(lens1[s]))
                t_end(ctx.h, ctypes.byref(dev_ms))
                acc += dev_ms.value
        elif k == "seg_sort_u64":
            for s in range(nseg2):
                t_beg(ctx.h)
                lib.pxd_sort(ctx.h, PX_U64, d2.addr(8 * int(off2[s])), int(off2[s + 1] - off2[s]))
                t_end(ctx.h, ctypes.byref(dev_ms))
                acc += dev_ms.value
        else:
            for s in range(nseg2):
                t_beg(ctx.h)
                lib.pxd_copy_if(ctx.h, ctypes.byref(ud), d2.addr(8 * int(off2[s])), u2.addr(8 * int(off2[s])),
                                int(off2[s + 1] - off2[s]), c2.addr(8 * s))
                t_end(ctx.h, ctypes.byref(dev_ms))
                acc += dev_ms.value
        return (time.perf_counter() - t0) * 1e3, acc

    pc = {k: [] for k in order}
    pcd = {k: [] for k in order}
    for _ in range(3):
        for k in order:
            if k == "seg_sort_u64":
                check(lib.pxd_copy(ctx.h, d2.ptr, p2.ptr, x2.nbytes), "restore keys")
                ctx.sync()
            w, dv = per_call(k)
            pc[k].append(w)
            pcd[k].append(dv)
    assert d2.download(np.uint64, x2.size).tobytes() == srt.tobytes(), "per-call seg sort"
    n1, n2 = int(off1[-1]), int(off2[-1])
    alg = {"seg_reduce_f32": n1 * 4 + (nseg1 + 1) * 8, "seg_scan_f32": 2 * n1 * 4 + (nseg1 + 1) * 8,
           "seg_sort_u64": 2 * n2 * 8 + (nseg2 + 1) * 8, "seg_unique_u64": n2 * 8 + kept_total * 8 + (nseg2 + 1) * 8}
    per = {}
    total_ms = total_pc = total_pcd = 0.0
    for k in order:
        ms = float(np.median(times[k]))
        pcm = float(np.median(pc[k]))
        pcdm = float(np.median(pcd[k]))
        total_ms += ms
        total_pc += pcm
        total_pcd += pcdm
        per[k] = dict(rate(alg[k], ms), per_call_ms=round(pcm, 3), per_call_device_ms=round(pcdm, 3),
                      speedup_vs_per_call=round(pcm / ms, 1), speedup_vs_per_call_device=round(pcdm / ms, 1))
    for p in plans.values():
        lib.pxb_plan_destroy(p)
    value = gbps(sum(alg.values()), total_ms)
    dom = max(per, key=lambda k: per[k]["ms"])
    nseg_total = nseg1 * 2 + nseg2 * 2
    return {"value": round(value, 1), "unit": "GB/s", "ms_per_step": round(total_ms, 3),
            "dtype": "mixed(f32,u64)",
            "config": {"workload": f"batched: B1 {nseg1} f32 segments log-uniform [1,2^18] (seg reduce + "
                                   f"inclusive scan), B2 {nseg2} u64 segments log-uniform [16,2^18] "
                                   "(seg sort + unique); hipGraph replays", "scale": args.scale},
            "roofline": {"bound": "hbm", "kernel": dom, "achieved": per[dom]["gbps"], "peak": HBM_PEAK_GBPS,
                         "unit": "GB/s", "frac": round(per[dom]["gbps"] / HBM_PEAK_GBPS, 4), "traffic": None,
                         "units_per_launch": f"{alg[dom]} algorithmic bytes per replay"},
            "segments_per_s": {"batched": round(nseg_total / (total_ms * 1e-3), 1),
                               "plain_per_call_device_resident": round(nseg_total / (total_pc * 1e-3), 1)},
            "speedup_vs_plain": {"per_call_wall": round(total_pc / total_ms, 2),
                                 "per_call_device_time_only": round(total_pcd / total_ms, 2)},
            "per_primitive": per, "autotuned_tiles": tiles, "parity": parity}


# ------------------------------------------------------------------ served form
def served(args):
    """Served mixed stream (seed 4): saturated req/s (all requests queued at once; median
    of --steps passes after --warmup, every result of the last pass checked against the
    oracle), p50/p99 latency under Poisson arrivals at half that rate, and the same
    requests as plain synchronous per-call calls (device-resident operands; C ABI with
    host operands) for the speed-up."""
    import pxo
    from parallax_amd import lib, check, make_desc, consts as C, datagen
    from parallax_amd._lib import PxsConfig, PxsRequest, PxsStats
    from parallax_amd.device import Context
    from parallax_amd.dtypes import NP_DTYPES, PX_F32, PX_F64, PX_I64

    nreq = max(8, int(400 * args.scale))
    t0 = time.time()
    gen = datagen.served_requests(seed=4, nreq=nreq)
    cfg = PxsConfig()
    lib.pxs_config_default(ctypes.byref(cfg))
    server = lib.pxs_create(ctypes.byref(cfg))
    assert server, "pxs_create failed"
    reqs = []
    for op, dt, a, c0 in gen:
        p = lib.pxs_alloc(server, a.nbytes)
        assert p
        check(lib.pxd_memcpy_h2d(p, a.ctypes.data, a.nbytes), "upload")
        reqs.append((op, dt, a, p, make_desc(C.OP_FLAGS, dt, fn=C.PRED_LT, c0=c0), c0))
    total_bytes = sum(a.nbytes * (1 if op == C.OP_REDUCE else 2) for op, _, a, _, _, _ in reqs)
    log(f"[bench] served: {nreq} requests, {total_bytes / 1e9:.2f} GB algorithmic, ready in {time.time() - t0:.1f}s")

    def restore_sorts():
        for op, dt, a, p, _, _ in reqs:
            if op == C.OP_SORT:
                check(lib.pxd_memcpy_h2d(p, a.ctypes.data, a.nbytes), "restore")

    # the request records, built once (the client's own cost, outside every pass)
    rq = (PxsRequest * nreq)(*[PxsRequest(op, dt, a.size, p, fl) for op, dt, a, p, fl, _ in reqs])
    rid_buf = (ctypes.c_int64 * nreq)()

    def run_served(arrivals, keep=False):
        """arrivals None: all requests arrive together (one pxs_submit_batch); else one
        pxs_submit per request at its arrival time (seconds from the start)."""
        lib.pxs_reset_stats(server)
        if arrivals is None:
            assert lib.pxs_submit_batch(server, rq, nreq, rid_buf) == 0
            ids = list(rid_buf)
        else:
            ids = []
            t0 = time.perf_counter()
            for i, t_arr in enumerate(arrivals):
                d = t0 + t_arr - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                rid = lib.pxs_submit(server, ctypes.byref(rq[i]))
                assert rid >= 0
                ids.append(rid)
        check(lib.pxs_drain(server), "drain")
        st = PxsStats()
        lib.pxs_get_stats(server, ctypes.byref(st))
        assert st.failed == 0, st.failed
        if not keep:
            for rid in ids:
                lib.pxs_release(server, rid)
        return st, ids

    for _ in range(max(1, args.warmup)):
        restore_sorts()
        run_served(None)
    passes = []
    last_ids = None
    for i in range(max(1, args.steps)):
        restore_sorts()
        keep = i == max(1, args.steps) - 1
        st, ids = run_served(None, keep=keep)
        passes.append(st)
        if keep:
            last_ids = ids
    # every result of the last saturated pass against the oracle
    t0 = time.time()
    got = []
    for rid, (op, dt, a, p, fl, c0) in zip(last_ids, reqs):
        out, cnt, lat = ctypes.c_void_p(), ctypes.c_uint64(), ctypes.c_double()
        check(lib.pxs_result(server, rid, ctypes.byref(out), ctypes.byref(cnt), ctypes.byref(lat)), "result")
        nb = a.itemsize if op == C.OP_REDUCE else (cnt.value * a.itemsize if op == C.OP_FLAGS else a.nbytes)
        buf = np.empty(max(nb, 1), dtype=np.uint8)
        if nb:
            check(lib.pxd_memcpy_d2h(buf.ctypes.data, out.value, nb), "d2h")
        got.append((buf[:nb].tobytes(), cnt.value))
    for rid in last_ids:
        lib.pxs_release(server, rid)

    def check_one(i):
        op, dt, a, p, fl, c0 = reqs[i]
        ob, cnt = got[i]
        npd = NP_DTYPES[dt]
        if op == C.OP_REDUCE:
            g = np.frombuffer(ob, dtype=npd)[0]
            if dt in (PX_F32, PX_F64):
                hi, lo, ab = pxo.sum_exact(a)
                b = 2 * max(1, math.ceil(math.log2(max(a.size, 2)))) * float(np.finfo(npd).eps) * ab
                return abs(float(g) - (hi + lo)) <= b
            return g == pxo.reduce(a)
        if op == C.OP_SCAN:
            g = np.frombuffer(ob, dtype=npd)
            if dt in (PX_F32, PX_F64):
                val, ab = pxo.scan_exact(a)
                b = 2 * max(1, math.ceil(math.log2(max(a.size, 2)))) * float(np.finfo(npd).eps)
                return bool(np.all(np.abs(g.astype(np.float64) - val) <= b * ab + 1e-300))
            return np.array_equal(g, pxo.inclusive_scan(a))
        if op == C.OP_SORT:
            return ob == pxo.sort(a).tobytes()
        ref = pxo.copy_if(fl, a)
        return cnt == ref.size and ob == ref.tobytes()

    with concurrent.futures.ThreadPoolExecutor(max_workers=8) as ex:
        ok = list(ex.map(check_one, range(nreq)))
    assert all(ok), [i for i, o in enumerate(ok) if not o][:8]
    log(f"[bench] served: every result of the last pass checked in {time.time() - t0:.1f}s")
    passes.sort(key=lambda s: s.req_per_s)
    sat = passes[len(passes) // 2]
    # Poisson arrivals at half the saturated rate
    lam = 0.5 * sat.req_per_s
    gaps = -np.log1p(-datagen.uniform01_f64(4, nreq, stream=9)) / lam
    restore_sorts()
    poi, _ = run_served(np.cumsum(gaps) - gaps[0])
    # plain per-call: the same requests, one synchronous device-resident call after another
    ctx = Context()
    out = lib.pxd_malloc(max(a.nbytes for _, _, a, _, _, _ in reqs) + 64)
    cnt = lib.pxd_malloc(64)

    def plain_all():
        for op, dt, a, p, fl, _ in reqs:
            if op == C.OP_REDUCE:
                check(lib.pxd_reduce(ctx.h, dt, p, a.size, out), "reduce")
            elif op == C.OP_SCAN:
                check(lib.pxd_inclusive_scan(ctx.h, dt, p, out, a.size), "scan")
            elif op == C.OP_SORT:
                check(lib.pxd_sort(ctx.h, dt, p, a.size), "sort")
            else:
                check(lib.pxd_copy_if(ctx.h, ctypes.byref(fl), p, out, a.size, cnt), "copy_if")
            ctx.sync()
    restore_sorts()
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
    K = {}
    for dt in set(r[1] for r in reqs):
        K[dt] = {"red": load_kernel(lib, make_desc(C.OP_REDUCE, dt)), "scan": load_kernel(lib, make_desc(C.OP_SCAN, dt)),
                 "add": load_kernel(lib, make_desc(C.OP_SCAN_ADD, dt)), "sort": load_kernel(lib, make_desc(C.OP_SORT, dt)),
                 "iscan": load_kernel(lib, make_desc(C.OP_SCAN, PX_I64)), "iadd": load_kernel(lib, make_desc(C.OP_SCAN_ADD, PX_I64)),
                 "scatter": load_kernel(lib, make_desc(C.OP_SCATTER, dt, aux=C.SCATTER_COMPACT))}
    flk = {i: load_kernel(lib, r[4]) for i, r in enumerate(reqs) if r[0] == C.OP_FLAGS}
    res = np.zeros(16, dtype=np.uint8)

    def abi_all(work, outs):
        for i, ((op, dt, a, p, fl, _), h) in enumerate(zip(reqs, work)):
            k, es = K[dt], a.itemsize
            if op == C.OP_REDUCE:
                lib.parallax_reduce(k["red"], h.ctypes.data, h.size, es, res.ctypes.data)
            elif op == C.OP_SCAN:
                lib.parallax_scan(k["scan"], k["add"], h.ctypes.data, h.size, es)
            elif op == C.OP_SORT:
                lib.parallax_sort(k["sort"], h.ctypes.data, h.size, es)
            else:
                lib.parallax_copy_if(flk[i], k["iscan"], k["iadd"], k["scatter"], h.ctypes.data, outs[i].ctypes.data,
                                     h.size, es, 0)
    outs = [np.zeros_like(a) for _, _, a, _, _, _ in reqs]
    abi_all([a.copy() for _, _, a, _, _, _ in reqs], outs)
    abi_t = []
    for _ in range(3):
        work = [a.copy() for _, _, a, _, _, _ in reqs]
        t0 = time.perf_counter()
        abi_all(work, outs)
        abi_t.append(time.perf_counter() - t0)
    abi_rps = nreq / sorted(abi_t)[1]
    pass_ms = 1e3 / sat.req_per_s * nreq
    return {"value": round(sat.req_per_s, 1), "unit": "req/s", "ms_per_step": round(pass_ms, 3),
            "dtype": "mixed(i32,u64,f32,f64)",
            "config": {"workload": f"served mixed stream: {nreq} requests (30% reduce, 30% scan, 25% sort, "
                                   "15% copy_if), sizes log-uniform [2^10,2^24], dtypes {i32,u64,f32,f64}, "
                                   f"seed 4, 64 GB pool, {cfg.slots} slots on {cfg.lanes} stream(s)", "scale": args.scale},
            "req_per_s": {"served_saturated": round(sat.req_per_s, 1), "served_poisson": round(poi.req_per_s, 1),
                          "saturated_passes": len(passes), "saturated_min": round(passes[0].req_per_s, 1),
                          "saturated_max": round(passes[-1].req_per_s, 1),
                          "plain_per_call_c_abi_host_operands": round(abi_rps, 1),
                          "plain_per_call_device_resident": round(plain_rps, 1)},
            "latency_ms": {"poisson_rate_req_s": round(lam, 1), "p50": round(poi.p50_ms, 3), "p99": round(poi.p99_ms, 3),
                           "saturated_p50": round(sat.p50_ms, 3), "saturated_p99": round(sat.p99_ms, 3)},
            "speedup_vs_plain": {"device_resident": round(sat.req_per_s / plain_rps, 2),
                                 "c_abi_host_operands": round(sat.req_per_s / abi_rps, 2)},
            "roofline": {"bound": "hbm", "kernel": "whole pass (mixed)", "achieved": round(gbps(total_bytes, pass_ms), 1),
                         "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                         "frac": round(gbps(total_bytes, pass_ms) / HBM_PEAK_GBPS, 4), "traffic": None},
            "batches_per_pass": sat.batches, "parity": f"every result of a saturated pass checked ({nreq})"}


# ------------------------------------------------------------------ counter runs
def kernels_once(args):
    """The plain suite's calls, argmin f32 and the fused count_if, each twice (for
    rocprofv3 --pmc runs: few dispatches, no verification or flushes)."""
    from parallax_amd import lib, make_desc, consts as C, datagen
    from parallax_amd.device import Context, DeviceBuffer
    from parallax_amd.dtypes import PX_F32, PX_F64, PX_I32, PX_I64, PX_U32
    X = datagen.plain_suite_inputs(seed=1, scale=args.scale)
    L = lambda d: load_kernel(lib, d)  # noqa: E731
    n28, n26, n25 = X["reduce_f32"].size, X["sort_u32"].size, X["sort_f64"].size
    D = {k: DeviceBuffer.from_array(v) for k, v in X.items()}
    out = DeviceBuffer(X["scan_u32"].nbytes)
    res = DeviceBuffer(64)
    ctx = Context()
    init = np.zeros(1, dtype=np.uint32)
    fl = make_desc(C.OP_FLAGS, PX_I32, fn=C.PRED_MOD_EQ, c0=2, c1=0)
    ev = make_desc(C.OP_TRANSFORM, PX_I32, aux=PX_I32, fn=C.MAP_PRED, fn2=C.PRED_MOD_EQ, c0=2, c1=0)
    ks, ka, kh = L(make_desc(C.OP_SCAN, PX_U32)), L(make_desc(C.OP_SCAN_ADD, PX_U32)), L(make_desc(C.OP_SCAN_SHIFT, PX_U32))
    for _ in range(2):
        lib.parallax_reduce(L(make_desc(C.OP_REDUCE, PX_F32)), D["reduce_f32"].ptr, n28, 4, res.ptr)
        lib.parallax_reduce(L(make_desc(C.OP_REDUCE, PX_I64)), D["reduce_i64"].ptr, n28, 8, res.ptr)
        lib.parallax_exclusive_scan(ks, ka, kh, D["scan_u32"].ptr, out.ptr, n28, 4, init.ctypes.data)
        lib.parallax_scan(ks, ka, out.ptr, n28, 4)
        D["sort_u32"].upload(X["sort_u32"])
        lib.parallax_sort(L(make_desc(C.OP_SORT, PX_U32)), D["sort_u32"].ptr, n26, 4)
        D["sort_f64"].upload(X["sort_f64"])
        lib.parallax_sort(L(make_desc(C.OP_SORT, PX_F64)), D["sort_f64"].ptr, n25, 8)
        lib.parallax_copy_if(L(fl), L(make_desc(C.OP_SCAN, PX_I32)), L(make_desc(C.OP_SCAN_ADD, PX_I32)),
                             L(make_desc(C.OP_SCATTER, PX_I32, aux=C.SCATTER_COMPACT)), D["copy_if_i32"].ptr,
                             out.ptr, n28, 4, 0)
        lib.pxd_argminmax(ctx.h, PX_F32, D["reduce_f32"].ptr, n28, 0, 0, res.ptr)
        lib.pxd_transform_reduce(ctx.h, ctypes.byref(ev), D["copy_if_i32"].ptr, n28, res.ptr)
        ctx.sync()
    lib.pxd_device_sync()
    log("[bench] kernels: each plain-suite kernel, argmin f32 and count_if i32 run twice")


# ------------------------------------------------------------------ counter reference
def copy_probe():
    """Three streaming copies of exactly 1 GiB (1 GiB read + 1 GiB written per dispatch):
    the reference that converts FETCH_SIZE / WRITE_SIZE counter readings to bytes."""
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
    log("[bench] copy probe: 3 x 1 GiB device-to-device copies (k_copy16)")


# ------------------------------------------------------------------ CPU baseline
def cpu_baseline_child():
    """Runs in a child process pinned to one core: the oracle's restatement of the
    reference's CPU fallback loops (stdpar.hpp:146-148, 178, 209-210, 237-238, 319-321;
    -O3 -DNDEBUG, serial like the fallback) on the full suite; prints one JSON line."""
    import pxo
    from parallax_amd import datagen, make_desc, consts as C
    from parallax_amd.dtypes import PX_I32

    L = pxo.lib
    X = datagen.plain_suite_inputs(seed=1, scale=1.0)  # generated on every core, then pinned
    affinity = None
    try:
        cpu0 = sorted(os.sched_getaffinity(0))[0]
        os.sched_setaffinity(0, {cpu0})
        affinity = len(os.sched_getaffinity(0))
    except Exception:
        pass
    ns, nu, nf = X["reduce_f32"].size, X["sort_u32"].size, X["sort_f64"].size
    out_u = np.zeros_like(X["scan_u32"])  # touched before timing (no page faults timed)
    out_i = np.zeros_like(X["copy_if_i32"])
    res = np.zeros(2, dtype=np.int64)
    zero = np.zeros(1, dtype=np.uint32)
    fl = pxo.desc(make_desc(C.OP_FLAGS, PX_I32, fn=C.PRED_MOD_EQ, c0=2, c1=0))
    kept_box = [0]
    calls = [
        ("reduce_f32", ns * 4, lambda: L.pxo_reduce(4, X["reduce_f32"].ctypes.data, ns, res.ctypes.data)),
        ("reduce_i64", ns * 8, lambda: L.pxo_reduce(2, X["reduce_i64"].ctypes.data, ns, res.ctypes.data)),
        ("inclusive_scan_u32", 2 * ns * 4, lambda: L.pxo_inclusive_scan(1, X["scan_u32"].ctypes.data, out_u.ctypes.data, ns)),
        ("exclusive_scan_u32", 2 * ns * 4, lambda: L.pxo_exclusive_scan(1, X["scan_u32"].ctypes.data, out_u.ctypes.data, ns,
                                                                        zero.ctypes.data)),
        ("sort_u32", 2 * nu * 4, None),
        ("sort_f64", 2 * nf * 8, None),
        ("copy_if_i32", None, lambda: kept_box.__setitem__(0, L.pxo_copy_if(ctypes.byref(fl), X["copy_if_i32"].ctypes.data,
                                                                            out_i.ctypes.data, ns))),
    ]
    times = {}
    total_bytes = 0
    for name, nb, fn in calls:
        reps = []
        for _ in range(1 if name.startswith("sort") else 3):
            if name in ("sort_u32", "sort_f64"):
                w = X[name].copy()
                t = time.perf_counter()
                L.pxo_sort(1 if name == "sort_u32" else 5, w.ctypes.data, w.size)
            else:
                t = time.perf_counter()
                fn()
            reps.append(time.perf_counter() - t)
        if name == "copy_if_i32":
            nb = ns * 4 + int(kept_box[0]) * 4
        times[name] = sorted(reps)[len(reps) // 2]
        total_bytes += nb
    total = sum(times.values())
    # C1: the fallback's sequential f32 fold over 2^20 uniform [0,1) (seed 0) against FP64
    c1 = datagen.uniform01_f32(0, 1 << 20)
    r = np.zeros(1, dtype=np.float32)
    L.pxo_reduce(4, c1.ctypes.data, c1.size, r.ctypes.data)
    exact = math.fsum(c1.astype(np.float64))
    bnd = 2 * math.ceil(math.log2(c1.size)) * float(np.finfo(np.float32).eps) * exact
    c1r = {"fallback_sum": float(r[0]), "fp64_sum": exact, "abs_err": abs(float(r[0]) - exact), "bound": bnd,
           "err_over_bound": abs(float(r[0]) - exact) / bnd}
    cpu = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except Exception:
        pass
    print(json.dumps({"value": round(total_bytes / total / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
                      "sample": (f"the oracle's restatement of the reference CPU fallback loops (g++ -O3 -DNDEBUG, "
                                 f"serial, pinned to 1 core; median of 3 for the streaming primitives, one run per "
                                 f"sort) on the full suite: streams 2^28, sort u32 2^26, sort f64 2^25, same inputs; "
                                 f"host {cpu}, {os.cpu_count()} logical CPUs, affinity {affinity}"),
                      "seconds": {k: round(v, 4) for k, v in times.items()}, "c1_reduce_2p20_f32": c1r}))


def cpu_baseline():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--cpu-baseline-child"],
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
    ap.add_argument("--workload", default="all", choices=["all", "plain", "batched", "served", "kernels"],
                    help="all (default): plain -> the headline line, with batched and served keys")
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
        print("bench.py: measures one GPU only (multi-GPU runs are not supported: the sharded "
              "form is a design, not code); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "kernels":
        kernels_once(args)
        return
    from parallax_amd import lib
    flush = Flush(lib)
    extra = {}
    if args.workload in ("all", "plain"):
        copy = copy_ceiling(lib, flush)
        log(f"[bench] copy ceiling: {copy}")
        r = plain_suite(args, flush, copy)
        if args.workload == "all":
            extra["batched"] = batched(args, flush)
            log(f"[bench] batched: {extra['batched']}")
            extra["served"] = served(args)
            log(f"[bench] served: {extra['served']}")
    elif args.workload == "batched":
        r = batched(args, flush)
    else:
        r = served(args)
    cpu = None if (args.no_cpu_baseline or args.workload not in ("all", "plain")) else cpu_baseline()
    line = {"metric": METRIC, "value": r["value"], "unit": r["unit"], "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["ms_per_step"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": r["dtype"], "data": "synthetic", "config": r["config"],
            "roofline": r.get("roofline"), "cpu_baseline": cpu}
    for k in ("per_primitive", "per_primitive_timed_steps", "kernels", "copy_ceiling", "autotuned_tiles", "parity",
              "c_abi_host_operands", "fused_map_reduce", "search", "latency_ms", "req_per_s", "speedup_vs_plain",
              "segments_per_s"):
        if k in r:
            line[k] = r[k]
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
