"""Deterministic synthetic inputs: counter-based SplitMix64(seed, index).

u(seed, stream, i) = mix64(key * 0xD1B54A32D192ED03 + (i + 1) * 0x9E3779B97F4A7C15), with
key = seed * 0x100000001B3 + stream * 0x9E37 and the SplitMix64 finaliser mix64(z): z ^= z>>30; z *= 0xBF58476D1CE4E5B9; z ^= z>>27;
z *= 0x94D049BB133111EB; z ^= z>>31 (all mod 2^64).  Every element depends only on
(seed, index, stream), so any slice can be regenerated and the bytes fed to the GPU and
to the oracle are the same.
"""
import concurrent.futures
import os

import numpy as np

_G = np.uint64(0x9E3779B97F4A7C15)
_S = np.uint64(0xD1B54A32D192ED03)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_CHUNK = 1 << 22


def mix64(z):
    z = z ^ (z >> np.uint64(30))
    z = z * _M1
    z = z ^ (z >> np.uint64(27))
    z = z * _M2
    return z ^ (z >> np.uint64(31))


_POOL = None


def _pool():
    global _POOL
    if _POOL is None:
        _POOL = concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1))
    return _POOL


def u64(seed, n, stream=0, start=0):
    """n raw 64-bit draws for indices start..start+n-1 of (seed, stream)."""
    out = np.empty(int(n), dtype=np.uint64)
    key = (int(seed) * 0x100000001B3 + int(stream) * 0x9E37) & (2**64 - 1)
    base = np.uint64((key * 0xD1B54A32D192ED03) & (2**64 - 1))

    def chunk(s):
        e = min(int(n), s + _CHUNK)
        with np.errstate(over="ignore"):
            idx = np.arange(start + s + 1, start + e + 1, dtype=np.uint64)
            out[s:e] = mix64(base + idx * _G)

    starts = range(0, int(n), _CHUNK)
    if int(n) > 4 * _CHUNK:
        list(_pool().map(chunk, starts))  # numpy releases the GIL on these arrays
    else:
        for s in starts:
            chunk(s)
    return out


def uniform01_f32(seed, n, stream=0):
    """float32 uniform in [0, 1): (u >> 40) * 2^-24 (exact)."""
    return ((u64(seed, n, stream) >> np.uint64(40)).astype(np.float32) * np.float32(2.0**-24))


def uniform01_f64(seed, n, stream=0):
    """float64 uniform in [0, 1): (u >> 11) * 2^-53 (exact)."""
    return (u64(seed, n, stream) >> np.uint64(11)).astype(np.float64) * (2.0**-53)


def normal_f32(seed, n, stream=0):
    """float32 normal(0, 1) by Box-Muller (cosine branch) in float64."""
    out = np.empty(int(n), dtype=np.float32)

    def chunk(s):
        e = min(int(n), s + _CHUNK)
        a = u64(seed, e - s, stream * 2 + 0, start=s)
        b = u64(seed, e - s, stream * 2 + 1, start=s)
        u1 = ((a >> np.uint64(11)).astype(np.float64) + 1.0) * (2.0**-53)  # (0, 1]
        u2 = (b >> np.uint64(11)).astype(np.float64) * (2.0**-53)
        out[s:e] = (np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)).astype(np.float32)

    list(_pool().map(chunk, range(0, int(n), _CHUNK)))
    return out


def int_range(seed, n, lo, hi, dtype, stream=0):
    """Integers uniform in [lo, hi) (hi - lo <= 2^32) by the multiply-shift map."""
    span = int(hi) - int(lo)
    assert 0 < span <= 2**32
    u = u64(seed, n, stream) >> np.uint64(32)
    v = (u * np.uint64(span)) >> np.uint64(32)
    return (v.astype(np.int64) + int(lo)).astype(dtype)


def full_u32(seed, n, stream=0):
    return (u64(seed, n, stream) >> np.uint64(32)).astype(np.uint32)


def full_i32(seed, n, stream=0):
    return full_u32(seed, n, stream).view(np.int32)


def with_duplicates(a, seed, frac_pct=1, stream=7):
    """Copy `frac_pct`% of slots (chosen by the draw) from a random earlier slot."""
    n = a.size
    if n < 2:
        return a
    d = u64(seed, n, stream)
    pick = (d % np.uint64(100)) < np.uint64(frac_pct)
    idx = np.nonzero(pick)[0]
    idx = idx[idx > 0]
    src = (u64(seed, idx.size, stream + 1) % idx.astype(np.uint64)).astype(np.int64)
    out = a.copy()
    out[idx] = a[src]
    return out


def log_uniform_lengths(seed, count, lo, hi, stream=0):
    """Lengths floor(lo * (hi/lo)^U), U uniform in [0,1), clamped to [lo, hi]."""
    u = uniform01_f64(seed, count, stream)
    lens = np.floor(lo * np.exp(u * np.log(hi / lo))).astype(np.int64)
    return np.clip(lens, lo, hi)


def offsets_from_lengths(lens):
    off = np.zeros(len(lens) + 1, dtype=np.uint64)
    np.cumsum(np.asarray(lens, dtype=np.uint64), out=off[1:])
    return off


# ---- the workloads of the headline configurations
def plain_suite_inputs(seed=1, scale=1.0):
    """Inputs of the plain per-call primitive suite (sizes times `scale`)."""
    n28 = int((1 << 28) * scale)
    n26 = int((1 << 26) * scale)
    n25 = int((1 << 25) * scale)
    return {
        "reduce_f32": normal_f32(seed, n28, stream=1),
        "reduce_i64": int_range(seed, n28, -(1 << 20), 1 << 20, np.int64, stream=2),
        "scan_u32": int_range(seed, n28, 0, 16, np.uint32, stream=3),
        "sort_u32": with_duplicates(full_u32(seed, n26, stream=4), seed, 1, stream=40),
        "sort_f64": with_duplicates(uniform01_f64(seed, n25, stream=5), seed, 1, stream=50),
        "copy_if_i32": full_i32(seed, n28, stream=6),
    }


def batched_b1(seed=2, nseg=16384):
    """B1: nseg float32 segments, lengths log-uniform in [1, 2^18], values normal(0, 1)."""
    lens = log_uniform_lengths(seed, nseg, 1, 1 << 18, stream=1)
    off = offsets_from_lengths(lens)
    return normal_f32(seed, int(off[-1]), stream=2), off


def batched_b2(seed=3, nseg=4096):
    """B2: nseg uint64 key segments, lengths log-uniform in [16, 2^18], keys uniform in
    [0, 2^20) (about 30% duplicates per segment)."""
    lens = log_uniform_lengths(seed, nseg, 16, 1 << 18, stream=1)
    off = offsets_from_lengths(lens)
    return int_range(seed, int(off[-1]), 0, 1 << 20, np.uint64, stream=2), off


def served_requests(seed=4, nreq=400):
    """The served mixed stream: per request (op, dtype, values, predicate constant).
    Ops 30% reduce, 30% scan, 25% sort, 15% copy_if (predicate x < c, c = 500 for
    uint64 and 0 otherwise); sizes log-uniform in [2^10, 2^24]; dtypes uniform over
    {int32, uint64, float32, float64}; values uniform in [-1000, 1000] ([0, 1000] for
    uint64)."""
    from . import consts as C
    from .dtypes import NP_DTYPES, PX_F32, PX_F64, PX_I32, PX_U64
    u = uniform01_f64(seed, 4 * nreq, stream=1)
    ops = np.array([C.OP_REDUCE, C.OP_SCAN, C.OP_SORT, C.OP_FLAGS])
    op_of = ops[np.searchsorted(np.cumsum([0.3, 0.3, 0.25, 0.15]), u[:nreq], side="right").clip(0, 3)]
    dts = [PX_I32, PX_U64, PX_F32, PX_F64]
    dt_of = [dts[int(x * 4) % 4] for x in u[nreq:2 * nreq]]
    sizes = np.floor(2.0 ** (10 + 14 * u[2 * nreq:3 * nreq])).astype(np.int64)
    def one(i):
        dt, n = dt_of[i], int(sizes[i])
        if dt == PX_U64:
            a = int_range(seed, n, 0, 1001, np.uint64, stream=100 + i)
        elif dt == PX_I32:
            a = int_range(seed, n, -1000, 1001, np.int32, stream=100 + i)
        else:
            a = (uniform01_f64(seed, n, stream=100 + i) * 2000.0 - 1000.0).astype(NP_DTYPES[dt])
        return (int(op_of[i]), dt, a, 500 if dt == PX_U64 else 0)

    # requests in parallel (numpy releases the GIL); a pool of its own, since large
    # requests use the chunk pool inside u64()
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        return list(ex.map(one, range(nreq)))
