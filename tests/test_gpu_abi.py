"""The reference's own unit tests (tests/unit/test_*.cpp), re-run through the C ABI of
parallax/runtime.h on the HIP engine: descriptors loaded with parallax_kernel_load take
the place of the SPIR-V files, arena pointers are host-visible on return."""
import ctypes

import numpy as np
import pytest

import pxo
from parallax_amd import lib, make_desc, descriptor_words, consts as C
from parallax_amd.device import DeviceBuffer, stats
from parallax_amd.dtypes import NP_DTYPES, PX_F32, PX_F64, PX_I32, PX_I64, PX_U32, PX_U64

pytestmark = pytest.mark.gpu


def load(desc):
    w, n = descriptor_words(desc)
    k = lib.parallax_kernel_load(w, n)
    assert k, "parallax_kernel_load failed"
    return k


def arena_array(a):
    p = lib.parallax_arena_alloc(a.nbytes, 16)
    assert p, "arena_alloc failed"
    ctypes.memmove(p, a.ctypes.data, a.nbytes)
    return p


def read(p, dtype, n):
    out = np.empty(n, dtype=dtype)
    ctypes.memmove(out.ctypes.data, p, out.nbytes)
    return out


def test_reduce_known_answer(gpu_ctx):
    # test_reduce.cpp:47-63 (and test_staging.cpp: the arena here is always device-local
    # with a host mirror, i.e. the "discrete" path the staging test forces)
    k = load(make_desc(C.OP_REDUCE, PX_F32))
    data = (np.arange(5000) % 8).astype(np.float32)
    p = arena_array(data)
    res = ctypes.c_float(-1.0)
    before = stats()["reduce"]
    lib.parallax_reduce(k, p, 5000, 4, ctypes.byref(res))
    assert res.value == 17500.0
    assert stats()["reduce"] == before + 1  # the HIP path ran
    lib.parallax_arena_free(p)


def test_scan_known_answer(gpu_ctx):
    # test_scan.cpp:46-61
    ks, ka = load(make_desc(C.OP_SCAN, PX_F32)), load(make_desc(C.OP_SCAN_ADD, PX_F32))
    p = arena_array(np.ones(1000, dtype=np.float32))
    lib.parallax_scan(ks, ka, p, 1000, 4)
    assert np.array_equal(read(p, np.float32, 1000), np.arange(1, 1001, dtype=np.float32))
    lib.parallax_arena_free(p)


def test_sort_known_answer(gpu_ctx):
    # test_sort.cpp:41-68
    N = 1024
    v = np.array([((N - 1 - i) * 2654435761) % N for i in range(N)], dtype=np.float32)
    k = load(make_desc(C.OP_SORT, PX_F32))
    p = arena_array(v)
    lib.parallax_sort(k, p, N, 4)
    got = read(p, np.float32, N)
    assert np.all(got[1:] >= got[:-1]) and got.sum() == v.sum()
    assert np.array_equal(got, np.arange(N, dtype=np.float32))
    lib.parallax_arena_free(p)


def test_copy_if_known_answer(gpu_ctx):
    # test_copy_if.cpp:54-74 (the flags keep values > 0.5)
    kf = load(make_desc(C.OP_FLAGS, PX_F32, fn=C.PRED_GT, c0=0.5))
    ks, ka = load(make_desc(C.OP_SCAN, PX_F32)), load(make_desc(C.OP_SCAN_ADD, PX_F32))
    kx = load(make_desc(C.OP_SCATTER, PX_F32, aux=C.SCATTER_COMPACT))
    i = np.arange(1000)
    pin = arena_array(np.where(i % 2 == 1, i, 0).astype(np.float32))
    pout = arena_array(np.full(1000, -1.0, dtype=np.float32))
    kept = lib.parallax_copy_if(kf, ks, ka, kx, pin, pout, 1000, 4, 1)
    assert kept == 500
    out = read(pout, np.float32, 1000)
    assert np.array_equal(out[:500], (2 * np.arange(500) + 1).astype(np.float32))
    assert np.all(out[500:] == -1.0)  # the unwritten tail keeps its host contents
    lib.parallax_arena_free(pin)
    lib.parallax_arena_free(pout)


def test_physptr_relocation(gpu_ctx):
    # test_physptr.cpp:50-122: the kernel relocates the arena HOST pointer itself
    p = arena_array(np.ones(64, dtype=np.float32))
    assert lib.pxr_relocate_probe(p, 64) == 0
    assert np.all(read(p, np.float32, 64) == 3.0)
    lib.parallax_arena_free(p)


def test_arena_behaviour(gpu_ctx):
    # test_arena.cpp:26-65 on the process arena
    a = lib.parallax_arena_alloc(1000 * 4, 16)
    b = lib.parallax_arena_alloc(500 * 4, 16)
    assert a and b and lib.parallax_arena_contains(a) and lib.parallax_arena_contains(b)
    assert b > a and a % 256 == 0 and b % 256 == 0
    arr = np.arange(1000, dtype=np.int32) * 3
    ctypes.memmove(a, arr.ctypes.data, arr.nbytes)
    assert np.array_equal(read(a, np.int32, 1000), arr)
    lib.parallax_arena_free(a)
    c = lib.parallax_arena_alloc(1000 * 4, 16)
    assert c == a  # first-fit reuse at the same offset
    on_stack = ctypes.c_int(42)
    assert lib.parallax_arena_contains(ctypes.addressof(on_stack)) == 0
    host, dev, cap, used = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_size_t(), ctypes.c_size_t()
    assert lib.pxr_arena_info(ctypes.byref(host), ctypes.byref(dev), ctypes.byref(cap), ctypes.byref(used)) == 0
    assert lib.parallax_arena_alloc(cap.value * 2, 16) is None  # over capacity -> NULL
    lib.parallax_arena_free(b)
    lib.parallax_arena_free(c)


@pytest.mark.parametrize("dt", [PX_I32, PX_U32, PX_I64, PX_U64, PX_F32, PX_F64])
def test_abi_all_dtypes_host_and_device_pointers(gpu_ctx, dt):
    npd = NP_DTYPES[dt]
    es = npd.itemsize
    rng = np.random.default_rng(dt)
    a = (rng.integers(0, 100, 100_003)).astype(npd)
    kr = load(make_desc(C.OP_REDUCE, dt))
    ks, ka, kh = (load(make_desc(C.OP_SCAN, dt)), load(make_desc(C.OP_SCAN_ADD, dt)),
                  load(make_desc(C.OP_SCAN_SHIFT, dt)))
    ksort = load(make_desc(C.OP_SORT, dt))
    # plain host (numpy) memory: staged in and out
    res = np.zeros(1, dtype=npd)
    lib.parallax_reduce(kr, a.ctypes.data, a.size, es, res.ctypes.data)
    assert res[0] == pxo.reduce(a)
    b = a.copy()
    lib.parallax_scan(ks, ka, b.ctypes.data, b.size, es)
    assert np.array_equal(b, pxo.inclusive_scan(a))
    out = np.zeros_like(a)
    init = np.array([11], dtype=npd)
    lib.parallax_exclusive_scan(ks, ka, kh, a.ctypes.data, out.ctypes.data, a.size, es, init.ctypes.data)
    assert np.array_equal(out, pxo.exclusive_scan(a, init[0]))
    c = a.copy()
    lib.parallax_sort(ksort, c.ctypes.data, c.size, es)
    assert np.array_equal(c, np.sort(a))
    # device memory: used in place
    d = DeviceBuffer.from_array(a)
    lib.parallax_sort(ksort, d.ptr, a.size, es)
    assert np.array_equal(d.download(npd, a.size), np.sort(a))


def test_abi_partition_unique_remove(gpu_ctx):
    a = np.random.default_rng(1).integers(0, 20, 50_000).astype(np.int64)
    kf = load(make_desc(C.OP_FLAGS, PX_I64, fn=C.PRED_LT, c0=7))
    kfn = load(make_desc(C.OP_FLAGS, PX_I64, fn=C.PRED_LT, c0=7, flags=C.FLAG_NEGATE))
    ks, ka = load(make_desc(C.OP_SCAN, PX_I64)), load(make_desc(C.OP_SCAN_ADD, PX_I64))
    kc = load(make_desc(C.OP_SCATTER, PX_I64, aux=C.SCATTER_COMPACT))
    kp = load(make_desc(C.OP_SCATTER, PX_I64, aux=C.SCATTER_PARTITION))
    out = np.zeros_like(a)
    nt = lib.parallax_copy_if(kf, ks, ka, kp, a.ctypes.data, out.ctypes.data, a.size, 8, 0)
    ref, rk = pxo.partition(make_desc(C.OP_FLAGS, PX_I64, fn=C.PRED_LT, c0=7), a)
    assert nt == rk and out.tobytes() == ref.tobytes()  # std::partition's order
    w = a.copy()  # in place (host memory staged once), as the partition funnel calls it
    nt = lib.parallax_copy_if(kf, ks, ka, kp, w.ctypes.data, w.ctypes.data, a.size, 8, 0)
    assert nt == rk and w.tobytes() == ref.tobytes()
    out[:] = 0
    k = lib.parallax_copy_if(kfn, ks, ka, kc, a.ctypes.data, out.ctypes.data, a.size, 8, 0)  # remove_if
    assert np.array_equal(out[:k], a[a >= 7])
    w = a.copy()
    k = lib.parallax_copy_if(kfn, ks, ka, kc, w.ctypes.data, w.ctypes.data, a.size, 8, 0)  # in place
    assert np.array_equal(w[:k], a[a >= 7]) and np.array_equal(w[k:], a[k:])
    s = np.sort(a)
    ku = load(make_desc(C.OP_FLAGS, PX_I64, fn=C.PRED_UNIQUE))
    k = lib.parallax_copy_if(ku, ks, ka, kc, s.ctypes.data, out.ctypes.data, a.size, 8, 0)
    assert np.array_equal(out[:k], np.unique(a))
    k = lib.parallax_copy_if(ku, ks, ka, kc, s.ctypes.data, s.ctypes.data, a.size, 8, 0)  # in place
    assert np.array_equal(s[:k], np.unique(a))


def test_abi_map_skeletons(gpu_ctx):
    a = np.arange(10_000, dtype=np.float32)
    km = load(make_desc(C.OP_MAP, PX_F32, fn=C.MAP_AXPB))
    p = arena_array(a)
    cap = np.array([2.0, 1.0], dtype=np.float32)
    lib.parallax_kernel_launch_with_captures(km, p, a.size, cap.ctypes.data, 8, 4)
    assert np.array_equal(read(p, np.float32, a.size), a * 2 + 1)
    kfill = load(make_desc(C.OP_MAP, PX_F32, fn=C.MAP_FILL, c0=4.0))
    lib.parallax_kernel_launch(ctypes.c_void_p(kfill), ctypes.c_void_p(p), ctypes.c_size_t(a.size), ctypes.c_size_t(4))
    assert np.all(read(p, np.float32, a.size) == 4.0)
    kt = load(make_desc(C.OP_TRANSFORM, PX_F32, aux=PX_F64, fn=C.MAP_SQUARE))
    out = np.zeros(a.size, dtype=np.float64)
    lib.parallax_kernel_launch_transform2(kt, a.ctypes.data, out.ctypes.data, a.size, 4, 8)
    assert np.array_equal(out, a.astype(np.float64) ** 2)
    # gather through a captured ARENA pointer, relocated on the device
    table = arena_array(np.array([10, 20, 30, 40], dtype=np.int32))
    idx = np.array([3, 0, 2, 9, 1], dtype=np.int32)
    kg = load(make_desc(C.OP_TRANSFORM, PX_I32, aux=PX_I32, fn=C.MAP_GATHER))
    capg = (ctypes.c_uint64 * 2)(table, 4)
    res = np.zeros(5, dtype=np.int32)
    lib.parallax_kernel_launch_transform2_captures(kg, idx.ctypes.data, res.ctypes.data, 5, 4, 4,
                                                   ctypes.addressof(capg), 16)
    assert res.tolist() == [40, 10, 30, 0, 20]
    lib.parallax_arena_free(p)
    lib.parallax_arena_free(table)


def test_builtin_funnel_lookup_hits_gpu(gpu_ctx):
    key = b"T parallax::detail::device_reduce(const T*, std::size_t, T) [with T = double; std::size_t = long unsigned int]"
    k = lib.parallax_kernel_lookup(key)
    assert k
    assert lib.parallax_kernel_lookup(key) == k  # loaded once
    a = np.arange(1, 1001, dtype=np.float64)
    r = ctypes.c_double()
    lib.parallax_reduce(k, a.ctypes.data, a.size, 8, ctypes.byref(r))
    assert r.value == 500500.0
    assert lib.parallax_kernel_lookup(b"some unregistered functor key") is None


def test_registered_descriptor_lookup(gpu_ctx):
    d = make_desc(C.OP_FLAGS, PX_I32, fn=C.PRED_MOD_EQ, c0=2, c1=0)
    w, n = descriptor_words(d)
    test_registered_descriptor_lookup.keep = w  # registration stores the pointer only
    lib.parallax_kernel_register(b"my copy_if key:flags", w, n)
    k = lib.parallax_kernel_lookup(b"my copy_if key:flags")
    assert k


def test_abi_results_to_device_pointers_back_to_back():
    """Reduce results written to device memory by consecutive C-ABI calls: each result
    word is copied out on the launcher's stream before the next call's kernel reuses the
    launcher's device word (a copy ordered only on the null stream could read the next
    call's value)."""
    kf, ki = load(make_desc(C.OP_REDUCE, PX_F32)), load(make_desc(C.OP_REDUCE, PX_I64))
    rng = np.random.default_rng(17)
    f = rng.integers(-100, 100, 1 << 20).astype(np.float32)  # exact sums in f32
    i = rng.integers(-(1 << 40), 1 << 40, 1 << 20).astype(np.int64)
    df, di = DeviceBuffer.from_array(f), DeviceBuffer.from_array(i)
    res = DeviceBuffer(16 * 32)
    for r in range(32):
        lib.parallax_reduce(kf, df.ptr, f.size, 4, res.addr(16 * r))
        lib.parallax_reduce(ki, di.ptr, i.size, 8, res.addr(16 * r + 8))
    raw = res.download(np.uint8, 16 * 32)
    got_f = raw.reshape(32, 16)[:, :4].copy().view(np.float32).ravel()
    got_i = raw.reshape(32, 16)[:, 8:].copy().view(np.int64).ravel()
    assert np.all(got_f == np.float32(f.astype(np.int64).sum()))
    assert np.all(got_i == pxo.reduce(i))


def test_variadic_launch_transform(gpu_ctx):
    """parallax_kernel_launch_transform(k, in, out, count, elem_size) (reference
    runtime.h:35, cache.cpp:147-190: the variadic (void*, void*, size_t, size_t) form)."""
    a = np.arange(-5000, 5000, dtype=np.int32)
    kt = load(make_desc(C.OP_TRANSFORM, PX_I32, aux=PX_I32, fn=C.MAP_AXPB, c0=3, c1=-7))
    out = np.zeros_like(a)
    lib.parallax_kernel_launch_transform(ctypes.c_void_p(kt), ctypes.c_void_p(a.ctypes.data),
                                         ctypes.c_void_p(out.ctypes.data), ctypes.c_size_t(a.size),
                                         ctypes.c_size_t(4))
    d = make_desc(C.OP_TRANSFORM, PX_I32, aux=PX_I32, fn=C.MAP_AXPB, c0=3, c1=-7)
    assert np.array_equal(out, pxo.transform(d, a, np.int32))
    # device memory in and out
    di, do = DeviceBuffer.from_array(a), DeviceBuffer(a.nbytes)
    lib.parallax_kernel_launch_transform(ctypes.c_void_p(kt), ctypes.c_void_p(di.ptr), ctypes.c_void_p(do.ptr),
                                         ctypes.c_size_t(a.size), ctypes.c_size_t(4))
    assert np.array_equal(do.download(np.int32, a.size), out)


def test_register_buffer_data_path(gpu_ctx):
    """parallax_register_buffer (reference runtime.h:145, C++ linkage; cache.cpp:394-402):
    calls on a registered host range see its current contents every time and write their
    results back to it (the reference copied a registered buffer only once and never
    copied results back, unified_buffer.cpp:266-310)."""
    reg = getattr(lib, "_Z24parallax_register_bufferPvm")
    reg.restype = ctypes.c_bool
    reg.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
    a = np.random.default_rng(4).integers(0, 1 << 30, 100_003).astype(np.uint32)
    assert reg(a.ctypes.data, a.nbytes)
    ksum = load(make_desc(C.OP_REDUCE, PX_U32))
    r = np.zeros(1, dtype=np.uint32)
    lib.parallax_reduce(ksum, a.ctypes.data, a.size, 4, r.ctypes.data)
    assert r[0] == pxo.reduce(a)
    a[::7] += 3  # rewritten after registration: the next call sees the new values
    lib.parallax_reduce(ksum, a.ctypes.data, a.size, 4, r.ctypes.data)
    assert r[0] == pxo.reduce(a)
    ref = pxo.sort(a)
    ksort = load(make_desc(C.OP_SORT, PX_U32))
    lib.parallax_sort(ksort, a.ctypes.data + 0, a.size, 4)  # in place: the result reaches a
    assert np.array_equal(a, ref)
    # a pointer inside the registered range
    b = a[1000:2000].copy()
    a[1000:2000] = b[::-1]
    lib.parallax_sort(ksort, a.ctypes.data + 4000, 1000, 4)
    assert np.array_equal(a[1000:2000], np.sort(b))
