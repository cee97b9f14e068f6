"""Code-object payloads (PX_OP_CODE, parallax/descriptor.h): a gfx950 code object carried
in the parallax_kernel_load / parallax_kernel_register payload runs for launch,
launch_with_captures and launch_transform2_captures, with captured constants and a
captured arena pointer relocated on the device.  The functors are tests/cpp/user_kernels.hip;
their results are checked against numpy on the same inputs."""
import ctypes
import pathlib

import numpy as np
import pytest

from parallax_amd import lib, make_desc, consts as C
from parallax_amd.dtypes import PX_F32, PX_F64, PX_I32, PX_U32

pytestmark = pytest.mark.gpu

HSACO = pathlib.Path(__file__).resolve().parent / "cpp" / "_bin" / "user_kernels.hsaco"
OP_CODE = 13
CODE_MAP, CODE_TRANSFORM = 0, 1


def payload(symbol, kind, dtype, aux):
    """Header + symbol name + code object, as 32-bit words."""
    code = HSACO.read_bytes()
    name = symbol.encode()
    d = make_desc(OP_CODE, dtype, aux=aux, fn=kind, fn2=len(name))
    d.c0 = len(code)  # a byte count, not a constant of the element type
    pad = lambda b: b + b"\0" * (-len(b) % 4)
    raw = bytes(d) + pad(name) + pad(code)
    words = np.frombuffer(raw, dtype=np.uint32).copy()
    return words


def load(words):
    k = lib.parallax_kernel_load(words.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), words.size)
    assert k, "code-object payload did not load"
    return k


def test_map_with_captures_in_place(gpu_ctx):
    k = load(payload("user_poly_f32", CODE_MAP, PX_F32, PX_F32))
    # integer-valued inputs: a*x, a*x*x and a*x*x + b are all exact in f32, so the
    # result does not depend on rounding or contraction
    x = np.random.default_rng(1).integers(-1000, 1001, 1_000_003).astype(np.float32)
    ref = x.astype(np.float64)
    cap = np.array([1.5, -0.25], dtype=np.float32)
    lib.parallax_kernel_launch_with_captures(ctypes.c_void_p(k), x.ctypes.data, x.size, cap.ctypes.data, cap.nbytes, 4)
    assert np.array_equal(x.astype(np.float64), 1.5 * ref * ref - 0.25)


def test_transform_with_captures(gpu_ctx):
    k = load(payload("user_sqrt_u32_f64", CODE_TRANSFORM, PX_U32, PX_F64))
    a = np.random.default_rng(2).integers(0, 1 << 31, 300_001, dtype=np.uint32)
    out = np.zeros(a.size, dtype=np.float64)
    cap = np.array([0.5], dtype=np.float64)
    lib.parallax_kernel_launch_transform2_captures(ctypes.c_void_p(k), a.ctypes.data, out.ctypes.data, a.size, 4, 8,
                                                   cap.ctypes.data, cap.nbytes)
    want = np.sqrt(a.astype(np.float64)) + 0.5
    assert np.all(np.abs(out - want) <= np.spacing(want))  # within one ulp


def test_captured_arena_pointer_is_relocated(gpu_ctx):
    k = load(payload("user_lookup_i32", CODE_TRANSFORM, PX_U32, PX_I32))
    size = 4099
    tp = lib.parallax_arena_alloc(size * 4, 256)
    assert tp, "arena allocation"
    table = np.ctypeslib.as_array(ctypes.cast(tp, ctypes.POINTER(ctypes.c_int32)), shape=(size,))
    table[:] = np.arange(size, dtype=np.int32) * 7 - 100
    keys = np.random.default_rng(3).integers(0, 1 << 30, 200_000, dtype=np.uint32)
    out = np.zeros(keys.size, dtype=np.int32)
    cap = np.array([tp, size], dtype=np.uint64)
    lib.parallax_kernel_launch_transform2_captures(ctypes.c_void_p(k), keys.ctypes.data, out.ctypes.data, keys.size,
                                                   4, 4, cap.ctypes.data, cap.nbytes)
    assert np.array_equal(out, table[keys % size])
    lib.parallax_arena_free(tp)


def test_registered_under_a_funnel_key(gpu_ctx):
    words = payload("user_poly_f32", CODE_MAP, PX_F32, PX_F32)
    key = b"void parallax::detail::device_invoke(T*, std::size_t, F) [with T = float; F = Poly]"
    lib.parallax_kernel_register(key, words.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), words.size)
    k = lib.parallax_kernel_lookup(key)
    assert k
    x = np.arange(1000, dtype=np.float32)
    lib.parallax_kernel_launch(ctypes.c_void_p(k), ctypes.c_void_p(x.ctypes.data), ctypes.c_size_t(x.size),
                               ctypes.c_size_t(4))
    assert np.array_equal(x, (np.arange(1000, dtype=np.float32) ** 2))  # a = 1, b = 0 without captures
    words_keep.append(words)  # the registry keeps the pointer: keep the words alive


words_keep = []


def test_bad_payloads_are_refused(gpu_ctx):
    w = payload("no_such_symbol", CODE_MAP, PX_F32, PX_F32)
    assert not lib.parallax_kernel_load(w.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), w.size)
    w = payload("user_poly_f32", CODE_MAP, PX_F32, PX_F32)
    assert not lib.parallax_kernel_load(w.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), 20)  # truncated
