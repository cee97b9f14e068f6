"""The C ABI library (CPU-side checks): it loads, exports every function the public
headers declare, parses funnel keys, and answers "no device" correctly without a GPU."""
import ctypes
import os
import pathlib
import re
import subprocess
import sys

import pytest

from parallax_amd import LIB_PATH, lib, make_desc, consts as C, PxDesc
from parallax_amd.dtypes import PX_F32, PX_F64, PX_I32, PX_I64, PX_U32, PX_U64

ROOT = pathlib.Path(__file__).resolve().parent.parent
HEADERS = sorted((ROOT / "include" / "parallax").glob("*.h"))


def declared_functions():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b((?:parallax|pxd|pxb|pxr|pxs)_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def exported():
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_every_declared_function_is_exported():
    decl = declared_functions()
    exp = exported()
    assert len(decl) > 60
    # parallax_register_buffer has C++ linkage (reference runtime.h:145)
    missing = sorted(n for n in decl if n not in exp and n != "parallax_register_buffer")
    assert not missing, missing
    assert any(s.startswith("_Z24parallax_register_buffer") for s in exp)


def test_runtime_h_matches_reference_signatures():
    # the 22 C entry points of the reference runtime.h:16-113 are all present
    ref = ["parallax_umalloc", "parallax_ufree", "parallax_arena_alloc", "parallax_arena_free",
           "parallax_arena_contains", "parallax_heap_contains", "parallax_kernel_load",
           "parallax_kernel_launch", "parallax_kernel_launch_transform",
           "parallax_kernel_launch_transform2", "parallax_kernel_launch_transform2_captures",
           "parallax_reduce", "parallax_argminmax", "parallax_find", "parallax_mismatch",
           "parallax_scan", "parallax_exclusive_scan", "parallax_sort", "parallax_copy_if",
           "parallax_kernel_register", "parallax_kernel_lookup",
           "parallax_kernel_launch_with_captures"]
    assert set(ref) <= exported()


def test_runtime_h_compiles_as_c_and_cpp(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "parallax/runtime.h"\n#include "parallax/px_device.h"\nint main(void){return 0;}\n')
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", str(ROOT / "include"), "-c", str(src),
                    "-o", str(tmp_path / "t.o")], check=True)
    cpp = tmp_path / "t.cpp"
    cpp.write_text('#include "parallax/runtime.h"\nint main(){ return parallax_register_buffer(nullptr, 0) ? 1 : 0; }\n')
    subprocess.run(["g++", "-std=c++20", "-Wall", "-Werror", "-I", str(ROOT / "include"), "-c", str(cpp),
                    "-o", str(tmp_path / "t2.o")], check=True)


def test_descriptor_layout():
    assert ctypes.sizeof(PxDesc) == 48
    d = make_desc(C.OP_FLAGS, PX_F32, fn=C.PRED_GT, c0=0.5)
    words = (ctypes.c_uint32 * 12).from_buffer_copy(bytes(d))
    assert words[0] == C.DESC_MAGIC and words[2] == C.OP_FLAGS and words[3] == PX_F32
    assert words[8] == 0x3F000000  # 0.5f


GCC_KEYS = {
    "T parallax::detail::device_reduce(const T*, std::size_t, T) [with T = float; std::size_t = long unsigned int]":
        (C.OP_REDUCE, PX_F32, 0),
    "void parallax::detail::device_sort(T*, std::size_t) [with T = long unsigned int; std::size_t = long unsigned int]":
        (C.OP_SORT, PX_U64, 0),
    "void parallax::detail::device_scan(const T*, T*, std::size_t) [with T = unsigned int; std::size_t = long unsigned int]:scan":
        (C.OP_SCAN, PX_U32, 0),
    "void parallax::detail::device_scan(const T*, T*, std::size_t) [with T = unsigned int; std::size_t = long unsigned int]:add":
        (C.OP_SCAN_ADD, PX_U32, 0),
    "void parallax::detail::device_exclusive_scan(const T*, T*, std::size_t, T) [with T = long int; std::size_t = long unsigned int]:shift":
        (C.OP_SCAN_SHIFT, PX_I64, 0),
    "std::size_t parallax::detail::device_unique(T*, std::size_t) [with T = double; std::size_t = long unsigned int]:flags":
        (C.OP_FLAGS, PX_F64, C.PRED_UNIQUE),
    "std::size_t parallax::detail::device_unique(T*, std::size_t) [with T = int; std::size_t = long unsigned int]:scatter":
        (C.OP_SCATTER, PX_I32, 0),
}
CLANG_KEYS = {
    "T parallax::detail::device_reduce(const T *, std::size_t, T) [T = double]": (C.OP_REDUCE, PX_F64, 0),
    "void parallax::detail::device_sort(T *, std::size_t) [T = unsigned long]": (C.OP_SORT, PX_U64, 0),
}


@pytest.mark.parametrize("key,want", list(GCC_KEYS.items()) + list(CLANG_KEYS.items()))
def test_builtin_funnel_keys(key, want):
    w = (ctypes.c_uint32 * 12)()
    assert lib.pxr_builtin_descriptor(key.encode(), w) == 0
    op, dt, fn = want
    assert (w[2], w[3], w[5]) == (op, dt, fn)


def test_functor_funnel_keys_are_not_builtin():
    w = (ctypes.c_uint32 * 12)()
    key = ("std::size_t parallax::detail::device_copy_if(const T*, T*, std::size_t, Pred) "
           "[with T = int; Pred = main()::<lambda(int)>; std::size_t = long unsigned int]:flags")
    assert lib.pxr_builtin_descriptor(key.encode(), w) == C.E_UNSUPPORTED
    assert lib.pxr_builtin_descriptor(b"not a funnel key", w) == C.E_UNSUPPORTED


def test_key_parsing_matches_real_pretty_function(tmp_path):
    # compile a probe that prints __PRETTY_FUNCTION__ of funnel-shaped templates with g++
    src = tmp_path / "k.cpp"
    src.write_text(r'''
#include <cstddef>
#include <cstdio>
namespace parallax { namespace detail {
template <class T> T device_reduce(const T* d, std::size_t n, T init) { std::puts(__PRETTY_FUNCTION__); return init; }
template <class T> void device_sort(T* d, std::size_t n) { std::puts(__PRETTY_FUNCTION__); }
}}
int main() { float f = 0; long l = 0; unsigned long u = 0; double x = 0;
  parallax::detail::device_reduce(&f, 1, f); parallax::detail::device_reduce(&l, 1, l);
  parallax::detail::device_sort(&u, 1); parallax::detail::device_reduce(&x, 1, x); }
''')
    exe = tmp_path / "k"
    subprocess.run(["g++", "-std=c++20", str(src), "-o", str(exe)], check=True)
    keys = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines()
    want = [(C.OP_REDUCE, PX_F32), (C.OP_REDUCE, PX_I64), (C.OP_SORT, PX_U64), (C.OP_REDUCE, PX_F64)]
    for key, (op, dt) in zip(keys, want):
        w = (ctypes.c_uint32 * 12)()
        assert lib.pxr_builtin_descriptor(key.encode(), w) == 0, key
        assert (w[2], w[3]) == (op, dt), key


def test_no_device_paths_in_a_cpu_process():
    # In a process with PARALLAX_FORCE_CPU the runtime reports no device: lookups miss
    # (the funnel then runs its CPU loop) and arena allocation returns NULL.
    code = r'''
import ctypes, sys
sys.path.insert(0, sys.argv[1])
from parallax_amd import lib, make_desc, descriptor_words, consts as C
assert lib.parallax_kernel_lookup(b"T parallax::detail::device_reduce(const T *, std::size_t, T) [T = float]") is None
assert lib.parallax_arena_alloc(1024, 16) is None
assert lib.parallax_arena_contains(None) == 0
w, n = descriptor_words(make_desc(C.OP_REDUCE, 4))
assert lib.parallax_kernel_load(w, n) is None
print("ok")
'''
    env = dict(os.environ, PARALLAX_FORCE_CPU="1")
    r = subprocess.run([sys.executable, "-c", code, str(ROOT / "parallax-runtime_amd")], env=env,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "ok" in r.stdout, r.stderr
