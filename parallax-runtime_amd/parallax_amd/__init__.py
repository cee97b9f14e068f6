"""Python bindings of libparallax_amd.so (the MI355X engine of the Parallax primitive
library).  The library itself is C++/HIP; this package only loads it with ctypes and
offers numpy-friendly helpers for tests and benchmarks.  Importing never falls back to
anything else: a missing library raises."""
from .dtypes import DTYPES, NP_DTYPES, PX_I32, PX_U32, PX_I64, PX_U64, PX_F32, PX_F64, dtype_code
from .desc import PxDesc, make_desc, descriptor_words
from . import consts
from ._lib import lib, LIB_PATH, last_error, check

__all__ = [
    "DTYPES", "NP_DTYPES", "PX_I32", "PX_U32", "PX_I64", "PX_U64", "PX_F32", "PX_F64",
    "dtype_code", "PxDesc", "make_desc", "descriptor_words", "consts", "lib", "LIB_PATH",
    "last_error", "check",
]
