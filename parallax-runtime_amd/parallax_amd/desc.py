"""The primitive descriptor (parallax/descriptor.h struct px_desc) as a ctypes struct."""
import ctypes

import numpy as np

from . import consts
from .dtypes import NP_DTYPES, dtype_code


class PxDesc(ctypes.Structure):
    _fields_ = [
        ("magic", ctypes.c_uint32),
        ("version", ctypes.c_uint32),
        ("op", ctypes.c_uint32),
        ("dtype", ctypes.c_uint32),
        ("aux", ctypes.c_uint32),
        ("fn", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("fn2", ctypes.c_uint32),
        ("c0", ctypes.c_uint64),
        ("c1", ctypes.c_uint64),
    ]


assert ctypes.sizeof(PxDesc) == 4 * consts.DESC_WORDS


def const_bits(value, dtype):
    """Little-endian bits of `value` in element type `dtype`, as uint64."""
    if value is None:
        return 0
    arr = np.array([value], dtype=NP_DTYPES[dtype_code(dtype)])
    raw = arr.tobytes() + b"\0" * (8 - arr.itemsize)
    return int.from_bytes(raw, "little")


def make_desc(op, dtype, aux=0, fn=0, flags=0, fn2=0, c0=None, c1=None, const_dtype=None):
    """Build a descriptor; constants are given as numbers of the element type (or of
    `const_dtype`, e.g. the output type of a transform)."""
    d = dtype_code(dtype)
    cd = d if const_dtype is None else dtype_code(const_dtype)
    return PxDesc(consts.DESC_MAGIC, consts.DESC_VERSION, op, d, aux, fn, flags, fn2,
                  const_bits(c0, cd), const_bits(c1, cd))


def descriptor_words(desc):
    """The descriptor as the (words, count) pair that parallax_kernel_load takes."""
    arr = (ctypes.c_uint32 * consts.DESC_WORDS).from_buffer_copy(bytes(desc))
    return arr, consts.DESC_WORDS
