"""Element type codes of parallax/descriptor.h and their numpy counterparts."""
import numpy as np

PX_I32, PX_U32, PX_I64, PX_U64, PX_F32, PX_F64 = range(6)
DTYPES = {"i32": PX_I32, "u32": PX_U32, "i64": PX_I64, "u64": PX_U64, "f32": PX_F32, "f64": PX_F64}
NP_DTYPES = {
    PX_I32: np.dtype(np.int32),
    PX_U32: np.dtype(np.uint32),
    PX_I64: np.dtype(np.int64),
    PX_U64: np.dtype(np.uint64),
    PX_F32: np.dtype(np.float32),
    PX_F64: np.dtype(np.float64),
}
_BY_NP = {v: k for k, v in NP_DTYPES.items()}


def dtype_code(x):
    """Code of a numpy dtype, array, name ("f32") or code."""
    if isinstance(x, int):
        return x
    if isinstance(x, str) and x in DTYPES:
        return DTYPES[x]
    if isinstance(x, np.ndarray):
        x = x.dtype
    return _BY_NP[np.dtype(x)]
