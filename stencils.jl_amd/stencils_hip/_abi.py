"""ctypes mirror of include/stencils_hip.h (descriptors, enums, dtype tables)."""
from __future__ import annotations

import ctypes

import numpy as np

STH_OK = 0
STH_EDEVICE = 5
STH_ENOMEM = 12
STH_EINVAL = 22
STH_ENOTSUP = 95

# sth_dtype
BOOL, U8, I8, U16, I16, U32, I32, U64, I64, F16, BF16, F32, F64 = range(13)
DTYPE_NAMES = ["bool", "u8", "i8", "u16", "i16", "u32", "i32", "u64", "i64",
               "f16", "bf16", "f32", "f64"]
DTYPE_SIZE = [1, 1, 1, 2, 2, 4, 4, 8, 8, 2, 2, 4, 8]
FLOAT_DTYPES = {F16, BF16, F32, F64}
SIGNED_DTYPES = {I8, I16, I32, I64}

NUMPY_TO_STH = {
    np.dtype(np.bool_): BOOL, np.dtype(np.uint8): U8, np.dtype(np.int8): I8,
    np.dtype(np.uint16): U16, np.dtype(np.int16): I16,
    np.dtype(np.uint32): U32, np.dtype(np.int32): I32,
    np.dtype(np.uint64): U64, np.dtype(np.int64): I64,
    np.dtype(np.float16): F16, np.dtype(np.float32): F32, np.dtype(np.float64): F64,
}
STH_TO_NUMPY = {v: k for k, v in NUMPY_TO_STH.items()}
STH_TO_NUMPY[BF16] = np.dtype(np.uint16)  # raw bf16 bits on the host

# sth_boundary / sth_padding
REMOVE, WRAP, REFLECT, USE = range(4)
CONDITIONAL, HALO_OUT, HALO_IN = range(3)

# sth_opkind
OP_SUM, OP_MEAN, OP_MIN, OP_MAX, OP_KERNELPRODUCT, OP_LIFE, OP_LAPLACIAN = range(7)
OP_NAMES = ["sum", "mean", "minimum", "maximum", "kernelproduct", "life", "laplacian"]
OPF_ACC_F32 = 1
OPF_INT_WRAP = 2

# call flags
ASYNC = 1
GRAPH = 2
GENERIC = 4
BATCH_SERIAL = 8
STATIC_BATCH_MAX_JOBS = 4096  # sth_map_batch_static's fixed job table
NO_TUNE = 16
NO_FUSE = 32


class SthArray(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("dtype", ctypes.c_int32),
        ("ndim", ctypes.c_int32),
        ("size", ctypes.c_int64 * 3),
        ("stride", ctypes.c_int64 * 3),
        ("halo", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class SthStencil(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("radius", ctypes.c_int32),
        ("ndim", ctypes.c_int32),
        ("length", ctypes.c_int32),
        ("offsets", ctypes.POINTER(ctypes.c_int32)),
        ("weights", ctypes.c_void_p),
        ("weights_dtype", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class SthOp(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("flags", ctypes.c_uint32),
        ("alpha", ctypes.c_double),
        ("birth_mask", ctypes.c_uint64),
        ("survive_mask", ctypes.c_uint64),
    ]


class SthJob(ctypes.Structure):
    _fields_ = [
        ("src", SthArray),
        ("dst", SthArray),
        ("stencil", SthStencil),
        ("op", SthOp),
        ("padval", ctypes.c_void_p),
        ("boundary", ctypes.c_int32),
        ("padding", ctypes.c_int32),
    ]


class SthTerm(ctypes.Structure):
    _fields_ = [
        ("src", SthArray),
        ("stencil", SthStencil),
        ("op", SthOp),
        ("padval", ctypes.c_void_p),
        ("boundary", ctypes.c_int32),
        ("padding", ctypes.c_int32),
        ("coef", ctypes.c_double),
        ("scaled", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class SthScatterRule(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("combine", ctypes.c_int32), ("coef", ctypes.c_double)]


SCATTER_ZERO_DEST = 64
COMBINE_ADD = 0
COMBINE_MUL = 1
COMBINE_MAX = 2
COMBINE_MIN = 3
MAX_TERMS = 8
MULTI_DTYPES = {F32, F64, U8, I32, I64}  # element types of sth_map_multi


def out_dtype(src_dt: int, opkind: int, flags: int = 0, wdtype: int = -1) -> int:
    """Result element type of `op` over `src_dt` (the reference's _return_type,
    src/gatherstencil.jl:41-56), or -1 when the GPU path does not cover it."""
    storage = src_dt in (BF16, F16)
    if storage and not (flags & OPF_ACC_F32):
        return -1
    if opkind == OP_SUM:
        return I64 if src_dt == BOOL else src_dt
    if opkind == OP_MEAN:
        return src_dt if src_dt in FLOAT_DTYPES else F64
    if opkind in (OP_MIN, OP_MAX):
        return src_dt
    if opkind == OP_KERNELPRODUCT:
        if src_dt == BOOL:
            return -1
        if storage:
            return src_dt if wdtype == F32 else -1
        return src_dt if wdtype == src_dt else -1
    if opkind == OP_LIFE:
        return -1 if src_dt in FLOAT_DTYPES else src_dt
    if opkind == OP_LAPLACIAN:
        return src_dt if src_dt in FLOAT_DTYPES else -1
    return -1


def int_fold_declined(src_dt: int, opkind: int, flags: int, length: int) -> bool:
    """True where the library declines an integer sum / mean because the
    result depends on how the caller's `sum` folds (wrapping in the element
    type or widening) and the caller did not say (OPF_INT_WRAP)."""
    if length < 2 or (flags & OPF_INT_WRAP) or src_dt in FLOAT_DTYPES or src_dt == BOOL:
        return False
    if opkind == OP_SUM:
        return DTYPE_SIZE[src_dt] < 8
    return opkind == OP_MEAN


def make_array(data_ptr: int, dtype: int, shape, strides_elems, halo: int) -> SthArray:
    """Descriptor of a logical array (sizes/strides in Julia dim order)."""
    a = SthArray()
    a.data = data_ptr
    a.dtype = dtype
    a.ndim = len(shape)
    for d in range(3):
        a.size[d] = int(shape[d]) if d < len(shape) else 1
        a.stride[d] = int(strides_elems[d]) if d < len(shape) else 0
    a.halo = int(halo)
    return a


# serving loop (sth_server_*)
SERVER_SERIAL = 1


class SthServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_size_t), ("max_active", ctypes.c_int32),
                ("round_steps", ctypes.c_int32), ("flags", ctypes.c_uint32),
                ("reserved", ctypes.c_int32)]


class SthRequest(ctypes.Structure):
    _fields_ = [("state", SthArray), ("stencil", SthStencil), ("op", SthOp),
                ("padval", ctypes.c_void_p), ("boundary", ctypes.c_int32),
                ("padding", ctypes.c_int32), ("steps", ctypes.c_int64),
                ("result", ctypes.c_void_p), ("arrival", ctypes.c_double),
                ("tag", ctypes.c_uint64)]


class SthCompletion(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_uint64), ("status", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("arrival", ctypes.c_double), ("admitted", ctypes.c_double),
                ("finished", ctypes.c_double), ("steps", ctypes.c_int64), ("cells", ctypes.c_int64)]


class SthServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64),
                ("active", ctypes.c_int64), ("pending", ctypes.c_int64),
                ("rounds", ctypes.c_int64), ("launches", ctypes.c_int64),
                ("cell_steps", ctypes.c_int64), ("hbm_in_use", ctypes.c_size_t),
                ("hbm_peak", ctypes.c_size_t), ("started", ctypes.c_double),
                ("solo_launches", ctypes.c_int64), ("fused_launches", ctypes.c_int64)]
