"""ORACLE — test infrastructure only.

ctypes front end of oracle/liboracle.so, the plain-C restatement of
Stencils.jl's gather path (see oracle/stencil_oracle.c for the reference
file:line of each step).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module, and only as the checker or as the
reported CPU baseline; the product never routes through it.

All arrays here are host numpy arrays in Julia index order (dim 1 first).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parent
LIB_PATH = HERE / "liboracle.so"
sys.path.insert(0, str(REPO / "stencils.jl_amd"))

from stencils_hip import _abi  # noqa: E402  (descriptor structs only)

MODE_NATIVE = 0
MODE_F64 = 1

_lib = None


def build(force: bool = False):
    """Compile liboracle.so with the Makefile beside this file."""
    args = ["make", "-s", "-C", str(HERE)]
    if force:
        args.insert(2, "-B")
    subprocess.run(args, check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        P = ctypes.POINTER
        L.orc_map.restype = ctypes.c_int
        L.orc_map.argtypes = [P(_abi.SthArray), P(_abi.SthArray), ctypes.c_int, ctypes.c_int,
                              P(ctypes.c_int32), ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                              P(_abi.SthOp), ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                              ctypes.c_int]
        L.orc_iterate.restype = ctypes.c_int
        L.orc_iterate.argtypes = [P(_abi.SthArray), P(_abi.SthArray), ctypes.c_int, ctypes.c_int,
                                  P(ctypes.c_int32), ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                  P(_abi.SthOp), ctypes.c_void_p, ctypes.c_int, ctypes.c_int64,
                                  ctypes.c_int, P(ctypes.c_int32)]
        L.orc_update_boundary.restype = ctypes.c_int
        L.orc_update_boundary.argtypes = [P(_abi.SthArray), ctypes.c_int, ctypes.c_void_p]
        L.orc_out_dtype.restype = ctypes.c_int
        L.orc_out_dtype.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint, ctypes.c_int,
                                    ctypes.c_int]
        for name in ("orc_f32_to_bf16", "orc_f32_to_f16"):
            getattr(L, name).restype = ctypes.c_uint16
            getattr(L, name).argtypes = [ctypes.c_float]
        for name in ("orc_f16_to_f32", "orc_bf16_to_f32"):
            getattr(L, name).restype = ctypes.c_float
            getattr(L, name).argtypes = [ctypes.c_uint16]
        _lib = L
    return _lib


def host_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def _desc(a: np.ndarray, dtype_code: int, logical_shape, halo: int) -> _abi.SthArray:
    return _abi.make_array(a.ctypes.data, dtype_code, logical_shape,
                           tuple(s // a.itemsize for s in a.strides), halo)


def _op(kind, alpha=0.0, birth=(3,), survive=(2, 3), flags=0):
    o = _abi.SthOp()
    o.kind = kind
    o.flags = flags
    o.alpha = alpha
    o.birth_mask = sum(1 << b for b in set(birth))
    o.survive_mask = sum(1 << s for s in set(survive))
    return o


def map_raw(parent: np.ndarray, dtype_code: int, logical_shape, halo: int, offsets, st_ndim: int,
            bc: int, padval_bytes, padding: int, op: _abi.SthOp, weights=None,
            wdtype: int = -1, mode: int = MODE_NATIVE, nthreads: int = 1,
            out_dtype: int | None = None) -> np.ndarray:
    """One gatherstencil!(f, dest, source) on host arrays.  `parent` is the
    source parent (halo included) and is refreshed in place for Halo padding,
    as update_boundary! does.  Returns the result (logical shape)."""
    L = lib()
    offs = np.ascontiguousarray(np.asarray(offsets, dtype=np.int32).reshape(-1))
    nL = offs.size // st_ndim
    if out_dtype is None:
        out_dtype = L.orc_out_dtype(dtype_code, op.kind, op.flags, wdtype, mode)
        if out_dtype < 0:
            raise ValueError("op / dtype combination not covered")
    out = np.zeros(tuple(logical_shape), dtype=_abi.STH_TO_NUMPY[out_dtype], order="F")
    src = _desc(parent, dtype_code, logical_shape, halo)
    dst = _desc(out, out_dtype, logical_shape, 0)
    pv = ctypes.create_string_buffer(padval_bytes, len(padval_bytes)) if padval_bytes else None
    w = None if weights is None else np.ascontiguousarray(weights)
    rc = L.orc_map(ctypes.byref(src), ctypes.byref(dst), nL, st_ndim,
                   offs.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), bc, pv, padding,
                   ctypes.byref(op), None if w is None else w.ctypes.data, wdtype, mode, nthreads)
    if rc != 0:
        raise ValueError(f"oracle status {rc}")
    return out


def iterate_raw(a: np.ndarray, b: np.ndarray, dtype_code: int, logical_shape, halo: int,
                offsets, st_ndim: int, bc: int, padval_bytes, padding: int, op: _abi.SthOp,
                steps: int, nthreads: int = 1) -> int:
    """`steps` mapstencil! calls on the host buffers a (source) and b (dest);
    returns 0 if `a` holds the newest state, 1 if `b` does."""
    L = lib()
    offs = np.ascontiguousarray(np.asarray(offsets, dtype=np.int32).reshape(-1))
    nL = offs.size // st_ndim
    da = _desc(a, dtype_code, logical_shape, halo)
    db = _desc(b, dtype_code, logical_shape, halo)
    pv = ctypes.create_string_buffer(padval_bytes, len(padval_bytes)) if padval_bytes else None
    cur = ctypes.c_int32(0)
    rc = L.orc_iterate(ctypes.byref(da), ctypes.byref(db), nL, st_ndim,
                       offs.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), bc, pv, padding,
                       ctypes.byref(op), None, -1, int(steps), nthreads, ctypes.byref(cur))
    if rc != 0:
        raise ValueError(f"oracle status {rc}")
    return cur.value


# ------------------------------------------------- front end over the mirror

def _parts(sa):
    """(dtype code, padval bytes, op flags) of a host-parent StencilArray."""
    from stencils_hip.gather import _op_flags, _padval_bytes
    dt = sa.sth_dtype()
    return dt, _padval_bytes(sa), _op_flags(dt)


def mapstencil(f, sa, mode: int = MODE_NATIVE, nthreads: int = 1) -> np.ndarray:
    """Reference result of mapstencil(f, sa) for a host-parent StencilArray
    (numpy parent, refreshed in place for Halo padding)."""
    from stencils_hip.gather import _weights
    from stencils_hip.ops import resolve
    red = resolve(f)
    dt, pv, flags = _parts(sa)
    op = red.op(flags)
    w, wdt = _weights(sa.stencil, dt)
    if w is not None and mode == MODE_F64:
        w = np.asarray(sa.stencil.weights, dtype=np.float64)
        wdt = _abi.F64
    return map_raw(sa.parent(), dt, sa.size(), sa.halo(), sa.stencil.offsets(), sa.stencil.N,
                   sa.boundary.code, pv, sa.padding.code, op, w, wdt, mode, nthreads)


def mapstencil_steps(f, A, steps: int, nthreads: int = 1):
    """`steps` mapstencil!(f, A) on a host-parent SwitchingStencilArray, in place;
    returns the switched array the last call would return."""
    from stencils_hip.ops import resolve
    red = resolve(f)
    dt, pv, flags = _parts(A)
    cur = iterate_raw(A.source, A.dest, dt, A.size(), A.halo(), A.stencil.offsets(), A.stencil.N,
                      A.boundary.code, pv, A.padding.code, red.op(flags), steps, nthreads)
    return A.switch() if cur == 1 else A


def f32_to_bf16(x: float) -> int:
    return lib().orc_f32_to_bf16(x)


def bf16_to_f32(b: int) -> float:
    return lib().orc_bf16_to_f32(b)


def _interior(parent: np.ndarray, halo: int) -> np.ndarray:
    return parent[tuple(slice(halo, s - halo) for s in parent.shape)] if halo else parent


def mapstencil_multi(f, args) -> np.ndarray:
    """Reference result of gatherstencil!(f, dest, args...) for a `Terms`
    function over host-parent StencilArrays and plain numpy arrays
    (src/gatherstencil.jl:89-113): every StencilArray argument's halo is
    refreshed (update_boundary!, :93-95; Halo parents change in place), then
    per cell the terms fold left in the element type, `+` or `*` each
    rounded (integers wrap).  A term's fold over a stencil is the single-array
    restatement above stored in the element type -- the value `f` receives
    from `sum(hood)` etc. -- so composing the per-term results cell by cell
    is the reference's expression."""
    from stencils_hip.array import AbstractStencilArray
    from stencils_hip.boundary import Halo, Use
    from stencils_hip.ops import Center, Scaled
    vals = []
    dt = None
    for g, a in zip(f.terms, args):
        coef = None
        if isinstance(g, Scaled):
            coef, g = g.coef, g.g
        if isinstance(a, AbstractStencilArray):
            if isinstance(g, Center):
                if isinstance(a.padding, Halo) and not isinstance(a.boundary, Use):
                    d, pv, _ = _parts(a)
                    desc = _desc(a.parent(), d, a.size(), a.halo())
                    pvb = ctypes.create_string_buffer(pv, len(pv)) if pv else None
                    rc = lib().orc_update_boundary(ctypes.byref(desc), a.boundary.code, pvb)
                    if rc != 0:
                        raise ValueError(f"oracle status {rc}")
                v = np.array(_interior(a.parent(), a.halo()), order="F")
            else:
                v = mapstencil(g, a)
        else:
            v = np.asarray(a)
        if dt is None:
            dt = v.dtype
        if v.dtype != dt:
            raise ValueError("terms of different element types")
        if coef is not None:
            v = dt.type(coef) * v
        vals.append(v)
    acc = np.array(vals[0], order="F")
    with np.errstate(over="ignore"):
        for v in vals[1:]:
            acc = acc * v if f.combine == "*" else acc + v
    return np.asfortranarray(acc.astype(dt, copy=False))


def mapstencil_multi_switching(f, A, rest):
    """mapstencil!(f, A::SwitchingStencilArray, rest...) on host parents
    (src/gatherstencil.jl:84-88): A's dest interior <- f(A, rest...), then
    the switched array."""
    out = mapstencil_multi(f, (A,) + tuple(rest))
    _interior(A.dest, A.halo())[...] = out
    return A.switch()


def scatterstencil(f, op, dest: np.ndarray, source, aux=None, zero_dest: bool = False) -> np.ndarray:
    """Reference result of scatterstencil!(f, op, dest, source[, aux]) for a
    `Scatter` rule on host arrays (src/scatterstencil.jl:37-114), written
    into `dest` in place (Halo parents refreshed in place)."""
    from stencils_hip.scatter import combine_code
    L = lib()
    L.orc_scatter.restype = ctypes.c_int
    dt = source.sth_dtype()
    offs = np.ascontiguousarray(np.asarray(source.stencil.offsets(), dtype=np.int32).reshape(-1))
    src = _desc(source.parent(), dt, source.size(), source.halo())
    _, pv, _ = _parts(source)
    pvb = ctypes.create_string_buffer(pv, len(pv)) if pv else None
    dsc = _desc(dest, dt, dest.shape, 0)
    if aux is not None:
        adesc = _desc(aux.parent(), aux.sth_dtype(), aux.size(), aux.halo())
        _, apv, _ = _parts(aux)
        apvb = ctypes.create_string_buffer(apv, len(apv)) if apv else None
        aref, abc, apad = ctypes.byref(adesc), aux.boundary.code, aux.padding.code
    else:
        aref, apvb, abc, apad = None, None, 0, 0
    rc = L.orc_scatter(ctypes.byref(src), ctypes.c_int(len(source.stencil)),
                       offs.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ctypes.c_int(source.stencil.radius()),
                       ctypes.c_int(source.boundary.code), pvb, ctypes.c_int(source.padding.code), aref,
                       ctypes.c_int(abc), apvb, ctypes.c_int(apad), ctypes.c_int(f.kind), ctypes.c_double(float(f.coef)),
                       ctypes.c_int(combine_code(op)), ctypes.byref(dsc), ctypes.c_int(1 if zero_dest else 0))
    if rc != 0:
        raise ValueError(f"oracle status {rc}")
    return dest
