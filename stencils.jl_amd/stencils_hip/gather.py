"""mapstencil / mapstencil_ (= gatherstencil / gatherstencil!) over the C ABI.

Mirrors src/gatherstencil.jl:
  gatherstencil(f, A::StencilArray)                :14-21
  gatherstencil(f, hood, A; boundary, padding)     :22-39
  gatherstencil!(f, A::SwitchingStencilArray)      :77-83
  gatherstencil!(f, dest, source)                  :89-103
  mapstencil / mapstencil! deprecated aliases      :127-128
Python has no `!`, so the mutating forms end in an underscore.

Device parents (torch tensors) are used in place.  Host parents (numpy) are
copied to the device for the call and the results copied back, including
the Halo refresh the reference performs on the source (update_boundary!).
There is no CPU path: without the HIP library or a device the call raises.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _abi, _lib
from .array import (AbstractStencilArray, StencilArray, SwitchingStencilArray,
                    colmajor_copy, colmajor_empty, is_device, sth_dtype_of,
                    torch_dtype, to_numpy)
from .boundary import Conditional, Halo, Remove, Use
from .ops import Center, Reduction, Scaled, Terms, Unsupported, integer_fold_wraps, resolve
from .stencil import Kernel, Stencil


def _torch():
    import torch
    return torch


# --------------------------------------------------------------- descriptors

class _Call:
    """Keeps ctypes buffers alive for the duration of one library call."""

    def __init__(self):
        self.keep = []

    def hold(self, x):
        self.keep.append(x)
        return x


def _padval_bytes(sa: AbstractStencilArray):
    bc = sa.boundary
    if not isinstance(bc, Remove):
        return None
    pv = bc.padval
    if pv is None:
        raise Unsupported("Remove() with padval=nothing cannot be folded (reference MethodError)")
    dt = sa.sth_dtype()
    if isinstance(pv, np.generic) and _abi.NUMPY_TO_STH.get(np.dtype(type(pv))) != dt:
        # The reference promotes the stencil element type (src/gatherstencil.jl:50).
        raise Unsupported(f"padval of type {type(pv).__name__} promotes the element type")
    if dt == _abi.BF16 or dt == _abi.F16:
        torch = _torch()
        t = torch.tensor([float(pv)], dtype=torch_dtype(dt))
        if float(t[0]) != float(pv):
            raise Unsupported("padval is not representable in the storage type")
        return t.view(torch.int16).numpy().tobytes()
    npdt = _abi.STH_TO_NUMPY[dt]
    v = np.array(pv).astype(npdt)
    if not (v == np.array(pv)).all() or (dt not in _abi.FLOAT_DTYPES and isinstance(pv, float)):
        raise Unsupported("padval is not exactly representable in the element type")
    return v.tobytes()


def _weights(st: Stencil, src_dt: int):
    if not isinstance(st, Kernel):
        return None, -1
    w = np.asarray(st.weights)
    if src_dt in (_abi.BF16, _abi.F16):
        return np.ascontiguousarray(w, dtype=np.float32), _abi.F32
    wdt = _abi.NUMPY_TO_STH.get(w.dtype, -1)
    if w.dtype.kind == "i" and src_dt in _abi.FLOAT_DTYPES:
        # integer weights promote to the data's float type without loss here
        w = w.astype(_abi.STH_TO_NUMPY[src_dt])
        wdt = src_dt
    return np.ascontiguousarray(w), wdt


# Descriptors of stencils already described, per (stencil object, element
# type): the entry keeps the stencil alive, so its id is not reused while
# cached, and keeps the offset / weight arrays the descriptor points to.
_STENCIL_DESCS: dict = {}


def stencil_descriptor(st: Stencil, src_dt: int, call: _Call) -> _abi.SthStencil:
    hit = _STENCIL_DESCS.get((id(st), src_dt))
    if hit is not None and hit[0] is st:
        return hit[1]
    d, arrays = _build_stencil_descriptor(st, src_dt)
    for a in arrays:
        call.hold(a)
    if len(_STENCIL_DESCS) >= 256:
        _STENCIL_DESCS.clear()
    _STENCIL_DESCS[(id(st), src_dt)] = (st, d, arrays)
    return d


def _build_stencil_descriptor(st: Stencil, src_dt: int):
    """(descriptor, the arrays it points to)."""
    if not st.gpu_supported():
        raise Unsupported(f"{st!r} is not supported on the GPU path")
    offs = np.ascontiguousarray(np.array(st.offsets(), dtype=np.int32).reshape(-1))
    d = _abi.SthStencil()
    d.kind = st.kind
    d.radius = st.radius()
    d.ndim = st.N
    d.length = len(st)
    d.offsets = offs.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    w, wdt = _weights(st, src_dt)
    if w is not None:
        d.weights = w.ctypes.data
        d.weights_dtype = wdt
        return d, (offs, w)
    d.weights = None
    d.weights_dtype = -1
    return d, (offs,)


def _op_flags(src_dt: int) -> int:
    fl = _abi.OPF_ACC_F32 if src_dt in (_abi.BF16, _abi.F16) else 0
    if integer_fold_wraps() and src_dt not in _abi.FLOAT_DTYPES:
        fl |= _abi.OPF_INT_WRAP
    return fl


def result_dtype(f, sa: AbstractStencilArray) -> int:
    red = resolve(f)
    dt = sa.sth_dtype()
    _, wdt = _weights(sa.stencil, dt)
    fl = _op_flags(dt)
    o = _abi.out_dtype(dt, red.kind, fl, wdt)
    if o < 0:
        raise Unsupported(f"{red!r} over {_abi.DTYPE_NAMES[dt]} is not accelerated")
    if _abi.int_fold_declined(dt, red.kind, fl, len(sa.stencil)):
        raise Unsupported(f"{red!r} over {_abi.DTYPE_NAMES[dt]} with {len(sa.stencil)} offsets: integer "
                          "overflow semantics unpinned (set_integer_fold_wraps)")
    return o


def _device_index(x) -> int:
    return x.device.index if x.device.index is not None else _torch().cuda.current_device()


def _stream(dev: int):
    return ctypes.c_void_p(_torch().cuda.current_stream(dev).cuda_stream)


def _raw_stream(dev: int):
    """The current stream of `dev` as a hipStream_t (torch's raw getter when
    it has one: no Stream object per call)."""
    torch = _torch()
    get = getattr(torch._C, "_cuda_getCurrentRawStream", None)
    if get is None:
        return _stream(dev)
    return ctypes.c_void_p(get(dev))


def _desc(x, shape, halo) -> _abi.SthArray:
    from .array import data_ptr, elem_strides
    return _abi.make_array(data_ptr(x), sth_dtype_of(x), shape, elem_strides(x), halo)


# ---------------------------------------------------------------- the calls

def _map_device(f, dst, dst_halo, sa: AbstractStencilArray, src_buf, flags: int = 0):
    """One sth_map call on device buffers `src_buf` (parent) -> `dst`."""
    red = resolve(f)
    dt = sa.sth_dtype()
    call = _Call()
    st = stencil_descriptor(sa.stencil, dt, call)
    op = red.op(_op_flags(dt))
    pv = _padval_bytes(sa)
    pvbuf = call.hold(ctypes.create_string_buffer(pv, len(pv))) if pv is not None else None
    size = sa.size()
    src = _desc(src_buf, size, sa.halo())
    dsc = _desc(dst, size, dst_halo)
    dev = _device_index(src_buf)
    L = _lib.lib()
    ctx = _lib.context(dev)
    _lib.check(L.sth_map(ctx, ctypes.byref(src), ctypes.byref(dsc), ctypes.byref(st),
                         sa.boundary.code, pvbuf, sa.padding.code, ctypes.byref(op),
                         flags, _stream(dev)))


def _term_parts(g):
    """(fold, coef or None) of one term of a Terms function."""
    if isinstance(g, Scaled):
        return g.g, g.coef
    return g, None


def _multi_check(f: Terms, args):
    if not isinstance(f, Terms):
        raise Unsupported("a multi-array gather is accelerated for Terms functions only")
    if len(f) != len(args):
        raise TypeError(f"{f!r} takes {len(f)} arguments, got {len(args)}")
    if len(args) > _abi.MAX_TERMS:
        raise Unsupported(f"more than {_abi.MAX_TERMS} arguments")


def _map_multi_device(f: Terms, dst, dst_halo, args, bufs, flags: int = 0):
    """One sth_map_multi call: argument k is `args[k]` (a StencilArray or a
    plain array) whose device parent is `bufs[k]`."""
    call = _Call()
    terms = (_abi.SthTerm * len(args))()
    dt = sth_dtype_of(dst)
    size = None
    for k, (a, buf) in enumerate(zip(args, bufs)):
        g, coef = _term_parts(f.terms[k])
        t = terms[k]
        if isinstance(a, AbstractStencilArray):
            size = a.size()
            t.src = _desc(buf, a.size(), a.halo())
            t.boundary = a.boundary.code
            t.padding = a.padding.code
            pv = _padval_bytes(a) if (not isinstance(g, Center) or not isinstance(a.padding, Conditional)) else None
            if pv is not None:
                t.padval = ctypes.cast(call.hold(ctypes.create_string_buffer(pv, len(pv))), ctypes.c_void_p)
            adt = a.sth_dtype()
            if isinstance(g, Center):
                t.stencil.length = 0
            else:
                t.stencil = stencil_descriptor(a.stencil, adt, call)
                t.op = g.op(_op_flags(adt))
        else:
            if not isinstance(g, Center):
                raise TypeError(f"argument {k} is a plain array (read at I): its term must be `center`, not {g!r}")
            t.src = _desc(buf, tuple(buf.shape), 0)
            t.boundary = _abi.WRAP
            t.padding = _abi.CONDITIONAL
            t.stencil.length = 0
        if coef is not None:
            t.scaled = 1
            t.coef = float(coef)
    dsc = _desc(dst, size if size is not None else tuple(dst.shape), dst_halo)
    dev = _device_index(dst)
    L = _lib.lib()
    _lib.check(L.sth_map_multi(_lib.context(dev), len(args), terms, ctypes.byref(dsc), f.code(), flags,
                               _stream(dev)))


def _multi_parent(a):
    return a.parent() if isinstance(a, AbstractStencilArray) else a


def gatherstencil_multi_(f: Terms, dest, args, dest_halo: int = 0, flags: int = 0):
    """gatherstencil!(f, dest, source, args...) with several arguments
    (src/gatherstencil.jl:89-103): every StencilArray argument's halo is
    refreshed, then dest[I] = f(args[0] at I, args[1] at I, ...)."""
    _multi_check(f, args)
    _require_backend()
    sizes = [tuple(a.size()) if isinstance(a, AbstractStencilArray) else tuple(np.shape(a)) for a in args]
    dsize = tuple(np.shape(dest)) if dest_halo == 0 else tuple(s - 2 * dest_halo for s in np.shape(dest))
    for s in sizes:
        if s != dsize:
            raise ValueError(f"Source array sizes must match. Found: {dsize} and {s}")
    parents = [_multi_parent(a) for a in args]
    if all(is_device(p) for p in parents) and is_device(dest):
        _map_multi_device(f, dest, dest_halo, args, parents, flags)
        return dest
    # host round trip: copy every argument in, the result and the refreshed
    # halos back
    torch = _torch()
    dev = torch.cuda.current_device()
    bufs = [p if is_device(p) else colmajor_copy(p, dev) for p in parents]
    dbuf = dest if is_device(dest) else colmajor_copy(dest, dev)
    _map_multi_device(f, dbuf, dest_halo, args, bufs, flags)
    if not is_device(dest):
        dest[...] = to_numpy(dbuf)
    for a, p, b in zip(args, parents, bufs):
        if not is_device(p) and isinstance(a, AbstractStencilArray) and isinstance(a.padding, Halo):
            p[...] = to_numpy(b)
    return dest


def _switching_multi_step(f: Terms, A: SwitchingStencilArray, rest, flags: int = 0):
    """mapstencil!(f, A::SwitchingStencilArray, B, ...) (src/gatherstencil.jl:84-88):
    A's dest interior <- f(A, B, ...), then switch."""
    _multi_check(f, (A,) + tuple(rest))
    _require_backend()
    torch = _torch()
    if is_device(A.source):
        gatherstencil_multi_(f, A.dest, (A,) + tuple(rest), A.halo(), flags)
        return A.switch()
    dev = torch.cuda.current_device()
    s = colmajor_copy(A.source, dev)
    d = colmajor_copy(A.dest, dev)
    proxy = SwitchingStencilArray(stencil=A.stencil, boundary=A.boundary, padding=A.padding, _buffers=(s, d))
    gatherstencil_multi_(f, d, (proxy,) + tuple(rest), A.halo(), flags)
    A.source[...] = to_numpy(s)
    A.dest[...] = to_numpy(d)
    return A.switch()


def autotune(f, dest, sa: AbstractStencilArray) -> str:
    """Time every kernel variant that can run gatherstencil!(f, dest, sa) on
    device buffers (sth_autotune), keep the fastest for this call's family,
    run the call with it and return its name.  A variant forced with
    `_lib.set_variant` is returned (and run) without timing."""
    red = resolve(f)
    _require_backend()
    dt = sa.sth_dtype()
    call = _Call()
    st = stencil_descriptor(sa.stencil, dt, call)
    op = red.op(_op_flags(dt))
    pv = _padval_bytes(sa)
    pvbuf = call.hold(ctypes.create_string_buffer(pv, len(pv))) if pv is not None else None
    size = sa.size()
    src = _desc(sa.parent(), size, sa.halo())
    dsc = _desc(dest, size, 0)
    dev = _device_index(sa.parent())
    name = ctypes.create_string_buffer(256)
    _lib.check(_lib.lib().sth_autotune(_lib.context(dev), ctypes.byref(src), ctypes.byref(dsc),
                                       ctypes.byref(st), sa.boundary.code, pvbuf, sa.padding.code,
                                       ctypes.byref(op), _stream(dev), name, len(name)))
    return name.value.decode()


_BACKEND_OK = False


def _require_backend():
    global _BACKEND_OK
    if _BACKEND_OK:
        return
    torch = _torch()
    if not torch.cuda.is_available():
        raise _lib.BackendUnavailable("no HIP device is visible; the GPU path has no CPU fallback")
    _BACKEND_OK = True


def gatherstencil_(f, dest, source=None, *args, flags: int = 0):
    """gatherstencil!(f, dest, source) or gatherstencil!(f, A::SwitchingStencilArray)."""
    if isinstance(dest, SwitchingStencilArray):
        if source is None:
            return _switching_step(f, dest, flags)
        return _switching_multi_step(f, dest, (source,) + args, flags)
    if args:
        if isinstance(dest, AbstractStencilArray):
            gatherstencil_multi_(f, dest.parent(), (source,) + args, dest.halo(), flags)
            return dest
        return gatherstencil_multi_(f, dest, (source,) + args, 0, flags)
    if not isinstance(source, AbstractStencilArray):
        raise TypeError("source must be a StencilArray")
    resolve(f)
    _require_backend()
    if isinstance(dest, AbstractStencilArray):
        # a StencilArray destination is written through its halo offset
        # (setindex! with add_halo, src/array.jl:355-362)
        if tuple(dest.size()) != tuple(source.size()):
            raise ValueError(f"Source array sizes must match. Found: {tuple(dest.size())} and "
                             f"{tuple(source.size())}")
        if not (is_device(dest.parent()) and is_device(source.parent())):
            raise Unsupported("StencilArray destinations must live on the device")
        _map_device(f, dest.parent(), dest.halo(), source, source.parent(), flags)
        return dest
    if tuple(np.shape(dest)) != tuple(source.size()):
        raise ValueError(f"Source array sizes must match. Found: {tuple(np.shape(dest))} and "
                         f"{tuple(source.size())}")
    src_parent = source.parent()
    halo_refresh = isinstance(source.padding, Halo) and not isinstance(source.boundary, Use)
    if is_device(src_parent) and is_device(dest):
        _map_device(f, dest, 0, source, src_parent, flags)
        return dest
    # host round trip
    torch = _torch()
    dev = torch.cuda.current_device()
    sbuf = src_parent if is_device(src_parent) else colmajor_copy(src_parent, dev)
    dbuf = colmajor_empty(np.shape(dest), torch_dtype(sth_dtype_of(dest)), dev)
    _map_device(f, dbuf, 0, source, sbuf, flags)
    if is_device(dest):
        dest.copy_(dbuf)
    else:
        dest[...] = to_numpy(dbuf)
    if halo_refresh and not is_device(src_parent):
        src_parent[...] = to_numpy(sbuf)
    return dest


def _alloc_like(parent, dt_code: int, shape):
    if is_device(parent):
        return colmajor_empty(shape, torch_dtype(dt_code), parent.device)
    return np.zeros(shape, dtype=_abi.STH_TO_NUMPY[dt_code], order="F")


def gatherstencil(f, *args, boundary=None, padding=None):
    """gatherstencil(f, A::StencilArray) or gatherstencil(f, hood, A; boundary, padding)."""
    if not args:
        raise TypeError("gatherstencil needs an array")
    if isinstance(args[0], Stencil):
        hood, A = args[0], args[1]
        rest = args[2:]
        sa = StencilArray(A, hood, boundary=boundary, padding=padding)
        T = result_dtype(f, sa)
        size = sa.size()
        dest = _alloc_like(A, T, size)
        return gatherstencil_(f, dest, sa, *rest)
    sa = args[0]
    if not isinstance(sa, AbstractStencilArray):
        raise TypeError("gatherstencil(f, A) needs a StencilArray")
    if len(args) > 1:
        # several arguments: the result has the arguments' element type
        _multi_check(f, args)
        dest = _alloc_like(sa.parent(), sa.sth_dtype(), sa.size())
        return gatherstencil_multi_(f, dest, args)
    T = result_dtype(f, sa)
    dest = _alloc_like(sa.parent(), T, sa.size())
    return gatherstencil_(f, dest, sa, *args[1:])


def _inner(sa, buf):
    h = sa.halo()
    return buf[tuple(slice(h, s - h) for s in buf.shape)]


def _switching_step(f, A: SwitchingStencilArray, flags: int = 0):
    """gatherstencil!(f, A::SwitchingStencilArray): dest <- f(source), then switch."""
    resolve(f)
    _require_backend()
    if result_dtype(f, A) != A.sth_dtype():
        raise ValueError("the result type differs from the array's element type")
    if is_device(A.source):
        _map_device(f, A.dest, A.halo(), A, A.source, flags)
        return A.switch()
    torch = _torch()
    dev = torch.cuda.current_device()
    s = colmajor_copy(A.source, dev)
    d = colmajor_copy(A.dest, dev)
    _map_device(f, d, A.halo(), A, s, flags)
    A.source[...] = to_numpy(s)
    A.dest[...] = to_numpy(d)
    return A.switch()


def _swap_iterate_args(c):
    """The cached sth_iterate arguments of the switched array."""
    a, b, dev, ctx, da, db, st, op, pvbuf, keep = c[6]
    return c[:6] + ((b, a, dev, ctx, db, da, st, op, pvbuf, keep),)


def _iterate_args(red, A: SwitchingStencilArray, dev: int, copy: bool):
    """Device buffers and ABI descriptors of sth_iterate over A's buffers
    (copied to the device first when `copy`)."""
    if result_dtype(red, A) != A.sth_dtype():
        raise ValueError("the result type differs from the array's element type")
    a = colmajor_copy(A.source, dev) if copy else A.source
    b = colmajor_copy(A.dest, dev) if copy else A.dest
    dt = A.sth_dtype()
    call = _Call()
    st = stencil_descriptor(A.stencil, dt, call)
    op = red.op(_op_flags(dt))
    pv = _padval_bytes(A)
    pvbuf = call.hold(ctypes.create_string_buffer(pv, len(pv))) if pv is not None else None
    size = A.size()
    da = _desc(a, size, A.halo())
    db = _desc(b, size, A.halo())
    return a, b, dev, _lib.context(dev), da, db, st, op, pvbuf, call


def mapstencil_steps(f, A: SwitchingStencilArray, steps: int, graph: bool = True,
                     flags: int = 0):
    """`steps` calls of mapstencil!(f, A) on device in one library call
    (sth_iterate); with `graph` the steps replay as a captured hipGraph.
    Returns the array switched as the last call of the loop would return it."""
    red = resolve(f)
    _require_backend()
    host = not is_device(A.source)
    if host:
        args = _iterate_args(red, A, _torch().cuda.current_device(), copy=True)
    else:
        # device buffers: the descriptors ride on the array (and, roles
        # swapped, on the arrays switch() returns), valid while its stencil,
        # boundary, padding and the function are the ones they were made for
        c = A.__dict__.get("_sth_args")
        if (c is not None and c[1] is red and c[2] is A.stencil and c[3] is A.boundary and c[4] is A.padding
                and c[5] == integer_fold_wraps() and c[6][0] is A.source and c[6][1] is A.dest):
            args = c[6]
        else:
            args = _iterate_args(red, A, _device_index(A.source), copy=False)
            A._sth_args = (_swap_iterate_args, red, A.stencil, A.boundary, A.padding, integer_fold_wraps(), args)
    a, b, dev, ctx, da, db, st, op, pvbuf, _keep = args
    cur = ctypes.c_int32(0)
    fl = flags | (_abi.GRAPH if graph else 0)
    _lib.check(_lib.lib().sth_iterate(ctx, ctypes.byref(da), ctypes.byref(db), ctypes.byref(st),
                                      A.boundary.code, pvbuf, A.padding.code, ctypes.byref(op),
                                      int(steps), fl, ctypes.byref(cur), _raw_stream(dev)))
    if host:
        A.source[...] = to_numpy(a)
        A.dest[...] = to_numpy(b)
    return A.switch() if cur.value == 1 else A


class BatchPlan:
    """Descriptors of many gatherstencil!(f, dest, source) calls, built once
    (device tensors only) and run with sth_map_batch as often as needed."""

    def __init__(self, calls):
        _require_backend()
        calls = list(calls)
        self.n = len(calls)
        self.jobs = (_abi.SthJob * max(self.n, 1))()
        self._keep = _Call()
        self.dev = None
        self.cells = 0
        for i, (f, dest, sa) in enumerate(calls):
            red = resolve(f)
            if not (is_device(dest) and is_device(sa.parent())):
                raise Unsupported("batched jobs must live on the device")
            dt = sa.sth_dtype()
            j = self.jobs[i]
            j.src = _desc(sa.parent(), sa.size(), sa.halo())
            j.dst = _desc(dest, sa.size(), 0)
            j.stencil = stencil_descriptor(sa.stencil, dt, self._keep)
            j.op = red.op(_op_flags(dt))
            pv = _padval_bytes(sa)
            if pv is not None:
                buf = self._keep.hold(ctypes.create_string_buffer(pv, len(pv)))
                j.padval = ctypes.cast(buf, ctypes.c_void_p)
            j.boundary = sa.boundary.code
            j.padding = sa.padding.code
            self._keep.hold((dest, sa))
            c = 1
            for s in sa.size():
                c *= s
            self.cells += c
            d = _device_index(sa.parent())
            if self.dev is None:
                self.dev = d
            elif d != self.dev:
                raise ValueError("all jobs of a batch must be on one device")

    def run(self, serial: bool = False, flags: int = 0):
        if self.n == 0:
            return
        L = _lib.lib()
        fl = flags | (_abi.BATCH_SERIAL if serial else 0)
        _lib.check(L.sth_map_batch(_lib.context(self.dev), self.n, self.jobs, fl,
                                   _stream(self.dev)))

    def run_static(self, flags: int = 0):
        """The same calls through sth_map_batch_static, which allocates no
        memory (its job tables live in a fixed per-device array of
        STATIC_BATCH_MAX_JOBS entries shared by every such call)."""
        if self.n == 0:
            return
        if self.n > _abi.STATIC_BATCH_MAX_JOBS:
            raise ValueError(f"{self.n} jobs exceed the fixed table of {_abi.STATIC_BATCH_MAX_JOBS}")
        L = _lib.lib()
        _lib.check(L.sth_map_batch_static(_lib.context(self.dev), self.n, self.jobs, flags,
                                          _stream(self.dev)))


def mapstencil_batch(calls, serial: bool = False, flags: int = 0):
    """Many gatherstencil!(f, dest, source) calls in one library call
    (sth_map_batch): jobs of mixed sizes, shapes and reductions share one
    launch per (element type, shape, reduction).  `calls` is a sequence of
    (f, dest, source) with device tensors; `serial` runs one launch per job
    instead (the unbatched path).  Results equal the one-call-per-job path."""
    BatchPlan(calls).run(serial=serial, flags=flags)


def mapstencil_steps_batch(items, graph: bool = True, serial: bool = False, flags: int = 0):
    """Many iterations with their own step counts in one library call
    (sth_iterate_batch): `items` is a sequence of (f, A, steps), A a
    device-resident SwitchingStencilArray; each A gets `steps` calls of
    mapstencil!(f, A), all arrays advancing together one step per round (one
    launch per element type, shape and reduction over the arrays still
    running).  With `graph` the whole schedule replays as a captured hipGraph
    on later calls with the same arrays, functions and step counts.  Returns
    the arrays switched as each one's own loop would return it; each equals
    mapstencil_steps(f, A, steps) on that array alone."""
    _require_backend()
    items = list(items)
    n = len(items)
    jobs = (_abi.SthJob * max(n, 1))()
    steps = (ctypes.c_int64 * max(n, 1))()
    cur = (ctypes.c_int32 * max(n, 1))()
    keep = _Call()
    dev = None
    for i, (f, A, k) in enumerate(items):
        red = resolve(f)
        if not isinstance(A, SwitchingStencilArray) or not is_device(A.source):
            raise Unsupported("batched iterations take device-resident SwitchingStencilArrays")
        if result_dtype(red, A) != A.sth_dtype():
            raise ValueError("the result type differs from the array's element type")
        if int(k) < 0:
            raise ValueError("negative step count")
        dt = A.sth_dtype()
        j = jobs[i]
        j.src = _desc(A.source, A.size(), A.halo())
        j.dst = _desc(A.dest, A.size(), A.halo())
        j.stencil = stencil_descriptor(A.stencil, dt, keep)
        j.op = red.op(_op_flags(dt))
        pv = _padval_bytes(A)
        if pv is not None:
            j.padval = ctypes.cast(keep.hold(ctypes.create_string_buffer(pv, len(pv))), ctypes.c_void_p)
        j.boundary = A.boundary.code
        j.padding = A.padding.code
        steps[i] = int(k)
        d = _device_index(A.source)
        if dev is None:
            dev = d
        elif d != dev:
            raise ValueError("all arrays of a batch must be on one device")
    if n == 0:
        return []
    fl = flags | (_abi.GRAPH if graph else 0) | (_abi.BATCH_SERIAL if serial else 0)
    _lib.check(_lib.lib().sth_iterate_batch(_lib.context(dev), n, jobs, steps, fl, cur, _raw_stream(dev)))
    return [A.switch() if cur[i] == 1 else A for i, (_, A, _k) in enumerate(items)]


# Deprecated aliases of the reference (src/gatherstencil.jl:127-128).
mapstencil = gatherstencil
mapstencil_ = gatherstencil_
