"""scatterstencil! (src/scatterstencil.jl) over the C ABI (sth_scatter).

Every source cell sends one value per stencil offset to dest at the cell
plus that offset, folded with `op` (`+`, `max` or `min`).  The reference
takes an arbitrary `f(hood, args...) -> SVector`; the GPU path covers a fixed
family of value rules (`Scatter`), the shapes of the reference's own
examples and tests (test/array.jl:385-492).  Results equal the reference's
loop bit for bit: each dest cell receives its values in that loop's order.
Python has no `!`: the mutating form is `scatterstencil_`.
"""
from __future__ import annotations

import builtins
import ctypes
import operator

import numpy as np

from . import _abi, _lib
from .array import AbstractStencilArray, StencilArray, SwitchingStencilArray, colmajor_copy, is_device, to_numpy
from .boundary import Halo
from .ops import Unsupported


class Scatter:
    """The value each cell sends along each offset (A: the source's value at
    the cell, B: the second array, B_k its neighbour k, c the coefficient in
    the element type):

    - ``Scatter.constant(c)``: ``map(_ -> c, neighbors(a))``
    - ``Scatter.center()``: ``map(_ -> center(a), neighbors(a))``
    - ``Scatter.scaled(c)``: ``A * c`` to every neighbour
    - ``Scatter.downhill(c)``: ``Δ = center(b) - B_k; Δ > 0 ? A * c : 0``
    - ``Scatter.downhill_flow(c)``: ``Δ > 0 ? A * Δ * c : 0``
    """

    CONST, CENTER, SCALED, DOWNHILL, DOWNHILL_FLOW = range(5)

    def __init__(self, kind: int, coef=0.0):
        self.kind = int(kind)
        self.coef = coef

    @classmethod
    def constant(cls, c):
        return cls(cls.CONST, c)

    @classmethod
    def center(cls):
        return cls(cls.CENTER, 0.0)

    @classmethod
    def scaled(cls, c):
        return cls(cls.SCALED, c)

    @classmethod
    def downhill(cls, c):
        return cls(cls.DOWNHILL, c)

    @classmethod
    def downhill_flow(cls, c):
        return cls(cls.DOWNHILL_FLOW, c)

    def needs_aux(self) -> bool:
        return self.kind >= self.DOWNHILL

    def __repr__(self):
        names = ["constant", "center", "scaled", "downhill", "downhill_flow"]
        return f"Scatter.{names[self.kind]}({'' if self.kind == self.CENTER else repr(self.coef)})"


_OPS = {id(operator.add): _abi.COMBINE_ADD, id(np.add): _abi.COMBINE_ADD,
        id(builtins.max): _abi.COMBINE_MAX, id(np.maximum): _abi.COMBINE_MAX,
        id(builtins.min): _abi.COMBINE_MIN, id(np.minimum): _abi.COMBINE_MIN}


def combine_code(op) -> int:
    """`+` (operator.add), `max` or `min` as the ABI's combine code."""
    if isinstance(op, str) and op in ("+", "max", "min"):
        return {"+": _abi.COMBINE_ADD, "max": _abi.COMBINE_MAX, "min": _abi.COMBINE_MIN}[op]
    c = _OPS.get(id(op))
    if c is None:
        raise Unsupported(f"scatter op {op!r} is not accelerated (+, max, min)")
    return c


def _term(sa: AbstractStencilArray, buf, call):
    from .gather import _desc, _padval_bytes, stencil_descriptor
    t = _abi.SthTerm()
    t.src = _desc(buf, sa.size(), sa.halo())
    t.stencil = stencil_descriptor(sa.stencil, sa.sth_dtype(), call)
    t.boundary = sa.boundary.code
    t.padding = sa.padding.code
    pv = _padval_bytes(sa)
    if pv is not None:
        t.padval = ctypes.cast(call.hold(ctypes.create_string_buffer(pv, len(pv))), ctypes.c_void_p)
    return t


def _scatter_device(f: Scatter, op, dest, source, sbuf, aux, abuf, flags: int):
    from .gather import _Call, _desc, _device_index, _stream
    call = _Call()
    ts = _term(source, sbuf, call)
    ta = _term(aux, abuf, call) if aux is not None else None
    rule = _abi.SthScatterRule()
    rule.kind = f.kind
    rule.combine = combine_code(op)
    rule.coef = float(f.coef)
    dsc = _desc(dest, tuple(dest.shape), 0)
    dev = _device_index(dest)
    _lib.check(_lib.lib().sth_scatter(_lib.context(dev), ctypes.byref(ts), None if ta is None else ctypes.byref(ta),
                                      ctypes.byref(dsc), ctypes.byref(rule), flags, _stream(dev)))


def scatterstencil_(f: Scatter, op, dest, source=None, *args, flags: int = 0):
    """scatterstencil!(f, op, dest, source, args...) (src/scatterstencil.jl:37-47),
    scatterstencil!(f, op, A::SwitchingStencilArray, args...) (:127-136) and
    scatterstencil!(f, op, A::SwitchingStencilArray, source, args...)
    (:117-124); the SwitchingStencilArray forms clear dest, scatter into it
    and return the switched array."""
    from .gather import _require_backend
    if not isinstance(f, Scatter):
        raise Unsupported("scatterstencil is accelerated for Scatter value rules only")
    _require_backend()
    if isinstance(dest, SwitchingStencilArray):
        A = dest
        if isinstance(source, AbstractStencilArray):
            src = source  # the explicit-source form (the more specific method)
        else:
            # the array scatters its own source (Halo padding re-wrapped as Halo{:in})
            if source is not None:
                args = (source,) + args
            pad = Halo("in") if isinstance(A.padding, Halo) else A.padding
            src = StencilArray(A.source, A.stencil, boundary=A.boundary, padding=pad)
        _scatter_into(f, op, A.dest, src, args, flags | _abi.SCATTER_ZERO_DEST)
        return A.switch()
    _scatter_into(f, op, dest, source, args, flags)
    return dest


def _scatter_into(f, op, dest, source, args, flags):
    if not isinstance(source, AbstractStencilArray):
        raise TypeError("source must be a StencilArray")
    aux = None
    if f.needs_aux():
        if len(args) != 1 or not isinstance(args[0], AbstractStencilArray):
            raise TypeError(f"{f!r} reads a second StencilArray argument")
        aux = args[0]
    elif args:
        raise TypeError(f"{f!r} takes no further arguments")
    sizes = [tuple(np.shape(dest)), tuple(source.size())] + ([tuple(aux.size())] if aux is not None else [])
    if any(s != sizes[0] for s in sizes):
        raise ValueError(f"Source array sizes must match. Found: {sizes}")
    parents = [source.parent()] + ([aux.parent()] if aux is not None else [])
    if is_device(dest) and all(is_device(p) for p in parents):
        _scatter_device(f, op, dest, source, source.parent(), aux, None if aux is None else aux.parent(), flags)
        return
    # host round trip (Halo parents get their refreshed halos back)
    import torch
    dev = torch.cuda.current_device()
    bufs = [p if is_device(p) else colmajor_copy(p, dev) for p in parents]
    dbuf = dest if is_device(dest) else colmajor_copy(dest, dev)
    _scatter_device(f, op, dbuf, source, bufs[0], aux, bufs[1] if aux is not None else None, flags)
    if not is_device(dest):
        dest[...] = to_numpy(dbuf)
    for sa, p, b in zip([source, aux], parents, bufs):
        if sa is not None and not is_device(p) and isinstance(sa.padding, Halo):
            p[...] = to_numpy(b)
