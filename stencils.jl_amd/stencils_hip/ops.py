"""The fixed reduction set the GPU path accelerates.

``sum``, ``mean``, ``minimum``, ``maximum`` and ``kernelproduct`` are the
functions a Stencils.jl user passes to ``mapstencil`` (StaticArrays /
Statistics reductions and src/stencils/kernel.jl:34-50).  ``LifeRule`` and
``Laplacian`` are canonical callables the reference does not define: their
CPU semantics are stated here and restated by the oracle, so every path
evaluates the same formula.  Any other function is not accelerated
(``resolve`` raises ``Unsupported``).
"""
from __future__ import annotations

import builtins
import statistics

import numpy as np

from . import _abi


class Unsupported(NotImplementedError):
    """The GPU path does not cover this call (the Julia shim falls back to the
    reference method in this case)."""


class Reduction:
    def __init__(self, name: str, kind: int):
        self.name = name
        self.kind = kind

    def __repr__(self):
        return self.name

    def op(self, flags: int = 0) -> _abi.SthOp:
        o = _abi.SthOp()
        o.kind = self.kind
        o.flags = flags
        return o


sum = Reduction("sum", _abi.OP_SUM)
mean = Reduction("mean", _abi.OP_MEAN)
minimum = Reduction("minimum", _abi.OP_MIN)
maximum = Reduction("maximum", _abi.OP_MAX)
kernelproduct = Reduction("kernelproduct", _abi.OP_KERNELPRODUCT)


class LifeRule(Reduction):
    """Life-like automaton: with ``s`` the number of nonzero neighbours and
    ``c`` the centre, the new cell is ``T(s in birth)`` if ``c == 0`` and
    ``T(s in survive)`` otherwise.  Conway's Life is B3/S23, the default.
    For 0/1 data this is ``s = sum(neighbors(hood))``."""

    def __init__(self, birth=(3,), survive=(2, 3)):
        super().__init__(f"LifeRule(B{''.join(map(str, birth))}/S{''.join(map(str, survive))})",
                         _abi.OP_LIFE)
        self.birth = tuple(int(b) for b in birth)
        self.survive = tuple(int(s) for s in survive)
        if any(not 0 <= v < 64 for v in self.birth + self.survive):
            raise ValueError("neighbour counts must be in 0..63")

    def op(self, flags: int = 0) -> _abi.SthOp:
        o = super().op(flags)
        o.birth_mask = builtins.sum(1 << b for b in set(self.birth))
        o.survive_mask = builtins.sum(1 << s for s in set(self.survive))
        return o


class Laplacian(Reduction):
    """Explicit diffusion step ``c + alpha * (sum(neighbors) - L * c)``,
    evaluated in that order in the element type (L = stencil length, 2N for
    VonNeumann(1, N))."""

    def __init__(self, alpha: float):
        super().__init__(f"Laplacian({alpha})", _abi.OP_LAPLACIAN)
        self.alpha = float(alpha)

    def op(self, flags: int = 0) -> _abi.SthOp:
        o = super().op(flags)
        o.alpha = self.alpha
        return o


class Center(Reduction):
    """``center(hood)``: the value at the centre index (src/array.jl:32); for
    a plain array argument of a multi-array gather, the value at I
    (``_getarg(a::AbstractArray, I) = a[I]``, src/gatherstencil.jl:112)."""

    def __init__(self):
        super().__init__("center", -1)

    def op(self, flags: int = 0) -> _abi.SthOp:
        raise Unsupported("center is a term of a multi-array gather, not a reduction of its own")


center = Center()


class Scaled:
    """``coef * g(hood)``: a term scaled by a constant of the element type
    (Julia evaluates ``0.1 * sum(neighbors(h))`` as one rounded product)."""

    def __init__(self, coef, g):
        self.coef = coef
        self.g = g if isinstance(g, Reduction) else resolve(g)

    def __repr__(self):
        return f"{self.coef!r} * {self.g!r}"


class Terms:
    """A function of several hoods, ``f(h_0, h_1, ...)``, of the form
    ``g_0(h_0) + g_1(h_1) + ...`` (``combine="+"``) or the product
    (``combine="*"``), folded left to right in the element type, each
    operation rounded: the multi-array gather's accelerated family
    (src/gatherstencil.jl:84-88, 105-113).  Each ``g_k`` is ``center``, a
    reduction (``sum``, ``mean``, ``minimum``, ``maximum``,
    ``kernelproduct``, ``Laplacian``, ``LifeRule``) or ``Scaled(coef, g)``.
    Examples from test/array.jl:312-383:
    ``(a, b) -> center(a) + center(b)`` is ``Terms(center, center)``;
    ``(m, r) -> center(m) + 0.1 * sum(neighbors(r))`` is
    ``Terms(center, Scaled(0.1, sum))``."""

    def __init__(self, *gs, combine: str = "+"):
        if not gs:
            raise ValueError("Terms needs at least one term")
        if combine not in ("+", "*"):
            raise ValueError("combine is '+' or '*'")
        self.terms = [g if isinstance(g, (Reduction, Scaled)) else resolve(g) for g in gs]
        self.combine = combine

    def __len__(self):
        return len(self.terms)

    def __repr__(self):
        return "Terms(" + f" {self.combine} ".join(map(repr, self.terms)) + ")"

    def code(self) -> int:
        return _abi.COMBINE_MUL if self.combine == "*" else _abi.COMBINE_ADD


_ALIASES = {
    id(builtins.sum): sum, id(np.sum): sum,
    id(np.mean): mean, id(statistics.mean): mean, id(statistics.fmean): mean,
    id(builtins.min): minimum, id(np.min): minimum, id(np.amin): minimum,
    id(builtins.max): maximum, id(np.max): maximum, id(np.amax): maximum,
}


_INT_FOLD_WRAPS = False


def set_integer_fold_wraps(flag: bool) -> bool:
    """State how this host's `sum` over a Stencil of an integer type folds:
    True when it uses the element type's own wrapping `+` (StaticArrays'
    reduction; `mean` is then that sum / L in Float64).  The reference does
    not pin it (a Stencil is a StaticVector, src/stencil.jl:17), so integer
    `sum` over Int8..UInt32 and integer `mean` are accelerated only after a
    caller has said so; otherwise they raise Unsupported.  Returns the
    previous setting."""
    global _INT_FOLD_WRAPS
    old = _INT_FOLD_WRAPS
    _INT_FOLD_WRAPS = bool(flag)
    return old


def integer_fold_wraps() -> bool:
    return _INT_FOLD_WRAPS


def resolve(f) -> Reduction:
    """Map a user function onto the accelerated set, or raise Unsupported."""
    if isinstance(f, Reduction):
        return f
    r = _ALIASES.get(id(f))
    if r is None:
        raise Unsupported(f"function {f!r} is not in the accelerated reduction set")
    return r
