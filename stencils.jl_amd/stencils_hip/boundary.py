"""Boundary conditions (src/boundary.jl:7-52) and padding modes (src/padding.jl:7-39)."""
from __future__ import annotations

from . import _abi


class BoundaryCondition:
    code = -1

    def __eq__(self, other):
        return type(self) is type(other) and self.__dict__ == other.__dict__

    def __hash__(self):
        return hash(type(self))

    def __repr__(self):
        return f"{type(self).__name__}()"


class Wrap(BoundaryCondition):
    """Out-of-range index i -> i + s or i - s (src/boundary.jl:17, src/array.jl:167-179)."""
    code = _abi.WRAP


class Reflect(BoundaryCondition):
    """Mirror without repeating the edge: i < 1 -> 2 - i, i > s -> 2s - i
    (src/boundary.jl:52, src/array.jl:154-166)."""
    code = _abi.REFLECT


class Remove(BoundaryCondition):
    """Out-of-range neighbours read `padval` (src/boundary.jl:27-32).
    ``Remove()`` has ``padval = None`` as in the reference."""
    code = _abi.REMOVE

    def __init__(self, padval=None):
        self.padval = padval

    def __repr__(self):
        return f"Remove({self.padval!r})"


class Use(BoundaryCondition):
    """Use the halo as it is; only meaningful with Halo{:in} (src/boundary.jl:42)."""
    code = _abi.USE


def padval(bc):
    return bc.padval if isinstance(bc, Remove) else None


class Padding:
    code = -1

    def __eq__(self, other):
        return type(self) is type(other) and self.__dict__ == other.__dict__

    def __hash__(self):
        return hash((type(self), tuple(sorted(self.__dict__.items()))))


class Conditional(Padding):
    """No halo; boundaries resolved per neighbour (src/padding.jl:39)."""
    code = _abi.CONDITIONAL

    def __repr__(self):
        return "Conditional()"


class Halo(Padding):
    """In-memory halo of R cells (src/padding.jl:28-30).  ``Halo("out")``
    (the default) allocates a larger parent; ``Halo("in")`` uses the outer R
    cells of the given array as the halo."""

    def __init__(self, mode: str = "out"):
        mode = str(mode).lstrip(":")
        if mode not in ("out", "in"):
            raise ValueError("Halo mode must be 'out' or 'in'")
        self.mode = mode

    @property
    def code(self):
        return _abi.HALO_OUT if self.mode == "out" else _abi.HALO_IN

    def __repr__(self):
        return f"Halo{{:{self.mode}}}()"
