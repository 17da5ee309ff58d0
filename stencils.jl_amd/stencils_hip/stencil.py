"""Stencil shape types: ordered offset tables, radius, length.

Mirrors the stencil types of Stencils.jl v0.3.6.  Offsets are tuples
``(d1, d2[, d3])`` with ``d1`` along the contiguous (first, column-major)
axis, listed in Julia ``CartesianIndices`` order (first component varies
fastest).  The order matters: every reduction folds left over it.

Reference offset generators:
  Window      src/stencils/window.jl:4-8
  Moore       src/stencils/moore.jl:5-18
  VonNeumann  src/stencils/vonneumman.jl:5-15
  Circle etc. src/stencils/shapes.jl:1-175
  Rectangle   src/stencils/rectangle.jl:13-56
  Positional  src/stencils/positional.jl:31-76
  NamedStencil src/stencils/named.jl:43-120
  Kernel      src/stencils/kernel.jl:93-123
  indices     src/stencil.jl:90-96 (trailing dims pass through)
  distances   src/stencil.jl:116-120
"""
from __future__ import annotations

import itertools
import math

import numpy as np

# Shape kinds shared with the C ABI (include/stencils_hip.h, STH_SHAPE_*).
SHAPE_POSITIONAL = 0
SHAPE_WINDOW = 1
SHAPE_MOORE = 2
SHAPE_VONNEUMANN = 3
SHAPE_CIRCLE = 4
SHAPE_CROSS = 5
SHAPE_ANGLEDCROSS = 6
SHAPE_FORWARDSLASH = 7
SHAPE_BACKSLASH = 8
SHAPE_VERTICAL = 9
SHAPE_HORIZONTAL = 10
SHAPE_DIAMOND = 11
SHAPE_ANNULUS = 12
SHAPE_CARDINAL = 13
SHAPE_ORDINAL = 14
SHAPE_RECTANGLE = 15
SHAPE_NAMED = 16


def _cartesian(R: int, N: int):
    """CartesianIndices(ntuple(_ -> -R:R, N)) in Julia order (first fastest)."""
    rng = range(-R, R + 1)
    for t in itertools.product(rng, repeat=N):
        yield tuple(reversed(t))


class Stencil:
    """Abstract stencil: an ordered list of offsets around a centre cell.

    Subclasses define ``_offsets()``.  ``R`` is the reference's radius type
    parameter and ``N`` its dimensionality (src/stencil.jl:17).
    """

    kind = SHAPE_POSITIONAL

    def __init__(self, R: int, N: int):
        if N < 1:
            raise ValueError("stencil ndims must be >= 1")
        self.R = int(R)
        self.N = int(N)
        self._offs = None

    # -- reference API -------------------------------------------------
    def offsets(self):
        if self._offs is None:
            self._offs = tuple(tuple(int(v) for v in o) for o in self._offsets())
        return self._offs

    def radius(self) -> int:
        return self.R

    def diameter(self) -> int:
        return 2 * self.radius() + 1

    def ndims(self) -> int:
        return self.N

    def __len__(self) -> int:
        return len(self.offsets())

    def indices(self, I):
        """Offsets added to centre ``I``; trailing array dims pass through."""
        I = tuple(I)
        N1 = self.N
        if len(I) < N1:
            raise ValueError("index has fewer dims than the stencil")
        head, tail = I[:N1], I[N1:]
        return [tuple(a + b for a, b in zip(o, head)) + tail for o in self.offsets()]

    def distances(self):
        return [math.sqrt(sum(v * v for v in o)) for o in self.offsets()]

    def distance_zones(self):
        return [sum(abs(v) for v in o) for o in self.offsets()]

    # -- what the GPU path needs ----------------------------------------
    def reach(self):
        """Largest |offset| per stencil dim (the memory actually touched)."""
        offs = self.offsets()
        return tuple(max(abs(o[d]) for o in offs) for d in range(self.N))

    def gpu_supported(self) -> bool:
        return True

    def __eq__(self, other):
        return (isinstance(other, Stencil) and self.N == other.N
                and self.radius() == other.radius()
                and self.offsets() == other.offsets())

    def __hash__(self):
        return hash((self.N, self.radius(), self.offsets()))

    def __repr__(self):
        return f"{type(self).__name__}{{{self.R},{self.N}}}"


class Window(Stencil):
    """(2R+1)^N box including the centre (src/stencils/window.jl:4-8).

    The reference generates 2-tuples ``(rem(i-1, D) - R, (i-1) div D - R)``
    for ``i in 1:D^N``.  For N=1 the second component is dropped when added
    to a 1-tuple index, so Window{R,1} is the 1-D segment -R..R.  For N=3 the
    reference table is malformed (2-tuples indexing a 3-D array), so the GPU
    path declines Window{R,3}.
    """

    kind = SHAPE_WINDOW

    def __init__(self, radius: int = 1, ndims: int = 2):
        super().__init__(radius, ndims)

    def _offsets(self):
        R, N = self.R, self.N
        D = 2 * R + 1
        raw = [((i - 1) % D - R, (i - 1) // D - R) for i in range(1, D ** N + 1)]
        if N == 1:
            return [(a,) for a, _ in raw]
        if N == 2:
            return raw
        # Reference-malformed for N >= 3: keep the table as the reference
        # defines it, padded with zeros so the object is still inspectable.
        return [o + (0,) * (N - 2) for o in raw]

    def gpu_supported(self) -> bool:
        return self.N <= 2


class Moore(Stencil):
    """Box without the centre (src/stencils/moore.jl:5-18)."""

    kind = SHAPE_MOORE

    def __init__(self, radius: int = 1, ndims: int = 2):
        super().__init__(radius, ndims)

    def _offsets(self):
        allc = list(_cartesian(self.R, self.N))
        L = len(allc) - 1
        return allc[: L // 2] + allc[L // 2 + 1:]


class VonNeumann(Stencil):
    """Manhattan distance 1..R, no centre (src/stencils/vonneumman.jl:5-15)."""

    kind = SHAPE_VONNEUMANN

    def __init__(self, radius: int = 1, ndims: int = 2):
        super().__init__(radius, ndims)

    def _offsets(self):
        return [I for I in _cartesian(self.R, self.N)
                if 1 <= sum(abs(v) for v in I) <= self.R]


class Circle(Stencil):
    """Euclidean distance < R + 0.5, centre included (src/stencils/shapes.jl:57-67)."""

    kind = SHAPE_CIRCLE

    def __init__(self, radius: int = 1, ndims: int = 2):
        super().__init__(radius, ndims)

    def _offsets(self):
        return [I for I in _cartesian(self.R, self.N)
                if math.sqrt(sum(v * v for v in I)) < self.R + 0.5]


class Cross(Stencil):
    """Offsets with >= N-1 zero components (src/stencils/shapes.jl:1-10)."""

    kind = SHAPE_CROSS

    def __init__(self, radius: int = 1, ndims: int = 2):
        super().__init__(radius, ndims)

    def _offsets(self):
        return [I for I in _cartesian(self.R, self.N)
                if sum(1 for v in I if v == 0) >= self.N - 1]


class AngledCross(Stencil):
    """All diagonals (src/stencils/shapes.jl:12-25)."""

    kind = SHAPE_ANGLEDCROSS

    def __init__(self, radius: int = 1, ndims: int = 2):
        super().__init__(radius, ndims)

    def _offsets(self):
        return [I for I in _cartesian(self.R, self.N)
                if sum(1 for v in I[1:] if abs(v) == abs(I[0])) == self.N - 1]


class ForwardSlash(Stencil):
    """'Forward' diagonal (src/stencils/shapes.jl:27-40)."""

    kind = SHAPE_FORWARDSLASH

    def __init__(self, radius: int = 1, ndims: int = 2):
        super().__init__(radius, ndims)

    def _offsets(self):
        return [I for I in _cartesian(self.R, self.N)
                if sum(1 for v in I[1:] if v == -I[0]) == self.N - 1]


class BackSlash(Stencil):
    """'Backward' diagonal (src/stencils/shapes.jl:42-55)."""

    kind = SHAPE_BACKSLASH

    def __init__(self, radius: int = 1, ndims: int = 2):
        super().__init__(radius, ndims)

    def _offsets(self):
        return [I for I in _cartesian(self.R, self.N)
                if sum(1 for v in I[1:] if v == I[0]) == self.N - 1]


class Vertical(Stencil):
    """Bar along dim 1 (src/stencils/shapes.jl:69-79)."""

    kind = SHAPE_VERTICAL

    def __init__(self, radius: int = 1, ndims: int = 2):
        super().__init__(radius, ndims)

    def _offsets(self):
        out = []
        for I in _cartesian(self.R, self.N):
            if (len(I) > 1 and I[1] == 0) or (len(I) == 1 and I[0] == 0):
                out.append(I)
        return out


class Horizontal(Stencil):
    """Bar along dim 2 (src/stencils/shapes.jl:81-91)."""

    kind = SHAPE_HORIZONTAL

    def __init__(self, radius: int = 1, ndims: int = 2):
        super().__init__(radius, ndims)

    def _offsets(self):
        return [I for I in _cartesian(self.R, self.N) if len(I) > 1 and I[0] == 0]


class Diamond(Stencil):
    """Manhattan distance 0..R (src/stencils/shapes.jl:93-104)."""

    kind = SHAPE_DIAMOND

    def __init__(self, radius: int = 1, ndims: int = 2):
        super().__init__(radius, ndims)

    def _offsets(self):
        return [I for I in _cartesian(self.R, self.N)
                if 0 <= sum(abs(v) for v in I) <= self.R]


class Annulus(Stencil):
    """RI + 0.5 <= distance < RO + 0.5 (src/stencils/shapes.jl:106-151)."""

    kind = SHAPE_ANNULUS

    def __init__(self, outer_radius: int = 2, inner_radius: int | None = None,
                 ndims: int = 2):
        super().__init__(outer_radius, ndims)
        self.RI = outer_radius - 1 if inner_radius is None else int(inner_radius)

    def _offsets(self):
        out = []
        for I in _cartesian(self.R, self.N):
            d = math.sqrt(sum(v * v for v in I))
            if d < self.R + 0.5 and d >= self.RI + 0.5:
                out.append(I)
        return out


class Cardinal(Stencil):
    """N, S, E, W at distance R (src/stencils/shapes.jl:153-163)."""

    kind = SHAPE_CARDINAL

    def __init__(self, radius: int = 1, ndims: int = 2):
        super().__init__(radius, ndims)

    def _offsets(self):
        out = []
        for I in _cartesian(self.R, self.N):
            a = [abs(v) for v in I]
            if sum(a) == self.R and max(a) == self.R:
                out.append(I)
        return out


class Ordinal(Stencil):
    """Diagonal corners at distance R (src/stencils/shapes.jl:165-175)."""

    kind = SHAPE_ORDINAL

    def __init__(self, radius: int = 1, ndims: int = 2):
        super().__init__(radius, ndims)

    def _offsets(self):
        out = []
        for I in _cartesian(self.R, self.N):
            a = [abs(v) for v in I]
            if sum(a) == self.R * self.N and max(a) == self.R:
                out.append(I)
        return out


def _positional_radius(offs) -> int:
    # Signed maximum, as src/stencils/positional.jl:74-76 computes it.
    return max(max(o) for o in offs)


class Positional(Stencil):
    """User offset list in user order (src/stencils/positional.jl:31-76).

    ``radius()`` reproduces the reference's signed maximum; ``reach()`` is the
    true extent the GPU path uses for its tiles and boundary handling.
    """

    kind = SHAPE_POSITIONAL

    def __init__(self, *offsets):
        if len(offsets) == 1 and isinstance(offsets[0], (tuple, list)) \
                and len(offsets[0]) and isinstance(offsets[0][0], (tuple, list)):
            offsets = tuple(offsets[0])
        offs = tuple(tuple(int(v) for v in o) for o in offsets)
        if not offs:
            raise ValueError("Positional needs at least one offset")
        N = len(offs[0])
        if any(len(o) != N for o in offs):
            raise ValueError(f"All offsets must be the length `N` of {N}, got {offs}")
        super().__init__(_positional_radius(offs), N)
        self._user = offs

    def _offsets(self):
        return list(self._user)


class NamedStencil(Positional):
    """Named offsets (src/stencils/named.jl:43-120); order = user order."""

    kind = SHAPE_NAMED

    def __init__(self, *args, names=None, **kw):
        if args and isinstance(args[0], Stencil):
            offs = args[0].offsets()
            if names is None or len(names) != len(offs):
                raise ValueError("Length of keys must match length of offsets and L parameter")
            names = tuple(names)
        elif kw:
            names = tuple(kw.keys())
            offs = tuple(kw.values())
        else:
            raise ValueError("NamedStencil needs name=offset keywords or (stencil, names=...)")
        super().__init__(*offs)
        self.names = names

    def name_index(self, name: str) -> int:
        return self.names.index(name)


def named_cardinal(s: Cardinal) -> NamedStencil:
    """NamedStencil(::Cardinal) (src/stencils/named.jl:87)."""
    return NamedStencil(s, names=("E", "S", "N", "W"))


def named_ordinal(s: Ordinal) -> NamedStencil:
    """NamedStencil(::Ordinal) (src/stencils/named.jl:88)."""
    return NamedStencil(s, names=("SE", "NE", "SW", "NW"))


class Rectangle(Stencil):
    """Per-axis offset ranges (src/stencils/rectangle.jl:13-56)."""

    kind = SHAPE_RECTANGLE

    def __init__(self, *ranges):
        if len(ranges) == 1 and isinstance(ranges[0][0], (tuple, list)):
            ranges = tuple(ranges[0])
        ranges = tuple((int(a), int(b)) for a, b in ranges)
        R = max(max(abs(a), abs(b)) for a, b in ranges)
        super().__init__(R, len(ranges))
        self.ranges = ranges

    def _offsets(self):
        axes = [range(a, b + 1) for a, b in self.ranges]
        return [tuple(reversed(t)) for t in itertools.product(*reversed(axes))]


class Kernel(Stencil):
    """A stencil plus one weight per offset (src/stencils/kernel.jl:93-123).

    ``Kernel(weights_matrix)`` builds ``Window{size(A,1) div 2, ndims(A)}``
    and takes weights in column-major linear order (kernel.jl:111), so
    ``W[a, b]`` multiplies the neighbour at offset ``(a - R, b - R)``
    (0-based a, b): correlation, not convolution.
    """

    def __init__(self, stencil_or_weights, weights=None):
        if weights is None:
            W = np.asarray(stencil_or_weights)
            st = Window(W.shape[0] // 2, W.ndim)
            w = W.reshape(-1, order="F")
        elif callable(stencil_or_weights) and isinstance(weights, Stencil):
            st = weights
            w = np.array([stencil_or_weights(d) for d in st.distances()])
        else:
            st = stencil_or_weights
            w = np.asarray(weights)
            w = w.reshape(-1, order="F") if w.ndim > 1 else w
        if len(st) != w.size:
            raise ValueError(f"Stencil length {len(st)} does not match kernel length {w.size}")
        super().__init__(st.R, st.N)
        self.stencil = st
        self.weights = w
        self.kind = st.kind

    def _offsets(self):
        return list(self.stencil.offsets())

    def gpu_supported(self) -> bool:
        return self.stencil.gpu_supported()

    def __repr__(self):
        return f"Kernel({self.stencil!r})"


def offsets(s: Stencil):
    return s.offsets()


def radius(s) -> int:
    return s.radius()


def diameter(s) -> int:
    return s.diameter()


def indices(s, I):
    return s.indices(I)


def distances(s):
    return s.distances()
