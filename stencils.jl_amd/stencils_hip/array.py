"""StencilArray / SwitchingStencilArray (src/array.jl:441-611, src/padding.jl:102-110).

A parent is either a host numpy array or a device torch tensor, indexed like
the Julia array it mirrors: ``A[i, j]`` (0-based) is Julia ``A[i+1, j+1]`` and
dim 1 is the contiguous one.  Device parents must be column-major in dim 1
(stride 1); ``colmajor_empty`` allocates them.  Halo{:out} parents are views
into a pitched allocation whose interior rows start 16-byte aligned, so
kernels can use aligned vector loads; the parent's logical shape is still
``size + 2R`` as in the reference.
"""
from __future__ import annotations

import numpy as np

from . import _abi
from .boundary import Conditional, Halo, Remove, Use, BoundaryCondition, Padding
from .stencil import Stencil, Window


def _torch():
    import torch
    return torch


def is_device(x) -> bool:
    try:
        torch = _torch()
    except ImportError:  # pragma: no cover
        return False
    return isinstance(x, torch.Tensor)


TORCH_DTYPES = None


def sth_dtype_of(x) -> int:
    if is_device(x):
        torch = _torch()
        global TORCH_DTYPES
        if TORCH_DTYPES is None:
            TORCH_DTYPES = {
                torch.bool: _abi.BOOL, torch.uint8: _abi.U8, torch.int8: _abi.I8,
                torch.uint16: _abi.U16, torch.int16: _abi.I16, torch.uint32: _abi.U32,
                torch.int32: _abi.I32, torch.uint64: _abi.U64, torch.int64: _abi.I64,
                torch.float16: _abi.F16, torch.bfloat16: _abi.BF16,
                torch.float32: _abi.F32, torch.float64: _abi.F64,
            }
        return TORCH_DTYPES[x.dtype]
    return _abi.NUMPY_TO_STH[np.dtype(x.dtype)]


def torch_dtype(code: int):
    torch = _torch()
    sth_dtype_of(torch.empty(0))  # fill the table
    inv = {v: k for k, v in TORCH_DTYPES.items()}
    return inv[code]


def colmajor_empty(shape, dtype, device, pitch=None, lead=0):
    """Column-major device tensor of Julia shape `shape`.

    `pitch` (elements) pads the contiguous dim of the allocation; `lead`
    elements precede the first element (used to align Halo interiors)."""
    torch = _torch()
    shape = tuple(int(s) for s in shape)
    p0 = shape[0] if pitch is None else int(pitch)
    rest = 1
    for s in shape[1:]:
        rest *= s
    total = lead + p0 * rest
    buf = torch.zeros(max(total, 1), dtype=dtype, device=device)
    strides = [1]
    acc = p0
    for s in shape[1:]:
        strides.append(acc)
        acc *= s
    return buf.as_strided(shape, tuple(strides), lead)


def colmajor_copy(x, device):
    """Device copy of a host array (or tensor) in column-major layout."""
    torch = _torch()
    if is_device(x):
        t = colmajor_empty(x.shape, x.dtype, device)
        t.copy_(x)
        return t
    a = np.asarray(x)
    t = colmajor_empty(a.shape, torch_dtype(_abi.NUMPY_TO_STH[a.dtype]), device)
    t.copy_(torch.from_numpy(np.ascontiguousarray(a)))
    return t


def to_numpy(t):
    return t.detach().cpu().numpy().copy(order="F") if is_device(t) else np.asarray(t)


def elem_strides(x):
    if is_device(x):
        return tuple(int(s) for s in x.stride())
    return tuple(s // x.itemsize for s in x.strides)


def data_ptr(x) -> int:
    return int(x.data_ptr()) if is_device(x) else int(x.ctypes.data)


def halo_pitch(n0: int, R: int, esize: int, align: int = 64):
    """Pitch (elements) and lead (elements) that put interior rows of a
    Halo{:out} parent on 16-byte boundaries with a 64-byte-multiple pitch."""
    pe = max(1, align // esize)
    pitch = -(-(n0 + 2 * R) // pe) * pe
    lead = ((-R * esize) % 16) // esize
    return pitch, lead


def pad_array(padding: Padding, bc, st: Stencil, parent):
    """src/padding.jl:102-110."""
    if isinstance(padding, Conditional) or (isinstance(padding, Halo) and padding.mode == "in"):
        return parent
    R = st.radius()
    shape = tuple(s + 2 * R for s in parent.shape)
    if is_device(parent):
        esize = parent.element_size()
        pitch, lead = halo_pitch(shape[0] - 2 * R, R, esize)
        out = colmajor_empty(shape, parent.dtype, parent.device, pitch=pitch, lead=lead)
        inner = tuple(slice(R, R + s) for s in parent.shape)
        out[inner] = parent
        return out
    out = np.zeros(shape, dtype=parent.dtype, order="F")
    out[tuple(slice(R, R + s) for s in parent.shape)] = parent
    return out


def _check_radius(parent, R):
    # src/array.jl:451-453: R must be smaller than every parent axis
    for s in parent.shape:
        if not abs(R) < s:
            raise ValueError(f"stencil radius is larger than array axis {s}")


class AbstractStencilArray:
    stencil: Stencil
    boundary: BoundaryCondition
    padding: Padding

    def radius(self) -> int:
        return self.stencil.radius()

    def halo(self) -> int:
        return self.radius() if isinstance(self.padding, Halo) else 0

    @property
    def ndim(self):
        return self.parent().ndim

    @property
    def shape(self):
        return self.size()

    def size(self):
        h = self.halo()
        return tuple(s - 2 * h for s in self.parent().shape)

    def sth_dtype(self) -> int:
        return sth_dtype_of(self.parent())

    def is_device(self) -> bool:
        return is_device(self.parent())

    def interior(self, buf=None):
        """View of the logical array inside a parent buffer."""
        p = self.parent() if buf is None else buf
        h = self.halo()
        return p[tuple(slice(h, s - h) for s in p.shape)]

    def __array__(self, dtype=None, copy=None):
        a = to_numpy(self.interior())
        return a.astype(dtype) if dtype is not None else a

    def to_numpy(self):
        return np.array(to_numpy(self.interior()), order="F")

    def descriptor(self, buf=None) -> _abi.SthArray:
        p = self.parent() if buf is None else buf
        return _abi.make_array(data_ptr(p), sth_dtype_of(p), self.size(),
                               elem_strides(p), self.halo())


def _defaults(parent, stencil, boundary, padding):
    if stencil is None:
        stencil = Window(1, parent.ndim)
    if boundary is None:
        boundary = Remove(_zero_of(parent))
    if padding is None:
        padding = Conditional()
    return stencil, boundary, padding


def _zero_of(parent):
    if is_device(parent):
        return 0 if not parent.is_floating_point() else 0.0
    return parent.dtype.type(0)


class StencilArray(AbstractStencilArray):
    """StencilArray(A, stencil=Window(1, ndims(A)); boundary=Remove(zero), padding=Conditional())
    (src/array.jl:441-468)."""

    def __init__(self, A, stencil=None, boundary=None, padding=None, *, _padded=False):
        stencil, boundary, padding = _defaults(A, stencil, boundary, padding)
        if not isinstance(stencil, Stencil):
            raise TypeError("stencil must be a Stencil")
        self.stencil = stencil
        self.boundary = boundary
        self.padding = padding
        self._parent = A if _padded else pad_array(padding, boundary, stencil, A)
        _check_radius(self._parent, stencil.radius())

    def parent(self):
        return self._parent

    def __repr__(self):
        return (f"StencilArray(size={self.size()}, stencil={self.stencil!r}, "
                f"boundary={self.boundary!r}, padding={self.padding!r})")


class SwitchingStencilArray(AbstractStencilArray):
    """Two buffers that switch roles on every mapstencil! (src/array.jl:564-611).

    ``mapstencil_`` returns the switched array, as the reference does; the
    caller rebinds (``A = mapstencil_(f, A)``)."""

    def __init__(self, A=None, stencil=None, boundary=None, padding=None, *,
                 _buffers=None):
        if _buffers is not None:
            self.source, self.dest = _buffers
            self.stencil, self.boundary, self.padding = stencil, boundary, padding
        else:
            stencil, boundary, padding = _defaults(A, stencil, boundary, padding)
            self.stencil, self.boundary, self.padding = stencil, boundary, padding
            src = pad_array(padding, boundary, stencil, A)
            dst = pad_array(padding, boundary, stencil, A)
            if dst is src:
                dst = src.clone() if is_device(src) else src.copy(order="K")
                if is_device(src) and tuple(dst.stride()) != tuple(src.stride()):
                    dst = colmajor_copy(src, src.device)
            self.source, self.dest = src, dst
        if tuple(self.source.shape) != tuple(self.dest.shape):
            raise ValueError("source and dest arrays must be the same size, got "
                             f"{tuple(self.source.shape)} and {tuple(self.dest.shape)}")
        _check_radius(self.source, self.stencil.radius())

    def parent(self):
        return self.source

    def switch(self) -> "SwitchingStencilArray":
        # the buffers were checked when this array was made; the switched
        # array takes over the library-call descriptors with the roles swapped
        out = object.__new__(SwitchingStencilArray)
        out.source, out.dest = self.dest, self.source
        out.stencil, out.boundary, out.padding = self.stencil, self.boundary, self.padding
        c = self.__dict__.get("_sth_args")
        if c is not None:
            out._sth_args = c[0](c)
        return out

    def __repr__(self):
        return (f"SwitchingStencilArray(size={self.size()}, stencil={self.stencil!r}, "
                f"boundary={self.boundary!r}, padding={self.padding!r})")


def switch(A: SwitchingStencilArray) -> SwitchingStencilArray:
    return A.switch()


def parent(A):
    return A.parent()


def source(A):
    return A.source if isinstance(A, SwitchingStencilArray) else A.parent()


def dest(A):
    return A.dest if isinstance(A, SwitchingStencilArray) else A.parent()
