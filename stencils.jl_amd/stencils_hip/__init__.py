"""stencils_hip — MI355X-native `mapstencil` for Stencils.jl users.

Host mirror of the Stencils.jl v0.3.6 API over libstencils_hip.so (C ABI in
include/stencils_hip.h).  Arrays are indexed as in Julia (dim 1 contiguous)
with 0-based indices.
"""
from . import _abi
from ._lib import BackendUnavailable, StencilsHipError, load_library
from .array import (StencilArray, SwitchingStencilArray, colmajor_empty, dest,
                    parent, source, switch)
from .boundary import Conditional, Halo, Reflect, Remove, Use, Wrap, padval
from .gather import (BatchPlan, autotune, gatherstencil, gatherstencil_, mapstencil, mapstencil_,
                     mapstencil_batch, mapstencil_steps, mapstencil_steps_batch, result_dtype)
from .scatter import Scatter, scatterstencil_
from .server import Completion, Server, serve_latency_summary
from .ops import (Laplacian, LifeRule, Scaled, Terms, Unsupported, center, integer_fold_wraps,
                  kernelproduct, maximum, mean, minimum, set_integer_fold_wraps, sum)
from .stencil import (AngledCross, Annulus, BackSlash, Cardinal, Circle, Cross,
                      Diamond, ForwardSlash, Horizontal, Kernel, Moore, NamedStencil,
                      Ordinal, Positional, Rectangle, Stencil, Vertical, VonNeumann,
                      Window, diameter, distances, indices, named_cardinal,
                      named_ordinal, offsets, radius)

__all__ = [n for n in dir() if not n.startswith("_")]
