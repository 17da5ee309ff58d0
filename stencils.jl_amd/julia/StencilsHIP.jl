"""
    StencilsHIP

`ccall` shim that routes Stencils.jl's `gatherstencil!` (and so `mapstencil!`)
to `libstencils_hip.so` on an MI355X. It applies to `StencilArray`s whose
parent is a `ROCArray` and whose function is one of the fixed reductions:
`sum`, `mean`, `minimum`, `maximum`, `kernelproduct`, [`LifeRule`] and
[`Laplacian`]. Every other call, and every combination the library reports
as `STH_ENOTSUP`, is `invoke`d on the reference method
(src/gatherstencil.jl:89-103). Argument errors come back as `ArgumentError`,
as in the reference.

Load it after Stencils and AMDGPU:

    using Stencils, AMDGPU, Statistics
    include("StencilsHIP.jl"); using .StencilsHIP

The C structs below must match `include/stencils_hip.h` field for field.
"""
module StencilsHIP

using Stencils
using Statistics: mean
using Stencils: AbstractStencilArray, StencilArray, SwitchingStencilArray, Stencil,
    AbstractKernelStencil, Remove, Wrap, Reflect, Use, Halo, Conditional,
    kernelproduct, offsets, radius, padding, boundary, stencil, source, dest, switch,
    neighbors, center, kernel
using AMDGPU: ROCArray
import AMDGPU
using StaticArrays: SVector

export LifeRule, Laplacian, Terms, Scaled, mapstencil_steps!, mapstencil_batch!, mapstencil_steps_batch!, sth_available
export ScatterConstant, ScatterCenter, ScatterScaled, Downhill, DownhillFlow
export Server, submit!, start!, completions!, drain!, close!

const libsth = get(ENV, "STENCILS_HIP_LIB",
                   joinpath(@__DIR__, "..", "lib", "libstencils_hip.so"))

# ---------------------------------------------------------------- status / enums
const STH_OK, STH_EDEVICE, STH_ENOMEM, STH_EINVAL, STH_ENOTSUP = 0, 5, 12, 22, 95
const STH_ASYNC, STH_GRAPH = UInt32(1), UInt32(2)
const STH_OPF_ACC_F32 = UInt32(1)
const STH_OPF_INT_WRAP = UInt32(2)

const DTYPES = Dict{DataType,Int32}(
    Bool => 0, UInt8 => 1, Int8 => 2, UInt16 => 3, Int16 => 4, UInt32 => 5, Int32 => 6,
    UInt64 => 7, Int64 => 8, Float16 => 9, Float32 => 11, Float64 => 12)

_bc(::Remove) = Int32(0)
_bc(::Wrap) = Int32(1)
_bc(::Reflect) = Int32(2)
_bc(::Use) = Int32(3)
_pad(::Conditional) = Int32(0)
_pad(::Halo{:out}) = Int32(1)
_pad(::Halo{:in}) = Int32(2)

# ------------------------------------------------------------------ C structs
struct SthArray
    data::Ptr{Cvoid}
    dtype::Int32
    ndim::Int32
    size::NTuple{3,Int64}
    stride::NTuple{3,Int64}
    halo::Int32
    reserved::Int32
end

struct SthStencil
    kind::Int32
    radius::Int32
    ndim::Int32
    length::Int32
    offsets::Ptr{Int32}
    weights::Ptr{Cvoid}
    weights_dtype::Int32
    reserved::Int32
end

struct SthOp
    kind::Int32
    flags::UInt32
    alpha::Float64
    birth_mask::UInt64
    survive_mask::UInt64
end

struct SthJob
    src::SthArray
    dst::SthArray
    stencil::SthStencil
    op::SthOp
    padval::Ptr{Cvoid}
    boundary::Int32
    padding::Int32
end

struct SthTerm
    src::SthArray
    stencil::SthStencil
    op::SthOp
    padval::Ptr{Cvoid}
    boundary::Int32
    padding::Int32
    coef::Float64
    scaled::Int32
    reserved::Int32
end

struct SthScatterRule
    kind::Int32
    combine::Int32
    coef::Float64
end

struct SthServerConfig
    hbm_budget::Csize_t
    max_active::Int32
    round_steps::Int32
    flags::UInt32
    reserved::Int32
end

struct SthRequest
    state::SthArray
    stencil::SthStencil
    op::SthOp
    padval::Ptr{Cvoid}
    boundary::Int32
    padding::Int32
    steps::Int64
    result::Ptr{Cvoid}
    arrival::Float64
    tag::UInt64
end

struct SthCompletion
    tag::UInt64
    status::Int32
    reserved::Int32
    arrival::Float64
    admitted::Float64
    finished::Float64
    steps::Int64
    cells::Int64
end

# ---------------------------------------------------- canonical rule callables
"""
    LifeRule(; birth=(3,), survive=(2, 3))

Life-like rule over Moore(1): a cell with `s` nonzero neighbours becomes
`one(T)` if it is dead and `s ∈ birth`, or alive and `s ∈ survive`.
"""
struct LifeRule
    birth::UInt64
    survive::UInt64
end
LifeRule(; birth=(3,), survive=(2, 3)) =
    LifeRule(reduce(|, (UInt64(1) << b for b in birth); init=UInt64(0)),
             reduce(|, (UInt64(1) << s for s in survive); init=UInt64(0)))
function (r::LifeRule)(hood)
    s = count(!iszero, neighbors(hood))
    mask = iszero(center(hood)) ? r.birth : r.survive
    T = typeof(center(hood))
    return T((mask >> s) & 1)
end

"""
    Laplacian(α)

Explicit diffusion step `c + α * (sum(neighbours) - L * c)` over a
VonNeumann(1, N) stencil, each operation rounded in the element type.
"""
struct Laplacian
    alpha::Float64
end
function (l::Laplacian)(hood)
    c = center(hood)
    T = typeof(c)
    return c + T(l.alpha) * (sum(neighbors(hood)) - length(hood) * c)
end

"""
    Scaled(coef, g)

The term `coef * g(hood)` of a [`Terms`] function (`coef` of the element
type: a Float64 constant on Float32 data would promote the result).
"""
struct Scaled{C,G}
    coef::C
    g::G
end

"""
    Terms(g_1, g_2, ...; combine=+)

A function of several hoods, `(h_1, h_2, ...) -> g_1(h_1) + g_2(h_2) + ...`
(or `*`), folded left to right: the family the multi-array gather
accelerates (`mapstencil(f, A, B, ...)`, `mapstencil!(f, S, B, ...)`,
src/gatherstencil.jl:84-113).  Each `g_k` is `center`, `sum`, `mean`,
`minimum`, `maximum`, `kernelproduct`, a [`LifeRule`], a [`Laplacian`] or a
[`Scaled`] term; a plain array argument arrives as its value at `I`, so its
term is `center`.  On the CPU (the reference path) it evaluates exactly that.
"""
struct Terms{G<:Tuple,C}
    terms::G
    combine::C
end
Terms(gs...; combine=+) = Terms(gs, combine)
_termval(::typeof(center), x::Number) = x
_termval(g, h) = g(h)
_termval(s::Scaled, h) = s.coef * _termval(s.g, h)
(t::Terms)(hs...) = foldl(t.combine, map(_termval, t.terms, hs))

# --------------------------------------------------------- op descriptors
_op(::typeof(sum), T) = SthOp(0, _accflag(T) | _intflag(T), 0.0, 0, 0)
_op(::typeof(mean), T) = SthOp(1, _accflag(T) | _intflag(T), 0.0, 0, 0)
_op(::typeof(minimum), T) = SthOp(2, _accflag(T), 0.0, 0, 0)
_op(::typeof(maximum), T) = SthOp(3, _accflag(T), 0.0, 0, 0)
_op(::typeof(kernelproduct), T) = SthOp(4, _accflag(T), 0.0, 0, 0)
_op(r::LifeRule, T) = SthOp(5, 0, 0.0, r.birth, r.survive)
_op(l::Laplacian, T) = SthOp(6, _accflag(T), l.alpha, 0, 0)
_op(f, T) = nothing
_accflag(::Type{Float16}) = STH_OPF_ACC_F32
_accflag(::Type) = UInt32(0)

# How `sum` and `mean` fold a Stencil (a StaticVector) of an integer type is
# decided by the installed StaticArrays / Statistics, which the reference
# does not pin.  The library accelerates integer sum / mean where the
# readings differ only when told that the sum wraps in the element type and
# the mean is that sum / L; this checks it once per type on an SVector whose
# sum overflows.  Otherwise the call falls back to the reference method.
const _INTWRAP = Dict{DataType,Bool}()
function _intwraps(::Type{T}) where {T<:Integer}
    get!(_INTWRAP, T) do
        v = SVector{2,T}(typemax(T), one(T))
        w = typemax(T) + one(T)   # wraps in T
        s = sum(v)
        s isa T && s == w && mean(v) == Float64(w) / 2
    end
end
_intflag(::Type{Bool}) = UInt32(0)
_intflag(::Type{T}) where {T<:Integer} = _intwraps(T) ? STH_OPF_INT_WRAP : UInt32(0)
_intflag(::Type) = UInt32(0)

# The task-local AMDGPU stream: library work is ordered with the caller's
# kernels and copies on it (AMDGPU.HIP.HIPStream wraps the hipStream_t).
_stream() = Ptr{Cvoid}(AMDGPU.stream().stream)

const AccelOp = Union{typeof(sum),typeof(mean),typeof(minimum),typeof(maximum),
                      typeof(kernelproduct),LifeRule,Laplacian}

# ------------------------------------------------------------------ context
const _ctx = Ref{Ptr{Cvoid}}(C_NULL)

sth_available() = isfile(libsth)

function _context()
    if _ctx[] == C_NULL
        h = Ref{Ptr{Cvoid}}(C_NULL)
        st = ccall((:sth_ctx_create, libsth), Cint, (Cint, Ref{Ptr{Cvoid}}), 0, h)
        st == STH_OK || _throw(st)
        _ctx[] = h[]
    end
    return _ctx[]
end

_lasterror() = unsafe_string(ccall((:sth_last_error, libsth), Cstring, ()))
function _throw(st)
    msg = _lasterror()
    st == STH_EINVAL && throw(ArgumentError(msg))
    error("libstencils_hip status $st: $msg")
end

# ------------------------------------------------------------ descriptors
function _array(p::ROCArray{T,N}, logical::NTuple{N,Int}, halo::Int) where {T,N}
    sz = ntuple(d -> d <= N ? logical[d] : 1, 3)
    sd = ntuple(d -> d <= N ? stride(p, d) : 0, 3)
    SthArray(Ptr{Cvoid}(UInt(pointer(p))), DTYPES[T], N, sz, sd, halo, 0)
end


function _stencil_desc(st::Stencil{R,N,L}) where {R,N,L}
    base = st isa AbstractKernelStencil ? stencil(st) : st
    offs = Int32[o[d] for o in offsets(typeof(base)) for d in 1:N]
    return offs, Int32(0), Int32(R), Int32(N), Int32(L)
end

_weights(st::AbstractKernelStencil, T) = collect(T, kernel(st))
_weights(::Stencil, T) = nothing

_padval(b::Remove, ::Type{T}) where T = b.padval isa T ? Ref{T}(b.padval) : nothing
_padval(b, ::Type{T}) where T = Ref{T}(zero(T))

# ------------------------------------------------------ the accelerated call
function _map(f, dest_parent::ROCArray, dest_halo::Int, src::StencilArray{R,T}) where {R,T}
    haskey(DTYPES, T) || return STH_ENOTSUP
    op = _op(f, T)
    op === nothing && return STH_ENOTSUP
    pv = _padval(boundary(src), T)
    pv === nothing && return STH_ENOTSUP   # padval of another type promotes (src/gatherstencil.jl:50)
    offs, kind, r, n, l = _stencil_desc(stencil(src))
    w = _weights(stencil(src), T)
    GC.@preserve offs w pv begin
        sd = SthStencil(kind, r, n, l, pointer(offs),
                        w === nothing ? C_NULL : Ptr{Cvoid}(pointer(w)),
                        w === nothing ? Int32(-1) : DTYPES[eltype(w)], 0)
        sa = _array(parent(src), size(src), _halo(src))
        da = _array(dest_parent, size(src), dest_halo)
        return ccall((:sth_map, libsth), Cint,
                     (Ptr{Cvoid}, Ref{SthArray}, Ref{SthArray}, Ref{SthStencil}, Int32,
                      Ptr{Cvoid}, Int32, Ref{SthOp}, UInt32, Ptr{Cvoid}),
                     _context(), sa, da, sd, _bc(boundary(src)),
                     Base.unsafe_convert(Ptr{Cvoid}, pv), _pad(padding(src)), op,
                     UInt32(0), _stream())
    end
end

# dest a plain device array of the logical size
function Stencils.gatherstencil!(f::AccelOp, dest::ROCArray,
                                 source::StencilArray{<:Any,<:Any,<:Any,<:ROCArray})
    st = _map(f, dest, 0, source)
    st == STH_OK && return dest
    st == STH_ENOTSUP || _throw(st)
    return invoke(Stencils.gatherstencil!,
                  Tuple{Any,Any,AbstractStencilArray}, f, dest, source)
end

# dest a StencilArray over a device array (the SwitchingStencilArray step)
function Stencils.gatherstencil!(f::AccelOp,
                                 dest::StencilArray{<:Any,<:Any,<:Any,<:ROCArray},
                                 source::StencilArray{<:Any,<:Any,<:Any,<:ROCArray})
    st = _map(f, parent(dest), _halo(dest), source)
    st == STH_OK && return dest
    st == STH_ENOTSUP || _throw(st)
    return invoke(Stencils.gatherstencil!,
                  Tuple{Any,Any,AbstractStencilArray}, f, dest, source)
end

# ------------------------------------------------------ multi-array gather
_halo(A::AbstractStencilArray) = padding(A) isa Halo ? radius(stencil(A)) : 0
_devparent(A::AbstractStencilArray) = parent(A)
_devparent(A::ROCArray) = A

# One sth_term per argument, or nothing when the call is outside the family.
function _term(g, A, ::Type{T}, keep) where T
    coef, scaled = 0.0, Int32(0)
    if g isa Scaled
        T <: AbstractFloat || return nothing
        g.coef isa T || return nothing   # the reference result would promote
        coef, scaled, g = Float64(g.coef), Int32(1), g.g
    end
    eltype(A) === T || return nothing
    if A isa ROCArray
        g === center || return nothing
        return SthTerm(_array(A, size(A), 0), SthStencil(0, 0, 0, 0, C_NULL, C_NULL, -1, 0),
                       SthOp(0, 0, 0.0, 0, 0), C_NULL, Int32(1), Int32(0), coef, scaled, 0)
    end
    pv = _padval(boundary(A), T)
    pv === nothing && return nothing
    push!(keep, pv)
    sa = _array(_devparent(A), size(A), _halo(A))
    if g === center
        sd = SthStencil(0, 0, 0, 0, C_NULL, C_NULL, -1, 0)
        op = SthOp(0, 0, 0.0, 0, 0)
    else
        op = _op(g, T)
        op === nothing && return nothing
        offs, kind, r, n, l = _stencil_desc(stencil(A))
        w = _weights(stencil(A), T)
        push!(keep, offs, w)
        sd = SthStencil(kind, r, n, l, pointer(offs),
                        w === nothing ? C_NULL : Ptr{Cvoid}(pointer(w)),
                        w === nothing ? Int32(-1) : DTYPES[eltype(w)], 0)
    end
    return SthTerm(sa, sd, op, Base.unsafe_convert(Ptr{Cvoid}, pv), _bc(boundary(A)), _pad(padding(A)),
                   coef, scaled, 0)
end

function _map_multi(f::Terms, dest_parent::ROCArray{T}, dest_halo::Int, size_, args) where T
    haskey(DTYPES, T) || return STH_ENOTSUP
    length(args) == length(f.terms) || return STH_ENOTSUP
    (f.combine === (+) || f.combine === (*)) || return STH_ENOTSUP
    keep = Any[]
    terms = SthTerm[]
    for (g, A) in zip(f.terms, args)
        t = _term(g, A, T, keep)
        t === nothing && return STH_ENOTSUP
        push!(terms, t)
    end
    da = _array(dest_parent, size_, dest_halo)
    return GC.@preserve keep terms ccall((:sth_map_multi, libsth), Cint,
        (Ptr{Cvoid}, Int32, Ptr{SthTerm}, Ref{SthArray}, Int32, UInt32, Ptr{Cvoid}),
        _context(), Int32(length(terms)), terms, da, f.combine === (*) ? Int32(1) : Int32(0),
        UInt32(0), _stream())
end

const DevArg = Union{ROCArray,AbstractStencilArray{<:Any,<:Any,<:Any,<:ROCArray}}

# gatherstencil!(f, dest, source, args...) with device arguments
function Stencils.gatherstencil!(f::Terms, dest::ROCArray,
                                 source::AbstractStencilArray{<:Any,<:Any,<:Any,<:ROCArray},
                                 arg1::DevArg, args::DevArg...)
    st = _map_multi(f, dest, 0, size(source), (source, arg1, args...))
    st == STH_OK && return dest
    st == STH_ENOTSUP || _throw(st)
    return invoke(Stencils.gatherstencil!, Tuple{Any,Any,AbstractStencilArray,Vararg{AbstractArray}},
                  f, dest, source, arg1, args...)
end

# dest a StencilArray over a device array: mapstencil!(f, S::SwitchingStencilArray, B, ...)
function Stencils.gatherstencil!(f::Terms, dest::StencilArray{<:Any,<:Any,<:Any,<:ROCArray},
                                 source::AbstractStencilArray{<:Any,<:Any,<:Any,<:ROCArray},
                                 arg1::DevArg, args::DevArg...)
    st = _map_multi(f, parent(dest), _halo(dest), size(source), (source, arg1, args...))
    st == STH_OK && return dest
    st == STH_ENOTSUP || _throw(st)
    return invoke(Stencils.gatherstencil!, Tuple{Any,Any,AbstractStencilArray,Vararg{AbstractArray}},
                  f, dest, source, arg1, args...)
end

# ------------------------------------------------------------ scatter rules
"""
    ScatterConstant(c), ScatterCenter(), ScatterScaled(c), Downhill(c), DownhillFlow(c)

The `f(hood, args...) -> SVector` functions of `scatterstencil!`
(src/scatterstencil.jl:1-47) that the device path covers: every neighbour
receives `c`, the centre value, the centre value times `c`; and, with a
second StencilArray `dem` of heights, `Δ = center(dem) - n` per neighbour
height `n` and `Δ > 0 ? center(hood) * c : 0` (`Downhill`) or
`Δ > 0 ? center(hood) * Δ * c : 0` (`DownhillFlow`, the reference's water
example).  Their CPU methods are what the reference path runs.
"""
abstract type ScatterRule end
struct ScatterConstant{C} <: ScatterRule
    c::C
end
struct ScatterCenter <: ScatterRule end
struct ScatterScaled{C} <: ScatterRule
    c::C
end
struct Downhill{C} <: ScatterRule
    c::C
end
struct DownhillFlow{C} <: ScatterRule
    c::C
end
(s::ScatterConstant)(hood) = map(_ -> s.c, neighbors(hood))
(s::ScatterCenter)(hood) = map(_ -> center(hood), neighbors(hood))
(s::ScatterScaled)(hood) = map(_ -> center(hood) * s.c, neighbors(hood))
function (s::Downhill)(hood, dem)
    a, z = center(hood), center(dem)
    map(n -> (Δ = z - n; Δ > 0 ? a * s.c : zero(a)), neighbors(dem))
end
function (s::DownhillFlow)(hood, dem)
    a, z = center(hood), center(dem)
    map(n -> (Δ = z - n; Δ > 0 ? a * Δ * s.c : zero(a)), neighbors(dem))
end

_rule_kind(::ScatterConstant) = Int32(0)
_rule_kind(::ScatterCenter) = Int32(1)
_rule_kind(::ScatterScaled) = Int32(2)
_rule_kind(::Downhill) = Int32(3)
_rule_kind(::DownhillFlow) = Int32(4)
_rule_coef(::ScatterCenter, ::Type) = 0.0
_rule_coef(r::ScatterRule, ::Type{T}) where T = r.c isa T ? Float64(r.c) : nothing  # else the values promote
_combine(::typeof(+)) = Int32(0)
_combine(::typeof(max)) = Int32(2)
_combine(::typeof(min)) = Int32(3)
_combine(_) = nothing

# sth_scatter for a device source (and second array); STH_ENOTSUP when the
# call is outside the family
function _scatter(f::ScatterRule, op, dest::ROCArray{T,2}, source, args, flags::UInt32) where T
    T <: Union{Float32,Float64} || return STH_ENOTSUP
    comb = _combine(op)
    coef = _rule_coef(f, T)
    (comb === nothing || coef === nothing) && return STH_ENOTSUP
    needs_aux = f isa Union{Downhill,DownhillFlow}
    length(args) == (needs_aux ? 1 : 0) || return STH_ENOTSUP
    needs_aux && !(args[1] isa AbstractStencilArray{<:Any,<:Any,2,<:ROCArray}) && return STH_ENOTSUP
    keep = Any[]
    ts = _term(sum, source, T, keep)
    ts === nothing && return STH_ENOTSUP
    ta = needs_aux ? _term(sum, args[1], T, keep) : nothing
    needs_aux && ta === nothing && return STH_ENOTSUP
    rule = SthScatterRule(_rule_kind(f), comb, coef)
    da = _array(dest, size(dest), 0)
    return GC.@preserve keep ccall((:sth_scatter, libsth), Cint,
        (Ptr{Cvoid}, Ref{SthTerm}, Ptr{SthTerm}, Ref{SthArray}, Ref{SthScatterRule}, UInt32, Ptr{Cvoid}),
        _context(), ts, ta === nothing ? C_NULL : Ref(ta), da, rule, flags, _stream())
end

const DevSA2 = AbstractStencilArray{<:Any,<:Any,2,<:ROCArray}

function Stencils.scatterstencil!(f::ScatterRule, op::OP, dest::ROCArray{<:Any,2}, source::DevSA2,
                                  args::AbstractArray...) where OP
    st = _scatter(f, op, dest, source, args, UInt32(0))
    st == STH_OK && return dest
    st == STH_ENOTSUP || _throw(st)
    return invoke(Stencils.scatterstencil!, Tuple{Any,Any,AbstractArray,AbstractStencilArray,Vararg{AbstractArray}},
                  f, op, dest, source, args...)
end

# the SwitchingStencilArray forms (src/scatterstencil.jl:117-136): dest is
# cleared in the same call (STH_SCATTER_ZERO_DEST) and the array switched
const STH_SCATTER_ZERO_DEST = UInt32(64)
function Stencils.scatterstencil!(f::ScatterRule, op::OP,
                                  A::SwitchingStencilArray{<:Any,<:Any,2,<:ROCArray},
                                  src::DevSA2, args::AbstractArray...) where OP
    st = _scatter(f, op, dest(A), src, args, STH_SCATTER_ZERO_DEST)
    st == STH_OK && return switch(A)
    st == STH_ENOTSUP || _throw(st)
    return invoke(Stencils.scatterstencil!, Tuple{Any,Any,SwitchingStencilArray,AbstractStencilArray,
                                                  Vararg{AbstractArray}}, f, op, A, src, args...)
end
function Stencils.scatterstencil!(f::ScatterRule, op::OP,
                                  A::SwitchingStencilArray{<:Any,<:Any,2,<:ROCArray},
                                  args::AbstractArray...) where OP
    pd = padding(A) isa Halo ? Halo{:in}() : padding(A)
    src = StencilArray(source(A), stencil(A), boundary(A), pd)
    st = _scatter(f, op, dest(A), src, args, STH_SCATTER_ZERO_DEST)
    st == STH_OK && return switch(A)
    st == STH_ENOTSUP || _throw(st)
    return invoke(Stencils.scatterstencil!, Tuple{Any,Any,SwitchingStencilArray,Vararg{AbstractArray}},
                  f, op, A, args...)
end

"""
    mapstencil_steps!(f, A::SwitchingStencilArray, steps; graph=true)

`steps` calls of `A = mapstencil!(f, A)`, ping-ponged on device and, with
`graph`, replayed from a captured hipGraph. Returns the switched array.
"""
function mapstencil_steps!(f::AccelOp, A::SwitchingStencilArray{R,T,N,<:ROCArray},
                           steps::Integer; graph::Bool=true) where {R,T,N}
    op = _op(f, T)
    pv = _padval(boundary(A), T)
    (op === nothing || pv === nothing) && return _fallback_steps(f, A, steps)
    offs, kind, r, n, l = _stencil_desc(stencil(A))
    w = _weights(stencil(A), T)
    h = padding(A) isa Halo ? R : 0
    lsz = size(source(A)) .- 2h
    cur = Ref{Int32}(0)
    st = GC.@preserve offs w pv begin
        sd = SthStencil(kind, r, n, l, pointer(offs),
                        w === nothing ? C_NULL : Ptr{Cvoid}(pointer(w)),
                        w === nothing ? Int32(-1) : DTYPES[eltype(w)], 0)
        a = _array(source(A), lsz, h)
        b = _array(dest(A), lsz, h)
        ccall((:sth_iterate, libsth), Cint,
              (Ptr{Cvoid}, Ref{SthArray}, Ref{SthArray}, Ref{SthStencil}, Int32, Ptr{Cvoid},
               Int32, Ref{SthOp}, Int64, UInt32, Ref{Int32}, Ptr{Cvoid}),
              _context(), a, b, sd, _bc(boundary(A)), Base.unsafe_convert(Ptr{Cvoid}, pv),
              _pad(padding(A)), op, Int64(steps),
              graph ? STH_GRAPH : UInt32(0), cur, _stream())
    end
    st == STH_ENOTSUP && return _fallback_steps(f, A, steps)
    st == STH_OK || _throw(st)
    return cur[] == 0 ? A : switch(A)
end

function _fallback_steps(f, A, steps)
    for _ in 1:steps
        A = Stencils.gatherstencil!(f, A)
    end
    return A
end

# ------------------------------------------------------------ batched calls
"""
    mapstencil_batch!(calls)

Many `gatherstencil!(f, dest, source)` calls, `calls` a vector of
`(f, dest::ROCArray, source::StencilArray)` over device arrays, in one
library call (`sth_map_batch`): one launch per (element type, shape,
reduction) instead of one per call.  Each `dest` ends exactly as the single
call leaves it.  When any call is outside the accelerated set, every call
runs through `gatherstencil!` one by one.
"""
function mapstencil_batch!(calls::AbstractVector)
    keep = Any[]
    jobs = SthJob[]
    for (f, dest, src) in calls
        T = eltype(src)
        op = haskey(DTYPES, T) ? _op(f, T) : nothing
        pv = _padval(boundary(src), T)
        if op === nothing || pv === nothing
            foreach(c -> Stencils.gatherstencil!(c...), calls)
            return calls
        end
        offs, kind, r, n, l = _stencil_desc(stencil(src))
        w = _weights(stencil(src), T)
        push!(keep, offs, w, pv)
        sd = SthStencil(kind, r, n, l, pointer(offs),
                        w === nothing ? C_NULL : Ptr{Cvoid}(pointer(w)),
                        w === nothing ? Int32(-1) : DTYPES[eltype(w)], 0)
        push!(jobs, SthJob(_array(parent(src), size(src), _halo(src)),
                           _array(dest, size(src), 0), sd, op,
                           Base.unsafe_convert(Ptr{Cvoid}, pv),
                           _bc(boundary(src)), _pad(padding(src))))
    end
    st = GC.@preserve keep jobs ccall((:sth_map_batch, libsth), Cint,
                                      (Ptr{Cvoid}, Int32, Ptr{SthJob}, UInt32, Ptr{Cvoid}),
                                      _context(), Int32(length(jobs)), jobs, UInt32(0), _stream())
    if st == STH_ENOTSUP
        foreach(c -> Stencils.gatherstencil!(c...), calls)
    elseif st != STH_OK
        _throw(st)
    end
    return calls
end

"""
    mapstencil_steps_batch!(items; graph=true)

Many iterations with their own step counts in one call (`sth_iterate_batch`):
`items` is a vector of `(f, A::SwitchingStencilArray, steps)` over device
arrays; every array advances one step per round together with the others
still running, and with `graph` the whole schedule is captured once and
replayed on later calls with the same arrays and counts.  Returns the
switched arrays, each as its own loop `for _ in 1:steps; A = mapstencil!(f,
A); end` (src/gatherstencil.jl:77-83) would; calls outside the accelerated
family run that loop.
"""
function mapstencil_steps_batch!(items::AbstractVector; graph::Bool=true)
    keep = Any[]
    jobs = SthJob[]
    steps = Int64[]
    fallback() = map(items) do (f, A, k)
        for _ in 1:k
            A = mapstencil!(f, A)
        end
        A
    end
    for (f, A, k) in items
        (A isa SwitchingStencilArray && source(A) isa ROCArray && k >= 0) || return fallback()
        T = eltype(A)
        op = haskey(DTYPES, T) ? _op(f, T) : nothing
        pv = _padval(boundary(A), T)
        (op === nothing || pv === nothing) && return fallback()
        offs, kind, r, n, l = _stencil_desc(stencil(A))
        w = _weights(stencil(A), T)
        push!(keep, offs, w, pv)
        sd = SthStencil(kind, r, n, l, pointer(offs),
                        w === nothing ? C_NULL : Ptr{Cvoid}(pointer(w)),
                        w === nothing ? Int32(-1) : DTYPES[eltype(w)], 0)
        push!(jobs, SthJob(_array(source(A), size(A), _halo(A)), _array(dest(A), size(A), _halo(A)),
                           sd, op, Base.unsafe_convert(Ptr{Cvoid}, pv), _bc(boundary(A)), _pad(padding(A))))
        push!(steps, Int64(k))
    end
    cur = zeros(Int32, length(jobs))
    st = GC.@preserve keep jobs ccall((:sth_iterate_batch, libsth), Cint,
        (Ptr{Cvoid}, Int32, Ptr{SthJob}, Ptr{Int64}, UInt32, Ptr{Int32}, Ptr{Cvoid}),
        _context(), Int32(length(jobs)), jobs, steps, graph ? STH_GRAPH : UInt32(0), cur, _stream())
    st == STH_ENOTSUP && return fallback()
    st == STH_OK || _throw(st)
    return [c == 1 ? switch(A) : A for ((_, A, _), c) in zip(items, cur)]
end

# ------------------------------------------------------------ serving loop
"""
    Server(hbm_budget; round_steps=8, serial=false)

Continuous batching of iterated solves (`sth_server_*`): `submit!` queues
`steps` calls of `A = mapstencil!(f, A)` on a SwitchingStencilArray over
device arrays, `start!` starts the scheduler, `completions!` returns the
finished requests (tag, times in server-clock seconds) and `drain!` waits for
all of them.  Each request's final state, halo included, is copied into its
result array.
"""
mutable struct Server
    h::Ptr{Cvoid}
    keep::Dict{UInt64,Any}
    next::UInt64
end

function Server(hbm_budget::Integer; round_steps::Integer=8, serial::Bool=false)
    cfg = Ref(SthServerConfig(Csize_t(hbm_budget), 0, Int32(round_steps), serial ? UInt32(1) : UInt32(0), 0))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    st = ccall((:sth_server_create, libsth), Cint,
               (Ptr{Cvoid}, Ref{SthServerConfig}, Ptr{Cvoid}, Ref{Ptr{Cvoid}}), _context(), cfg, _stream(), h)
    st == STH_OK || _throw(st)
    srv = Server(h[], Dict{UInt64,Any}(), UInt64(1))
    finalizer(close!, srv)
    return srv
end

"""
    submit!(srv, f, A::SwitchingStencilArray, steps; arrival=0.0, result=similar(source(A)))

Queue a request; returns its tag.  `result` receives the final state (the
parent array, halo included).
"""
function submit!(srv::Server, f::AccelOp, A::SwitchingStencilArray{R,T,N,<:ROCArray}, steps::Integer;
                 arrival::Real=0.0, result::ROCArray=similar(source(A))) where {R,T,N}
    op = _op(f, T)
    pv = _padval(boundary(A), T)
    (op === nothing || pv === nothing) && throw(ArgumentError("$f is not served on $T"))
    offs, kind, r, n, l = _stencil_desc(stencil(A))
    w = _weights(stencil(A), T)
    h = padding(A) isa Halo ? R : 0
    lsz = size(source(A)) .- 2h
    tag = srv.next
    srv.next += 1
    st = GC.@preserve offs w pv begin
        sd = SthStencil(kind, r, n, l, pointer(offs),
                        w === nothing ? C_NULL : Ptr{Cvoid}(pointer(w)),
                        w === nothing ? Int32(-1) : DTYPES[eltype(w)], 0)
        q = Ref(SthRequest(_array(source(A), lsz, h), sd, op, Base.unsafe_convert(Ptr{Cvoid}, pv),
                           _bc(boundary(A)), _pad(padding(A)), Int64(steps),
                           Ptr{Cvoid}(UInt(pointer(result))), Float64(arrival), tag))
        ccall((:sth_server_submit, libsth), Cint, (Ptr{Cvoid}, Ref{SthRequest}), srv.h, q)
    end
    st == STH_OK || _throw(st)
    srv.keep[tag] = (A, result)   # the state is copied at admission, the result at the end
    return tag
end

start!(srv::Server) = (st = ccall((:sth_server_start, libsth), Cint, (Ptr{Cvoid},), srv.h);
                       st == STH_OK || _throw(st); srv)

"""
    completions!(srv; wait_ms=0) -> Vector{SthCompletion}

Completions not returned before; waits up to `wait_ms` for the first.
"""
function completions!(srv::Server; wait_ms::Integer=0)
    out = Vector{SthCompletion}(undef, 256)
    n = Ref{Int32}(0)
    st = ccall((:sth_server_poll, libsth), Cint,
               (Ptr{Cvoid}, Ptr{SthCompletion}, Int32, Ref{Int32}, Int32),
               srv.h, out, Int32(length(out)), n, Int32(wait_ms))
    st == STH_OK || _throw(st)
    done = out[1:n[]]
    foreach(c -> delete!(srv.keep, c.tag), done)
    return done
end

function drain!(srv::Server; timeout::Real=600.0)
    st = ccall((:sth_server_drain, libsth), Cint, (Ptr{Cvoid}, Float64), srv.h, Float64(timeout))
    st == STH_OK || _throw(st)
    return srv
end

function close!(srv::Server)
    if srv.h != C_NULL
        ccall((:sth_server_destroy, libsth), Cint, (Ptr{Cvoid},), srv.h)
        srv.h = C_NULL
    end
    return nothing
end

end # module
