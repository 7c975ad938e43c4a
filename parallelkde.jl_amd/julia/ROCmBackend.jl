# :rocm device for ParallelKDE.jl v1.0.1, backed by libparallelkde_hip.
#
# This file is the reference-side binding: it is `include`d from
# src/ParallelKDE.jl after the existing submodules (it extends their generic
# functions) and adds one ccall per estimate.  It cannot be exercised in the
# build container (no Julia toolchain); the same C ABI is exercised from
# Python through ctypes by the repository's tests.
#
# Hooks used (ParallelKDE.jl v1.0.1):
#   Devices.jl:54-82       AVAILABLE_DEVICES / DEVICE_IMPLEMENTATIONS / DEFAULT_IMPLEMENTATIONS
#   Devices.jl:91-101      get_device
#   Grids.jl:38-50         initialize_grid(::AbstractDevice, ranges)
#   KDEs.jl:62-111         initialize_kde(::AbstractDevice, data, dims)
#   DensityEstimators.jl:58-115  estimate!(::Symbol) -> initialize_estimator / estimate!
#   GradeProEstimator.jl:695     the :gradepro registry entry (AbstractGradeProEstimator)
module ROCmBackend

using ..Devices
using ..Grids
using ..KDEs
using ..DensityEstimators

const libpkde = get(ENV, "PKDE_LIB", "libparallelkde_hip")

# ------------------------------------------------------------------ device
struct IsROCm <: Devices.AbstractDevice end
const GPU_ROCM = :rocm

function __init__()
    Devices.AVAILABLE_DEVICES[:rocm] = IsROCm()
    Devices.DEVICE_IMPLEMENTATIONS[IsROCm()] = Set([GPU_ROCM])
    Devices.DEFAULT_IMPLEMENTATIONS[IsROCm()] = GPU_ROCM
    return nothing
end

Devices.get_device(::IsROCm) = IsROCm()

const PKDE_OK = Cint(0)
const PKDE_NO_DEVICE = Cint(3)

function last_error()
    unsafe_string(ccall((:pkde_last_error, libpkde), Cstring, ()))
end

# status codes -> the reference's exceptions (DirectSpace.jl:489-492,
# GradeProEstimator.jl:1105-1108)
function check(rc::Cint)
    rc == PKDE_OK && return nothing
    rc == PKDE_NO_DEVICE && throw(ErrorException("No functional ROCm device: " * last_error()))
    rc in (Cint(1), Cint(2), Cint(4), Cint(6)) && throw(ArgumentError(last_error()))
    throw(ErrorException("libparallelkde_hip error $rc: " * last_error()))
end

const _ctx = Ref{Ptr{Cvoid}}(C_NULL)
function context()
    if _ctx[] == C_NULL
        out = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:pkde_context_create, libpkde), Cint, (Cint, Ref{Ptr{Cvoid}}), 0, out))
        _ctx[] = out[]
    end
    return _ctx[]
end

# true when the library loads and a device context can be created; the
# device-typed constructors below fall back to :cpu otherwise, as the
# reference does for :cuda (ParallelKDE.jl:120-128, Grids.jl:41-44,
# KDEs.jl:68-71)
function functional()
    try
        context()
        return true
    catch
        return false
    end
end

function warn_fallback()
    @warn "No functional ROCm device detected. Falling back to ':cpu'."
    return nothing
end

# ------------------------------------------------------------------- grids
# The grid stays on the host: lo, hi, spacing and n per axis are all the
# library needs (Grids.jl:67-87).
struct RocGrid{N,T<:Real,M} <: Grids.AbstractGrid{N,T,M}
    grid::Grids.Grid{N,T,M}
end
Devices.get_device(::RocGrid) = IsROCm()
Base.size(g::RocGrid) = size(g.grid)
Base.ndims(::RocGrid{N}) where {N} = N
Grids.spacings(g::RocGrid) = Grids.spacings(g.grid)
Grids.bounds(g::RocGrid) = Grids.bounds(g.grid)
Grids.get_coordinates(g::RocGrid) = Grids.get_coordinates(g.grid)
Grids.fftgrid(g::RocGrid) = Grids.fftgrid(g.grid)

function Grids.initialize_grid(::IsROCm, ranges; kwargs...)
    if !functional()
        warn_fallback()
        return Grids.initialize_grid(Devices.IsCPU(), ranges)
    end
    return RocGrid(Grids.Grid(ranges))
end

# --------------------------------------------------------------------- KDEs
# FP64 samples on the host (an FP32 copy could not reproduce the FP64 cell
# index of the CPU path), density returned as a host Array{Float64,N}.
struct RocKDE{N,T<:Real,S<:Real} <: KDEs.AbstractKDE{N,T,S}
    data::Matrix{S}
    density::Array{T,N}
end
Devices.get_device(::RocKDE) = IsROCm()
KDEs.get_data(k::RocKDE) = k.data
KDEs.get_density(k::RocKDE) = k.density
KDEs.get_nsamples(k::RocKDE) = size(k.data, 2)
KDEs.set_nan_density!(k::RocKDE) = (fill!(k.density, NaN); nothing)

function KDEs.initialize_kde(::IsROCm, data, dims::NTuple{N}; T=Float64) where {N}
    if !functional()
        warn_fallback()
        return KDEs.initialize_kde(Devices.IsCPU(), data, dims)
    end
    m = data isa AbstractMatrix ? Matrix{Float64}(data) : reduce(hcat, Vector{Float64}.(data))
    size(m, 1) == N || throw(ArgumentError("Data must have $N dimensions"))
    return RocKDE{N,T,Float64}(m, fill(T(NaN), dims))
end

# --------------------------------------------------------------- estimator
# Mirrors pkde_gradepro_params in include/parallelkde_hip.h (field order and
# padding are those of the C struct).
struct PkdeParams
    n_bootstraps::Int32
    n_steps::Int32
    bw_step::Float64
    bw_final::NTuple{6,Float64}
    alpha_s::Float64
    alpha_os::Float64
    eps::Float64
    rng_seed::UInt64
    deterministic_binning::Int32
    precision::Int32
end

# Mutable so that the finalizer is tied to the estimator object itself: the
# plan is destroyed only once nothing references the estimator.  M is the
# rank of the bootstrapped statistics (the grid's N axes plus the bootstrap
# axis), a type parameter of its own as in GradeProEstimator{N,T,M,P,S}
# (GradeProEstimator.jl:673); constructors pass M = N + 1.
mutable struct RocGradeProEstimator{N,M} <: DensityEstimators.AbstractGradeProEstimator{N,Float64,M}
    plan::Ptr{Cvoid}
end

function destroy!(e::RocGradeProEstimator)
    if e.plan != C_NULL
        ccall((:pkde_gradepro_plan_destroy, libpkde), Cvoid, (Ptr{Cvoid},), e.plan)
        e.plan = C_NULL
    end
    return nothing
end

function DensityEstimators.initialize_estimator(
    ::Type{<:DensityEstimators.AbstractGradeProEstimator},
    kde::RocKDE{N};
    method::Symbol,
    grid::Union{RocGrid,Nothing}=nothing,
    n_bootstraps::Integer=100,
    bw_step=nothing,
    n_steps=nothing,
    bw_final=nothing,
    alpha_s=nothing,
    alpha_os=nothing,
    eps=nothing,
    seed::Integer=0x5eed,
    precision::Symbol=:fp64,
    kwargs...,
) where {N}
    grid === nothing && throw(ArgumentError("KDE device IsROCm() does not match Grid device nothing."))
    N <= 6 || throw(ArgumentError("ROCm implementation for dimensions higher than 6 is not supported."))
    g = grid.grid
    n = Int64[length(g.coordinates[i]) for i in 1:N]
    lo = Float64[Grids.low_bounds(g)[i] for i in 1:N]
    hi = Float64[Grids.high_bounds(g)[i] for i in 1:N]
    h = Float64[Grids.spacings(g)[i] for i in 1:N]
    bwf = bw_final === nothing ? ntuple(_ -> NaN, 6) :
          ntuple(i -> Float64(bw_final isa Real ? bw_final : get(bw_final, i, NaN)), 6)
    p = PkdeParams(
        Int32(n_bootstraps),
        Int32(n_steps === nothing ? (bw_step === nothing ? 250 : 0) : n_steps),
        bw_step === nothing ? NaN : Float64(bw_step),
        bwf,
        alpha_s === nothing ? NaN : Float64(alpha_s),
        alpha_os === nothing ? NaN : Float64(alpha_os),
        eps === nothing ? NaN : Float64(eps),
        UInt64(seed), Int32(1), Int32(precision === :fp64 ? 1 : 0),
    )
    plan = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:pkde_gradepro_plan_create, libpkde), Cint,
        (Ptr{Cvoid}, Cint, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ref{PkdeParams}, Ref{Ptr{Cvoid}}),
        context(), N, n, lo, hi, h, p, plan))
    return finalizer(destroy!, RocGradeProEstimator{N,N + 1}(plan[]))
end

# One ccall per estimate (GradeProEstimator.jl:968-1072 on the device).
# autotune=true (no counterpart in the reference): after the estimate, time the
# ladder's launch shapes for this grid on the device (pkde_gradepro_autotune);
# later estimates of the same grid, bootstraps, ladder and precision use the
# fastest.  The density does not change.
function DensityEstimators.estimate!(est::RocGradeProEstimator{N,M}, kde::RocKDE{N}; method::Symbol,
                                     autotune::Bool=false, kwargs...) where {N,M}
    n = size(kde.data, 2)
    chosen = Array{Int16}(undef, size(kde.density))
    est.plan == C_NULL && throw(ArgumentError("estimator plan was destroyed"))
    GC.@preserve est kde chosen begin
        check(ccall((:pkde_gradepro_estimate, libpkde), Cint,
            (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Int32}, Ptr{Float64}, Ptr{Int16}, Ptr{Cvoid}),
            est.plan, kde.data, n, C_NULL, kde.density, chosen, C_NULL))
    end
    if autotune
        check(ccall((:pkde_gradepro_autotune, libpkde), Cint,
            (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}), est.plan, Int32(3), C_NULL, C_NULL))
    end
    return nothing
end

# ------------------------------------------------------------------- :rot
# RoTEstimator.jl:116-265 on the device: one plan per (grid, rule).
mutable struct RocRoTEstimator{N} <: DensityEstimators.AbstractRoTEstimator
    plan::Ptr{Cvoid}
end

function destroy!(e::RocRoTEstimator)
    if e.plan != C_NULL
        ccall((:pkde_gradepro_plan_destroy, libpkde), Cvoid, (Ptr{Cvoid},), e.plan)
        e.plan = C_NULL
    end
    return nothing
end

function DensityEstimators.initialize_estimator(
    ::Type{<:DensityEstimators.AbstractRoTEstimator},
    kde::RocKDE{N};
    method::Symbol,
    grid::Union{RocGrid,Nothing}=nothing,
    rule_of_thumb::Union{Nothing,Symbol}=:scott,
    precision::Symbol=:fp64,
    kwargs...,
) where {N}
    grid === nothing && throw(ArgumentError("KDE device IsROCm() does not match Grid device nothing."))
    rule = rule_of_thumb === :silverman ? Cint(1) :
           rule_of_thumb === :scott ? Cint(0) :
           throw(ArgumentError("Rule of thumb $rule_of_thumb not implemented."))
    g = grid.grid
    n = Int64[length(g.coordinates[i]) for i in 1:N]
    lo = Float64[Grids.low_bounds(g)[i] for i in 1:N]
    hi = Float64[Grids.high_bounds(g)[i] for i in 1:N]
    h = Float64[Grids.spacings(g)[i] for i in 1:N]
    plan = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:pkde_rot_plan_create, libpkde), Cint,
        (Ptr{Cvoid}, Cint, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cint, Cint, Ref{Ptr{Cvoid}}),
        context(), N, n, lo, hi, h, rule, Cint(precision === :fp64 ? 1 : 0), plan))
    return finalizer(destroy!, RocRoTEstimator{N}(plan[]))
end

function DensityEstimators.estimate!(est::RocRoTEstimator{N}, kde::RocKDE{N}; method::Symbol, kwargs...) where {N}
    est.plan == C_NULL && throw(ArgumentError("estimator plan was destroyed"))
    n = size(kde.data, 2)
    GC.@preserve est kde begin
        check(ccall((:pkde_gradepro_estimate, libpkde), Cint,
            (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Int32}, Ptr{Float64}, Ptr{Int16}, Ptr{Cvoid}),
            est.plan, kde.data, n, C_NULL, kde.density, C_NULL, C_NULL))
    end
    return nothing
end

end # module
