"""
    PartialFSHTHIPExt

MI355X (gfx950) leaves for `PartialFSHT.Plan`, loaded by `using AMDGPU`. A plan whose node arrays
are `ROCVector`s builds a `HIPStretchLeaf` (or, when the planner picks Direct, a `HIPDirectLeaf`):
every apply runs on the device through `libpfsht_hip.so`'s C ABI (`include/pfsht.h`), with
coefficients and fields left in HBM.

    using PartialFSHT, AMDGPU
    plan = PartialFSHT.Plan(ROCArray(θ), ROCArray(φ); lmax = 2048, tol = 1e-8,
                            method = PartialFSHT.StretchAlgorithm(),
                            nufft = PartialFSHT.DirectSumBackend())   # any resolvable backend
    mul!(f, plan, alm)            # f, alm :: ROCArray{ComplexF64}
    mul!(alm, plan', f)

Float32 nodes make a Float32 plan (`T = float(eltype(θ))`, `src/Plan.jl:104`) whose applies take and
return `ROCArray{ComplexF32}`; at `tol >= 1e-6` it keeps the FP64 recurrence with FP32 grids. The
leaf keywords `oversample_θ`, `oversample_φ`, `εapo_max` and `maxboxes` reach the library
(`src/Plan.jl:120-135`).

The NUFFT lives inside the library, so the `nufft` keyword only has to resolve
(`src/NUFFT.jl:131-142`); `DirectSumBackend()` always does. Dispatch is on the node array type,
through the leaf seam `_build_leaf` / `_synthesize!` / `_adjoint_apply!` (`src/Plan.jl:118-135,
181-188`), the same seam `PartialFSHTNUFSHTExt` uses to add a leaf: nothing in the package is
overwritten and no `method=` tag is added.
"""
module PartialFSHTHIPExt

using PartialFSHT: PartialFSHT
using AMDGPU: AMDGPU, ROCArray, ROCVector
using LinearAlgebra: LinearAlgebra

const libpfsht = get(ENV, "PFSHT_HIP_LIB",
                     normpath(joinpath(@__DIR__, "..", "..", "lib", "libpfsht_hip.so")))

const PFSHT_OK = Cint(0)
const PFSHT_E_DIM = Cint(1)
const PFSHT_E_ARG = Cint(2)
const PFSHT_E_GEOM_REFUSED = Cint(3)
const PFSHT_METHOD_DIRECT = Cint(1)
const PFSHT_METHOD_STRETCH = Cint(2)
const PFSHT_METHOD_MULTIBOX = Cint(3)
const PFSHT_PREC_FP64 = Cint(0)
const PFSHT_PREC_MIXED = Cint(1)
const PFSHT_WHERE_DEVICE = Cint(1)

struct PfshtPlanOpts          # pfsht_plan_opts, include/pfsht.h
    method::Int32
    precision::Int32
    spin::Int32
    maxboxes::Int32
    oversample_theta::Float64
    oversample_phi::Float64
    epsapo_max::Float64
    packed::Int32
    tune::Int32
    grid_fft::Int32           # PFSHT_GRID_FFT_* (0: auto)
end

struct PfshtInfo              # pfsht_info, include/pfsht.h
    method::Int32
    kind::Int32
    flip::Int32
    lmax::Int32
    ntrans::Int32
    precision::Int32
    nth_half::Int32
    nth::Int32
    nph::Int32
    mmax::Int32
    es_w::Int32
    es_n1::Int32
    es_n2::Int32
    npts::Int64
    ring_work::Int64
    device_bytes::Int64
    tol::Float64
    es_beta::Float64
    es_rows::Int32
    nboxes::Int32
    spin::Int32
    seg_blocks::Int32
    adj_tiles::Int32
    mfma_tiles::Int32
    grid_fft::Int32
end

_last_error() = unsafe_string(ccall((:pfsht_last_error, libpfsht), Cstring, ()))

# Status codes map onto the exceptions the host path throws (src/Layout.jl:35-56;
# src/Plan.jl:98-102; src/Windows.jl:66-74).
function _check(rc::Cint)
    rc == PFSHT_OK && return nothing
    msg = _last_error()
    rc == PFSHT_E_DIM && throw(DimensionMismatch(msg))
    (rc == PFSHT_E_ARG || rc == PFSHT_E_GEOM_REFUSED) && throw(ArgumentError(msg))
    error("libpfsht_hip: $msg (code $rc)")
end

"""
    HIPStretchLeaf, HIPDirectLeaf

Handles of device plans. A stretched plan holds its geometry, ring tables, bin-sorted points, NUFFT
kernel tables, rocFFT plans and work buffers in HBM (the device analogue of `StretchLeaf`,
`src/Stretch.jl:94-127`). A Direct plan holds the colatitude-sorted points and the ring engine at
ε = 0 (`DirectLeaf`, `src/Direct.jl:15-51`). Both are freed by a finalizer.
"""
abstract type AbstractHIPLeaf end

mutable struct HIPStretchLeaf <: AbstractHIPLeaf
    handle::Ptr{Cvoid}
    info::PfshtInfo
end

mutable struct HIPDirectLeaf <: AbstractHIPLeaf
    handle::Ptr{Cvoid}
    info::PfshtInfo
end

mutable struct HIPMultiBoxLeaf <: AbstractHIPLeaf
    handle::Ptr{Cvoid}
    info::PfshtInfo
end

# pfsht_plan_create_opts: spin = 0 is the scalar plan; a nonzero spin builds SpinDirectLeaf /
# StretchLeaf over SpinRings / per-box spin plans (src/Plan.jl:118-133), coefficients in the
# spin_coeff_index layout (src/Layout.jl:28-29). `opts` is the reference's leaf keyword tuple
# (src/Plan.jl:108-110); a Float32 plan at tol >= 1e-6 keeps FP32 grids over the FP64 recurrence.
function _create(::Type{LT}, method::Cint, θ::AbstractVector, φ::AbstractVector, lmax::Int,
                 ntrans::Int, tol::Real, spin::Integer = 0, opts = nothing) where {LT<:AbstractHIPLeaf}
    θh = Vector{Float64}(PartialFSHT._host(θ))
    φh = Vector{Float64}(PartialFSHT._host(φ))
    length(θh) == length(φh) || throw(DimensionMismatch("θ and φ must have the same length"))
    T = float(promote_type(eltype(θ), eltype(φ)))
    prec = (T === Float32 && tol >= 1e-6) ? PFSHT_PREC_MIXED : PFSHT_PREC_FP64
    get_opt(k, d) = (opts !== nothing && hasproperty(opts, k)) ? getproperty(opts, k) : d
    o = PfshtPlanOpts(method, prec, spin, Int32(get_opt(:maxboxes, 8)),
                      Float64(get_opt(:oversample_θ, 1)), Float64(get_opt(:oversample_φ, 1)),
                      Float64(get_opt(:εapo_max, 1.0)), Int32(0), Int32(0), Int32(0))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    _check(ccall((:pfsht_plan_create_opts, libpfsht), Cint,
                 (Ptr{Float64}, Ptr{Float64}, Int64, Int32, Float64, Int32, Ref{PfshtPlanOpts},
                  Ref{Ptr{Cvoid}}),
                 θh, φh, length(θh), lmax, Float64(tol), ntrans, o, h))
    info = Ref{PfshtInfo}()
    _check(ccall((:pfsht_plan_info, libpfsht), Cint, (Ptr{Cvoid}, Ref{PfshtInfo}), h[], info))
    leaf = LT(h[], info[])
    finalizer(leaf) do l
        l.handle == C_NULL || ccall((:pfsht_plan_destroy, libpfsht), Cvoid, (Ptr{Cvoid},), l.handle)
        l.handle = C_NULL
    end
    return leaf
end

HIPStretchLeaf(θ, φ, lmax::Int, ntrans::Int, tol::Real, spin::Integer = 0, opts = nothing) =
    _create(HIPStretchLeaf, PFSHT_METHOD_STRETCH, θ, φ, lmax, ntrans, tol, spin, opts)
HIPDirectLeaf(θ, φ, lmax::Int, ntrans::Int, spin::Integer = 0) =
    _create(HIPDirectLeaf, PFSHT_METHOD_DIRECT, θ, φ, lmax, ntrans, 1e-10, spin)
HIPMultiBoxLeaf(θ, φ, lmax::Int, ntrans::Int, tol::Real, spin::Integer = 0, opts = nothing) =
    _create(HIPMultiBoxLeaf, PFSHT_METHOD_MULTIBOX, θ, φ, lmax, ntrans, tol, spin, opts)

PartialFSHT.algorithm(::HIPStretchLeaf) = PartialFSHT.StretchAlgorithm()
PartialFSHT.algorithm(::HIPDirectLeaf) = PartialFSHT.DirectAlgorithm()
PartialFSHT.algorithm(::HIPMultiBoxLeaf) = PartialFSHT.MultiBoxAlgorithm()
PartialFSHT.ring_work(leaf::AbstractHIPLeaf) = Int(leaf.info.ring_work)

# Leaf seam (src/Plan.jl:118-129): a plan on ROCVector nodes builds the device leaf of the chosen
# algorithm.
PartialFSHT._build_leaf(::PartialFSHT.StretchAlgorithm, θ::ROCVector, φ::ROCVector, lmax, ntrans,
                        spin, tol, nufft_b, opts) =
    HIPStretchLeaf(θ, φ, lmax, ntrans, tol, spin, opts)
PartialFSHT._build_leaf(::PartialFSHT.DirectAlgorithm, θ::ROCVector, φ::ROCVector, lmax, ntrans,
                        spin, tol, nufft_b, opts) =
    HIPDirectLeaf(θ, φ, lmax, ntrans, spin)
# src/Plan.jl:131-132: the library repeats partition_boxes on the host copy of the nodes
PartialFSHT._build_leaf(::PartialFSHT.MultiBoxAlgorithm, θ::ROCVector, φ::ROCVector, lmax, ntrans,
                        spin, tol, nufft_b, opts) =
    HIPMultiBoxLeaf(θ, φ, lmax, ntrans, tol, spin, opts)

_stream() = AMDGPU.stream().stream   # hipStream_t of the task-local AMDGPU stream

function _device_ptr(x::ROCArray{<:Union{ComplexF64, ComplexF32}})
    return reinterpret(Ptr{Cvoid}, pointer(x))
end

# src/Plan.jl:181-188 → synthesize_stretch!/adjoint_stretch! (src/Stretch.jl:262-295) or
# synthesize_direct!/adjoint_direct! (src/Direct.jl:60-117)
function PartialFSHT._synthesize!(f::ROCArray{ComplexF64}, alm::ROCArray{ComplexF64},
                                  leaf::AbstractHIPLeaf, ::PartialFSHT.Plan)
    _check(ccall((:pfsht_synth, libpfsht), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ptr{Cvoid}),
                 leaf.handle, _device_ptr(alm), _device_ptr(f), PFSHT_WHERE_DEVICE, _stream()))
    return f
end

function PartialFSHT._adjoint_apply!(alm::ROCArray{ComplexF64}, f::ROCArray{ComplexF64},
                                     leaf::AbstractHIPLeaf, ::PartialFSHT.Plan)
    _check(ccall((:pfsht_adjoint, libpfsht), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ptr{Cvoid}),
                 leaf.handle, _device_ptr(f), _device_ptr(alm), PFSHT_WHERE_DEVICE, _stream()))
    return alm
end

# Float32 plans (gpu/runtests.jl:122-139): ComplexF32 coefficients and fields in HBM
function PartialFSHT._synthesize!(f::ROCArray{ComplexF32}, alm::ROCArray{ComplexF32},
                                  leaf::AbstractHIPLeaf, ::PartialFSHT.Plan)
    _check(ccall((:pfsht_synth_f32, libpfsht), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ptr{Cvoid}),
                 leaf.handle, _device_ptr(alm), _device_ptr(f), PFSHT_WHERE_DEVICE, _stream()))
    return f
end

function PartialFSHT._adjoint_apply!(alm::ROCArray{ComplexF32}, f::ROCArray{ComplexF32},
                                     leaf::AbstractHIPLeaf, ::PartialFSHT.Plan)
    _check(ccall((:pfsht_adjoint_f32, libpfsht), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ptr{Cvoid}),
                 leaf.handle, _device_ptr(f), _device_ptr(alm), PFSHT_WHERE_DEVICE, _stream()))
    return alm
end

# ---------------------------------------------------------------------------------------------
# Analysis on a device plan (src/Analysis.jl): the iterative callers keep every vector in HBM.
# ---------------------------------------------------------------------------------------------
const HIPPlan = PartialFSHT.Plan{<:Any, <:AbstractHIPLeaf}
const _SOLVE_MODE = Dict(:cg => Int32(0), :wiener => Int32(1), :minnorm => Int32(2))

struct PfshtSolveInfo         # pfsht_solve_info, include/pfsht.h
    shannon::Float64
    dl::Float64
    relres::Float64
    iters::Int32
    mode::Int32
end

_weights_dev(w::Nothing) = nothing
_weights_dev(w::ROCArray{Float64}) = w
_weights_dev(w::AbstractVector) = ROCArray{Float64}(Vector{Float64}(w))
_wptr(w::Nothing) = Ptr{Float64}(C_NULL)
_wptr(w::ROCArray{Float64}) = reinterpret(Ptr{Float64}, pointer(w))

# apply_resolution!(out, plan, alm, w, ws) — src/Analysis.jl:96-104
function PartialFSHT.apply_resolution!(out::ROCArray{ComplexF64}, plan::HIPPlan, alm::ROCArray{ComplexF64},
                                       w::AbstractVector, ws = nothing)
    wd = _weights_dev(w)
    GC.@preserve wd _check(ccall((:pfsht_apply_resolution, libpfsht), Cint,
                                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Cvoid}, Int32, Ptr{Cvoid}),
                                 plan.leaf.handle, _device_ptr(alm), _wptr(wd), _device_ptr(out),
                                 PFSHT_WHERE_DEVICE, _stream()))
    return out
end

# pseudo_spectrum!(C, plan, f, w, ws; batch) — src/Analysis.jl:286-308
function PartialFSHT.pseudo_spectrum!(C::AbstractVector, plan::HIPPlan, f::ROCArray{ComplexF64},
                                      w::AbstractVector, ws = nothing; batch::Integer = 1)
    length(C) == plan.lmax + 1 ||
        throw(DimensionMismatch("C has $(length(C)) entries, expected $(plan.lmax + 1)"))
    wd = _weights_dev(w)
    Cd = ROCArray{Float64}(undef, plan.lmax + 1)
    GC.@preserve wd _check(ccall((:pfsht_pseudo_spectrum, libpfsht), Cint,
                                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Int32, Ptr{Float64}, Int32, Ptr{Cvoid}),
                                 plan.leaf.handle, _device_ptr(f), _wptr(wd), batch,
                                 reinterpret(Ptr{Float64}, pointer(Cd)), PFSHT_WHERE_DEVICE, _stream()))
    copyto!(C, Array(Cd))
    return C
end

# solve!(alm, plan, f; mode, rtol, maxiter, prior, w, ws) — src/Analysis.jl:375-436; info carries
# shannon, Δℓ, mode, iters and relres as the reference's does
function PartialFSHT.solve!(alm::ROCArray{ComplexF64}, plan::HIPPlan, f::ROCArray{ComplexF64};
                            mode::Symbol = :cg, rtol::Real = 1e-6, maxiter::Integer = 500,
                            prior = nothing, w = nothing, ws = nothing)
    haskey(_SOLVE_MODE, mode) ||
        throw(ArgumentError("mode must be :cg, :wiener or :minnorm, got :$mode"))
    mode === :wiener && prior === nothing &&
        throw(ArgumentError(":wiener needs a prior C_ℓ — that prior is what makes the " *
                            "problem well posed, so it cannot be defaulted."))
    wd = _weights_dev(w)
    pr = prior === nothing ? nothing : Vector{Float64}(prior)
    pp = pr === nothing ? Ptr{Float64}(C_NULL) : pointer(pr)
    info = Ref{PfshtSolveInfo}()
    GC.@preserve wd pr _check(ccall((:pfsht_solve, libpfsht), Cint,
                                    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Int32, Float64, Int32, Ptr{Float64},
                                     Ptr{Cvoid}, Int32, Ptr{Cvoid}, Ref{PfshtSolveInfo}),
                                    plan.leaf.handle, _device_ptr(f), _wptr(wd), _SOLVE_MODE[mode], rtol,
                                    maxiter, pp, _device_ptr(alm), PFSHT_WHERE_DEVICE, _stream(), info))
    i = info[]
    return alm, (shannon = i.shannon, Δℓ = i.dl, mode = mode, iters = Int(i.iters), relres = i.relres)
end

# A host array handed to a device leaf is refused, as the reference refuses the converse
# (ext/PartialFSHTKernelAbstractionsExt.jl:266-271).
PartialFSHT._synthesize!(f, alm, ::AbstractHIPLeaf, ::PartialFSHT.Plan) = throw(ArgumentError(
    "a HIP plan applies to ROCArray{ComplexF64} or ROCArray{ComplexF32} coefficients and fields " *
    "of one element type"))
PartialFSHT._adjoint_apply!(alm, f, ::AbstractHIPLeaf, ::PartialFSHT.Plan) = throw(ArgumentError(
    "a HIP plan applies to ROCArray{ComplexF64} or ROCArray{ComplexF32} coefficients and fields " *
    "of one element type"))

end
