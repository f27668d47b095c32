# StructureFunctionsMI355XExt.jl -- backend shim plugging libsfmi355x.so into StructureFunctions.jl.
#
# Adds an execution backend `MI355XBackend` and the four methods of the reference's backend hook
#   _dispatch_execution_backend(backend, shape::PointField{D}, sf, x, u, distance_bins, CT; ...)
#   _dispatch_execution_backend(backend, shape::PointField{D}, sf, x, u, distance_bins, value_bins; ...)
#   _dispatch_execution_backend!(backend, shape::PointField{D}, sums, counts, sf, x, u, distance_bins; ...)
#   _dispatch_execution_backend!(backend, shape::PointField{D}, sums, counts, sf, x, u, distance_bins, value_bins; ...)
# (src/Calculations/dispatch.jl:377-477, 510-594), the same four for fields with auxiliary axes
# (shape::SharedPositionField / VaryingPositionField, batch-leading semantics), and the single-pass hooks
#   _dispatch_single_pass(backend, shape::PointField, x, u, distance_bins; ...)
#   _dispatch_single_pass!(backend, sums, counts, x, u, distance_bins; ...)
# (src/Calculations/serial_single_pass.jl:400-535), and the 2D single-pass hooks
#   _dispatch_single_pass_2d(backend, shape::PointField, x, u, distance_bins, value_bins; ...)
#   _dispatch_single_pass_2d!(backend, sums_3d, counts_3d, x, u, distance_bins, value_bins; ...)
# (serial_single_pass.jl:935-1105).  The public entry points validate shapes and
# count width and apply `_finalize` around these calls (dispatch.jl:64-136, 346-373), exactly as
# for the Serial/Threaded/GPU backends.
#
# MI355XBackend subtypes CB.AbstractExecutionBackend, not CB.AbstractGPUBackend, so the generic
# GPU methods (which reach into `backend.backend` and KernelAbstractions) never match it.
#
# Status: written against the reference's interfaces; there is no Julia toolchain in the build
# environment, so this file is reviewed, not executed.  The identical C ABI is exercised by the
# Python harness and the pytest suite.
module StructureFunctionsMI355XExt

using ComputationalBackends: ComputationalBackends as CB
using Distances: Distances as DI
using StructureFunctions:
    StructureFunctions as SF, Calculations as SFC, StructureFunctionObjects as SFO,
    StructureFunctionTypes as SFT

export MI355XBackend

const libsfmi355x = get(ENV, "SFMI355X_LIB", joinpath(@__DIR__, "..", "libsfmi355x.so"))

"""
    MI355XBackend(; device = 0, precision = :fp64)

Run the O(N^2) pair histogram on one AMD MI355X (gfx950).  `precision = :fp64` is the drop-in
binary64 kernel (counts bit-identical to SerialBackend); `precision = :mixed` runs the pair
arithmetic in binary32 with binary64 re-binning of pairs near bin edges (counts still
bit-identical) and binary64 accumulation; `precision = :mixed_fixed` additionally keeps joint
cells of uniform value bins as 64-bit fixed-point words on the device (each value's extra
absolute error at most 2^-23 of its value bin's width).
"""
struct MI355XBackend <: CB.AbstractExecutionBackend
    device::Int
    precision::Symbol
end
MI355XBackend(; device::Integer = 0, precision::Symbol = :fp64) = MI355XBackend(device, precision)

# ---- C ABI mirror (include/sfmi355x.h) -----------------------------------------------------
# sf_bin_rule: the fields of a LinearBinEdges (BinEdges.jl:94-100), optionally inside InfPaddedBinEdges
struct SfBinRule
    kind::Int32      # 1 = SF_RULE_LINEAR
    padded::Int32
    n::Int64
    first_edge::Float64
    last_edge::Float64
    inv_step::Float64
    step_val::Float64
end

struct SfProblem
    D::Int32
    W::Int32
    N::Int64
    x::Ptr{Float64}
    u::Ptr{Float64}
    geometry::Int32
    radius::Float64
    op::Int32
    nl::Int32
    nt::Int32
    n_dist::Int32
    dist_edges::Ptr{Float64}
    n_val::Int32
    val_edges::Ptr{Float64}
    precision::Int32
    count_bits::Int32
    semantics::Int32   # SF_SEM_POINT (point-field kernels) here; SF_SEM_BATCH for auxiliary axes
    reserved::Int32
    dist_rule::Ptr{SfBinRule}  # C_NULL: the distance edges are searched
    val_rule::Ptr{SfBinRule}   # C_NULL: the value edges are searched
end

const _CONTEXTS = Dict{Int, Ptr{Cvoid}}()

function _context(dev::Int)
    get!(_CONTEXTS, dev) do
        h = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:sf_create, libsfmi355x), Cint, (Cint, Ref{Ptr{Cvoid}}), dev, h)
        rc == 0 || error("sf_create(device=$dev) failed with code $rc")
        h[]
    end
end

function _check(ctx, rc)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:sf_last_error, libsfmi355x), Cstring, (Ptr{Cvoid},), ctx))
    rc == -2 && throw(DimensionMismatch(msg))
    rc in (-1, -3, -5) && throw(ArgumentError(msg))
    error("libsfmi355x: " * msg)
end

# operator codes (SF_OP_*)
_opcode(::SFT.ProjectedStructureFunctionType{NL, NT}) where {NL, NT} = (Int32(0), Int32(NL), Int32(NT))
_opcode(::SFT.SecondOrderStructureFunctionType) = (Int32(1), Int32(0), Int32(0))
_opcode(::SFT.ThirdOrderStructureFunctionType) = (Int32(2), Int32(0), Int32(0))
_opcode(::SFT.TransverseComponentSecondOrderStructureFunctionType) = (Int32(3), Int32(0), Int32(0))
_opcode(::SFT.LongitudinalTransverseComponentThirdOrderStructureFunctionType) = (Int32(4), Int32(0), Int32(0))
_opcode(::SFT.FullVectorStructureFunctionType{NF}) where {NF} = (Int32(5), Int32(NF), Int32(0))
struct _SinglePass end  # SF_OP_SINGLE_PASS: rows S2, L2, T2, S3, L3, L1T2 of a (6, n_bins) output
_opcode(::_SinglePass) = (Int32(6), Int32(0), Int32(0))
_opcode(sf) = throw(ArgumentError("$(typeof(sf)) is not supported by MI355XBackend"))

# geometry codes (SF_GEOM_*), from pair_geometry_for (HelperFunctions.jl:390-393)
_geometry(::DI.Euclidean) = (Int32(0), 1.0)
_geometry(m::DI.Haversine) = (Int32(1), Float64(m.radius))
_geometry(::DI.SphericalAngle) = (Int32(2), 1.0)
_geometry(m) = throw(ArgumentError("no pair geometry is defined for distance_metric=$(typeof(m))"))

# explicit binary64 edge vectors for every bin specification (BinEdges.jl)
_edges(b::SF.LogBinEdges) = SF.physical_edges_vector(b)
_edges(b::SF.InfPaddedBinEdges) = vcat(-Inf, collect(Float64, b.edges), Inf)
_edges(b::AbstractVector) = collect(Float64, b)

# A LinearBinEdges axis (a range is wrapped as one, BinEdges.jl:363) is digitized by its O(1)
# rule (BinEdges.jl:114-131; SquaredLinearPlan's sqrt(r^2) key on the flat path, :506, 516); the
# library applies the same rule to these exact fields.  Other axes are searched (nothing).
_linrule(b::SF.LinearBinEdges{Float64}, padded) =
    Ref(SfBinRule(Int32(1), Int32(padded), Int64(length(b.edges)), b.first_edge, b.last_edge,
                  b.inv_step, b.step_val))
_linrule(b::SF.LinearBinEdges, padded) =
    throw(ArgumentError("MI355XBackend digitizes LinearBinEdges in Float64; got $(eltype(b)) edges"))
_rule(b::SF.LinearBinEdges) = _linrule(b, 0)
_rule(b::AbstractRange) = _rule(SF.LinearBinEdges(b))
_rule(b::SF.InfPaddedBinEdges) = b.edges isa SF.LinearBinEdges ? _linrule(b.edges, 1) : nothing
_rule(b) = nothing
_rptr(::Nothing) = Ptr{SfBinRule}(C_NULL)
_rptr(r::Base.RefValue{SfBinRule}) = Base.unsafe_convert(Ptr{SfBinRule}, r)

# The reference accumulates in OT = promote_type(float(eltype(x)), float(eltype(u)))
# (src/Calculations/batch.jl:177-178) and runs its pair arithmetic in the inputs' type; the
# library computes in binary64 only, so other element types are refused rather than silently
# computed differently.
function _require_float64(x, u)
    eltype(x) === Float64 && eltype(u) === Float64 && return nothing
    throw(ArgumentError("MI355XBackend computes in Float64: x and u must be Float64 arrays " *
                        "(got $(eltype(x)), $(eltype(u)))"))
end

_count_bits(::Type{CT}) where {CT} = sizeof(CT) <= 4 ? Int32(32) : Int32(64)
_prec(b::MI355XBackend) = b.precision === :mixed ? Int32(1) : b.precision === :mixed_fixed ? Int32(2) : Int32(0)

function _run!(b::MI355XBackend, sums::Vector{Float64}, counts::Vector{UInt64}, sf, x, u,
               dbins, vbins, metric, CT)
    _require_float64(x, u)
    ctx = _context(b.device)
    xc = convert(Matrix{Float64}, x)
    uc = convert(Matrix{Float64}, u)
    de = _edges(dbins)
    ve = vbins === nothing ? Float64[] : _edges(vbins)
    dr = _rule(dbins)
    vr = vbins === nothing ? nothing : _rule(vbins)
    op, nl, nt = _opcode(sf)
    geom, radius = _geometry(metric)
    GC.@preserve xc uc de ve dr vr sums counts begin
        p = SfProblem(Int32(size(uc, 1)), Int32(size(xc, 1)), Int64(size(xc, 2)),
                      pointer(xc), pointer(uc), geom, radius, op, nl, nt,
                      Int32(length(de) - 1), pointer(de),
                      Int32(vbins === nothing ? 0 : length(ve) - 1),
                      vbins === nothing ? Ptr{Float64}(C_NULL) : pointer(ve),
                      _prec(b), _count_bits(CT), Int32(0), Int32(0), _rptr(dr), _rptr(vr))
        rc = ccall((:sf_hist, libsfmi355x), Cint,
                   (Ptr{Cvoid}, Ref{SfProblem}, Ptr{Float64}, Ptr{UInt64}),
                   ctx, p, sums, counts)
        _check(ctx, rc)
    end
    return nothing
end

# ---- allocating methods ----------------------------------------------------------------------
function SFC._dispatch_execution_backend(
    b::MI355XBackend, ::SFC.PointField{D}, sf::SFT.AbstractPairwiseStructureFunctionType,
    x, u, distance_bins, count_eltype::Type{CT} = UInt32;
    distance_metric::DI.PreMetric = DI.Euclidean(), kwargs...,
) where {D, CT}
    nb = length(distance_bins) - 1
    sums = zeros(Float64, nb)
    counts = zeros(UInt64, nb)
    _run!(b, sums, counts, sf, x, u, distance_bins, nothing, distance_metric, CT)
    return SFO.StructureFunctionSumsAndCounts(sf, distance_bins, sums, CT.(counts))
end

function SFC._dispatch_execution_backend(
    b::MI355XBackend, ::SFC.PointField{D}, sf::SFT.AbstractPairwiseStructureFunctionType,
    x, u, distance_bins, value_bins::AbstractVector;
    distance_metric::DI.PreMetric = DI.Euclidean(), count_eltype::Type{CT} = UInt32, kwargs...,
) where {D, CT}
    nd, nv = length(distance_bins) - 1, length(value_bins) - 1
    sums = zeros(Float64, nd * nv)
    counts = zeros(UInt64, nd * nv)
    _run!(b, sums, counts, sf, x, u, distance_bins, value_bins, distance_metric, CT)
    return SFO.StructureFunction2DSumsAndCounts(sf, distance_bins, value_bins,
        reshape(sums, nd, nv), reshape(CT.(counts), nd, nv))
end

# ---- mutating methods: accumulate, like calculate_structure_function! -------------------------
function SFC._dispatch_execution_backend!(
    b::MI355XBackend, ::SFC.PointField{D}, sums::AbstractVector, counts::AbstractVector,
    sf::SFT.AbstractPairwiseStructureFunctionType, x, u, distance_bins;
    distance_metric::DI.PreMetric = DI.Euclidean(), kwargs...,
) where {D}
    s = zeros(Float64, length(sums))
    c = zeros(UInt64, length(counts))
    _run!(b, s, c, sf, x, u, distance_bins, nothing, distance_metric, eltype(counts))
    sums .+= s
    counts .+= c
    return nothing
end

function SFC._dispatch_execution_backend!(
    b::MI355XBackend, ::SFC.PointField{D}, sums::AbstractMatrix, counts::AbstractMatrix,
    sf::SFT.AbstractPairwiseStructureFunctionType, x, u, distance_bins, value_bins::AbstractVector;
    distance_metric::DI.PreMetric = DI.Euclidean(), kwargs...,
) where {D}
    s = zeros(Float64, length(sums))
    c = zeros(UInt64, length(counts))
    _run!(b, s, c, sf, x, u, distance_bins, value_bins, distance_metric, eltype(counts))
    sums .+= reshape(s, size(sums))
    counts .+= reshape(c, size(counts))
    return nothing
end

# ---- auxiliary axes: u (D, N, B...) over shared x (W, N) or per-slice x (W, N, B...) -----------
# The reference runs these through its batch-leading kernels (batch.jl:201-217,
# batch_leading.jl:105-330), whose flat pairs bin r = sqrt(dx.dx) against the edges: the library
# follows them with semantics = SF_SEM_BATCH.  Slices share x and the edges; the library runs one
# pair pass per field (its B-strips, which evaluate each pair's geometry once for four fields,
# are an opt-in tuning variant).  Outputs are (n_bins[, n_val], aux...), column-major, like the
# reference's batch results.
function _run_aux!(b::MI355XBackend, sums::Vector{Float64}, counts::Vector{UInt64}, sf, x, u,
                   dbins, vbins, metric, CT)
    _require_float64(x, u)
    ctx = _context(b.device)
    D, N = size(u, 1), size(u, 2)
    B = prod(size(u)[3:end])
    uc = reshape(convert(Array{Float64}, u), D, N, B)
    shared = ndims(x) == 2
    xc = shared ? convert(Matrix{Float64}, x) : reshape(convert(Array{Float64}, x), size(x, 1), N, B)
    W = size(xc, 1)
    de = _edges(dbins)
    ve = vbins === nothing ? Float64[] : _edges(vbins)
    dr = _rule(dbins)
    vr = vbins === nothing ? nothing : _rule(vbins)
    op, nl, nt = _opcode(sf)
    geom, radius = _geometry(metric)
    GC.@preserve xc uc de ve dr vr sums counts begin
        probs = [SfProblem(Int32(D), Int32(W), Int64(N),
                           shared ? pointer(xc) : pointer(xc, 1 + (k - 1) * W * N),
                           pointer(uc, 1 + (k - 1) * D * N), geom, radius, op, nl, nt,
                           Int32(length(de) - 1), pointer(de),
                           Int32(vbins === nothing ? 0 : length(ve) - 1),
                           vbins === nothing ? Ptr{Float64}(C_NULL) : pointer(ve),
                           _prec(b), _count_bits(CT), Int32(1), Int32(0), _rptr(dr), _rptr(vr)) for k in 1:B]
        rc = ccall((:sf_hist_batch, libsfmi355x), Cint,
                   (Ptr{Cvoid}, Int32, Ptr{SfProblem}, Ptr{Float64}, Ptr{UInt64}),
                   ctx, Int32(B), probs, sums, counts)
        _check(ctx, rc)
    end
    return nothing
end

const _AuxShape = Union{SFC.SharedPositionField, SFC.VaryingPositionField}

function SFC._dispatch_execution_backend(
    b::MI355XBackend, ::_AuxShape, sf::SFT.AbstractPairwiseStructureFunctionType,
    x, u, distance_bins, count_eltype::Type{CT} = UInt32;
    distance_metric::DI.PreMetric = DI.Euclidean(), kwargs...,
) where {CT}
    nb = length(distance_bins) - 1
    aux = size(u)[3:end]
    sums = zeros(Float64, nb * prod(aux))
    counts = zeros(UInt64, nb * prod(aux))
    _run_aux!(b, sums, counts, sf, x, u, distance_bins, nothing, distance_metric, CT)
    return SFO.StructureFunctionSumsAndCounts(sf, distance_bins, reshape(sums, nb, aux...),
        CT.(reshape(counts, nb, aux...)))
end

function SFC._dispatch_execution_backend(
    b::MI355XBackend, ::_AuxShape, sf::SFT.AbstractPairwiseStructureFunctionType,
    x, u, distance_bins, value_bins::AbstractVector;
    distance_metric::DI.PreMetric = DI.Euclidean(), count_eltype::Type{CT} = UInt32, kwargs...,
) where {CT}
    nd, nv = length(distance_bins) - 1, length(value_bins) - 1
    aux = size(u)[3:end]
    sums = zeros(Float64, nd * nv * prod(aux))
    counts = zeros(UInt64, nd * nv * prod(aux))
    _run_aux!(b, sums, counts, sf, x, u, distance_bins, value_bins, distance_metric, CT)
    return SFO.StructureFunction2DSumsAndCounts(sf, distance_bins, value_bins,
        reshape(sums, nd, nv, aux...), CT.(reshape(counts, nd, nv, aux...)))
end

function SFC._dispatch_execution_backend!(
    b::MI355XBackend, ::_AuxShape, sums::AbstractArray, counts::AbstractArray,
    sf::SFT.AbstractPairwiseStructureFunctionType, x, u, distance_bins;
    distance_metric::DI.PreMetric = DI.Euclidean(), kwargs...,
)
    s = zeros(Float64, length(sums))
    c = zeros(UInt64, length(counts))
    _run_aux!(b, s, c, sf, x, u, distance_bins, nothing, distance_metric, eltype(counts))
    sums .+= reshape(s, size(sums))
    counts .+= reshape(c, size(counts))
    return nothing
end

function SFC._dispatch_execution_backend!(
    b::MI355XBackend, ::_AuxShape, sums::AbstractArray, counts::AbstractArray,
    sf::SFT.AbstractPairwiseStructureFunctionType, x, u, distance_bins, value_bins::AbstractVector;
    distance_metric::DI.PreMetric = DI.Euclidean(), kwargs...,
)
    s = zeros(Float64, length(sums))
    c = zeros(UInt64, length(counts))
    _run_aux!(b, s, c, sf, x, u, distance_bins, value_bins, distance_metric, eltype(counts))
    sums .+= reshape(s, size(sums))
    counts .+= reshape(c, size(counts))
    return nothing
end

# ---- single pass: six invariants in one pair pass (serial_single_pass.jl:400-535) -------------
# The library accumulates rows 1, 2, 4, 5 and count row 1, then rewrites rows 3, 6 and count rows
# 2..6 over the whole buffer, exactly as _sp1d_derive_rows! (serial_single_pass.jl:208-219).
function SFC._dispatch_single_pass(
    b::MI355XBackend, ::SFC.PointField, x::AbstractMatrix, u::AbstractMatrix, distance_bins::AbstractVector;
    count_eltype::Type{CT} = UInt32, distance_metric::DI.PreMetric = DI.Euclidean(), kwargs...,
) where {CT}
    nb = length(distance_bins) - 1
    sums = zeros(Float64, 6 * nb)
    counts = zeros(UInt64, 6 * nb)
    _run!(b, sums, counts, _SinglePass(), x, u, distance_bins, nothing, distance_metric, CT)
    return (sums = reshape(sums, 6, nb), counts = CT.(reshape(counts, 6, nb)))
end

function SFC._dispatch_single_pass!(
    b::MI355XBackend, sums::AbstractMatrix, counts::AbstractMatrix, x::AbstractMatrix, u::AbstractMatrix,
    distance_bins::AbstractVector; distance_metric::DI.PreMetric = DI.Euclidean(), kwargs...,
)
    size(sums, 1) == 6 && size(counts, 1) == 6 ||
        throw(DimensionMismatch("sums and counts must have shape (6, n_bins)"))
    s = zeros(Float64, length(sums))
    c = zeros(UInt64, length(counts))
    _run!(b, s, c, _SinglePass(), x, u, distance_bins, nothing, distance_metric, eltype(counts))
    s = reshape(s, size(sums))
    c = reshape(c, size(counts))
    @inbounds for bin in axes(sums, 2)
        for r in (1, 2, 4, 5)
            sums[r, bin] += s[r, bin]
        end
        counts[1, bin] += c[1, bin]
        sums[3, bin] = sums[1, bin] - sums[2, bin]
        sums[6, bin] = sums[4, bin] - sums[5, bin]
        for r in 2:6
            counts[r, bin] = counts[1, bin]
        end
    end
    return sums, counts
end

# ---- single pass, 2D form: six joint (distance, value) histograms (serial_single_pass.jl:694-1181)
# value_bins is one vector shared by the invariants or a 6-tuple (SinglePass2DValueBins); only the
# first n_val + 1 edges of each matter (a value above them falls in a bin > n_val, which the
# reference drops, _sp2d_scatter!, :852-864), so the library gets those, back to back.
_sp2d_vector(value_bins, t) = value_bins isa Tuple ? value_bins[t] : value_bins

function _run_sp2d!(b::MI355XBackend, sums::Vector{Float64}, counts::Vector{UInt64}, x, u,
                    dbins, value_bins, n_val::Int, metric, CT)
    _require_float64(x, u)
    ctx = _context(b.device)
    xc = convert(Matrix{Float64}, x)
    uc = convert(Matrix{Float64}, u)
    de = _edges(dbins)
    # the six value axes are searched (the library takes explicit vectors for this form)
    ve = reduce(vcat, (_edges(_sp2d_vector(value_bins, t))[1:(n_val + 1)] for t in 1:6))
    dr = _rule(dbins)
    geom, radius = _geometry(metric)
    GC.@preserve xc uc de ve dr sums counts begin
        p = SfProblem(Int32(size(uc, 1)), Int32(size(xc, 1)), Int64(size(xc, 2)),
                      pointer(xc), pointer(uc), geom, radius, Int32(6), Int32(0), Int32(0),
                      Int32(length(de) - 1), pointer(de), Int32(n_val), pointer(ve),
                      _prec(b), _count_bits(CT), Int32(0), Int32(0), _rptr(dr), _rptr(nothing))
        rc = ccall((:sf_hist, libsfmi355x), Cint,
                   (Ptr{Cvoid}, Ref{SfProblem}, Ptr{Float64}, Ptr{UInt64}),
                   ctx, p, sums, counts)
        _check(ctx, rc)
    end
    return nothing
end

function SFC._dispatch_single_pass_2d(
    b::MI355XBackend, ::SFC.PointField, x::AbstractMatrix, u::AbstractMatrix,
    distance_bins::AbstractVector, value_bins::SFC.SinglePass2DValueBins;
    count_eltype::Type{CT} = UInt32, distance_metric::DI.PreMetric = DI.Euclidean(), kwargs...,
) where {CT}
    nb = length(distance_bins) - 1
    n_val = length(_sp2d_vector(value_bins, 1)) - 1
    SFC._validate_value_bins!(value_bins, n_val)
    sums = zeros(Float64, 6 * nb * n_val)
    counts = zeros(UInt64, 6 * nb * n_val)
    _run_sp2d!(b, sums, counts, x, u, distance_bins, value_bins, n_val, distance_metric, CT)
    return (sums = reshape(sums, 6, nb, n_val), counts = CT.(reshape(counts, 6, nb, n_val)))
end

function SFC._dispatch_single_pass_2d!(
    b::MI355XBackend, sums_3d::AbstractArray, counts_3d::AbstractArray, x::AbstractMatrix,
    u::AbstractMatrix, distance_bins::AbstractVector, value_bins::SFC.SinglePass2DValueBins;
    distance_metric::DI.PreMetric = DI.Euclidean(), kwargs...,
)
    # the shape checks of _accumulate_single_pass_2d! (serial_single_pass.jl:740-747)
    nb = length(distance_bins) - 1
    size(sums_3d, 1) == 6 && size(sums_3d, 2) == nb ||
        throw(DimensionMismatch("sums must have shape (6, n_bins, n_val); got $(size(sums_3d))"))
    size(counts_3d) == size(sums_3d) ||
        throw(DimensionMismatch("counts and sums must have the same shape"))
    n_val = size(sums_3d, 3)
    SFC._validate_value_bins!(value_bins, n_val)
    s = zeros(Float64, length(sums_3d))
    c = zeros(UInt64, length(counts_3d))
    _run_sp2d!(b, s, c, x, u, distance_bins, value_bins, n_val, distance_metric, eltype(counts_3d))
    sums_3d .+= reshape(s, size(sums_3d))
    counts_3d .+= reshape(c, size(counts_3d))
    return sums_3d, counts_3d
end

end # module
