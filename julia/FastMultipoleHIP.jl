# FastMultipoleHIP.jl - the ccall binding of libfmmhip (include/fmmhip.h) inside FastMultipole.jl.
#
# Included from src/FastMultipole.jl after the package's own files (it uses Branch, Cache,
# DerivativesSwitch, the error-method and list-method types and the user-system interface of
# src/compatibility.jl). It keeps the package's public calls - fmm! (src/fmm.jl:807-1176), direct!
# (src/direct.jl:20-75), FastGaussSeidel / solve! (src/solve.jl:538-946) - and routes a call to the
# GPU only when the caller passes a `hip` handle and the system declares a built-in kernel through
# `hip_kernel`; every other call keeps the reference's CPU path.
#
# Status codes never unwind through C: `check` turns them into the exceptions the reference throws
# (ArgumentError for FMMHIP_EINVAL, an ErrorException otherwise); FMMHIP_ETOL becomes the
# reference's `error_success = false`.

const libfmmhip = get(ENV, "FMMHIP_LIB", "libfmmhip.so")

#--- status, kernel, precision and method codes (include/fmmhip.h) ---#

const FMMHIP_OK, FMMHIP_EINVAL, FMMHIP_ENOMEM, FMMHIP_EDEVICE, FMMHIP_ETOL, FMMHIP_EDEPTH = Cint.(0:5)
const FMMHIP_POINT_SOURCE, FMMHIP_POINT_VORTEX, FMMHIP_POINT_DIPOLE, FMMHIP_VORTEX_GAUSSIAN = Cint.(0:3)
const FMMHIP_PANEL_SOURCE_TRI, FMMHIP_PANEL_DIPOLE_TRI, FMMHIP_PANEL_SOURCE_QUAD, FMMHIP_PANEL_DIPOLE_QUAD = Cint.(4:7)
const FMMHIP_POINT_SOURCE_VORTEX = Cint(8)   # Point{SourceVortex}: rows [x, y, z, sigma, q, Gx, Gy, Gz]
# Panel{SourceDipole}: rows [centroid(3), radius, source strength, dipole strength, vertices]
const FMMHIP_PANEL_SOURCE_DIPOLE_TRI, FMMHIP_PANEL_SOURCE_DIPOLE_QUAD = Cint(9), Cint(10)
const FMMHIP_FP64, FMMHIP_MIXED = Cint(0), Cint(1)
const FMMHIP_INFLUENCE_POTENTIAL, FMMHIP_INFLUENCE_NORMAL_GRADIENT = Cint(0), Cint(1)
const FMMHIP_REQUEST_FMM, FMMHIP_REQUEST_SOLVE = Cint(0), Cint(1)

"""
    hip_kernel(system)

Kernel tag of a system whose `body_to_multipole!` / `direct!` are one of libfmmhip's built-in kernels,
or `nothing` (the default: the system stays on the CPU path). Users opt in next to their
`body_to_multipole!` overload, e.g. `FastMultipole.hip_kernel(::Gravitational) = FMMHIP_POINT_SOURCE`.
"""
hip_kernel(system) = nothing

hip_precision(p::Symbol) = p === :fp64 ? FMMHIP_FP64 : p === :mixed ? FMMHIP_MIXED :
    throw(ArgumentError("precision must be :fp64 or :mixed, got $p"))

#--- mirrors of the C structs (field order, types and padding as in include/fmmhip.h) ---#

struct FmmHipOptions                  # fmmhip_options
    expansion_order::Int32
    leaf_size_source::Int32
    leaf_size_target::Int32
    multipole_acceptance::Float64
    list_method::Int32
    shrink_recenter::Int32
    error_kind::Int32
    error_relative::Int32
    error_tolerance::Float64
    error_bonus::Int32
    scalar_potential::Int32
    gradient::Int32
    hessian::Int32
    farfield::Int32
    nearfield::Int32
    self_induced::Int32
    precision::Int32
end

mutable struct FmmHipStats            # fmmhip_stats
    pmax::Int32
    error_success::Int32
    n_target_branches::Int64
    n_source_branches::Int64
    n_m2l::Int64
    n_direct::Int64
    n_p2p_pairs::Int64
    ms_tree::Float64
    ms_lists::Float64
    ms_p2p::Float64
    ms_upward::Float64
    ms_m2l::Float64
    ms_downward::Float64
    ms_total::Float64
    ms_estimate::Float64
    ms_gather::Float64
    FmmHipStats() = new(0, 1, 0, 0, 0, 0, 0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
end

struct FmmHipRequest                  # fmmhip_request
    kind::Int32
    kernel::Int32
    n::Int64
    source_buffer::Ptr{Float64}
    options::FmmHipOptions
    influence::Int32
    max_iterations::Int32
    tolerance::Float64
    external::Ptr{Float64}
end

struct FmmHipRequestStats             # fmmhip_request_stats
    status::Int32
    sweeps::Int32
    pmax::Int32
    error_success::Int32
    converged::Int32
    queue_ms::Float64
    run_ms::Float64
    latency_ms::Float64
    start_ms::Float64
    end_ms::Float64
    footprint_bytes::Float64
end
FmmHipRequestStats() = FmmHipRequestStats(0, 0, 0, 1, 1, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)

struct FmmHipServerStats              # fmmhip_server_stats_t
    completed::Int64
    queued::Int64
    running::Int32
    p50_ms::Float64
    p99_ms::Float64
    mean_queue_ms::Float64
    span_ms::Float64
end
FmmHipServerStats() = FmmHipServerStats(0, 0, 0, 0.0, 0.0, 0.0, 0.0)

#--- keyword arguments -> FmmHipOptions (defaults of src/fmm.jl:862-923) ---#

# error_tolerance -> (kind, relative, epsilon, BE): the types of src/containers.jl:64-104
hip_error(::Nothing) = (Int32(0), Int32(0), 0.0, Int32(1))
hip_error(::RotatedCoefficientsAbsoluteGradient{ε,BE}) where {ε,BE} = (Int32(1), Int32(0), Float64(ε), Int32(BE))
hip_error(::RotatedCoefficientsRelativeGradient{ε,BE}) where {ε,BE} = (Int32(1), Int32(1), Float64(ε), Int32(BE))
hip_error(::PowerAbsolutePotential{ε,BE}) where {ε,BE} = (Int32(2), Int32(0), Float64(ε), Int32(BE))
hip_error(::PowerRelativePotential{ε,BE}) where {ε,BE} = (Int32(2), Int32(1), Float64(ε), Int32(BE))
hip_error(::PowerAbsoluteGradient{ε,BE}) where {ε,BE} = (Int32(3), Int32(0), Float64(ε), Int32(BE))
hip_error(::PowerRelativeGradient{ε,BE}) where {ε,BE} = (Int32(3), Int32(1), Float64(ε), Int32(BE))
hip_error(e) = throw(ArgumentError("error_tolerance $(typeof(e)) has no GPU counterpart"))

hip_method(::Barba) = Int32(0)
hip_method(::SelfTuning) = Int32(1)
hip_method(::SelfTuningTreeStop) = Int32(2)

# a Bool or a one-element Bool vector (per-system switches of a single system)
_flag(x::Bool) = Int32(x)
_flag(x::AbstractVector{Bool}) = length(x) == 1 ? Int32(x[1]) :
    throw(ArgumentError("libfmmhip evaluates one target system; got $(length(x)) switches"))
_size(x::Integer) = Int32(x)
_size(x::AbstractVector) = length(x) == 1 ? Int32(x[1]) :
    throw(ArgumentError("libfmmhip evaluates one source system; got $(length(x)) leaf sizes"))

function FmmHipOptions(; expansion_order=5, leaf_size_source=20, leaf_size_target=nothing, multipole_acceptance=0.4,
                       interaction_list_method=SelfTuningTreeStop(), shrink_recenter=true, error_tolerance=nothing,
                       scalar_potential=false, gradient=true, hessian=false, farfield=true, nearfield=true,
                       self_induced=true, precision=:fp64, ignored...)
    kind, relative, eps, be = hip_error(error_tolerance)
    ls = _size(leaf_size_source)
    lt = leaf_size_target === nothing ? ls : _size(leaf_size_target)
    return FmmHipOptions(Int32(expansion_order), ls, lt, Float64(multipole_acceptance), hip_method(interaction_list_method),
                         Int32(shrink_recenter), kind, relative, eps, be, _flag(scalar_potential), _flag(gradient),
                         _flag(hessian), Int32(farfield), Int32(nearfield), Int32(self_induced), hip_precision(precision))
end

# the reference's optimized_args (src/fmm.jl:1162-1166), from the options of the call and, with
# tune = true, the leaf size its phase times suggest (src/fmm.jl:1073-1089)
function optargs_of(opt::FmmHipOptions, stats::FmmHipStats; tune=false)
    leaf = Int(opt.leaf_size_source)
    order = Int(opt.expansion_order)
    if tune
        order = Int(stats.pmax)
        if stats.n_m2l > 0 && stats.n_p2p_pairs > 0 && stats.ms_p2p > 0
            t_m2l = stats.ms_m2l / stats.n_m2l
            t_direct = stats.ms_p2p / stats.n_p2p_pairs
            leaf = ceil(Int, sqrt(t_m2l / t_direct))
        else
            leaf = max(1, leaf ÷ 2)
        end
    end
    return (leaf_size_source = SVector{1}(leaf), expansion_order = max(order, 1),
            multipole_acceptance = opt.multipole_acceptance)
end

#--- context and evaluation handles ---#

function check(status::Cint, ctx::Ptr{Cvoid})
    status == FMMHIP_OK && return :ok
    status == FMMHIP_ETOL && return :tolerance_not_met
    msg = ctx == C_NULL ? "no context" :
        unsafe_string(ccall((:fmmhip_last_error, libfmmhip), Cstring, (Ptr{Cvoid},), ctx))
    status == FMMHIP_EINVAL && throw(ArgumentError("libfmmhip: " * msg))
    status == FMMHIP_ENOMEM && throw(OutOfMemoryError())
    error("libfmmhip (status $status): " * msg)            # EDEVICE, EDEPTH
end

"""
    HipEval(device=-1)

One HIP context (one stream; use one per Julia task) and one evaluation handle, which keeps trees,
lists, expansions and buffers resident in HBM between calls: the GPU counterpart of the reference's
`Cache` and trees (src/containers.jl:264-292).
"""
mutable struct HipEval
    ctx::Ptr{Cvoid}
    ev::Ptr{Cvoid}
    function HipEval(device::Integer=-1)
        c = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:fmmhip_ctx_create, libfmmhip), Cint, (Cint, Ptr{Ptr{Cvoid}}), device, c), C_NULL)
        e = Ref{Ptr{Cvoid}}(C_NULL)
        st = ccall((:fmmhip_eval_create, libfmmhip), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}), c[], e)
        if st != FMMHIP_OK
            ccall((:fmmhip_ctx_destroy, libfmmhip), Cvoid, (Ptr{Cvoid},), c[])
            check(st, C_NULL)
        end
        h = new(c[], e[])
        finalizer(close!, h)
        return h
    end
end

# launch shape of later evaluations (set by the GPU-side autotuner): workgroup slots per compute unit
# the nearfield leaves to the far-field stream, 0..3
set_p2p_free_slots!(h::HipEval, slots::Integer) =
    check(ccall((:fmmhip_ctx_set_p2p_free_slots, libfmmhip), Cint, (Ptr{Cvoid}, Cint), h.ctx, slots), h.ctx)

function close!(h::HipEval)
    if h.ev != C_NULL
        ccall((:fmmhip_eval_destroy, libfmmhip), Cvoid, (Ptr{Cvoid},), h.ev)
        h.ev = C_NULL
    end
    if h.ctx != C_NULL
        ccall((:fmmhip_ctx_destroy, libfmmhip), Cvoid, (Ptr{Cvoid},), h.ctx)
        h.ctx = C_NULL
    end
    return nothing
end

#--- gathering through the user-system interface (src/compatibility.jl:5-20, 300-399) ---#

function source_buffer(system)
    n = get_n_bodies(system)
    sbuf = Matrix{Float64}(undef, data_per_body(system), n)
    for i in 1:n
        source_system_to_buffer!(sbuf, i, system, i)
    end
    return sbuf
end

function target_positions(system)
    n = get_n_bodies(system)
    txyz = Matrix{Float64}(undef, 3, n)
    for i in 1:n
        txyz[:, i] .= get_position(system, i)
    end
    return txyz
end

function scatter_targets!(system, tbuf::Matrix{Float64}, sw::DerivativesSwitch)
    for i in 1:size(tbuf, 2)
        buffer_to_target_system!(system, i, sw, tbuf, i)
    end
    return system
end

#--- tree and list views (materialised from the device on first use) ---#

"""
    HipTreeView(hip, which)

The target (`which = 0`) or source (`which = 1`) tree of the last evaluation as the reference's
objects: `branches::Vector{Branch{Float64,1}}`, `levels_index`, `leaf_index`, `sort_index`,
`inverse_sort_index` (1-based, src/containers.jl:150-195); `expansions(view)` reads the expansions
as the reference's `Array{Float64,4}(2, 2, H, nbranch)` (src/tree.jl:1334-1339).
"""
struct HipTreeView
    branches::Vector{Branch{Float64,1}}
    levels_index::Vector{UnitRange{Int64}}
    leaf_index::Vector{Int}
    sort_index::Vector{Int}
    inverse_sort_index::Vector{Int}
    hip::HipEval
    which::Int32
end

function HipTreeView(hip::HipEval, which::Integer)
    w = Int32(which)
    nb = ccall((:fmmhip_eval_n_branches, libfmmhip), Int64, (Ptr{Cvoid}, Cint), hip.ev, w)
    ints = Matrix{Int64}(undef, 6, nb)          # [n_bodies, first, n_branches, first_child, parent, leaf] per branch
    reals = Matrix{Float64}(undef, 17, nb)      # [source_center(3), target_center(3), radii(2), boxes(6), max_influence, 0, 0]
    nb > 0 && check(ccall((:fmmhip_eval_get_branches, libfmmhip), Cint, (Ptr{Cvoid}, Cint, Ptr{Int64}, Ptr{Float64}),
                          hip.ev, w, ints, reals), hip.ctx)
    branches = Vector{Branch{Float64,1}}(undef, nb)
    for b in 1:nb
        n, first, nch, child, parent, leaf = view(ints, :, b)
        r = view(reals, :, b)
        branches[b] = Branch(SVector{1}(n), SVector{1}((first + 1):(first + n)), nch, (child + 1):(child + nch),
                             parent + 1, leaf >= 0 ? leaf + 1 : -1, SVector{3}(r[1], r[2], r[3]), SVector{3}(r[4], r[5], r[6]),
                             r[7], r[8], SVector{3}(r[9], r[10], r[11]), SVector{3}(r[12], r[13], r[14]), r[15])
    end
    nl = ccall((:fmmhip_eval_n_levels, libfmmhip), Int64, (Ptr{Cvoid}, Cint), hip.ev, w)
    lv = Matrix{Int64}(undef, 2, nl)
    nl > 0 && check(ccall((:fmmhip_eval_get_levels, libfmmhip), Cint, (Ptr{Cvoid}, Cint, Ptr{Int64}), hip.ev, w, lv), hip.ctx)
    levels = [(lv[1, k] + 1):(lv[2, k] + 1) for k in 1:nl]
    nlf = ccall((:fmmhip_eval_n_leaves, libfmmhip), Int64, (Ptr{Cvoid}, Cint), hip.ev, w)
    leaves = Vector{Int64}(undef, nlf)
    nlf > 0 && check(ccall((:fmmhip_eval_get_leaf_index, libfmmhip), Cint, (Ptr{Cvoid}, Cint, Ptr{Int64}), hip.ev, w, leaves),
                     hip.ctx)
    nbod = nb > 0 ? Int(ints[1, 1]) : 0
    si = Vector{Int64}(undef, nbod)
    nbod > 0 && check(ccall((:fmmhip_eval_get_sort_index, libfmmhip), Cint, (Ptr{Cvoid}, Cint, Ptr{Int64}), hip.ev, w, si),
                      hip.ctx)
    sort_index = si .+ 1
    inverse = similar(sort_index)
    inverse[sort_index] .= 1:nbod
    return HipTreeView(branches, levels, leaves .+ 1, sort_index, inverse, hip, w)
end

function expansions(t::HipTreeView, expansion_order::Integer)
    H = ((expansion_order + 1) * (expansion_order + 2)) ÷ 2
    e = Array{Float64,4}(undef, 2, 2, H, length(t.branches))     # [re/im, phi/chi, H, branch]
    length(t.branches) > 0 && check(ccall((:fmmhip_eval_get_expansions, libfmmhip), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}),
                                          t.hip.ev, t.which, e), t.hip.ctx)
    return e
end

"""
    HipListView(hip, which)

The m2l (`which = 0`) or direct (`which = 1`) list of the last evaluation as the reference's
`Vector{SVector{2,Int32}}` of (target, source) branch indices sorted by target (1-based).
"""
function HipListView(hip::HipEval, which::Integer)
    n = ccall((:fmmhip_eval_n_list, libfmmhip), Int64, (Ptr{Cvoid}, Cint), hip.ev, Int32(which))
    raw = Matrix{Int32}(undef, 2, n)
    n > 0 && check(ccall((:fmmhip_eval_get_list, libfmmhip), Cint, (Ptr{Cvoid}, Cint, Ptr{Int32}), hip.ev, Int32(which), raw),
                   hip.ctx)
    return [SVector{2,Int32}(raw[1, k] + Int32(1), raw[2, k] + Int32(1)) for k in 1:n]
end

"""
    hip_m2l_orders(hip)

Expansion order chosen for each M2L interaction of the last evaluation, in m2l-list order (the
reference computes them but does not return them: src/fmm.jl:468,476).
"""
function hip_m2l_orders(hip::HipEval)
    n = ccall((:fmmhip_eval_n_list, libfmmhip), Int64, (Ptr{Cvoid}, Cint), hip.ev, Int32(0))
    out = Vector{Int32}(undef, n)
    n > 0 && check(ccall((:fmmhip_eval_get_m2l_orders, libfmmhip), Cint, (Ptr{Cvoid}, Ptr{Int32}), hip.ev, out), hip.ctx)
    return out
end

#--- fmm! (src/fmm.jl:862-1176) ---#

"""
    hip_fmm!(hip, target_system, source_system=target_system; tune=false, kwargs...)

`fmm!` on the GPU for one target and one source system. Returns the reference's 8-tuple
`(optimized_args, cache, target_tree, source_tree, m2l_list, direct_list, derivatives_switches,
error_success)`, with the trees and lists as `HipTreeView`s and `SVector{2,Int32}` vectors and
`cache` the `HipEval` itself. Results accumulate into the target system through
`buffer_to_target_system!`, as the reference leaves them.
"""
function hip_fmm!(hip::HipEval, t, s=t; tune=false, update_target_systems=true, kwargs...)
    k = hip_kernel(s)
    k === nothing && throw(ArgumentError("$(typeof(s)) declares no hip_kernel"))
    sbuf = source_buffer(s)
    txyz = s === t ? nothing : target_positions(t)
    GC.@preserve sbuf txyz begin
        check(ccall((:fmmhip_eval_set_systems, libfmmhip), Cint,
                    (Ptr{Cvoid}, Cint, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Cint),
                    hip.ev, k, txyz === nothing ? size(sbuf, 2) : size(txyz, 2),
                    txyz === nothing ? Ptr{Float64}(C_NULL) : pointer(txyz), size(sbuf, 2), sbuf, 0), hip.ctx)
        opt = FmmHipOptions(; kwargs...)
        stats = FmmHipStats()
        st = check(ccall((:fmmhip_fmm, libfmmhip), Cint, (Ptr{Cvoid}, Ref{FmmHipOptions}, Ref{FmmHipStats}),
                         hip.ev, opt, stats), hip.ctx)
    end
    nt = txyz === nothing ? size(sbuf, 2) : size(txyz, 2)
    sw = DerivativesSwitch(opt.scalar_potential != 0, opt.gradient != 0, opt.hessian != 0)
    if update_target_systems
        tbuf = Matrix{Float64}(undef, 16, nt)                  # src/compatibility.jl:155-172
        check(ccall((:fmmhip_eval_get_targets, libfmmhip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cint), hip.ev, tbuf, 0), hip.ctx)
        scatter_targets!(t, tbuf, sw)
    end
    error_success = st !== :tolerance_not_met && stats.error_success != 0
    return (optargs_of(opt, stats; tune), hip, HipTreeView(hip, 0), HipTreeView(hip, 1), HipListView(hip, 0),
            HipListView(hip, 1), (sw,), error_success)
end

"""
    hip_fmm_batch!(hip, systems; update_target_systems=true, kwargs...)

`fmm!(system; kwargs...)` for each of several independent systems of one kernel (no counterpart in the
reference, whose callers loop over `fmm!`): one tree per system, lists within each system, every pass of
all systems in the same launches. Each system's results equal its own `fmm!` call. Returns the batch's
`FmmHipStats` and per system the `(pmax, error_success)` its own `fmm!` call reports.
"""
function hip_fmm_batch!(hip::HipEval, systems::AbstractVector; update_target_systems=true, kwargs...)
    isempty(systems) && return FmmHipStats()
    k = hip_kernel(systems[1])
    k === nothing && throw(ArgumentError("$(typeof(systems[1])) declares no hip_kernel"))
    all(s -> hip_kernel(s) == k, systems) || throw(ArgumentError("the systems of one batch share one kernel"))
    bufs = [source_buffer(s) for s in systems]
    counts = Int64[size(b, 2) for b in bufs]
    sbuf = reduce(hcat, bufs)
    opt = FmmHipOptions(; kwargs...)
    stats = FmmHipStats()
    GC.@preserve sbuf counts begin
        check(ccall((:fmmhip_eval_set_batch, libfmmhip), Cint,
                    (Ptr{Cvoid}, Cint, Int64, Ptr{Int64}, Ptr{Float64}, Cint),
                    hip.ev, k, length(counts), counts, sbuf, 0), hip.ctx)
        check(ccall((:fmmhip_fmm, libfmmhip), Cint, (Ptr{Cvoid}, Ref{FmmHipOptions}, Ref{FmmHipStats}),
                    hip.ev, opt, stats), hip.ctx)
    end
    if update_target_systems
        tbuf = Matrix{Float64}(undef, 16, sum(counts))
        check(ccall((:fmmhip_eval_get_targets, libfmmhip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cint), hip.ev, tbuf, 0), hip.ctx)
        sw = DerivativesSwitch(opt.scalar_potential != 0, opt.gradient != 0, opt.hessian != 0)
        first = 1
        for (s, n) in zip(systems, counts)
            scatter_targets!(s, tbuf[:, first:first+n-1], sw)
            first += n
        end
    end
    pmax = zeros(Int32, length(counts)); ok = zeros(Int32, length(counts))
    check(ccall((:fmmhip_eval_get_batch_stats, libfmmhip), Cint, (Ptr{Cvoid}, Ptr{Int32}, Ptr{Int32}), hip.ev, pmax, ok),
          hip.ctx)
    return stats, [(Int(pmax[i]), ok[i] != 0) for i in eachindex(counts)]
end

"""
    hip_fmm_tuple!(hip, systems::Tuple; update_target_systems=true, kwargs...)

`fmm!(systems)` for a tuple of systems that share one tree, e.g. panels and vortex particles
(src/fmm.jl:807-884, src/tree.jl:3-122): every system is a target of all of them and a source with its
own kernel. `leaf_size_source`: one leaf size, or one per system (an `SVector`, as the reference takes it);
`SelfTuningTreeStop` or `Barba` lists. Returns the reference's 8-tuple as `hip_fmm!` does, with one
`DerivativesSwitch` per system.
"""
function hip_fmm_tuple!(hip::HipEval, systems::Tuple; update_target_systems=true, leaf_size_source=20, kwargs...)
    ks = [hip_kernel(s) for s in systems]
    any(k -> k === nothing, ks) && throw(ArgumentError("every system of a tuple declares a hip_kernel"))
    kernels = Int32.(ks)
    bufs = [source_buffer(s) for s in systems]
    counts = Int64[size(b, 2) for b in bufs]
    R = max(7, maximum(size(b, 1) for b in bufs))      # the largest data_per_body of the systems
    sbuf = zeros(Float64, R, sum(counts))               # each system's rows first, zeros after them
    first = 1
    for b in bufs
        sbuf[1:size(b, 1), first:first+size(b, 2)-1] .= b
        first += size(b, 2)
    end
    per_system = leaf_size_source isa AbstractVector
    ls = per_system ? Int64.(collect(leaf_size_source)) : Int64[]
    opt = FmmHipOptions(; leaf_size_source=per_system ? minimum(ls) : leaf_size_source, kwargs...)
    stats = FmmHipStats()
    st = GC.@preserve sbuf counts kernels ls begin
        check(ccall((:fmmhip_eval_set_tuple, libfmmhip), Cint,
                    (Ptr{Cvoid}, Int64, Ptr{Int32}, Ptr{Int64}, Ptr{Float64}, Cint),
                    hip.ev, length(counts), kernels, counts, sbuf, 0), hip.ctx)
        per_system && check(ccall((:fmmhip_eval_set_tuple_leaf_sizes, libfmmhip), Cint, (Ptr{Cvoid}, Int64, Ptr{Int64}),
                                  hip.ev, length(ls), ls), hip.ctx)
        check(ccall((:fmmhip_fmm, libfmmhip), Cint, (Ptr{Cvoid}, Ref{FmmHipOptions}, Ref{FmmHipStats}),
                    hip.ev, opt, stats), hip.ctx)
    end
    sw = DerivativesSwitch(opt.scalar_potential != 0, opt.gradient != 0, opt.hessian != 0)
    if update_target_systems
        tbuf = Matrix{Float64}(undef, 16, sum(counts))
        check(ccall((:fmmhip_eval_get_targets, libfmmhip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cint), hip.ev, tbuf, 0), hip.ctx)
        first = 1
        for (s, n) in zip(systems, counts)
            scatter_targets!(s, tbuf[:, first:first+n-1], sw)
            first += n
        end
    end
    error_success = st !== :tolerance_not_met && stats.error_success != 0
    return (optargs_of(opt, stats; tune=false), hip, HipTreeView(hip, 0), HipTreeView(hip, 1), HipListView(hip, 0),
            HipListView(hip, 1), ntuple(_ -> sw, length(systems)), error_success)
end

# fmm! with an optional `hip` keyword: GPU when routable, the reference's path otherwise
function fmm!(target_systems::Tuple, source_systems::Tuple, hip::HipEval; kwargs...)
    if length(target_systems) == 1 && length(source_systems) == 1 && hip_kernel(source_systems[1]) !== nothing
        return hip_fmm!(hip, target_systems[1], source_systems[1]; kwargs...)
    end
    if target_systems === source_systems && all(s -> hip_kernel(s) !== nothing, source_systems) &&
       !(get(kwargs, :interaction_list_method, SelfTuningTreeStop()) isa SelfTuning)
        return hip_fmm_tuple!(hip, source_systems; kwargs...)
    end
    return fmm!(target_systems, source_systems; kwargs...)
end

#--- direct! (src/direct.jl:20-75) ---#

function hip_direct!(hip::HipEval, t, s=t; scalar_potential=false, gradient=true, hessian=false, precision=:fp64)
    k = hip_kernel(s)
    k === nothing && return direct!(t, s; scalar_potential, gradient, hessian)
    sbuf = source_buffer(s)
    txyz = target_positions(t)
    tbuf = zeros(16, size(txyz, 2))
    tbuf[1:3, :] .= txyz
    GC.@preserve sbuf txyz tbuf begin
        check(ccall((:fmmhip_direct, libfmmhip), Cint,
                    (Ptr{Cvoid}, Cint, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Cint, Cint, Cint, Cint, Ptr{Float64}, Cint),
                    hip.ctx, k, size(txyz, 2), txyz, size(sbuf, 2), sbuf, Cint(scalar_potential), Cint(gradient),
                    Cint(hessian), hip_precision(precision), tbuf, 0), hip.ctx)
    end
    # tbuf accumulated (+=) like the reference's set_* accessors (src/compatibility.jl:459-491)
    scatter_targets!(t, tbuf, DerivativesSwitch(scalar_potential, gradient, hessian))
    return t
end

# direct!(targets::Tuple, sources::Tuple) (src/direct.jl:43-75): every source system on every target system,
# sources outermost, the switches per target system; systems without a hip_kernel take the reference's path
function hip_direct!(hip::HipEval, targets::Tuple, sources::Tuple=targets; scalar_potential=false, gradient=true,
                     hessian=false, precision=:fp64)
    nt = length(targets)
    ps, gs, hs = (v isa AbstractVector ? v : fill(v, nt) for v in (scalar_potential, gradient, hessian))
    for s in sources, (i, t) in enumerate(targets)
        hip_direct!(hip, t, s; scalar_potential=ps[i], gradient=gs[i], hessian=hs[i], precision)
    end
    return targets
end

#--- FastGaussSeidel / solve! (src/solve.jl:538-946) ---#

# solved scalar strengths back into the user system: value_to_strength! into a source buffer, then
# buffer_to_system_strength! (src/compatibility.jl:222-262), as the reference's solver does
function store_strengths!(system, strengths::Vector{Float64})
    sbuf = source_buffer(system)
    for i in 1:length(strengths)
        value_to_strength!(sbuf, system, i, strengths[i])
        buffer_to_system_strength!(system, sbuf, i)
    end
    return system
end

struct HipFGS
    hip::HipEval
    influence::Cint
end

function hip_fast_gauss_seidel(hip::HipEval, system; influence=FMMHIP_INFLUENCE_POTENTIAL, expansion_order=4,
                               multipole_acceptance=0.5, leaf_size=30, shrink_recenter=true)
    k = hip_kernel(system)
    k in (FMMHIP_POINT_SOURCE, FMMHIP_PANEL_SOURCE_TRI, FMMHIP_PANEL_SOURCE_QUAD) ||
        throw(ArgumentError("FastGaussSeidel on the GPU takes point sources or source panels"))
    sbuf = source_buffer(system)
    GC.@preserve sbuf begin
        check(ccall((:fmmhip_eval_set_systems, libfmmhip), Cint,
                    (Ptr{Cvoid}, Cint, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Cint),
                    hip.ev, k, size(sbuf, 2), Ptr{Float64}(C_NULL), size(sbuf, 2), sbuf, 0), hip.ctx)
        opt = FmmHipOptions(; expansion_order, leaf_size_source=leaf_size, multipole_acceptance, shrink_recenter,
                            interaction_list_method=Barba())
        check(ccall((:fmmhip_fgs_setup, libfmmhip), Cint, (Ptr{Cvoid}, Ref{FmmHipOptions}, Cint), hip.ev, opt,
                    Cint(influence)), hip.ctx)
    end
    return HipFGS(hip, Cint(influence))
end

# external: the external influence per body (what target_influence_to_buffer! + influence! leave in
# the reference's right-hand side); returns (sweeps, mean squared residual per iteration, converged)
function hip_solve!(system, s::HipFGS, external::AbstractVector{Float64}; max_iterations=10, tolerance=1e-3)
    n = get_n_bodies(system)
    length(external) == n || throw(ArgumentError("one external influence per body"))
    ext = Vector{Float64}(external)
    strengths = Vector{Float64}(undef, n)
    sweeps = Ref{Cint}(0)
    mse = zeros(Float64, max(max_iterations, 1))
    check(ccall((:fmmhip_fgs_solve, libfmmhip), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Cint, Cint, Cdouble, Ptr{Float64}, Ref{Cint}, Ptr{Float64}),
                s.hip.ev, ext, 0, max_iterations, tolerance, strengths, sweeps, mse), s.hip.ctx)
    store_strengths!(system, strengths)
    tbuf = Matrix{Float64}(undef, 16, n)
    check(ccall((:fmmhip_eval_get_targets, libfmmhip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cint), s.hip.ev, tbuf, 0), s.hip.ctx)
    scatter_targets!(system, tbuf, DerivativesSwitch(true, true, false))
    return Int(sweeps[]), mse[1:max(Int(sweeps[]), 1)], sweeps[] < max_iterations
end

#--- request server ---#

"""
    HipServer(; slots=4, hbm_budget=64 * 2^30, device=-1, batch_requests=64, batch_bodies=2 << 20)

Many `fmm!(system)` evaluations and FastGaussSeidel solves of uneven sizes, submitted from any task,
run concurrently on one device (fmmhip_server_*). FMM requests with the same kernel and options that
wait together run as one batch evaluation (up to `batch_requests` requests and `batch_bodies` bodies;
`batch_requests=1`: each alone), each with its own call's results and status. `wait!` blocks in C:
call it from a task on a thread of its own so the Julia scheduler keeps running.
"""
mutable struct HipServer
    h::Ptr{Cvoid}
    function HipServer(; slots::Integer=4, hbm_budget::Real=64 * 2^30, device::Integer=-1,
                       batch_requests::Integer=64, batch_bodies::Integer=2 << 20)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        st = ccall((:fmmhip_server_create, libfmmhip), Cint, (Cint, Cint, Cdouble, Ref{Ptr{Cvoid}}), device, slots,
                   hbm_budget, h)
        check(st, C_NULL)
        s = new(h[])
        finalizer(close!, s)
        check(ccall((:fmmhip_server_set_batching, libfmmhip), Cint, (Ptr{Cvoid}, Int32, Int64), s.h, batch_requests,
                    batch_bodies), C_NULL)
        return s
    end
end

function close!(s::HipServer)
    if s.h != C_NULL
        ccall((:fmmhip_server_destroy, libfmmhip), Cvoid, (Ptr{Cvoid},), s.h)
        s.h = C_NULL
    end
    return nothing
end

function submit_fmm!(server::HipServer, system; kwargs...)
    k = hip_kernel(system)
    k === nothing && throw(ArgumentError("$(typeof(system)) declares no hip_kernel"))
    sbuf = source_buffer(system)
    ticket = Ref{Int64}(0)
    GC.@preserve sbuf begin          # the request is copied by submit
        req = FmmHipRequest(FMMHIP_REQUEST_FMM, k, size(sbuf, 2), pointer(sbuf), FmmHipOptions(; kwargs...),
                            Int32(0), Int32(0), 0.0, Ptr{Float64}(C_NULL))
        check(ccall((:fmmhip_server_submit, libfmmhip), Cint, (Ptr{Cvoid}, Ref{FmmHipRequest}, Ref{Int64}),
                    server.h, req, ticket), C_NULL)
    end
    return ticket[]
end

function submit_solve!(server::HipServer, system, external::AbstractVector{Float64}; influence=FMMHIP_INFLUENCE_POTENTIAL,
                       max_iterations=10, tolerance=1e-3, expansion_order=4, multipole_acceptance=0.5, leaf_size=30,
                       shrink_recenter=true)
    k = hip_kernel(system)
    k === nothing && throw(ArgumentError("$(typeof(system)) declares no hip_kernel"))
    sbuf = source_buffer(system)
    ext = Vector{Float64}(external)
    ticket = Ref{Int64}(0)
    GC.@preserve sbuf ext begin
        opt = FmmHipOptions(; expansion_order, leaf_size_source=leaf_size, multipole_acceptance, shrink_recenter,
                            interaction_list_method=Barba(), scalar_potential=true)
        req = FmmHipRequest(FMMHIP_REQUEST_SOLVE, k, size(sbuf, 2), pointer(sbuf), opt, Int32(influence),
                            Int32(max_iterations), Float64(tolerance), pointer(ext))
        check(ccall((:fmmhip_server_submit, libfmmhip), Cint, (Ptr{Cvoid}, Ref{FmmHipRequest}, Ref{Int64}),
                    server.h, req, ticket), C_NULL)
    end
    return ticket[]
end

# results into the system as fmm! (or solve!) would leave them; returns the request's statistics
# (status FMMHIP_ETOL: results delivered with error_success = false or an unconverged solve)
function wait!(server::HipServer, ticket::Int64, system; solve=false)
    n = get_n_bodies(system)
    tbuf = Matrix{Float64}(undef, 16, n)
    strengths = solve ? Vector{Float64}(undef, n) : Float64[]
    stats = Ref(FmmHipRequestStats())
    st = ccall((:fmmhip_server_wait, libfmmhip), Cint, (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ref{FmmHipRequestStats}),
               server.h, ticket, tbuf, solve ? pointer(strengths) : Ptr{Float64}(C_NULL), stats)
    check(st, C_NULL)
    solve && store_strengths!(system, strengths)
    scatter_targets!(system, tbuf, DerivativesSwitch(true, true, !solve))
    return stats[]
end

function server_stats(server::HipServer)
    out = Ref(FmmHipServerStats())
    check(ccall((:fmmhip_server_stats, libfmmhip), Cint, (Ptr{Cvoid}, Ref{FmmHipServerStats}), server.h, out), C_NULL)
    return out[]
end
