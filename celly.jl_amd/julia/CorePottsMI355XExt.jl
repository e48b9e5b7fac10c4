# CorePotts package extension that runs CheckerboardSweepCPM and LotteryCPM on one MI355X
# through libcpm_mi355x.so (include/cpm_mi355x.h). To be added to lib/CorePotts as
# ext/CorePottsMI355XExt.jl with this package's library on the loader path (or CPM_MI355X_LIB).
#
# How it plugs in (paths relative to the reference checkout, lib/CorePotts/src/):
# * MI355XBackend is a KernelAbstractions backend whose arrays (MI355XArray) are host
#   mirrors: backend_capabilities / execution_adaptor (execution/contracts.jl:52-57,
#   sciml/interface.jl:596-599) let PottsProblem / init adapt the compiled state to it, and
#   _require_plan_backend (execution/contracts.jl:398-402) accepts the pair.
# * init_scientific for CheckerboardSweepCPM / LotteryCPM (algorithms/checkerboard.jl:197-219,
#   algorithms/lottery.jl:94-115) runs unchanged; it takes the execution plan as a keyword,
#   which Julia does not dispatch on, but passes it positionally to
#   _allocate_checkerboard_workspace / _allocate_lottery_workspace (checkerboard.jl:159,
#   lottery.jl:46). The methods below specialise those on ExecutionPlan{MI355XBackend} and
#   return an MI355XWorkspace, which becomes the integrator's algorithm_workspace: the ninth
#   type parameter Q of ScientificPottsIntegrator{S, C, R, A, G, P, W, M, Q, L, B}
#   (algorithms/sequential.jl:255-273).
# * perform_scientific_mcs! and _current_mcs_report (checkerboard.jl:564-586, lottery.jl:328,
#   397) and logical_state (sequential.jl:630-632) are specialised on A <: CheckerboardSweepCPM
#   or A <: LotteryCPM together with Q = MI355XWorkspace. Each signature is a strict subtype of
#   the reference method it replaces, so it is more specific and never ambiguous.
# * The library context is created at the first MCS, when the integrator holds everything
#   cpm_model_desc needs (components, algorithm temperature, seed) and init_scientific's
#   mechanics initialization request (_initialize_mechanics!, specialised below, defers it to
#   cpm_initialize_mechanics after the state upload).
module CorePottsMI355XExt

using CorePotts, KernelAbstractions, SciMLBase, Adapt
import CorePotts: backend_capabilities, execution_adaptor, _allocate_checkerboard_workspace,
    _allocate_lottery_workspace, perform_scientific_mcs!, _current_mcs_report, logical_state,
    _initialize_mechanics!

const LIB = get(ENV, "CPM_MI355X_LIB", "libcpm_mi355x.so")
const MAX_DIRS = 128
const MAX_TYPES = 16
const MAX_MEDIA = 16
const CPM_ALG_CHECKERBOARD = Int32(1)
const CPM_ALG_LOTTERY = Int32(2)
const CPM_TERM_VOLUME = Int32(0)
const CPM_TERM_CONTACT = Int32(1)
const CPM_TERM_BOUNDARY = Int32(2)
const CPM_FACE_PERIODIC = Int32(0)
const CPM_FACE_CLOSED = Int32(1)
const CPM_FACE_FIXED = Int32(2)

# ---------------------------------------------------------------------------------------
# Backend and its host-mirror arrays
# ---------------------------------------------------------------------------------------
struct MI355XBackend <: KernelAbstractions.Backend end

struct MI355XArray{T, N} <: DenseArray{T, N}
    data::Array{T, N}
end
Base.size(a::MI355XArray) = size(a.data)
Base.IndexStyle(::Type{<:MI355XArray}) = IndexLinear()
Base.getindex(a::MI355XArray, i::Int) = a.data[i]
Base.setindex!(a::MI355XArray, v, i::Int) = (a.data[i] = v)
Base.similar(a::MI355XArray, ::Type{T}, dims::Dims) where {T} = MI355XArray(similar(a.data, T, dims))
Base.unsafe_convert(::Type{Ptr{T}}, a::MI355XArray{T}) where {T} = Base.unsafe_convert(Ptr{T}, a.data)
Base.Array(a::MI355XArray) = copy(a.data)
KernelAbstractions.get_backend(::MI355XArray) = MI355XBackend()
KernelAbstractions.synchronize(::MI355XBackend) = nothing  # the library synchronizes on observation
Adapt.adapt_storage(::Type{MI355XArray}, x::Array) = MI355XArray(x)

backend_capabilities(::MI355XBackend) = CorePotts.BackendCapabilities(
    CorePotts.AMDGPUFamily, CorePotts.QualifiedBackend,
    ccall((:cpm_abi_version, LIB), Int32, ()) == 1, true, true, false, (v"1.0.0",))
execution_adaptor(::MI355XBackend) = MI355XArray

const MI355XPlan = CorePotts.ExecutionPlan{MI355XBackend}

# ---------------------------------------------------------------------------------------
# cpm_model_desc (include/cpm_mi355x.h), field for field
# ---------------------------------------------------------------------------------------
struct CpmRelation
    count::Int32
    offsets::NTuple{MAX_DIRS, NTuple{3, Int32}}
    weights::NTuple{MAX_DIRS, Float64}
end

struct CpmMechanics
    kind::Int32
    initialization::Int32
    instance_id::UInt32
    eta::Float32
    noise_is_temperature::Int32
    noise::Float32
end

struct CpmModelDesc
    abi_version::Int32
    algorithm::Int32
    ndim::Int32
    dims::NTuple{3, Int32}
    face_kind::NTuple{3, NTuple{2, Int32}}
    face_medium::NTuple{3, NTuple{2, UInt32}}
    mutable_mask::Ptr{UInt8}
    immutable_tags::Ptr{UInt8}
    immutable_ids::Ptr{UInt32}
    proposal::CpmRelation
    contact::CpmRelation
    surface::CpmRelation
    n_energies::Int32
    energy_order::NTuple{3, Int32}
    volume_f64::Int32
    contact_f64::Int32
    boundary_f64::Int32
    n_types::Int32
    J::NTuple{MAX_TYPES * MAX_TYPES, Float64}
    n_table::Int32
    table_owner::NTuple{MAX_MEDIA, UInt32}
    table_type::NTuple{MAX_MEDIA, UInt32}
    n_media::Int32
    medium_ids::NTuple{MAX_MEDIA, UInt32}
    temperature::Float32
    capacity::UInt32
    seed::UInt64
    retire_target_volume::Float64
    retire_volume_strength::Float64
    retire_target_boundary::Int64
    retire_boundary_strength::Float64
    n_mechanics::Int32
    mechanics::NTuple{2, CpmMechanics}
    tile_size::NTuple{3, Int32}
    switching_interval::Int32
end

pad(values, n, zero_value) = ntuple(i -> i <= length(values) ? values[i] : zero_value, n)

function relation_desc(rel)
    rel === nothing && return CpmRelation(Int32(0), pad((), MAX_DIRS, (Int32(0), Int32(0), Int32(0))),
        pad((), MAX_DIRS, 0.0))
    offsets = [ntuple(a -> a <= length(o) ? Int32(o[a]) : Int32(0), 3) for o in rel.offsets]
    weights = [Float64(w) for w in rel.weights]
    return CpmRelation(Int32(length(offsets)), pad(offsets, MAX_DIRS, (Int32(0), Int32(0), Int32(0))),
        pad(weights, MAX_DIRS, 0.0))
end

face_kind(b::CorePotts.PeriodicBoundary) = CPM_FACE_PERIODIC
face_kind(b::CorePotts.ClosedBoundary) = CPM_FACE_CLOSED
face_kind(b::CorePotts.FixedExterior) = CPM_FACE_FIXED
face_medium(b) = UInt32(0)
face_medium(b::CorePotts.FixedExterior) = UInt32(b.owner.value)

# A term is evaluated in Float64 iff the reference's energy_change promotes it to Float64
# (components/scientific_components.jl:112, 291, 549): the property columns the term reads (and,
# for the boundary, its relation weights) decide, not the component's number_type alone.
term_f64(x) = Int32(x === Float64)
columns(state) = state.potts.storage.properties
coltype(state, key::Symbol) = eltype(getproperty(columns(state), key))
target_key(c) = typeof(c).parameters[1]
strength_key(c) = typeof(c).parameters[2]
energy_type(state, c::CorePotts.QuadraticVolumeHamiltonian) =
    float(promote_type(coltype(state, target_key(c)), coltype(state, strength_key(c))))
energy_type(state, c::CorePotts.UnorderedContactHamiltonian) = eltype(c.interactions)
energy_type(state, c::CorePotts.QuadraticBoundaryHamiltonian) =
    float(promote_type(coltype(state, target_key(c)), coltype(state, strength_key(c)),
        eltype(c.relation.weights)))

# cpm_model_desc of an integrator's model: what CompiledScientificState, the three relations,
# ScientificComponentSet and the algorithm hold (mirrors celly.jl_amd/corepotts/spec.py).
# `keep` roots the arrays the descriptor points to until cpm_create has copied them.
function model_descriptor(integrator, algorithm_id::Int32, keep::Vector{Any})
    state = integrator.state
    domain = state.domain
    dims = Int.(domain.descriptor.dims)
    N = length(dims)
    bounds = domain.descriptor.boundaries
    mask = Array(domain.storage.mutable_mask)
    has_obstacles = !all(!iszero, mask)
    tags = Array(domain.storage.immutable_tags)
    ids = Array(domain.storage.immutable_ids)
    push!(keep, mask, tags, ids)
    energies = integrator.components.energies
    order = Int32[]
    f64 = Dict{Int32, Int32}()
    contact = nothing
    J = zeros(Float64, MAX_TYPES * MAX_TYPES)
    ntypes = 0
    table_owner, table_type = UInt32[], UInt32[]
    for c in energies
        if c isa CorePotts.QuadraticVolumeHamiltonian
            push!(order, CPM_TERM_VOLUME); f64[CPM_TERM_VOLUME] = term_f64(energy_type(state, c))
        elseif c isa CorePotts.UnorderedContactHamiltonian
            push!(order, CPM_TERM_CONTACT); f64[CPM_TERM_CONTACT] = term_f64(energy_type(state, c))
            contact = c.relation
            ntypes = size(c.interactions, 1)
            for i in 1:ntypes, k in 1:ntypes
                J[(i - 1) * ntypes + k] = Float64(c.interactions[i, k])
            end
            table_owner = collect(c.medium_types.owner_ids)
            table_type = collect(c.medium_types.type_ids)
        elseif c isa CorePotts.QuadraticBoundaryHamiltonian
            push!(order, CPM_TERM_BOUNDARY); f64[CPM_TERM_BOUNDARY] = term_f64(energy_type(state, c))
        else
            throw(ArgumentError("the MI355X engine does not evaluate $(typeof(c))"))
        end
    end
    isempty(integrator.components.mechanics) || throw(ArgumentError(
        "this binding forwards no mechanical components; use the Python host or extend mechanics_desc"))
    media = collect(UInt32.(state.potts.descriptor.medium_ids))
    rd = state.retirement_defaults
    no_mech = CpmMechanics(Int32(0), Int32(0), UInt32(0), 0f0, Int32(0), 0f0)
    return CpmModelDesc(
        Int32(1), algorithm_id, Int32(N), pad(Int32.(dims), 3, Int32(1)),
        ntuple(a -> a <= N ? (face_kind(bounds[a].negative), face_kind(bounds[a].positive)) :
            (CPM_FACE_PERIODIC, CPM_FACE_PERIODIC), 3),
        ntuple(a -> a <= N ? (face_medium(bounds[a].negative), face_medium(bounds[a].positive)) :
            (UInt32(0), UInt32(0)), 3),
        has_obstacles ? pointer(mask) : Ptr{UInt8}(C_NULL),
        has_obstacles ? pointer(tags) : Ptr{UInt8}(C_NULL),
        has_obstacles ? pointer(ids) : Ptr{UInt32}(C_NULL),
        relation_desc(integrator.proposal_relation), relation_desc(contact),
        relation_desc(state.boundary_tracker.relation),
        Int32(length(order)), pad(order, 3, Int32(0)),
        get(f64, CPM_TERM_VOLUME, Int32(0)), get(f64, CPM_TERM_CONTACT, Int32(0)),
        get(f64, CPM_TERM_BOUNDARY, Int32(0)),
        Int32(ntypes), pad(J, MAX_TYPES * MAX_TYPES, 0.0),
        Int32(length(table_owner)), pad(table_owner, MAX_MEDIA, UInt32(0)), pad(table_type, MAX_MEDIA, UInt32(0)),
        Int32(length(media)), pad(media, MAX_MEDIA, UInt32(0)),
        Float32(integrator.algorithm.temperature), UInt32(length(state.potts.storage.active)),
        integrator.seed,
        Float64(get(rd, :target_volume, 0.0)), Float64(get(rd, :volume_strength, 0.0)),
        Int64(get(rd, :target_boundary, 0)), Float64(get(rd, :boundary_strength, 0.0)),
        Int32(0), (no_mech, no_mech), (Int32(0), Int32(0), Int32(0)), Int32(0))
end

# ---------------------------------------------------------------------------------------
# Workspace: the library context, created at the first MCS
# ---------------------------------------------------------------------------------------
mutable struct MI355XWorkspace
    algorithm::Int32
    ctx::Ptr{Cvoid}
    seed::UInt64                 # seed the context was last given
    initialize_mechanics::Bool   # init_scientific asked for _initialize_mechanics!
end

function MI355XWorkspace(algorithm::Int32)
    ws = MI355XWorkspace(algorithm, C_NULL, UInt64(0), false)
    finalizer(w -> w.ctx == C_NULL || ccall((:cpm_destroy, LIB), Int32, (Ptr{Cvoid},), w.ctx), ws)
    return ws
end

function check(rc::Int32, ctx::Ptr{Cvoid})
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:cpm_last_error, LIB), Cstring, (Ptr{Cvoid},), ctx))
    throw(rc == 2 ? ArgumentError("unsupported by the MI355X engine: " * msg) : ErrorException("cpm error $rc: " * msg))
end

_allocate_checkerboard_workspace(state, proposal_relation, accesses, plan::MI355XPlan) =
    MI355XWorkspace(CPM_ALG_CHECKERBOARD)
_allocate_lottery_workspace(state, proposal_relation, accesses, plan::MI355XPlan) =
    MI355XWorkspace(CPM_ALG_LOTTERY)

const MI355XIntegrator{A} = CorePotts.ScientificPottsIntegrator{
    S, C, R, A, G, P, W, M, MI355XWorkspace} where {S, C, R, G, P, W, M}

function _initialize_mechanics!(integrator::MI355XIntegrator)
    integrator.algorithm_workspace.initialize_mechanics = !isempty(integrator.components.mechanics)
    return integrator
end

# The four property columns of the volume and boundary energies, keyed as the components name
# them, as the C-ABI's Float64 / Int64 host arrays (nothing when the model lacks the term).
function property_columns(integrator)
    keys = Dict{Symbol, Symbol}()
    for c in integrator.components.energies
        if c isa CorePotts.QuadraticVolumeHamiltonian
            keys[:tv] = target_key(c); keys[:vs] = strength_key(c)
        elseif c isa CorePotts.QuadraticBoundaryHamiltonian
            keys[:tb] = target_key(c); keys[:bs] = strength_key(c)
        end
    end
    col(k, T) = haskey(keys, k) ? Vector{T}(vec(getproperty(columns(integrator.state), keys[k]).data)) : nothing
    return (tv = col(:tv, Float64), vs = col(:vs, Float64), tb = col(:tb, Int64), bs = col(:bs, Float64)), keys
end
ptr(x::Nothing, ::Type{T}) where {T} = Ptr{T}(C_NULL)
ptr(x::Vector{T}, ::Type{T}) where {T} = pointer(x)

# Create the context and upload the host-mirror state (first MCS, or after a state change).
function ensure_context!(integrator)
    ws = integrator.algorithm_workspace
    ws.ctx != C_NULL && return ws
    keep = Any[]
    desc = Ref(model_descriptor(integrator, ws.algorithm, keep))
    ctx = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve keep check(ccall((:cpm_create, LIB), Int32, (Ref{CpmModelDesc}, Ref{Ptr{Cvoid}}), desc, ctx), C_NULL)
    ws.ctx = ctx[]
    st = integrator.state
    s = st.potts.storage
    p, _ = property_columns(integrator)
    GC.@preserve p check(ccall((:cpm_upload_state, LIB), Int32,
        (Ptr{Cvoid}, Ptr{UInt8}, Ptr{UInt32}, Ptr{Int32}, Ptr{Int32}, Ptr{Int64}, Ptr{UInt8},
         Ptr{UInt32}, Ptr{Float64}, Ptr{Float64}, Ptr{Int64}, Ptr{Float64}),
        ws.ctx, s.ownership.tags.data, s.ownership.ids.data, st.trackers.finite_volumes.data,
        st.trackers.medium_volumes.data, st.trackers.boundary_measures.data, s.active.data,
        s.cell_types.data, ptr(p.tv, Float64), ptr(p.vs, Float64), ptr(p.tb, Int64), ptr(p.bs, Float64)), ws.ctx)
    check(ccall((:cpm_set_mcs, LIB), Int32, (Ptr{Cvoid}, UInt64), ws.ctx, integrator.mcs), ws.ctx)
    ws.initialize_mechanics && check(ccall((:cpm_initialize_mechanics, LIB), Int32, (Ptr{Cvoid},), ws.ctx), ws.ctx)
    ws.seed = integrator.seed
    return ws
end

function step_one!(integrator)
    ws = ensure_context!(integrator)
    if ws.seed != integrator.seed
        check(ccall((:cpm_set_seed, LIB), Int32, (Ptr{Cvoid}, UInt64), ws.ctx, integrator.seed), ws.ctx)
        ws.seed = integrator.seed
    end
    check(ccall((:cpm_step, LIB), Int32, (Ptr{Cvoid}, UInt64), ws.ctx, UInt64(1)), ws.ctx)
    integrator.plan.metrics.launches += 1
    integrator.mcs += UInt64(1)
    return integrator
end

perform_scientific_mcs!(integrator::MI355XIntegrator{A}, ::CheckerboardSweepCPM) where {A <: CheckerboardSweepCPM} =
    step_one!(integrator)
perform_scientific_mcs!(integrator::MI355XIntegrator{A}, ::LotteryCPM) where {A <: LotteryCPM} =
    step_one!(integrator)

function report(integrator)
    integrator.mcs > 0 || return nothing
    ws = ensure_context!(integrator)
    out = zeros(UInt64, 12)
    check(ccall((:cpm_report, LIB), Int32, (Ptr{Cvoid}, Ptr{UInt64}), ws.ctx, out), ws.ctx)
    CorePotts.synchronize_observation!(integrator.plan)
    return CorePotts.ScientificMCSReport(out...)
end

_current_mcs_report(integrator::MI355XIntegrator{A}, ::CheckerboardSweepCPM) where {A <: CheckerboardSweepCPM} =
    report(integrator)
_current_mcs_report(integrator::MI355XIntegrator{A}, ::LotteryCPM) where {A <: LotteryCPM} =
    report(integrator)

# logical_state: the library's state downloaded into the host mirrors, then the reference's own
# snapshot of them (sequential.jl:630-632, execution/contracts.jl:420-431).
function logical_state(integrator::MI355XIntegrator)
    ws = ensure_context!(integrator)
    st = integrator.state
    s = st.potts.storage
    p, keys = property_columns(integrator)
    GC.@preserve p check(ccall((:cpm_download_state, LIB), Int32,
        (Ptr{Cvoid}, Ptr{UInt8}, Ptr{UInt32}, Ptr{Int32}, Ptr{Int32}, Ptr{Int64}, Ptr{UInt8},
         Ptr{UInt32}, Ptr{Float64}, Ptr{Float64}, Ptr{Int64}, Ptr{Float64}),
        ws.ctx, s.ownership.tags.data, s.ownership.ids.data, st.trackers.finite_volumes.data,
        st.trackers.medium_volumes.data, st.trackers.boundary_measures.data, s.active.data,
        s.cell_types.data, ptr(p.tv, Float64), ptr(p.vs, Float64), ptr(p.tb, Int64), ptr(p.bs, Float64)), ws.ctx)
    for (k, v) in pairs(p)  # retirement rewrites property columns: copy them back in their own types
        v === nothing && continue
        dst = getproperty(columns(st), keys[k]).data
        dst .= v
    end
    return CorePotts.logical_snapshot(integrator.plan, st.potts)
end

# Device-resident snapshots (cpm_snapshot_*): the state, MCS counter and seed of the integrator's
# context kept in HBM. restore! continues exactly with device copies only; the canonical
# checkpoint (persistence/scientific.jl:250-292) stays the host capture of CorePotts itself.
mutable struct MI355XSnapshot
    handle::Ptr{Cvoid}
    mcs::UInt64
    seed::UInt64
end

function snapshot(integrator::MI355XIntegrator)
    ws = ensure_context!(integrator)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:cpm_snapshot_capture, LIB), Int32, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}), ws.ctx, h), ws.ctx)
    mcs, seed, nbytes = Ref{UInt64}(0), Ref{UInt64}(0), Ref{UInt64}(0)
    check(ccall((:cpm_snapshot_info, LIB), Int32, (Ptr{Cvoid}, Ptr{UInt64}, Ptr{UInt64}, Ptr{UInt64}),
                h[], mcs, seed, nbytes), C_NULL)
    snap = MI355XSnapshot(h[], mcs[], seed[])
    finalizer(s -> s.handle == C_NULL || ccall((:cpm_snapshot_destroy, LIB), Cvoid, (Ptr{Cvoid},), s.handle), snap)
    return snap
end

function restore!(integrator::MI355XIntegrator, snap::MI355XSnapshot)
    ws = ensure_context!(integrator)
    check(ccall((:cpm_snapshot_restore, LIB), Int32, (Ptr{Cvoid}, Ptr{Cvoid}), ws.ctx, snap.handle), ws.ctx)
    integrator.mcs = snap.mcs
    integrator.seed = snap.seed
    ws.seed = snap.seed
    return integrator
end

end # module
