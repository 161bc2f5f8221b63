"""
    LBMHIP

Julia binding of liblbm_hip.so (include/lbm_hip.h): the collide-stream step of
LatticeBoltzmann.jl on one MI355X.  `HipLattice` is a device population array
of the reference's layout (nx, ny, Q); methods of the reference's own
`collide!`, `stream!`, `apply!` and `simulate` for it call the C ABI with
`ccall`, so reference code runs unchanged once its arrays are `HipLattice`s.

    using LatticeBoltzmann, LBMHIP
    f = HipLattice(q, Array(f_host))               # upload
    collide!(SRT(0.8), q, f, g)                    # srt.jl:18-24
    stream!(q, g, f)                               # stream.jl:18-19
    Array(f)                                       # download
"""
module LBMHIP

using LatticeBoltzmann
using LatticeBoltzmann: Quadrature, CollisionModel, SRT, TRT, MRT, BoundaryCondition,
    BounceBack, MovingWall, North, East, South, West, LatticeBoltzmannModel
import LatticeBoltzmann: collide!, stream!, apply!, simulate
import ChainRulesCore

export HipLattice, simulate_loss, adjoint_gradient, initialize_on_device

const libpath = get(ENV, "LBMHIP_LIB", joinpath(@__DIR__, "..", "lib", "liblbm_hip.so"))

const STORAGE = Dict(:f64 => 0, :f32 => 1, :f16dev => 2, :bf16dev => 3, :f16neq => 4,
                     :bf16neq => 5)
const FORCE_NONE, FORCE_CONST, FORCE_FIELD = 0, 1, 2

function check(rc::Cint, what)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:lbm_last_error, libpath), Cstring, ()))
    rc == -1 && throw(ArgumentError("$what: $msg"))
    throw(ErrorException("$what: $msg"))
end

quadrature_id(q::Quadrature) = length(q.weights) == 25 ? 21 : length(q.weights)

# ---- device population arrays ----------------------------------------------
mutable struct HipLattice{Q <: Quadrature} <: AbstractArray{Float64, 3}
    handle::Ptr{Cvoid}
    q::Q
    dims::NTuple{3, Int}
    storage::Symbol
    function HipLattice(q::Q, nx::Integer, ny::Integer; storage::Symbol = :f64,
                        rho0::Real = 1.0) where {Q <: Quadrature}
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:lbm_field_create, libpath), Cint,
                    (Ptr{Ptr{Cvoid}}, Cint, Cint, Cint, Cint, Cdouble),
                    h, quadrature_id(q), nx, ny, STORAGE[storage], rho0), "lbm_field_create")
        a = new{Q}(h[], q, (Int(nx), Int(ny), length(q.weights)), storage)
        finalizer(a) do x
            x.handle == C_NULL || ccall((:lbm_field_destroy, libpath), Cint, (Ptr{Cvoid},),
                                        x.handle)
            x.handle = C_NULL
        end
    end
end

function HipLattice(q::Quadrature, f::Array{Float64, 3}; kw...)
    a = HipLattice(q, size(f, 1), size(f, 2); kw...)
    copyto!(a, f)
end

Base.size(a::HipLattice) = a.dims
Base.IndexStyle(::Type{<:HipLattice}) = IndexLinear()
Base.getindex(::HipLattice, ::Int) =
    error("HipLattice lives in GPU memory: use Array(f) to read populations")
Base.similar(a::HipLattice) = HipLattice(a.q, a.dims[1], a.dims[2]; storage = a.storage)
Base.copy(a::HipLattice) = copyto!(similar(a), a)

function Base.copyto!(a::HipLattice, f::Array{Float64, 3})
    size(f) == a.dims || throw(DimensionMismatch("$(size(f)) vs $(a.dims)"))
    GC.@preserve f check(ccall((:lbm_field_upload, libpath), Cint, (Ptr{Cvoid}, Ptr{Cdouble}),
                               a.handle, f), "lbm_field_upload")
    a
end

function Base.copyto!(f::Array{Float64, 3}, a::HipLattice)
    size(f) == a.dims || throw(DimensionMismatch("$(size(f)) vs $(a.dims)"))
    GC.@preserve f check(ccall((:lbm_field_download, libpath), Cint,
                               (Ptr{Cvoid}, Ptr{Cdouble}), a.handle, f), "lbm_field_download")
    f
end

function Base.copyto!(dst::HipLattice, src::HipLattice)
    check(ccall((:lbm_field_copy, libpath), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), dst.handle,
                src.handle), "lbm_field_copy")
    dst
end

Base.Array(a::HipLattice) = copyto!(Array{Float64, 3}(undef, a.dims), a)

# ---- model handles: one per quadrature and thread ----------------------------
mutable struct NativeModel
    handle::Ptr{Cvoid}
    key::Any
end

const MODELS = Dict{Tuple{Int, Int}, NativeModel}()

function native_model(q::Quadrature)
    get!(MODELS, (quadrature_id(q), Threads.threadid())) do
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:lbm_model_create, libpath), Cint, (Ptr{Ptr{Cvoid}}, Cint), h,
                    quadrature_id(q)), "lbm_model_create")
        NativeModel(h[], nothing)
    end
end

set_collision!(m::NativeModel, cm::SRT) =
    check(ccall((:lbm_model_set_srt, libpath), Cint, (Ptr{Cvoid}, Cdouble), m.handle, cm.τ),
          "lbm_model_set_srt")
set_collision!(m::NativeModel, cm::TRT) =
    check(ccall((:lbm_model_set_trt, libpath), Cint, (Ptr{Cvoid}, Cdouble, Cdouble), m.handle,
                cm.τ_symmetric, cm.τ_asymmetric), "lbm_model_set_trt")
function set_collision!(m::NativeModel, cm::MRT)
    τs = Vector{Float64}(cm.τs)
    GC.@preserve τs check(ccall((:lbm_model_set_mrt, libpath), Cint,
                                (Ptr{Cvoid}, Ptr{Cdouble}, Cint), m.handle, τs, length(τs)),
                          "lbm_model_set_mrt")
end

# The reference's force is a per-node closure (x_idx, y_idx, t) -> F; it is
# evaluated here on the host and handed over as a FIELD buffer.
function set_force!(m::NativeModel, cm::CollisionModel, nx, ny, time)
    force = cm.force
    if force === nothing
        return check(ccall((:lbm_model_set_force, libpath), Cint,
                           (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Cint, Cint),
                           m.handle, FORCE_NONE, C_NULL, 0, 0), "lbm_model_set_force")
    end
    F = Array{Float64, 3}(undef, nx, ny, 2)
    for y in 1:ny, x in 1:nx
        v = force(x, y, time)
        F[x, y, 1], F[x, y, 2] = v[1], v[2]
    end
    GC.@preserve F check(ccall((:lbm_model_set_force, libpath), Cint,
                               (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Cint, Cint),
                               m.handle, FORCE_FIELD, F, nx, ny), "lbm_model_set_force")
end

# lbm_bc (include/lbm_hip.h)
struct CBC
    kind::Int32
    direction::Int32
    x0::Int32
    x1::Int32
    y0::Int32
    y1::Int32
    ux::Cdouble
    uy::Cdouble
    rho::Cdouble
end

direction_id(::North) = 0
direction_id(::East) = 1
direction_id(::South) = 2
direction_id(::West) = 3

cbc(bc::BounceBack) = CBC(0, direction_id(bc.direction), first(bc.xs), last(bc.xs),
                          first(bc.ys), last(bc.ys), 0.0, 0.0, 1.0)
cbc(bc::MovingWall) = CBC(1, direction_id(bc.direction), first(bc.xs), last(bc.xs),
                          first(bc.ys), last(bc.ys), bc.u[1], bc.u[2], bc.ρ)

function set_bcs!(m::NativeModel, bcs)
    list = CBC[cbc(bc) for bc in bcs]
    GC.@preserve list check(ccall((:lbm_model_set_bcs, libpath), Cint,
                                  (Ptr{Cvoid}, Ptr{CBC}, Cint), m.handle, list, length(list)),
                            "lbm_model_set_bcs")
end

# ---- the reference's operators, for device arrays ---------------------------
function collide!(cm::CollisionModel, q::Quadrature, f_in::HipLattice, f_out::HipLattice;
                  time = 0.0)
    m = native_model(q)
    set_collision!(m, cm)
    set_force!(m, cm, f_in.dims[1], f_in.dims[2], time)
    check(ccall((:lbm_collide, libpath), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                m.handle, f_in.handle, f_out.handle), "lbm_collide")
    nothing
end

collide!(cm::CollisionModel, q::Quadrature; time = 0.0, f_new::HipLattice, f_old::HipLattice) =
    collide!(cm, q, f_old, f_new; time = time)

function stream!(q::Quadrature, f::HipLattice, f_new::HipLattice)
    check(ccall((:lbm_stream, libpath), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                native_model(q).handle, f.handle, f_new.handle), "lbm_stream")
    nothing
end

stream!(q::Quadrature; f_new::HipLattice, f_old::HipLattice) = stream!(q, f_old, f_new)

function apply!(bcs::AbstractVector{<:BoundaryCondition}, q::Quadrature, f_new::HipLattice,
                f_old::HipLattice; time = 0.0)
    m = native_model(q)
    set_bcs!(m, bcs)
    check(ccall((:lbm_apply_bcs, libpath), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                m.handle, f_new.handle, f_old.handle), "lbm_apply_bcs")
    nothing
end

# Steps t at which next!(pm, q, f, t) reads the populations.  At every other
# step it returns false without touching them, so simulate may skip the call:
# TrackHydrodynamicErrors checks its stop criteria at multiples of 100 and
# records a row at n_steps, or at every step with should_process
# (track_hydrodynamic_errors.jl:52-66); any other processing method is called
# at every step.
function reads_state(pm, t)
    if pm isa LatticeBoltzmann.TrackHydrodynamicErrors
        return pm.should_process || mod(t, 100) == 0 || t == pm.n_steps
    end
    return true
end

processing_view(pm, f::HipLattice) =
    pm isa LatticeBoltzmann.TrackHydrodynamicErrors ? f : Array(f)

# simulate(model, time) (lattice_boltzmann_model.jl:60-77) with device arrays.
# Step t collides and applies the BCs at time t * Δt, Δt = delta_t(problem) of
# the processing method (0 without one), and is followed by next!(model, t + 1);
# the loop ends with the reference's closing next!(model, last(time) + 1).
# Runs of steps between the points where next! reads the state go to one
# lbm_simulate call (one fused collide-stream kernel per step).  A model with
# a force closure advances one step per call instead, the closure evaluated at
# that step's time, as the reference's per-node force(x, y, t * Δt)
# (srt.jl:52, trt.jl:77, mrt.jl:92) is.
function simulate(model::LatticeBoltzmannModel{<:Quadrature, <:HipLattice}, time)
    pm = model.processing_method
    q = model.quadrature
    Δt = isdefined(pm, :problem) ? LatticeBoltzmann.delta_t(pm.problem) : 0.0
    m = native_model(q)
    cm = model.collision_model
    set_collision!(m, cm)
    nx, ny, _ = size(model.f_stream)
    set_bcs!(m, model.boundary_conditions)
    forced = hasproperty(cm, :force) && cm.force !== nothing
    forced || set_force!(m, cm, nx, ny, 0.0)
    n = length(time)
    i = 1
    while i <= n
        j = i
        if !forced
            while j < n && !reads_state(pm, time[j] + 1)
                j += 1
            end
        else
            set_force!(m, cm, nx, ny, time[i] * Δt)
        end
        check(ccall((:lbm_simulate, libpath), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64),
                    m.handle, model.f_stream.handle, model.f_collision.handle, j - i + 1),
              "lbm_simulate")
        t_next = time[j] + 1
        if reads_state(pm, t_next) &&
           LatticeBoltzmann.next!(pm, q, processing_view(pm, model.f_stream), t_next)
            return model
        end
        i = j + 1
    end
    LatticeBoltzmann.next!(pm, q, processing_view(pm, model.f_stream), last(time) + 1)
    model
end

# ---- iterative initialization (Mei et al.) ----------------------------------
# initialize(::IterativeInitializationMeiEtAl, q, problem) (mei_et_al.jl:13-41)
# on the device: the collide-only model of the prescribed lattice velocity
# (lbm_model_set_mei), stream! and the problem's BCs, stopping on the
# DensityConvergence check of node (nx, ny); returns the HipLattice f_stream.
function initialize_on_device(strategy::LatticeBoltzmann.IterativeInitializationMeiEtAl,
                              q::Quadrature, problem; max_steps = 10000)
    nx, ny = problem.NX, problem.NY
    x_range, y_range = LatticeBoltzmann.range(problem)
    u = Array{Float64}(undef, nx, ny, 2)
    for x_idx in 1:nx, y_idx in 1:ny
        v = LatticeBoltzmann.lattice_velocity(q, problem, x_range[x_idx], y_range[y_idx])
        u[x_idx, y_idx, 1], u[x_idx, y_idx, 2] = v[1], v[2]
    end
    m = native_model(q)
    GC.@preserve u check(ccall((:lbm_model_set_mei, libpath), Cint,
                               (Ptr{Cvoid}, Cdouble, Ptr{Cdouble}, Cint, Cint),
                               m.handle, strategy.τ, u, nx, ny), "lbm_model_set_mei")
    set_bcs!(m, LatticeBoltzmann.boundary_conditions(problem))
    fs, fc = HipLattice(q, nx, ny), HipLattice(q, nx, ny)
    steps = Ref{Int64}(0)
    check(ccall((:lbm_mei_initialize, libpath), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Int64, Ptr{Int64}),
                m.handle, fs.handle, fc.handle, strategy.ϵ, max_steps, steps),
          "lbm_mei_initialize")
    fs
end

# initialize(AnalyticalEquilibrium() / AnalyticalEquilibriumAndOffEquilibrium(),
# q, problem) (initial_conditions.jl:7-22) with the per-node arithmetic on the
# device: the problem's own field functions are evaluated here, node by node,
# as initial_condition evaluates them (problems.jl:98-106), and
# lbm_initialize_analytical forms the Hermite-sum equilibrium and the
# off-equilibrium part (analytical_equilibrium.jl:8-17,
# analytical_offequilibrium.jl:10-87).
const INIT_EQUILIBRIUM, INIT_OFFEQ_TGV, INIT_OFFEQ = 0, 1, 2

function initialize_on_device(strategy::Union{LatticeBoltzmann.AnalyticalEquilibrium,
                                              LatticeBoltzmann.AnalyticalEquilibriumAndOffEquilibrium},
                              q::Quadrature, problem; storage::Symbol = :f64)
    nx, ny = problem.NX, problem.NY
    x_range, y_range = LatticeBoltzmann.range(problem)
    ρ, ux, uy, T = (Array{Float64}(undef, nx, ny) for _ in 1:4)
    offeq = strategy isa LatticeBoltzmann.AnalyticalEquilibriumAndOffEquilibrium
    G = offeq ? Array{Float64}(undef, nx, ny, 4) : nothing
    tgv = problem isa LatticeBoltzmann.TGV
    for x_idx in 1:nx, y_idx in 1:ny
        x, y = x_range[x_idx], y_range[y_idx]
        ρ[x_idx, y_idx] = LatticeBoltzmann.lattice_density(q, problem, x, y)
        u = LatticeBoltzmann.lattice_velocity(q, problem, x, y)
        ux[x_idx, y_idx], uy[x_idx, y_idx] = u[1], u[2]
        T[x_idx, y_idx] = LatticeBoltzmann.lattice_temperature(q, problem, x, y)
        if offeq
            ∇u = tgv ? LatticeBoltzmann.velocity_gradient(problem, x, y, 0.0) :
                 problem.u_max^2 * LatticeBoltzmann.velocity_gradient(problem, x, y, 0.0)
            for k in 1:4
                G[x_idx, y_idx, k] = ∇u[k]
            end
        end
    end
    mode = !offeq ? INIT_EQUILIBRIUM : (tgv ? INIT_OFFEQ_TGV : INIT_OFFEQ)
    τ = q.speed_of_sound_squared * LatticeBoltzmann.lattice_viscosity(problem)
    factor = problem.domain_size[1] * problem.domain_size[2]
    f = HipLattice(q, nx, ny; storage = storage)
    Gp = offeq ? G : C_NULL
    GC.@preserve ρ ux uy T G check(ccall((:lbm_initialize_analytical, libpath), Cint,
                                         (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble},
                                          Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble),
                                         f.handle, mode, ρ, ux, uy, T, Gp, τ, factor),
                                   "lbm_initialize_analytical")
    f
end

# ---- on-device processing ---------------------------------------------------
# next!(::TrackHydrodynamicErrors) (track_hydrodynamic_errors.jl:52-222) and
# should_stop! (stopping_criteria.jl:17-115) on a HipLattice: one kernel pass
# each (lbm_track_errors, lbm_velocity_stats); the populations stay on the GPU.
struct CTrackParams
    tau::Cdouble
    u_max::Cdouble
    delta::Cdouble
    affine::NTuple{16, Cdouble}
end

mutable struct ExpectedFields
    handle::Ptr{Cvoid}
end

function ExpectedFields(planes::Array{Float64, 3})
    nx, ny, k = size(planes)
    k == 8 || throw(ArgumentError("8 expected planes needed, got $k"))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:lbm_expected_create, libpath), Cint, (Ptr{Ptr{Cvoid}}, Cint, Cint), h, nx, ny),
          "lbm_expected_create")
    e = finalizer(ExpectedFields(h[])) do x
        ccall((:lbm_expected_destroy, libpath), Cint, (Ptr{Cvoid},), x.handle)
    end
    GC.@preserve planes check(ccall((:lbm_expected_upload, libpath), Cint,
                                    (Ptr{Cvoid}, Ptr{Cdouble}), e.handle, planes),
                              "lbm_expected_upload")
    e
end

# the analytic fields of the problem at `time`, in the planes' order
function expected_planes(q, problem, time)
    x_range, y_range = LatticeBoltzmann.range(problem)
    P = Array{Float64}(undef, length(x_range), length(y_range), 8)
    for x_idx in eachindex(x_range), y_idx in eachindex(y_range)
        x, y = x_range[x_idx], y_range[y_idx]
        u = LatticeBoltzmann.velocity(problem, x, y, time)
        σ = LatticeBoltzmann.deviatoric_tensor(q, problem, x, y, time)
        P[x_idx, y_idx, :] .= (LatticeBoltzmann.density(q, problem, x, y, time), u[1], u[2],
                               LatticeBoltzmann.pressure(q, problem, x, y, time),
                               σ[1, 1], σ[1, 2], σ[2, 1], σ[2, 2])
    end
    P
end

function cell_area(problem)
    x_range, y_range = LatticeBoltzmann.range(problem)
    nx, ny = length(x_range), length(y_range)
    nx == 1 && return ny == 1 ? 1.0 : Float64(y_range.step)
    ny == 1 && return Float64(x_range.step)
    Float64(y_range.step) * Float64(x_range.step)
end

function LatticeBoltzmann.next!(pm::LatticeBoltzmann.TrackHydrodynamicErrors, q,
                                f::HipLattice, t::Int64)
    should_stop = false
    if mod(t, 100) == 0 && LatticeBoltzmann.should_stop!(pm.stop_criteria, q, f)
        @info "Stopping after $t iterations"
        should_stop = true
    end
    (should_stop == false && t != pm.n_steps && !pm.should_process) && return false
    problem = pm.problem
    e = ExpectedFields(expected_planes(q, problem, t * LatticeBoltzmann.delta_t(problem)))
    prm = Ref(CTrackParams(q.speed_of_sound_squared * LatticeBoltzmann.lattice_viscosity(problem),
                           problem.u_max, cell_area(problem),
                           ntuple(k -> isodd(k) ? 0.0 : 1.0, 16)))
    out = zeros(11)
    GC.@preserve out check(ccall((:lbm_track_errors, libpath), Cint,
                                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{CTrackParams}, Ptr{Cdouble}),
                                 f.handle, e.handle, prm, out), "lbm_track_errors")
    push!(pm.df, (timestep = t, error_ρ = out[1], error_u = out[2], error_p = out[3],
                  error_σ_xx = out[4], error_σ_xy = out[5], error_σ_yy = out[6],
                  error_σ_yx = out[7], mass = out[8], momentum = out[9], energy = out[10]))
    should_stop
end

function velocity_stats(f::HipLattice, tracker::Ptr{Cvoid} = C_NULL)
    out = zeros(5)
    GC.@preserve out check(ccall((:lbm_velocity_stats, libpath), Cint,
                                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}), f.handle, tracker, out),
                           "lbm_velocity_stats")
    out
end

function LatticeBoltzmann.should_stop!(c::LatticeBoltzmann.MeanVelocityStoppingCriteria, q,
                                       f::HipLattice)
    u_mean = velocity_stats(f)[1]
    abs(u_mean / c.old_mean_velocity - 1) < c.tolerance && return true
    isnan(u_mean) && (@warn "nan in velocity profile"; return true)
    c.old_mean_velocity = u_mean
    false
end

# the previous velocity of each VelocityConvergenceStoppingCriteria, on the device
const TRACKERS = IdDict{Any, Ptr{Cvoid}}()

function LatticeBoltzmann.should_stop!(c::LatticeBoltzmann.VelocityConvergenceStoppingCriteria,
                                       q, f::HipLattice)
    tr = get!(TRACKERS, c) do
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:lbm_velocity_tracker_create, libpath), Cint, (Ptr{Ptr{Cvoid}}, Cint, Cint),
                    h, f.dims[1], f.dims[2]), "lbm_velocity_tracker_create")
        h[]
    end
    s = velocity_stats(f, tr)
    converged = sqrt(s[3]) / s[4]
    converged < c.tolerance && return true
    isnan(converged) && (@warn "nan in velocity profile"; return true)
    false
end

# ---- reverse mode -----------------------------------------------------------
# L(u0, τ) = 1/2 Σ |u(S_T) - u_target|^2, S_0 = feq(ρ0, u0), T steps of
# collide! -> stream! on a periodic lattice (lbm_adjoint_*).  u0, u_target:
# (nx, ny, 2) arrays.  τ: a number (SRT) or (τ_s, τ_a) (TRT).
function with_adjoint(f, q, u0, τ, nsteps; storage = :f64, rho0 = 1.0, checkpoint_every = 0)
    m = native_model(q)
    cm = τ isa Number ? SRT(τ) : TRT(τ[1], τ[2], nothing)
    set_collision!(m, cm)
    set_force!(m, cm, size(u0, 1), size(u0, 2), 0.0)
    set_bcs!(m, BoundaryCondition[])
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:lbm_adjoint_create, libpath), Cint,
                (Ptr{Ptr{Cvoid}}, Ptr{Cvoid}, Cint, Cint, Cint, Cdouble, Int64, Cint),
                h, m.handle, STORAGE[storage], size(u0, 1), size(u0, 2), rho0, nsteps,
                checkpoint_every), "lbm_adjoint_create")
    try
        f(h[])
    finally
        ccall((:lbm_adjoint_destroy, libpath), Cint, (Ptr{Cvoid},), h[])
    end
end

function simulate_loss(q::Quadrature, u0::Array{Float64, 3}, τ, u_target::Array{Float64, 3},
                       nsteps::Integer; kw...)
    with_adjoint(q, u0, τ, nsteps; kw...) do h
        L = Ref(0.0)
        GC.@preserve u0 u_target check(ccall((:lbm_adjoint_loss, libpath), Cint,
                                             (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble},
                                              Ptr{Cdouble}, Ptr{Cdouble}),
                                             h, u0, u_target, L, C_NULL), "lbm_adjoint_loss")
        L[]
    end
end

function adjoint_gradient(q::Quadrature, u0::Array{Float64, 3}, τ, u_target::Array{Float64, 3},
                          nsteps::Integer; kw...)
    with_adjoint(q, u0, τ, nsteps; kw...) do h
        L = Ref(0.0)
        g = similar(u0)
        gτ = zeros(2)
        GC.@preserve u0 u_target g gτ check(ccall((:lbm_adjoint_gradient, libpath), Cint,
                                                  (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble},
                                                   Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble},
                                                   Ptr{Cdouble}),
                                                  h, u0, u_target, L, g, gτ, C_NULL),
                                            "lbm_adjoint_gradient")
        L[], g, τ isa Number ? gτ[1] : (gτ[1], gτ[2])
    end
end

function ChainRulesCore.rrule(::typeof(simulate_loss), q::Quadrature, u0, τ, u_target,
                              nsteps::Integer; kw...)
    L, g, gτ = adjoint_gradient(q, u0, τ, u_target, nsteps; kw...)
    function simulate_loss_pullback(ΔL)
        NoT = ChainRulesCore.NoTangent()
        return (NoT, NoT, ΔL .* g, τ isa Number ? ΔL * gτ : (ΔL * gτ[1], ΔL * gτ[2]), NoT, NoT)
    end
    return L, simulate_loss_pullback
end

end # module
