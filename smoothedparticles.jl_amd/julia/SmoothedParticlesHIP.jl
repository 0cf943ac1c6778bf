# SmoothedParticlesHIP -- ccall binding of libsph_hip (include/sph.h) that keeps
# the call surface of SmoothedParticles.jl's hot path:
#
#   ParticleSystem(T, domain, h)        src/structs.jl:57-92
#   sys.particles, push!, ParticleField src/structs.jl:118-125, src/grids.jl:253-258
#   create_cell_list!(sys)              src/core.jl:51-90
#   apply!(sys, op!; self, params...)   src/core.jl:151-161
#   apply_binary!/apply_unary!          src/core.jl:125-142
#
# The particle state lives on the device between calls.  `sys.particles`
# downloads it into a host Vector{T} of the user's particle type and marks it
# dirty; the next device call uploads it again, so reference code that edits
# particles in a loop (examples/collapse_dry.jl:97-100) keeps working.
#
# Operators are registered per reference program, by the name the program
# gives them: several programs define find_pressure!, internal_force!, move!
# and accelerate! with different bodies (examples/collapse_dry.jl:123-159,
# tests/test_collision_2d.jl:71-94), so the system is told which program it
# runs (`program = :collapse_dry`).  A function that is not registered runs
# through the host fallback: the device finds the neighbour pairs, the
# closure runs over the host mirror in the reference's visiting order, and
# the result is uploaded.
#
# No Julia toolchain exists where this library is built, so this file is
# source only and has not been executed.
module SmoothedParticlesHIP

import SmoothedParticles
using SmoothedParticles: AbstractParticle, Shape, boundarybox, RealVector

export ParticleSystem, create_cell_list!, apply!, apply_binary!, apply_unary!,
       ParticleField, set_params!, energy, neighbour_pairs, step!, probe_sums

const LIB = joinpath(@__DIR__, "..", "smoothedparticles_hip", "libsph_hip.so")

# sph_params of include/sph.h
struct SphParams
    m::Float64
    rho0::Float64
    c::Float64
    alpha::Float64
    hs::Float64
    mu::Float64
    nu::Float64
    g::NTuple{3,Float64}
    dt::Float64
    coeff::Float64
end

# A registered operator: library id, binary or unary, and the parameters the
# reference function fixes in its body (the half step of a move!).
struct Op
    id::Int32
    binary::Bool
    fixed::Dict{Symbol,Float64}
end
Op(id, binary) = Op(Int32(id), binary, Dict{Symbol,Float64}())

const PROGRAMS = Dict{Symbol,Dict{Symbol,Op}}(
    # examples/collapse_dry.jl
    :collapse_dry => Dict(
        Symbol("balance_of_mass!") => Op(5, true),                           # :112-115
        Symbol("find_pressure!") => Op(20, false),                           # :123-127
        Symbol("internal_force!") => Op(3, true),                            # :135-141
        Symbol("move!") => Op(Int32(22), false, Dict(:coeff => 0.5)),        # :148-153
        Symbol("accelerate!") => Op(Int32(21), false, Dict(:coeff => 0.5))), # :155-159
    # tests/test_collision_2d.jl
    :test_collision_2d => Dict(
        Symbol("find_rho!") => Op(1, true),                                  # :63-65
        Symbol("find_rho0!") => Op(2, true),                                 # :67-69
        Symbol("find_pressure!") => Op(19, false),                           # :71-73
        Symbol("internal_force!") => Op(4, true),                            # :75-78
        Symbol("reset_a!") => Op(17, false),                                 # :80-82
        Symbol("reset_rho!") => Op(16, false),                               # :84-86
        Symbol("move!") => Op(Int32(23), false, Dict(:coeff => 1.0)),        # :88-90
        Symbol("accelerate!") => Op(Int32(24), false, Dict(:coeff => 0.5))), # :92-94
)

# device field ids (SPH_FIELD_*); Dv and a are the same device field
const FIELDS = Dict(:x => 0, :v => 1, :a => 2, :Dv => 2, :rho => 3, :P => 4, :Drho => 5,
                    :rho0 => 6, :type => 7)
const FIELD_KEY = 8

mutable struct ParticleSystem{T}
    handle::Ptr{Cvoid}
    h::Float64
    domain::Shape
    program::Union{Symbol,Nothing}
    params::Dict{Symbol,Any}
    mirror::Vector{T}       # host copy of the particles (valid when synced)
    host_dirty::Bool        # the mirror was handed out and may have been edited
    mirror_valid::Bool      # the mirror equals the device state
end

function check(rc::Cint)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:sph_last_error, LIB), Cstring, ()))
    error("libsph_hip error $(rc): $(msg)")
end

"""
    ParticleSystem(T, domain, h; precision = :fp64, program = nothing)

As src/structs.jl:57-92: `T` is the particle type (a subtype of
AbstractParticle with a field `x::RealVector`), `domain` a Shape whose
bounding box fixes the cell grid, `h` the neighbour radius.  `precision =
:mixed` selects FP32 relative-coordinate pair arithmetic with FP64 state.
`program` names the reference program whose operator names `apply!`
resolves (`:collapse_dry`, `:test_collision_2d`).
"""
function ParticleSystem(T::DataType, domain::Shape, h::Float64;
                        precision::Symbol = :fp64, program = nothing)
    @assert(h > 0.0, "invalid ParticleSystem declaration! (h must be a positive float)")
    @assert(T <: AbstractParticle,
            "invalid ParticleSystem declaration! (" * string(T) * " is not an AbstractParticle subtype)")
    @assert(hasfield(T, :x) && fieldtype(T, :x) == RealVector,
            "invalid ParticleSystem declaration! (particles must have a field `x::RealVector`)")
    program === nothing || haskey(PROGRAMS, program) || error("unknown program $(program)")
    box = boundarybox(domain)
    bmin = (box.x1_min, box.x2_min, box.x3_min)
    bmax = (box.x1_max, box.x2_max, box.x3_max)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    mode = precision == :mixed ? Cint(1) : Cint(0)
    check(ccall((:sph_create, LIB), Cint,
                (Ref{Ptr{Cvoid}}, Cdouble, Ref{NTuple{3,Float64}}, Ref{NTuple{3,Float64}}, Cint),
                out, h, bmin, bmax, mode))
    sys = ParticleSystem{T}(out[], h, box, program, Dict{Symbol,Any}(), T[], false, true)
    finalizer(s -> ccall((:sph_destroy, LIB), Cint, (Ptr{Cvoid},), s.handle), sys)
    return sys
end

device_count(sys::ParticleSystem) = (n = Ref{Int64}(0);
    check(ccall((:sph_count, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}), sys.handle, n)); n[])

# the device fields that particle type T has
fields_of(::Type{T}) where T = [f for f in fieldnames(T) if haskey(FIELDS, f)]

function get_field(sys::ParticleSystem, f::Symbol, n::Int)
    nc = f in (:x, :v, :a, :Dv) ? 3 : 1
    buf = Vector{Float64}(undef, nc * n)
    check(ccall((:sph_get_field, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Int64),
                sys.handle, FIELDS[f], buf, n))
    return buf
end

function set_field(sys::ParticleSystem, f::Symbol, buf::Vector{Float64}, n::Int)
    check(ccall((:sph_set_field, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Int64),
                sys.handle, FIELDS[f], buf, n))
end

# device -> mirror (the mirror keeps the user's structs, their other fields
# untouched; particles removed by a build are dropped from its end after the
# swap-with-end renumbering the device replays)
function download!(sys::ParticleSystem{T}) where T
    sys.mirror_valid && return sys.mirror
    n = device_count(sys)
    # the device holds no particles the mirror has not seen (they enter
    # through push! on the mirror), so it is at least as long
    @assert length(sys.mirror) >= n
    length(sys.mirror) > n && resize!(sys.mirror, n)
    for f in fields_of(T)
        buf = get_field(sys, f, n)
        if f in (:x, :v, :a, :Dv)
            for i in 1:n
                setproperty!(sys.mirror[i], f, RealVector(buf[3i-2], buf[3i-1], buf[3i]))
            end
        else
            for i in 1:n
                setproperty!(sys.mirror[i], f, buf[i])
            end
        end
    end
    sys.mirror_valid = true
    return sys.mirror
end

# mirror -> device (after the user may have edited or pushed particles)
function upload!(sys::ParticleSystem{T}) where T
    sys.host_dirty || return nothing
    n_dev = device_count(sys)
    n = length(sys.mirror)
    if n > n_dev   # push!ed particles: append their positions first
        xs = Float64[c for p in sys.mirror[n_dev+1:n] for c in (p.x[1], p.x[2], p.x[3])]
        check(ccall((:sph_append, LIB), Cint,
                    (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                    sys.handle, n - n_dev, xs, C_NULL, C_NULL, C_NULL, C_NULL))
    end
    for f in fields_of(T)
        buf = f in (:x, :v, :a, :Dv) ?
            Float64[c for p in sys.mirror for c in (getproperty(p, f)[1], getproperty(p, f)[2], getproperty(p, f)[3])] :
            Float64[getproperty(p, f) for p in sys.mirror]
        set_field(sys, f, buf, n)
    end
    sys.host_dirty = false
    return nothing
end

# sys.particles hands out the host mirror (downloaded, and from then on
# assumed edited); every other property is the struct's own.
function Base.getproperty(sys::ParticleSystem, s::Symbol)
    if s === :particles
        download!(sys)
        setfield!(sys, :host_dirty, true)
        return getfield(sys, :mirror)
    end
    return getfield(sys, s)
end

# any call that runs on the device first uploads the host edits and then
# invalidates the mirror
function device_call(f, sys::ParticleSystem)
    upload!(sys)
    r = f()
    sys.mirror_valid = false
    return r
end

"Parameters the registered operators read (m, rho0, c, alpha, hs, mu, nu, g, dt)."
set_params!(sys::ParticleSystem; kw...) = (merge!(sys.params, Dict(kw)); nothing)

function params(sys::ParticleSystem, op::Union{Op,Nothing} = nothing; kw...)
    p = merge(sys.params, Dict(kw))
    op === nothing || merge!(p, op.fixed)
    get0(k) = Float64(get(p, k, 0.0))
    g = get(p, :g, (0.0, 0.0, 0.0))
    SphParams(get0(:m), get0(:rho0), get0(:c), get0(:alpha), get0(:hs), get0(:mu), get0(:nu),
              (Float64(g[1]), Float64(g[2]), Float64(g[3])), get0(:dt), Float64(get(p, :coeff, 1.0)))
end

"create_cell_list!(sys) -- src/core.jl:51-90"
function create_cell_list!(sys::ParticleSystem)
    device_call(sys) do
        n = Ref{Int64}(0)
        check(ccall((:sph_build_cells, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}), sys.handle, n))
        # removals renumber the particles: drop the mirror's tail to the new count
        length(sys.mirror) > n[] && resize!(sys.mirror, n[])
    end
    return nothing
end

# The registered operator of a function in this system's program, or
# nothing (host fallback).
function lookup(sys::ParticleSystem, op!::Function)
    name = nameof(op!)
    if sys.program !== nothing
        return get(PROGRAMS[sys.program], name, nothing)
    end
    found = [ops[name] for ops in values(PROGRAMS) if haskey(ops, name)]
    isempty(found) && return nothing
    all(o -> o.id == found[1].id && o.fixed == found[1].fixed, found) && return found[1]
    error("$(name) names different operators in different reference programs; " *
          "create the system with program = :collapse_dry or :test_collision_2d")
end

# ---------------------------------------------------------- host fallback --

# the accepted pairs (1-based) of the current cell list in the order of
# _apply_binary! (src/core.jl:94-112): by p, by the key_diff entry of q's
# cell (di outermost), by descending q
function visit_order(sys::ParticleSystem)
    n = device_count(sys)
    cnt = Ref{Int64}(0)
    check(ccall((:sph_neighbour_pairs, LIB), Cint, (Ptr{Cvoid}, Ptr{Int64}, Int64, Ref{Int64}),
                sys.handle, C_NULL, 0, cnt))
    raw = Vector{Int64}(undef, 2 * cnt[])
    check(ccall((:sph_neighbour_pairs, LIB), Cint, (Ptr{Cvoid}, Ptr{Int64}, Int64, Ref{Int64}),
                sys.handle, raw, cnt[], cnt))
    keys = Vector{Float64}(undef, n)
    check(ccall((:sph_get_field, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Int64),
                sys.handle, FIELD_KEY, keys, n))
    kp = Ref{NTuple{3,Int64}}((0, 0, 0)); kl = Ref{NTuple{3,Int64}}((0, 0, 0))
    km = Ref{Int64}(0); kd = Ref{NTuple{27,Int64}}(ntuple(_ -> 0, 27)); nd = Ref{Cint}(0)
    check(ccall((:sph_key_geometry, LIB), Cint,
                (Ptr{Cvoid}, Ref{NTuple{3,Int64}}, Ref{NTuple{3,Int64}}, Ref{Int64},
                 Ref{NTuple{27,Int64}}, Ref{Cint}), sys.handle, kp, kl, km, kd, nd))
    diff = collect(kd[][1:nd[]])
    pairs = [(raw[2k-1] + 1, raw[2k] + 1) for k in 1:cnt[]]
    delta(pq) = Int64(keys[pq[2]]) - Int64(keys[pq[1]])
    sort!(pairs, by = pq -> (pq[1], delta(pq), -pq[2]))
    # a pair the device lists m times (aliased grids) takes the m entries of
    # key_diff equal to its key difference, in order
    rank = Vector{Int}(undef, length(pairs))
    k = 1
    while k <= length(pairs)
        e = k
        while e < length(pairs) && pairs[e+1] == pairs[k]
            e += 1
        end
        slots = findall(==(delta(pairs[k])), diff)
        rank[k:e] .= slots[1:e-k+1]
        k = e + 1
    end
    perm = sortperm(collect(zip(first.(pairs), rank, -last.(pairs))))
    return pairs[perm]
end

function host_apply!(sys::ParticleSystem, op!::Function; self::Bool = false, kw...)
    T = eltype(getfield(sys, :mirror))
    ps = sys.particles   # downloaded; marked dirty, so the next device call uploads it
    f = isempty(kw) ? op! : (args...) -> op!(args...; kw...)
    if hasmethod(op!, (T, T, Float64))
        for (p, q) in visit_order(sys)
            f(ps[p], ps[q], SmoothedParticles.dist(ps[p], ps[q]))
        end
        if self
            for p in ps
                f(p, p, 0.0)
            end
        end
    else
        for p in ps
            f(p)
        end
    end
    return nothing
end

# ---------------------------------------------------------------- apply ---

"apply!(sys, op!; self = false, params...) -- src/core.jl:151-161"
function apply!(sys::ParticleSystem, op!::Function; self::Bool = false, kw...)
    op = lookup(sys, op!)
    op === nothing && return host_apply!(sys, op!; self = self, kw...)
    device_call(sys) do
        check(ccall((:sph_apply, LIB), Cint, (Ptr{Cvoid}, Cint, Ref{SphParams}, Cint),
                    sys.handle, op.id, params(sys, op; kw...), self ? Cint(1) : Cint(0)))
    end
end

function apply_binary!(sys::ParticleSystem, op!::Function; kw...)
    op = lookup(sys, op!)
    op === nothing && return host_apply!(sys, op!; kw...)
    device_call(sys) do
        check(ccall((:sph_apply_binary, LIB), Cint, (Ptr{Cvoid}, Cint, Ref{SphParams}, Cint),
                    sys.handle, op.id, params(sys, op; kw...), Cint(0)))
    end
end

function apply_unary!(sys::ParticleSystem, op!::Function; kw...)
    op = lookup(sys, op!)
    op === nothing && return host_apply!(sys, op!; kw...)
    device_call(sys) do
        check(ccall((:sph_apply_unary, LIB), Cint, (Ptr{Cvoid}, Cint, Ref{SphParams}),
                    sys.handle, op.id, params(sys, op; kw...)))
    end
end

"""
    ParticleField(sys, :sym)

Field view in the reference's particle order (src/structs.jl:118-125),
read and written through the host mirror.
"""
struct ParticleField <: AbstractVector{Any}
    sys::ParticleSystem
    var::Symbol
end

Base.size(f::ParticleField) = (length(f.sys.particles),)
Base.getindex(f::ParticleField, i::Int) = getproperty(f.sys.particles[i], f.var)
Base.setindex!(f::ParticleField, val, i::Int) = setproperty!(f.sys.particles[i], f.var, val)

"Total energy (kind 0: examples/collapse_dry.jl:166-171, 1: tests/test_collision_2d.jl:96-100)."
function energy(sys::ParticleSystem, kind::Integer = 0; kw...)
    out = Ref{NTuple{3,Float64}}((0.0, 0.0, 0.0))
    device_call(sys) do
        check(ccall((:sph_reduce, LIB), Cint, (Ptr{Cvoid}, Cint, Ref{SphParams}, Ref{NTuple{3,Float64}}),
                    sys.handle, kind, params(sys; kw...), out))
    end
    return out[][1]
end

"""SmoothedParticles.sum(sys, func, x) (src/core.jl:240-258) at probe points
`xs` for the closures of examples/cavity_flow.jl:166-172: a 4 x np matrix
whose column i holds sum (phi m) W and the three sums (phi m) v_k W."""
function probe_sums(sys::ParticleSystem, xs::AbstractVector; fluid_only::Bool = false, kw...)
    np = length(xs)
    pts = Float64[c for x in xs for c in (x[1], x[2], x[3])]
    out = zeros(Float64, 4, np)
    device_call(sys) do
        check(ccall((:sph_probe_sum, LIB), Cint,
                    (Ptr{Cvoid}, Ref{SphParams}, Cint, Ptr{Float64}, Int64, Ptr{Float64}),
                    sys.handle, params(sys; kw...), fluid_only ? 1 : 0, pts, np, out))
    end
    return out
end

"""
    step!(sys, nsteps; scheme = 1, params...)

nsteps fused device steps.  scheme 0/1: density sum, EOS (constant rho0 or
the rho0 field), internal force, symplectic Euler kick and drift; scheme 2:
the time loop of examples/collapse_dry.jl:202-211.
"""
function step!(sys::ParticleSystem, nsteps::Integer; scheme::Integer = 1, kw...)
    pairs = Ref{Int64}(0)
    device_call(sys) do
        check(ccall((:sph_step, LIB), Cint, (Ptr{Cvoid}, Cint, Ref{SphParams}, Cint, Ref{Int64}),
                    sys.handle, scheme, params(sys; kw...), nsteps, pairs))
    end
    return pairs[]
end

end # module
