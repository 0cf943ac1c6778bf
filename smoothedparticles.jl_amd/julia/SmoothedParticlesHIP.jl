# SmoothedParticlesHIP -- ccall binding of libsph_hip (include/sph.h) that keeps
# the call surface of SmoothedParticles.jl's hot path:
#
#   ParticleSystem(T, domain, h)        src/structs.jl:57-92
#   create_cell_list!(sys)              src/core.jl:51-90
#   apply!(sys, op!; self, params...)   src/core.jl:151-161
#   apply_binary!/apply_unary!          src/core.jl:125-142
#   ParticleField(sys, :sym)            src/structs.jl:118-125
#
# Operators are registered by the name of the reference function they stand
# for (find_rho!, internal_force!, ...); an operator whose name is not
# registered raises, because a Julia closure cannot run on the device.
# No Julia toolchain exists where this library is built, so this file is
# source only and has not been executed.
module SmoothedParticlesHIP

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

const OPS = Dict{Symbol,Int32}(
    Symbol("find_rho!") => 1, Symbol("find_rho0!") => 2, Symbol("internal_force!") => 3,
    Symbol("internal_force_rho0!") => 4, Symbol("balance_of_mass!") => 5,
    Symbol("reset_rho!") => 16, Symbol("reset_a!") => 17, Symbol("find_pressure!") => 18,
    Symbol("find_pressure_rho0!") => 19, Symbol("find_pressure_continuity!") => 20,
    Symbol("accelerate!") => 21, Symbol("move!") => 22)

const FIELDS = Dict(:x => 0, :v => 1, :a => 2, :Dv => 2, :rho => 3, :P => 4, :Drho => 5,
                    :rho0 => 6, :type => 7)

mutable struct ParticleSystem
    handle::Ptr{Cvoid}
    h::Float64
    params::Dict{Symbol,Any}
end

function check(rc::Cint)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:sph_last_error, LIB), Cstring, ()))
    error("libsph_hip error $(rc): $(msg)")
end

"""
    ParticleSystem(bbox_min, bbox_max, h; precision = :fp64)

Device-resident particle system; only the domain's bounding box matters, as
in the reference (src/structs.jl:63).  `precision = :mixed` selects FP32
relative-coordinate pair arithmetic with FP64 state.
"""
function ParticleSystem(bmin::NTuple{3,Float64}, bmax::NTuple{3,Float64}, h::Float64;
                        precision::Symbol = :fp64)
    @assert(h > 0.0, "invalid ParticleSystem declaration! (h must be a positive float)")
    out = Ref{Ptr{Cvoid}}(C_NULL)
    mode = precision == :mixed ? Cint(1) : Cint(0)
    check(ccall((:sph_create, LIB), Cint,
                (Ref{Ptr{Cvoid}}, Cdouble, Ref{NTuple{3,Float64}}, Ref{NTuple{3,Float64}}, Cint),
                out, h, bmin, bmax, mode))
    sys = ParticleSystem(out[], h, Dict{Symbol,Any}())
    finalizer(s -> ccall((:sph_destroy, LIB), Cint, (Ptr{Cvoid},), s.handle), sys)
    return sys
end

"""
    push_particles!(sys, x; v, rho, P, type)

push! of particles (generate_particles!, src/grids.jl:253-258); `x` is a
vector of 3-vectors.
"""
function push_particles!(sys::ParticleSystem, x::Vector{NTuple{3,Float64}};
                         v = nothing, rho = nothing, P = nothing, type = nothing)
    n = length(x)
    ptr(a) = a === nothing ? C_NULL : pointer(a)
    check(ccall((:sph_append, LIB), Cint,
                (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                sys.handle, n, pointer(reinterpret(Float64, x)), ptr(v), ptr(rho), ptr(P), ptr(type)))
end

Base.length(sys::ParticleSystem) = (n = Ref{Int64}(0);
    check(ccall((:sph_count, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}), sys.handle, n)); n[])

"Parameters the registered operators read (m, rho0, c, alpha, hs, mu, nu, g, dt)."
set_params!(sys::ParticleSystem; kw...) = (merge!(sys.params, Dict(kw)); nothing)

function params(sys::ParticleSystem; kw...)
    p = merge(sys.params, Dict(kw))
    get0(k) = Float64(get(p, k, 0.0))
    g = get(p, :g, (0.0, 0.0, 0.0))
    SphParams(get0(:m), get0(:rho0), get0(:c), get0(:alpha), get0(:hs), get0(:mu), get0(:nu),
              (Float64(g[1]), Float64(g[2]), Float64(g[3])), get0(:dt), Float64(get(p, :coeff, 1.0)))
end

"create_cell_list!(sys) -- src/core.jl:51-90"
function create_cell_list!(sys::ParticleSystem)
    n = Ref{Int64}(0)
    check(ccall((:sph_build_cells, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}), sys.handle, n))
    return nothing
end

opid(op!::Function) = get(OPS, nameof(op!)) do
    error("$(nameof(op!)) is not a registered device operator")
end

"apply!(sys, op!; self = false, params...) -- src/core.jl:151-161"
function apply!(sys::ParticleSystem, op!::Function; self::Bool = false, kw...)
    check(ccall((:sph_apply, LIB), Cint, (Ptr{Cvoid}, Cint, Ref{SphParams}, Cint),
                sys.handle, opid(op!), params(sys; kw...), self ? Cint(1) : Cint(0)))
end

apply_binary!(sys::ParticleSystem, op!::Function; kw...) =
    check(ccall((:sph_apply_binary, LIB), Cint, (Ptr{Cvoid}, Cint, Ref{SphParams}, Cint),
                sys.handle, opid(op!), params(sys; kw...), Cint(0)))

apply_unary!(sys::ParticleSystem, op!::Function; kw...) =
    check(ccall((:sph_apply_unary, LIB), Cint, (Ptr{Cvoid}, Cint, Ref{SphParams}),
                sys.handle, opid(op!), params(sys; kw...)))

"""
    ParticleField(sys, :sym)

Field view in the reference's particle order (src/structs.jl:118-125);
reading downloads the field, writing uploads it.
"""
struct ParticleField <: AbstractVector{Any}
    sys::ParticleSystem
    var::Symbol
end

ncomp(f::ParticleField) = f.var in (:x, :v, :a, :Dv) ? 3 : 1
Base.size(f::ParticleField) = (length(f.sys),)

function values(f::ParticleField)
    n = length(f.sys)
    buf = Vector{Float64}(undef, ncomp(f) * n)
    check(ccall((:sph_get_field, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Int64),
                f.sys.handle, FIELDS[f.var], buf, n))
    return ncomp(f) == 3 ? [(buf[3i-2], buf[3i-1], buf[3i]) for i in 1:n] : buf
end

Base.getindex(f::ParticleField, i::Int) = values(f)[i]

function Base.setindex!(f::ParticleField, val, i::Int)
    a = values(f)
    a[i] = val
    buf = ncomp(f) == 3 ? collect(reinterpret(Float64, a)) : a
    check(ccall((:sph_set_field, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Int64),
                f.sys.handle, FIELDS[f.var], buf, length(f.sys)))
end

"Total energy (kind 0: examples/collapse_dry.jl:166-171, 1: tests/test_collision_2d.jl:96-100)."
function energy(sys::ParticleSystem, kind::Integer = 0; kw...)
    out = Ref{NTuple{3,Float64}}((0.0, 0.0, 0.0))
    check(ccall((:sph_reduce, LIB), Cint, (Ptr{Cvoid}, Cint, Ref{SphParams}, Ref{NTuple{3,Float64}}),
                sys.handle, kind, params(sys; kw...), out))
    return out[][1]
end

"""SmoothedParticles.sum(sys, func, x) (src/core.jl:240-258) at probe points
`xs` for the closures of examples/cavity_flow.jl:166-172: a 4 x np matrix
whose column i holds sum (phi m) W and the three sums (phi m) v_k W."""
function probe_sums(sys::ParticleSystem, xs::AbstractVector; fluid_only::Bool = false, kw...)
    np = length(xs)
    pts = Float64[c for x in xs for c in (x[1], x[2], x[3])]
    out = zeros(Float64, 4, np)
    check(ccall((:sph_probe_sum, LIB), Cint,
                (Ptr{Cvoid}, Ref{SphParams}, Cint, Ptr{Float64}, Int64, Ptr{Float64}),
                sys.handle, params(sys; kw...), fluid_only ? 1 : 0, pts, np, out))
    return out
end

"nsteps fused device steps (density sum, EOS, internal force, kick, drift)."
function step!(sys::ParticleSystem, nsteps::Integer; scheme::Integer = 1, kw...)
    pairs = Ref{Int64}(0)
    check(ccall((:sph_step, LIB), Cint, (Ptr{Cvoid}, Cint, Ref{SphParams}, Cint, Ref{Int64}),
                sys.handle, scheme, params(sys; kw...), nsteps, pairs))
    return pairs[]
end

end # module
