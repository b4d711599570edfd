"""
    BLDGPU

Julia binding of the MI355X time-step engine (`include/bld.h`) for
BoundaryLayerDynamics.jl.  `BLDGPU.evolve!(model, tspan; dt)` is a drop-in for
`BoundaryLayerDynamics.evolve!(model, tspan; dt, method = SSPRK33(), output = ())`
(src/BoundaryLayerDynamics.jl:190-219): the constants of the model's
discretised processes are handed to the engine once, the state is uploaded,
all steps run on the GPU, and the state is written back into `model.state`.

Not executed in the build environment (no Julia toolchain).
"""
module BLDGPU

using BoundaryLayerDynamics: BoundaryLayerDynamics as BLD

const libbld = get(ENV, "BLD_LIBRARY", joinpath(@__DIR__, "..", "libbld.so"))
const ABI_VERSION = Int32(2)

struct BldBC
    kind::Int32       # 0 Dirichlet, 1 Neumann
    reserved::Int32
    value::Float64
    difference::Float64
end

struct BldConfig
    abi_version::Int32
    precision::Int32   # 0 plain float64, 1 float32 fields + float64 pressure / mean mode
    k1max::Int32
    k2max::Int32
    n3::Int32
    reserved0::Int32
    L1::Float64
    L2::Float64
    alpha_c::Ptr{Float64}
    alpha_i::Ptr{Float64}
    bc::NTuple{6,BldBC}          # [vel1 lower, vel1 upper, vel2 lower, ..., vel3 upper]
    diffusion::NTuple{3,Int32}
    reserved1::Int32
    diffusivity::NTuple{3,Float64}
    source::NTuple{3,Float64}
    pressure::Int32
    advection::Int32
    adv_n1::Int32
    adv_n2::Int32
    sgs::Int32
    sgs_n1::Int32
    sgs_n2::Int32
    sgs_wall::NTuple{2,Int32}
    reserved2::Int32
    sgs_lc::Ptr{Float64}
    sgs_li::Ptr{Float64}
    sgs_factor_strain::NTuple{2,Float64}
    sgs_factor_stress::NTuple{2,Float64}
    mean::NTuple{3,Int32}
    reserved3::Int32
    mean_value::NTuple{3,Float64}
    mean_weight::Ptr{Float64}
end

mutable struct Engine
    handle::Ptr{Cvoid}
    keep::Vector{Any}   # arrays referenced by the config until bld_create returns
end

lasterror(h) = unsafe_string(ccall((:bld_last_error, libbld), Cstring, (Ptr{Cvoid},), h))
check(rc, h) = rc == 0 || error(lasterror(h))

# boundary conditions of one field (boundary_conditions.jl:53-95)
function bcpair(domain, grid, field)
    map(((domain.lower_boundary, 0.0), (domain.upper_boundary, 1.0))) do (b, ζ)
        t = BLD.BoundaryConditions.init_bctype(Float64, BLD.BoundaryConditions.bctype(b, field);
                                               dx3 = domain.Dvmap(ζ) / grid.n3global)
        t isa BLD.BoundaryConditions.ConstantValue ? BldBC(0, 0, t.value, 0.0) :
        BldBC(1, 0, t.gradient, t.difference)
    end
end

"""
    Engine(model; precision = :fp64)

Creates the GPU engine from a `Model` (single process, no MPI) whose processes
are MomentumAdvection, MolecularDiffusion, Pressure, ConstantSource and
StaticSmagorinskyModel.
"""
function Engine(model::BLD.Model; precision = :fp64)
    grid, domain = model.grid, model.domain
    grid.comm === nothing || error("BLDGPU runs one process per model")
    n3 = grid.n3global
    αc = BLD.Derivatives.dx3factors(domain, grid, BLD.NodeSet(:C))
    αi = Float64[n3 / domain.Dvmap(ζ) for ζ in 0:1//n3:1]
    keep = Any[αc, αi]
    diffusion, diffusivity, source = zeros(Int32, 3), zeros(3), zeros(3)
    pressure = advection = sgs = Int32(0)
    adv_n = sgs_n = (Int32(0), Int32(0))
    lc = li = Float64[]
    walls, fstrain, fstress = (Int32(0), Int32(0)), (0.0, 0.0), (0.0, 0.0)
    mean, mean_value, mean_weight = zeros(Int32, 3), zeros(3), Float64[]
    fidx = Dict(:vel1 => 1, :vel2 => 2, :vel3 => 3)
    P = BLD.Processes
    for p in model.processes
        if p isa P.DiscretizedMomentumAdvection
            advection, adv_n = Int32(1), Int32.(p.dims)
        elseif p isa P.DiscretizedMolecularDiffusion
            diffusion[fidx[p.field]] = 1
            diffusivity[fidx[p.field]] = p.diffusivity
        elseif p isa P.DiscretizedPressure
            pressure = Int32(1)
        elseif p isa P.DiscretizedConstantSource
            source[fidx[p.field]] += p.strength
        elseif p isa P.DiscretizedConstantMean
            mean[fidx[p.field]] = 1
            mean_value[fidx[p.field]] = p.mean_value
            mean_weight = collect(Float64, p.weight)
        elseif p isa P.DiscretizedStaticSmagorinskyModel
            sgs, sgs_n = Int32(1), Int32.(p.dims)
            lc, li = vec(p.lengthscale_c), vec(p.lengthscale_i)
            bcs = p.boundary_conditions
            walls = Tuple(Int32(bc.type isa BLD.BoundaryConditions.DynamicValues) for bc in bcs.wallstress)
            fstrain, fstress = Float64.(bcs.factors_strain), Float64.(bcs.factors_stress)
        else
            error("process $(nameof(p)) is not part of the GPU step engine")
        end
    end
    push!(keep, lc, li, mean_weight)
    bcs = (bcpair(domain, grid, :vel1)..., bcpair(domain, grid, :vel2)..., bcpair(domain, grid, :vel3)...)
    cfg = BldConfig(ABI_VERSION, precision == :fp32 ? 1 : 0, grid.k1max, grid.k2max, n3, 0,
                    domain.hsize[1], domain.hsize[2], pointer(αc), pointer(αi), bcs,
                    Tuple(diffusion), 0, Tuple(diffusivity), Tuple(source), pressure, advection, adv_n...,
                    sgs, sgs_n..., walls, 0, isempty(lc) ? C_NULL : pointer(lc),
                    isempty(li) ? C_NULL : pointer(li), fstrain, fstress,
                    Tuple(mean), 0, Tuple(mean_value), isempty(mean_weight) ? C_NULL : pointer(mean_weight))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = GC.@preserve keep ccall((:bld_create, libbld), Cint, (Ref{BldConfig}, Ref{Ptr{Cvoid}}), cfg, h)
    check(rc, C_NULL)
    e = Engine(h[], keep)
    finalizer(e -> ccall((:bld_destroy, libbld), Cvoid, (Ptr{Cvoid},), e.handle), e)
    e
end

statepointers(s) = (haskey(s, :vel1) ? pointer(s.vel1) : C_NULL, haskey(s, :vel2) ? pointer(s.vel2) : C_NULL,
                    haskey(s, :vel3) ? pointer(s.vel3) : C_NULL)

function upload!(e::Engine, model)
    s = model.state
    rc = GC.@preserve s ccall((:bld_set_state, libbld), Cint,
                              (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{ComplexF64}, Ptr{ComplexF64}),
                              e.handle, statepointers(s)...)
    check(rc, e.handle)
end

function download!(model, e::Engine)
    s = model.state
    rc = GC.@preserve s ccall((:bld_get_state, libbld), Cint,
                              (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{ComplexF64}, Ptr{ComplexF64}),
                              e.handle, statepointers(s)...)
    check(rc, e.handle)
end

"""
    evolve!(model, tspan; dt, precision = :fp64, engine = Engine(model; precision))

GPU counterpart of `BoundaryLayerDynamics.evolve!` for SSPRK33 without output
modules.  Returns the engine so that it can be reused for further calls.
"""
function evolve!(model::BLD.Model, tspan; dt, precision = :fp64, engine = nothing)
    e = isnothing(engine) ? Engine(model; precision = precision) : engine
    t1 = length(tspan) == 1 ? zero(first(tspan)) : first(tspan)
    t2 = last(tspan)
    upload!(e, model)
    check(ccall((:bld_init_problem, libbld), Cint, (Ptr{Cvoid}, Cdouble), e.handle, t1), e.handle)
    if t2 > t1
        nt = BLD.Helpers.approxdiv(t2 - t1, dt)
        t = Ref{Cdouble}(t1)
        rc = ccall((:bld_step_graph, libbld), Cint, (Ptr{Cvoid}, Cdouble, Int64, Ref{Cdouble}, Cdouble),
                   e.handle, dt, nt, t, t2)
        check(rc, e.handle)
    end
    download!(model, e)
    e
end

"""
    rates!(rates, model, engine)

`compute_rates!` of the model state on the GPU (Processes.jl:39-62), written
into `rates` (a NamedTuple like `model.state`).
"""
function rates!(rates, model::BLD.Model, e::Engine)
    upload!(e, model)
    rc = GC.@preserve rates ccall((:bld_rates, libbld), Cint,
                                  (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{ComplexF64}, Ptr{ComplexF64}),
                                  e.handle, statepointers(rates)...)
    check(rc, e.handle)
    rates
end

"""
    set_closure!(engine, params, Δ, wall_feature; guard = true)

Replace the stresses of the engine's StaticSmagorinskyModel by the learned
closure (`params`: the 6288 Float32 network parameters; `wall_feature`: one
value 1/(1 + d/Δ) per C-layer).  The Smagorinsky model stays the wall model
and the fallback; `params = nothing` restores it.
"""
function set_closure!(e::Engine, params, Δ, wall_feature; guard = true)
    p = params === nothing ? C_NULL : convert(Vector{Float32}, params)
    w = convert(Vector{Float64}, wall_feature)
    rc = GC.@preserve p w ccall((:bld_sgs_closure, libbld), Cint,
                                (Ptr{Cvoid}, Ptr{Float32}, Cdouble, Ptr{Cdouble}, Cint),
                                e.handle, p, Δ, w, guard ? 1 : 0)
    check(rc, e.handle)
    e
end

end # module BLDGPU
