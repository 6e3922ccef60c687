# TrixiHIP.jl -- Julia host binding of libtrixihip.so (the gfx950 DGSEM
# rhs! for Trixi.jl's TreeMesh).  HIP-named twins of TrixiCUDA.jl's exports
# (reference src/TrixiCUDA.jl:74-77): DGSEMHIP, SemidiscretizationHyperbolicHIP,
# semidiscretizeHIP, plus rhs_hip! (rhs_gpu!, src/solvers/solvers.jl:18-31).
#
# Shipped as source: it needs Julia, Trixi.jl and AMDGPU.jl (for ROCArray
# device buffers), none of which exist in the build container.  The C ABI it
# calls (include/trixi_hip.h) is exercised by the Python tests.
module TrixiHIP

using Trixi
using StaticArrays: SVector
using AMDGPU
using ChainRulesCore
using SciMLBase: ODEProblem, FullSpecialize

const libtrixihip = get(ENV, "TRIXI_HIP_LIBRARY",
                        joinpath(@__DIR__, "..", "libtrixihip.so"))

# ---- C structs (must match include/trixi_hip.h) ------------------------------
struct CConfig
    ndims::Int32
    equation::Int32
    polydeg::Int32
    volume_integral::Int32
    volume_flux::Int32
    surface_flux::Int32
    fv_flux::Int32
    precision::Int32
    alpha_smooth::Int32
    indicator_variable::Int32
    gamma::Float64
    c_h::Float64
    advection_velocity::NTuple{3, Float64}
    alpha_max::Float64
    alpha_min::Float64
    ic_length::Float64
    boundary_condition::Int32
    boundary_ic::Int32
    boundary_seed::UInt64
    source_terms::Int32
    boundary_outflow_faces::Int32
    nu::Float64
    Lr::Float64
    ncomponents::Int32
    reserved::Int32
    gammas::NTuple{4, Float64}
    gas_constants::NTuple{4, Float64}
    gravity::Float64
    H0::Float64
end

struct CMesh
    nelements::Int64
    ninterfaces::Int64
    neighbor_ids::Ptr{Int32}
    orientations::Ptr{Int32}
    inverse_jacobian::Ptr{Float64}
    node_coordinates::Ptr{Float64}
    total_volume::Float64
    nmortars::Int64
    mortar_neighbor_ids::Ptr{Int32}
    mortar_orientations::Ptr{Int32}
    mortar_large_sides::Ptr{Int32}
end

function check(rc)
    if rc != 0
        msg = unsafe_string(ccall((:trixi_hip_last_error, libtrixihip), Cstring, ()))
        error("trixi_hip error $rc: $msg")
    end
    nothing
end

# ---- identities of Trixi objects (a Julia function cannot cross a C ABI) -----
equation_id(::LinearScalarAdvectionEquation1D) = 0
equation_id(::LinearScalarAdvectionEquation2D) = 0
equation_id(::LinearScalarAdvectionEquation3D) = 0
equation_id(::CompressibleEulerEquations1D) = 1
equation_id(::CompressibleEulerEquations2D) = 1
equation_id(::CompressibleEulerEquations3D) = 1
equation_id(::IdealGlmMhdEquations3D) = 2
equation_id(::IdealGlmMhdEquations2D) = 2
equation_id(::IdealGlmMhdEquations1D) = 2
equation_id(::InviscidBurgersEquation1D) = 3
equation_id(::Union{HyperbolicDiffusionEquations1D, HyperbolicDiffusionEquations2D,
                    HyperbolicDiffusionEquations3D}) = 4
equation_id(::Union{CompressibleEulerMulticomponentEquations1D,
                    CompressibleEulerMulticomponentEquations2D}) = 5
equation_id(::CompressibleEulerEquationsQuasi1D) = 6
equation_id(::Union{ShallowWaterEquations1D, ShallowWaterEquations2D}) = 7
equation_id(eq) = throw(ArgumentError("equation $(typeof(eq)) is not built for HIP"))

flux_id(::typeof(flux_central)) = 0
flux_id(f::FluxLaxFriedrichs) = f.max_abs_speed === max_abs_speed_naive ? 2 : 1
flux_id(::typeof(flux_ranocha)) = 3
flux_id(::typeof(flux_shima_etal)) = 4
flux_id(::typeof(flux_hindenlang_gassner)) = 5
flux_id(::typeof(flux_ec)) = 6
flux_id(f::FluxHLL) = f.min_max_speed === min_max_speed_einfeldt ? 9 :
                      f.min_max_speed === min_max_speed_davis ? 7 :
    throw(ArgumentError("FluxHLL is built with min_max_speed_davis"))
flux_id(::typeof(flux_chan_etal)) = 8
flux_id(::typeof(flux_wintermeyer_etal)) = 10
flux_id(::typeof(flux_fjordholm_etal)) = 11
flux_id(f::FluxHydrostaticReconstruction) =
    f.numerical_flux === flux_lax_friedrichs && f.hydrostatic_reconstruction === hydrostatic_reconstruction_audusse_etal ?
    12 : throw(ArgumentError("FluxHydrostaticReconstruction is built for (flux_lax_friedrichs, audusse_etal)"))
flux_id(f::Tuple) = (f[2] in (flux_nonconservative_powell, flux_nonconservative_chan_etal,
                            flux_nonconservative_wintermeyer_etal, flux_nonconservative_fjordholm_etal,
                            flux_nonconservative_audusse_etal) ?
                     flux_id(f[1]) : throw(ArgumentError("unsupported nonconservative flux $(f[2])")))
flux_id(f) = throw(ArgumentError("flux $f is not built for HIP"))

ic_id(::typeof(initial_condition_constant)) = 0
ic_id(::typeof(initial_condition_convergence_test)) = 1
ic_id(::typeof(initial_condition_weak_blast_wave)) = 2
# the Burgers step states of test/tree_dgsem_1d/burgers_shock.jl and
# burgers_rarefraction.jl, as functions Trixi can evaluate too
initial_condition_shock(x, t, ::InviscidBurgersEquation1D) = SVector(x[1] < 0.5 ? 1.5 : 0.5)
initial_condition_rarefaction(x, t, ::InviscidBurgersEquation1D) = SVector(x[1] < 0.5 ? 0.5 : 1.5)
ic_id(::typeof(initial_condition_shock)) = 6
ic_id(::typeof(initial_condition_rarefaction)) = 7
ic_id(::typeof(initial_condition_eoc_test_coupled_euler_gravity)) = 8
ic_id(::typeof(initial_condition_poisson_nonperiodic)) = 9
# the harmonic state of test/tree_dgsem_1d/hypdiff_harmonic_nonperiodic.jl
function initial_condition_harmonic_nonperiodic(x, t, ::HyperbolicDiffusionEquations1D)
    t == 0 ? SVector(5.0, 0.0) : SVector(3.0 + exp(1.0) * x[1], exp(1.0))
end
ic_id(::typeof(initial_condition_harmonic_nonperiodic)) = 10
ic_id(f) = throw(ArgumentError("initial condition $f is not built for HIP"))
is_builtin_ic(f) = try ic_id(f); true catch; false end
# seeded benchmark states (ids 3, 4, 5 of trixi_hip_initial_condition) carry their seed
struct SeededInitialCondition
    id::Int32
    seed::UInt64
end
initial_condition_seeded_random(seed = 20261202) = SeededInitialCondition(3, seed)
initial_condition_smooth_periodic(seed = 20261202) = SeededInitialCondition(4, seed)
initial_condition_weak_blast_wave_seeded(seed = 20261202) = SeededInitialCondition(5, seed)
ic_id(ic::SeededInitialCondition) = ic.id
ic_seed(ic::SeededInitialCondition) = ic.seed
ic_seed(ic) = UInt64(0)

# ---- public API ----------------------------------------------------------------
struct DGSEMHIP{VI, SF}
    polydeg::Int
    surface_flux::SF
    volume_integral::VI
    precision::Symbol      # :fp64, :mixed, :mixed_f32log or :fp32 (Float32 state vectors)
end

# enum trixi_hip_precision
precision_id(p::Symbol) = p === :fp64 ? Int32(0) : p === :mixed ? Int32(1) : p === :mixed_f32log ? Int32(2) :
    p === :fp32 ? Int32(3) : throw(ArgumentError("precision must be :fp64, :mixed, :mixed_f32log or :fp32"))

function DGSEMHIP(; RealT = Float64, polydeg::Integer, surface_flux = flux_central,
                  volume_integral = VolumeIntegralWeakForm(), precision = nothing)
    prec = precision === nothing ? (RealT === Float32 ? :fp32 : :fp64) : precision
    DGSEMHIP(Int(polydeg), surface_flux, volume_integral, prec)
end

mutable struct SemidiscretizationHyperbolicHIP{Mesh, Equations, IC, Solver, Cache}
    mesh::Mesh
    equations::Equations
    initial_condition::IC
    solver::Solver
    ctx::Ptr{Cvoid}
    nvalues::Int
    dg_cpu::DGSEM        # host containers, for initial conditions written as functions
    cache_cpu::Cache
end

# The outflow boundary of the reference's Burgers tests: flux(u_inner).
function boundary_condition_outflow(u_inner, orientation, direction, x, t, surface_flux_function, equations)
    Trixi.flux(u_inner, orientation, equations)
end

# boundary_conditions: periodic, one BoundaryConditionDirichlet, or a NamedTuple
# (x_neg, x_pos, ...) of Dirichlet conditions that all carry the same state and
# boundary_condition_outflow sides (bit f of the outflow mask)
function boundary_config(bcs)
    bcs === boundary_condition_periodic && return (0, 0, UInt64(0), 0)
    bcs === boundary_condition_slip_wall && return (2, 0, UInt64(0), 0)  # shallow water
    sides = bcs isa NamedTuple ? collect(values(bcs)) : [bcs]
    outflow = 0
    if bcs isa NamedTuple
        names = (:x_neg, :x_pos, :y_neg, :y_pos, :z_neg, :z_pos)
        for (k, b) in pairs(bcs)
            b === boundary_condition_outflow && (outflow |= 1 << (findfirst(==(k), names) - 1))
        end
    end
    # boundary_condition_poisson_nonperiodic: the steady Poisson state at any time
    sides = [b === boundary_condition_poisson_nonperiodic ?
             BoundaryConditionDirichlet(initial_condition_poisson_nonperiodic) : b for b in sides]
    dir = [b for b in sides if b !== boundary_condition_periodic && b !== boundary_condition_outflow]
    all(b -> b isa BoundaryConditionDirichlet, dir) ||
        throw(ArgumentError("only BoundaryConditionDirichlet, outflow and periodic boundaries are built for HIP"))
    isempty(dir) && return (1, 0, UInt64(0), outflow)
    f = dir[1].boundary_value_function
    all(b -> b.boundary_value_function === f, dir) ||
        throw(ArgumentError("all Dirichlet boundary faces must carry the same state"))
    (1, ic_id(f), ic_seed(f), outflow)
end

indicator_variable_id(::typeof(density_pressure)) = 0
indicator_variable_id(::typeof(pressure)) = 1
indicator_variable_id(::typeof(first)) = 2
indicator_variable_id(::typeof(density)) = 2
indicator_variable_id(::typeof(waterheight_pressure)) = 0
indicator_variable_id(f) = throw(ArgumentError("indicator variable $f is not built for HIP"))

source_id(::Nothing) = 0
source_id(::typeof(source_terms_convergence_test)) = 1
source_id(::typeof(source_terms_eoc_test_euler)) = 2
source_id(::typeof(source_terms_poisson_nonperiodic)) = 3
source_id(::typeof(source_terms_harmonic)) = 4
source_id(f) = throw(ArgumentError("source terms $f are not built for HIP"))

function make_config(solver::DGSEMHIP, equations, mesh; boundary_conditions = boundary_condition_periodic,
                     source_terms = nothing)
    vi = solver.volume_integral
    volume_integral, vflux, fvflux, smooth, amax, amin, var = 0, 0, 0, 0, 0.5, 0.001, 0
    if vi isa VolumeIntegralFluxDifferencing
        volume_integral, vflux = 1, flux_id(vi.volume_flux)
    elseif vi isa VolumeIntegralShockCapturingHG
        volume_integral, vflux, fvflux = 2, flux_id(vi.volume_flux_dg), flux_id(vi.volume_flux_fv)
        smooth = vi.indicator.alpha_smooth ? 1 : 0
        amax, amin = vi.indicator.alpha_max, vi.indicator.alpha_min
        var = indicator_variable_id(vi.indicator.variable)
    end
    gamma = hasproperty(equations, :gamma) ? equations.gamma : 1.4
    c_h = hasproperty(equations, :c_h) ? equations.c_h : 0.0
    a = hasproperty(equations, :advection_velocity) ? Tuple(equations.advection_velocity) : ()
    vel = ntuple(d -> d <= length(a) ? Float64(a[d]) : 0.0, 3)
    len = mesh.tree.length_level_0
    bc, bc_ic, bc_seed, outflow = boundary_config(boundary_conditions)
    # multicomponent Euler: per-component gamma_k and R_k
    gammas = hasproperty(equations, :gammas) ? Tuple(equations.gammas) : ()
    gas_r = hasproperty(equations, :gas_constants) ? Tuple(equations.gas_constants) : ()
    CConfig(ndims(mesh), equation_id(equations), solver.polydeg, volume_integral, vflux,
            flux_id(solver.surface_flux), fvflux, precision_id(solver.precision),
            smooth, var, gamma, c_h, vel, amax, amin, len, bc, bc_ic, bc_seed, source_id(source_terms), outflow,
            hasproperty(equations, :nu) ? Float64(equations.nu) : 0.0,
            hasproperty(equations, :Lr) ? Float64(equations.Lr) : 0.0,
            Int32(length(gammas)), Int32(0),
            ntuple(k -> k <= length(gammas) ? Float64(gammas[k]) : 0.0, 4),
            ntuple(k -> k <= length(gas_r) ? Float64(gas_r[k]) : 0.0, 4),
            hasproperty(equations, :gravity) ? Float64(equations.gravity) : 0.0,
            hasproperty(equations, :H0) ? Float64(equations.H0) : 0.0)
end

function SemidiscretizationHyperbolicHIP(mesh::TreeMesh, equations, initial_condition,
                                         solver::DGSEMHIP;
                                         source_terms = nothing,
                                         boundary_conditions = boundary_condition_periodic)
    # Trixi builds the element, interface and L2 mortar containers on the host
    # (faces on a non-periodic boundary get no interface: the library treats
    # them as boundary faces).
    dg_cpu = DGSEM(polydeg = solver.polydeg, surface_flux = flux_central)
    cache = Trixi.create_cache(mesh, equations, dg_cpu, Float64, Float64)
    e, i = cache.elements, cache.interfaces
    nbr = Int32.(i.neighbor_ids .- 1)            # 0-based
    ori = Int32.(i.orientations)
    invj = Vector{Float64}(e.inverse_jacobian)
    xyz = Array{Float64}(e.node_coordinates)     # [d, i, j, k, e]
    if hasproperty(cache, :mortars)
        m = cache.mortars                        # neighbor_ids [2^(d-1)+1, nmortars]
        mids = Int32.(m.neighbor_ids .- 1)
        mori = Int32.(m.orientations)
        mls = Int32.(m.large_sides)
    else
        mids = Int32[]; mori = Int32[]; mls = Int32[]
    end
    cfg = make_config(solver, equations, mesh; boundary_conditions, source_terms)
    GC.@preserve nbr ori invj xyz mids mori mls begin
        cmesh = CMesh(nelements(e), ninterfaces(i), pointer(nbr), pointer(ori), pointer(invj),
                      pointer(xyz), mesh.tree.length_level_0^ndims(mesh), length(mori),
                      isempty(mori) ? C_NULL : pointer(mids), isempty(mori) ? C_NULL : pointer(mori),
                      isempty(mori) ? C_NULL : pointer(mls))
        ctx = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:trixi_hip_create, libtrixihip), Cint,
                    (Ref{CConfig}, Ref{CMesh}, Ref{Ptr{Cvoid}}), cfg, cmesh, ctx))
    end
    nv = Ref{Int32}(0); nn = Ref{Int32}(0); nval = Ref{Int64}(0)
    check(ccall((:trixi_hip_sizes, libtrixihip), Cint, (Ptr{Cvoid}, Ref{Int32}, Ref{Int32}, Ref{Int64}),
                ctx[], nv, nn, nval))
    semi = SemidiscretizationHyperbolicHIP(mesh, equations, initial_condition, solver, ctx[], Int(nval[]),
                                           dg_cpu, cache)
    finalizer(s -> ccall((:trixi_hip_destroy, libtrixihip), Cint, (Ptr{Cvoid},), s.ctx), semi)
    semi
end

stream_ptr() = AMDGPU.stream().stream   # hipStream_t of the task-local stream

"rhs_hip!(du_ode, u_ode, semi, t): the ODE right-hand side on the GPU (rhs_gpu!)."
function rhs_hip!(du_ode::ROCVector{Float64}, u_ode::ROCVector{Float64},
                  semi::SemidiscretizationHyperbolicHIP, t)
    check(ccall((:trixi_hip_rhs, libtrixihip), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Float64, Ptr{Cvoid}),
                semi.ctx, pointer(du_ode), pointer(u_ode), t, stream_ptr()))
    nothing
end

# RealT = Float32 (precision :fp32): Float32 state vectors
function rhs_hip!(du_ode::ROCVector{Float32}, u_ode::ROCVector{Float32},
                  semi::SemidiscretizationHyperbolicHIP, t)
    check(ccall((:trixi_hip_rhs_f32, libtrixihip), Cint,
                (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Float64, Ptr{Cvoid}),
                semi.ctx, pointer(du_ode), pointer(u_ode), t, stream_ptr()))
    nothing
end

function compute_coefficients_hip(t, semi::SemidiscretizationHyperbolicHIP)
    u = ROCVector{Float64}(undef, semi.nvalues)
    if !is_builtin_ic(semi.initial_condition)
        # an initial condition written as a Julia function: Trixi samples it
        # on the host (Trixi's layout), the library packs it on the device
        u_host = Trixi.compute_coefficients(semi.initial_condition, t, semi.mesh, semi.equations,
                                            semi.dg_cpu, semi.cache_cpu)
        u_trixi = ROCVector{Float64}(u_host)
        check(ccall((:trixi_hip_pack, libtrixihip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid}),
                    semi.ctx, pointer(u), pointer(u_trixi), stream_ptr()))
        return u
    end
    check(ccall((:trixi_hip_compute_coefficients, libtrixihip), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Int32, Float64, UInt64, Ptr{Cvoid}),
                semi.ctx, pointer(u), ic_id(semi.initial_condition), t, ic_seed(semi.initial_condition),
                stream_ptr()))
    u
end

"semidiscretizeHIP(semi, tspan): ODEProblem whose f is rhs_hip! (semidiscretizeGPU)."
# Trixi-layout host copy u[v, i, j, k, e] of a device state, unpacked on the
# device (trixi_hip_unpack): the state Trixi's SaveSolutionCallback reads
# through wrap_array_native (reference src/solvers/dg.jl:24-32).  Save it with
# Trixi.save_solution_file(wrap_array_native_hip(u, semi), t, dt, timestep,
# semi.mesh, semi.equations, semi.dg_cpu, semi.cache_cpu, save_solution).
function wrap_array_native_hip(u_ode::ROCVector{Float64}, semi::SemidiscretizationHyperbolicHIP)
    u_trixi = ROCVector{Float64}(undef, semi.nvalues)
    check(ccall((:trixi_hip_unpack, libtrixihip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid}),
                semi.ctx, pointer(u_trixi), pointer(u_ode), stream_ptr()))
    hip_synchronize()
    N = Trixi.nnodes(semi.dg_cpu)
    reshape(Array(u_trixi), Trixi.nvariables(semi.equations), ntuple(_ -> N, ndims(semi.mesh))...,
            Trixi.nelements(semi.dg_cpu, semi.cache_cpu))
end

function semidiscretizeHIP(semi::SemidiscretizationHyperbolicHIP, tspan)
    u0 = compute_coefficients_hip(first(tspan), semi)
    ODEProblem{true, FullSpecialize}(rhs_hip!, u0, tspan, semi)
end

"max_dt on the device (StepsizeCallback)."
function max_dt_hip(u_ode::ROCVector{Float64}, semi::SemidiscretizationHyperbolicHIP)
    out = Ref{Float64}(0.0)
    check(ccall((:trixi_hip_max_dt, libtrixihip), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ref{Float64}, Ptr{Cvoid}),
                semi.ctx, pointer(u_ode), out, stream_ptr()))
    out[]
end

"(l2, linf) against the initial condition at time t (AnalysisCallback)."
function calc_error_norms_hip(u_ode::ROCVector{Float64}, t, semi::SemidiscretizationHyperbolicHIP)
    nv = nvariables(semi.equations)
    l2 = zeros(nv); linf = zeros(nv)
    check(ccall((:trixi_hip_error_norms, libtrixihip), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Int32, Float64, UInt64, Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid}),
                semi.ctx, pointer(u_ode), ic_id(semi.initial_condition), t, 0, l2, linf, stream_ptr()))
    l2, linf
end

# ---- time stepping ---------------------------------------------------------------
"One CarpenterKennedy2N54 step (5 fused stages; hipGraph replay when use_graph)."
function rk_step_hip!(u::ROCVector{Float64}, k::ROCVector{Float64}, t, dt,
                      semi::SemidiscretizationHyperbolicHIP; use_graph = true)
    check(ccall((:trixi_hip_rk_step, libtrixihip), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Int32, Ptr{Cvoid}),
                semi.ctx, pointer(u), pointer(k), t, dt, use_graph ? 1 : 0, stream_ptr()))
    nothing
end

function rk_step_hip!(u::ROCVector{Float32}, k::ROCVector{Float32}, t, dt,
                      semi::SemidiscretizationHyperbolicHIP; use_graph = true)
    check(ccall((:trixi_hip_rk_step_f32, libtrixihip), Cint,
                (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Float64, Float64, Int32, Ptr{Cvoid}),
                semi.ctx, pointer(u), pointer(k), t, dt, use_graph ? 1 : 0, stream_ptr()))
    nothing
end

"Stage s (0..4) of the CK2N54 step from t: k = A_s k + dt rhs(u), u_out = u + B_s k."
function rk_stage_hip!(u_out::ROCVector{Float64}, k::ROCVector{Float64}, u::ROCVector{Float64},
                       stage::Integer, t, dt, semi::SemidiscretizationHyperbolicHIP)
    check(ccall((:trixi_hip_rk_stage, libtrixihip), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int32, Float64, Float64, Ptr{Cvoid}),
                semi.ctx, pointer(u), pointer(k), pointer(u_out), stage, t, dt, stream_ptr()))
    nothing
end

hip_synchronize() = check(ccall((:trixi_hip_sync, libtrixihip), Cint, (Ptr{Cvoid},), stream_ptr()))

# ---- adjoint -----------------------------------------------------------------------
"Out-of-place rhs for reverse-mode AD; its rrule pulls back with trixi_hip_rhs_vjp."
function rhs_hip(u_ode::ROCVector{Float64}, semi::SemidiscretizationHyperbolicHIP, t)
    du = similar(u_ode)
    rhs_hip!(du, u_ode, semi, t)
    du
end

function ChainRulesCore.rrule(::typeof(rhs_hip), u_ode::ROCVector{Float64},
                              semi::SemidiscretizationHyperbolicHIP, t)
    du = rhs_hip(u_ode, semi, t)
    function rhs_hip_pullback(dbar)
        vbar = convert(ROCVector{Float64}, unthunk(dbar))
        ubar = similar(u_ode)
        check(ccall((:trixi_hip_rhs_vjp, libtrixihip), Cint,
                    (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Float64, Ptr{Cvoid}),
                    semi.ctx, pointer(ubar), pointer(u_ode), pointer(vbar), t, 1.0, 0.0, stream_ptr()))
        return NoTangent(), ubar, NoTangent(), NoTangent()
    end
    du, rhs_hip_pullback
end

"Reverse one CK2N54 step from its checkpoint u_n: ubar (of u_{n+1}) becomes the cotangent of u_n."
function rk_step_vjp_hip!(ubar::ROCVector{Float64}, u_n::ROCVector{Float64}, t, dt,
                          semi::SemidiscretizationHyperbolicHIP)
    check(ccall((:trixi_hip_enable_adjoint, libtrixihip), Cint, (Ptr{Cvoid},), semi.ctx))
    check(ccall((:trixi_hip_rk_step_vjp, libtrixihip), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Ptr{Cvoid}),
                semi.ctx, pointer(u_n), pointer(ubar), t, dt, stream_ptr()))
    nothing
end

"Gradient of a loss of u(t0 + n dt) with respect to u0, keeping one checkpoint per step."
function gradient_checkpointed(u0::ROCVector{Float64}, ubar_final::ROCVector{Float64}, t0, dt, nsteps,
                               semi::SemidiscretizationHyperbolicHIP)
    u = copy(u0); k = similar(u0); cps = ROCVector{Float64}[]
    for n in 1:nsteps
        push!(cps, copy(u))
        rk_step_hip!(u, k, t0 + (n - 1) * dt, dt, semi; use_graph = false)
    end
    ubar = copy(ubar_final)
    for n in nsteps:-1:1
        rk_step_vjp_hip!(ubar, cps[n], t0 + (n - 1) * dt, dt, semi)
    end
    ubar
end

export DGSEMHIP, SemidiscretizationHyperbolicHIP, semidiscretizeHIP, rhs_hip!, rhs_hip, wrap_array_native_hip,
       initial_condition_shock, initial_condition_rarefaction, boundary_condition_outflow,
       initial_condition_harmonic_nonperiodic,
       max_dt_hip, calc_error_norms_hip, rk_step_hip!, rk_stage_hip!, hip_synchronize,
       rk_step_vjp_hip!, gradient_checkpointed

end # module
