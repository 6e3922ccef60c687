/*
 * trixi_hip.h -- C ABI of libtrixihip.so, the gfx950 (AMD Instinct MI355X)
 * DGSEM semidiscretization for Trixi.jl's TreeMesh.
 *
 * This library is the drop-in replacement for the GPU side of TrixiCUDA.jl
 * (reference: src/solvers/solvers.jl rhs_gpu!, src/solvers/dg_{1,2,3}d.jl,
 * src/solvers/dg_{1,2,3}d_kernel.jl and src/callbacks_step/).  A host binding
 * (Julia ccall, Python ctypes) maps Trixi's equation/flux/initial-condition
 * objects to the enums below and passes plain device pointers.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative TRIXI_HIP_E* code on
 *     failure; trixi_hip_last_error() then holds a message (thread-local).
 *     No C++ exception crosses this boundary.
 *   - State vectors live in device memory (HBM) in the library's native
 *     layout, element innermost:
 *         u[(v * N^d + node) * nelements + element],
 *         node = i + N*j + N^2*k  (i = x index, 0-based),
 *     with elements in Trixi's TreeMesh leaf (Morton) order.  trixi_hip_pack /
 *     trixi_hip_unpack convert from/to Trixi's u[v, i, j, k, element] layout.
 *   - All device work is asynchronous on the caller's stream unless stated.
 *   - A context is not thread safe; the caller serialises calls on it.
 *   - Every device buffer of the forward path (rhs!, RK stages and steps,
 *     callbacks) is allocated in trixi_hip_create; none of those calls
 *     allocates device memory.  The adjoint allocates its own buffers once:
 *     trixi_hip_enable_adjoint (the stage checkpoints and registers) and the
 *     first trixi_hip_rhs_vjp of a shock-capturing context (two per-element
 *     arrays); call trixi_hip_enable_adjoint before capturing adjoint work.
 *   - Supported: TreeMesh, uniform or refined (L2 mortars), periodic or with
 *     the boundary conditions of enum trixi_hip_boundary_condition, with the
 *     equations, fluxes and precisions listed below; any other combination is
 *     TRIXI_HIP_EUNSUPPORTED (the reference's GPU path asserts periodicity and
 *     no mortars, dg_3d.jl:463-466, 542-544).
 */
#ifndef TRIXI_HIP_H
#define TRIXI_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct trixi_hip_ctx trixi_hip_ctx;
typedef void* trixi_hip_stream; /* a hipStream_t; NULL = the default stream */

/* ---- error codes ----------------------------------------------------------- */
#define TRIXI_HIP_OK 0
#define TRIXI_HIP_EINVAL (-1)       /* bad argument / handle / size mismatch     */
#define TRIXI_HIP_EUNSUPPORTED (-2) /* combination of equation/fluxes/N not built */
#define TRIXI_HIP_EDEVICE (-3)      /* HIP runtime error                          */
#define TRIXI_HIP_ENOMEM (-4)       /* device allocation failed                   */
#define TRIXI_HIP_ESTATE (-5)       /* non-finite or non-positive state found     */

/* ---- enums (values shared with the host bindings) ----------------------------- */
enum trixi_hip_equation {
    TRIXI_HIP_LINEAR_SCALAR_ADVECTION = 0, /* LinearScalarAdvectionEquation{1,2,3}D */
    TRIXI_HIP_COMPRESSIBLE_EULER = 1,      /* CompressibleEulerEquations{1,2,3}D    */
    TRIXI_HIP_IDEAL_GLM_MHD = 2,           /* IdealGlmMhdEquations1D / 2D / 3D (1D:
                                              8 variables, no psi / Powell terms)  */
    TRIXI_HIP_INVISCID_BURGERS = 3,        /* InviscidBurgersEquation1D             */
    TRIXI_HIP_HYPERBOLIC_DIFFUSION = 4,    /* HyperbolicDiffusionEquations{1,2,3}D  */
    TRIXI_HIP_COMPRESSIBLE_EULER_MULTICOMPONENT = 5, /* CompressibleEulerMulticomponent-
                                              Equations{1,2}D: (rho v, rho e, rho_1..
                                              rho_ncomponents), 1 or 2 components  */
    TRIXI_HIP_COMPRESSIBLE_EULER_QUASI1D = 6, /* CompressibleEulerEquationsQuasi1D:
                                              (a rho, a rho v, a e, a)              */
    TRIXI_HIP_SHALLOW_WATER = 7            /* ShallowWaterEquations{1,2}D: (h, h v, b);
                                              volume (flux_wintermeyer_etal,
                                              flux_nonconservative_wintermeyer_etal),
                                              surface (any flux,
                                              flux_nonconservative_fjordholm_etal) */
};
enum trixi_hip_volume_integral {
    TRIXI_HIP_VOLUME_WEAK_FORM = 0,          /* VolumeIntegralWeakForm          */
    TRIXI_HIP_VOLUME_FLUX_DIFFERENCING = 1,  /* VolumeIntegralFluxDifferencing  */
    TRIXI_HIP_VOLUME_SHOCK_CAPTURING_HG = 2  /* VolumeIntegralShockCapturingHG  */
};
enum trixi_hip_flux {
    TRIXI_HIP_FLUX_CENTRAL = 0,
    TRIXI_HIP_FLUX_LAX_FRIEDRICHS = 1,       /* FluxLaxFriedrichs(max_abs_speed)       */
    TRIXI_HIP_FLUX_LAX_FRIEDRICHS_NAIVE = 2, /* FluxLaxFriedrichs(max_abs_speed_naive) */
    TRIXI_HIP_FLUX_RANOCHA = 3,
    TRIXI_HIP_FLUX_SHIMA_ETAL = 4,
    TRIXI_HIP_FLUX_HINDENLANG_GASSNER = 5,   /* with flux_nonconservative_powell       */
    TRIXI_HIP_FLUX_EC = 6,                   /* flux_ec (InviscidBurgersEquation1D)    */
    TRIXI_HIP_FLUX_HLL = 7,                  /* flux_hll = FluxHLL(min_max_speed_davis),
                                                compressible Euler                     */
    TRIXI_HIP_FLUX_CHAN_ETAL = 8,            /* (flux_chan_etal,
                                                flux_nonconservative_chan_etal), quasi-1D */
    TRIXI_HIP_FLUX_HLLE = 9,                 /* flux_hlle = FluxHLL(min_max_speed_einfeldt),
                                                GLM-MHD 2D/3D                          */
    TRIXI_HIP_FLUX_WINTERMEYER_ETAL = 10,    /* shallow water                          */
    TRIXI_HIP_FLUX_FJORDHOLM_ETAL = 11,      /* shallow water                          */
    TRIXI_HIP_FLUX_HYDROSTATIC_LLF = 12      /* (FluxHydrostaticReconstruction(flux_lax_friedrichs,
                                                hydrostatic_reconstruction_audusse_etal),
                                                flux_nonconservative_audusse_etal), shallow water */
};
enum trixi_hip_precision {
    TRIXI_HIP_FP64 = 0,  /* all arithmetic in FP64                                    */
    TRIXI_HIP_MIXED = 1, /* FP32 volume/surface arithmetic; FP64 state, the logarithms
                            of the ln-means (per node, and their differences), RK
                            update, shock indicator and reductions                   */
    TRIXI_HIP_MIXED_F32LOG = 2, /* as TRIXI_HIP_MIXED with the ln-mean logarithms in
                            FP32 as well (built for 3D Euler flux_ranocha + LLF, 2D
                            Euler flux_ranocha + LLF with flux differencing or HG
                            shock capturing, 3D GLM-MHD flux_hindenlang_gassner +
                            LLF; periodic meshes): faster, less accurate             */
    TRIXI_HIP_FP32 = 3   /* Trixi's RealT = uEltype = Float32: state, du, the RK
                            register and all rhs! arithmetic in FP32 (logarithms
                            included), the RK update and the shock indicator
                            evaluated in FP64 from the FP32 values.  State vectors
                            of such a context are float: trixi_hip_rhs_f32 and
                            trixi_hip_rk_step_f32 (the FP64-vector entry points
                            reject it; pack/unpack, compute_coefficients and the
                            callbacks still take FP64 vectors).  Built for 2D/3D
                            Euler flux differencing (flux_ranocha + LLF or
                            flux_ranocha) and 2D Euler HG shock capturing
                            (flux_ranocha or flux_shima_etal + LLF) on periodic
                            meshes: the reference's benchmark configurations  */
};
enum trixi_hip_initial_condition {
    TRIXI_HIP_IC_CONSTANT = 0,
    TRIXI_HIP_IC_CONVERGENCE_TEST = 1, /* per equation: LinAdv / Burgers sine wave, Euler
                                          manufactured state, GLM-MHD Alfven wave  */
    TRIXI_HIP_IC_WEAK_BLAST_WAVE = 2,
    TRIXI_HIP_IC_SEEDED_RANDOM = 3,
    TRIXI_HIP_IC_SMOOTH_PERIODIC = 4,
    TRIXI_HIP_IC_WEAK_BLAST_WAVE_SEEDED = 5, /* Euler: weak blast + seeded U(-0.01, 0.01)
                                               on rho, v, p per node               */
    TRIXI_HIP_IC_SHOCK = 6,            /* Burgers: 1.5 for x < 0.5, else 0.5
                                          (test/tree_dgsem_1d/burgers_shock.jl)    */
    TRIXI_HIP_IC_RAREFACTION = 7,      /* Burgers: 0.5 for x < 0.5, else 1.5
                                          (test/tree_dgsem_1d/burgers_rarefraction.jl) */
    TRIXI_HIP_IC_EOC_TEST_COUPLED_EULER_GRAVITY = 8, /* Euler 2D/3D: rho = 2 + 0.1 sin(pi (sum x - t)),
                                          v = 1, p = 2 rho^2 / (d pi) (gamma = 2) */
    TRIXI_HIP_IC_POISSON_NONPERIODIC = 9,  /* hyperbolic diffusion: the initial guess at
                                          t = 0, else the steady solution (as a
                                          boundary state: always the steady solution,
                                          boundary_condition_poisson_nonperiodic) */
    TRIXI_HIP_IC_HARMONIC_NONPERIODIC = 10 /* hyperbolic diffusion 1D: phi = 5, q = 0 at
                                          t = 0, else phi = 3 + e x, q = e
                                          (test/tree_dgsem_1d/hypdiff_harmonic_nonperiodic.jl) */
};

/* Boundary condition on the faces of a non-periodic TreeMesh that have no
 * interface (Trixi's boundary_conditions argument; reference cuda_boundary_flux!
 * dg_2d.jl:453-494 and boundary_flux_kernel!, the same condition on every side
 * as in test/tree_dgsem_2d/euler_source_terms_nonperiodic.jl:12-17). */
enum trixi_hip_boundary_condition {
    TRIXI_HIP_BC_PERIODIC = 0,  /* every face has an interface                     */
    TRIXI_HIP_BC_DIRICHLET = 1, /* BoundaryConditionDirichlet(boundary_ic): state
                                   boundary_ic(x_face, t); flux(u_inner, u_b) on
                                   + faces, flux(u_b, u_inner) on - faces        */
    TRIXI_HIP_BC_SLIP_WALL = 2  /* boundary_condition_slip_wall (shallow water): the
                                   inner state with the normal momentum reversed */
};

/* Source terms added after the Jacobian (reference cuda_sources!,
 * dg_2d.jl:746-757; source_terms_kernel!). */
enum trixi_hip_source_terms {
    TRIXI_HIP_SOURCE_NONE = 0,
    TRIXI_HIP_SOURCE_CONVERGENCE_TEST = 1, /* source_terms_convergence_test (Euler, Burgers,
                                              quasi-1D Euler)                            */
    TRIXI_HIP_SOURCE_EOC_TEST_EULER = 2,   /* source_terms_eoc_test_euler (Euler 2D/3D)     */
    TRIXI_HIP_SOURCE_POISSON_NONPERIODIC = 3, /* hyperbolic diffusion: f of the steady state,
                                             -q / T_r                                  */
    TRIXI_HIP_SOURCE_HARMONIC = 4          /* hyperbolic diffusion: 0, -q / T_r            */
};

/* IndicatorHennemannGassner(...; variable) (Trixi's density_pressure,
 * pressure and density functions of compressible Euler). */
enum trixi_hip_indicator_variable {
    TRIXI_HIP_INDICATOR_DENSITY_PRESSURE = 0,
    TRIXI_HIP_INDICATOR_PRESSURE = 1,
    TRIXI_HIP_INDICATOR_DENSITY = 2
};

/* Solver configuration: DGSEMGPU(; polydeg, surface_flux, volume_integral)
 * (reference src/solvers/dgsem_gpu.jl:43-53) plus the equation parameters. */
typedef struct trixi_hip_config {
    int32_t ndims;            /* 1, 2 or 3                                        */
    int32_t equation;         /* enum trixi_hip_equation                          */
    int32_t polydeg;          /* N - 1; 3 for every headline configuration        */
    int32_t volume_integral;  /* enum trixi_hip_volume_integral                   */
    int32_t volume_flux;      /* DG volume flux (flux differencing / HG)          */
    int32_t surface_flux;     /* numerical interface flux                         */
    int32_t fv_flux;          /* HG subcell finite-volume flux                    */
    int32_t precision;        /* enum trixi_hip_precision                         */
    int32_t alpha_smooth;     /* HG: smooth blending factors with neighbours      */
    int32_t indicator_variable; /* HG: enum trixi_hip_indicator_variable          */
    double gamma;             /* ratio of specific heats (Euler, MHD)             */
    double c_h;               /* GLM divergence-cleaning speed                    */
    double advection_velocity[3];
    double alpha_max, alpha_min; /* HG indicator parameters                       */
    double ic_length;         /* period used by TRIXI_HIP_IC_SMOOTH_PERIODIC      */
    int32_t boundary_condition; /* enum trixi_hip_boundary_condition              */
    int32_t boundary_ic;      /* Dirichlet state: enum trixi_hip_initial_condition */
    uint64_t boundary_seed;   /* seed of a seeded boundary_ic                     */
    int32_t source_terms;     /* enum trixi_hip_source_terms                      */
    int32_t boundary_outflow_faces; /* bit f (face direction -x,+x,-y,+y,-z,+z = 0..5)
                                       set: that side takes the outflow state
                                       (the inner state, flux(u_inner)) instead of
                                       the Dirichlet state                        */
    double nu;                /* hyperbolic diffusion: diffusion constant (1)     */
    double Lr;                /* hyperbolic diffusion: reference length 1/(2 pi); T_r = Lr^2 / nu */
    int32_t ncomponents;      /* multicomponent Euler: number of gas components   */
    int32_t reserved;         /* 0                                                */
    double gammas[4];         /* multicomponent Euler: gamma_k of each component  */
    double gas_constants[4];  /* multicomponent Euler: R_k of each component      */
    double gravity;           /* shallow water: gravitational acceleration g      */
    double H0;                /* shallow water: reference water height (unused by
                                 the built states)                                */
} trixi_hip_config;

/* Mesh and connectivity (host arrays, copied to the device by create): the
 * cache the reference builds in create_cache_gpu (src/solvers/cache.jl:130-158,
 * containers_3d.jl:6-96). */
typedef struct trixi_hip_mesh {
    int64_t nelements;
    int64_t ninterfaces;
    const int32_t* neighbor_ids;     /* [2*ninterfaces]: (left, right), 0-based        */
    const int32_t* orientations;     /* [ninterfaces]: 1..ndims                        */
    const double* inverse_jacobian;  /* [nelements]                                    */
    const double* node_coordinates;  /* Trixi layout [d, i, j, k, element]; may be NULL */
    double total_volume;
    /* L2 mortars of a refined mesh (Trixi's init_mortars! container, as the
     * reference copies it in create_cache_gpu, cache.jl:150-158): 0 for a
     * conforming mesh, then the pointers may be NULL.  2D and 3D only. */
    int64_t nmortars;
    const int32_t* mortar_neighbor_ids; /* [nmortars][2^(d-1)+1]: the small elements
                                           (lower, upper; 3D lower_left, lower_right,
                                           upper_left, upper_right), then the large */
    const int32_t* mortar_orientations; /* [nmortars]: 1..ndims                      */
    const int32_t* mortar_large_sides;  /* [nmortars]: 1 large element on the left
                                           (-side), 2 on the right                   */
} trixi_hip_mesh;

/* Library version string. */
const char* trixi_hip_version(void);

/* Message of the last failed call on this thread. */
const char* trixi_hip_last_error(void);

/* Lobatto-Legendre basis of `nnodes` nodes, computed on the host (no device
 * needed): the arrays of LobattoLegendreBasisGPU (basis_lobatto_legendre.jl:
 * 60-100).  Matrices are row-major N x N; boundary_interpolation is N x 2
 * (row-major: [j][0] = lhat_j(-1), [j][1] = lhat_j(+1)).  Any pointer may be
 * NULL. */
int trixi_hip_basis(int32_t nnodes, double* nodes, double* weights, double* derivative_matrix,
                    double* derivative_split, double* derivative_dhat,
                    double* boundary_interpolation, double* inverse_vandermonde_legendre);

/* Build a context: basis, device copies of the mesh, all scratch buffers.
 * Replaces LobattoLegendreBasisGPU + DGSEMGPU + SemidiscretizationHyperbolicGPU
 * + create_cache_gpu (basis_lobatto_legendre.jl:60-100, dgsem_gpu.jl:43-53,
 * semidiscretization_hyperbolic.jl:60-87, cache.jl:130-212).  Returns
 * TRIXI_HIP_EUNSUPPORTED for combinations that are not built. */
int trixi_hip_create(const trixi_hip_config* config, const trixi_hip_mesh* mesh,
                     trixi_hip_ctx** out);
int trixi_hip_destroy(trixi_hip_ctx* ctx);

/* Sizes: nvars, nnodes per direction, values per state vector (nv*N^d*E). */
int trixi_hip_sizes(const trixi_hip_ctx* ctx, int32_t* nvars, int32_t* nnodes,
                    int64_t* nvalues);

/* du = rhs(u, t), in place into du, asynchronous on `stream`.
 * Replaces rhs_gpu!(du_ode, u_ode, semi, t) (solvers.jl:18-31) and the stage
 * pipeline rhs_gpu!(du, u, t, mesh::TreeMesh{d}, ...) (dg_1d.jl:523-548,
 * dg_2d.jl:760-790, dg_3d.jl:895-925).  u and du are device pointers in the
 * native layout; they must not alias. */
int trixi_hip_rhs(trixi_hip_ctx* ctx, double* du, const double* u, double t,
                  trixi_hip_stream stream);

/* Layout conversion, device to device: Trixi u[v,i,j,k,e] <-> native.
 * Stands in for wrap_array / wrap_array_native (dg.jl:14-32). */
int trixi_hip_pack(trixi_hip_ctx* ctx, double* u_native, const double* u_trixi,
                   trixi_hip_stream stream);
int trixi_hip_unpack(trixi_hip_ctx* ctx, double* u_trixi, const double* u_native,
                     trixi_hip_stream stream);

/* u = initial_condition(x_node, t) at every node (native layout), incl. the
 * 1D end-node nudge of compute_coefficients_kernel! (dg.jl:46-68, 148-160). */
int trixi_hip_compute_coefficients(trixi_hip_ctx* ctx, double* u, int32_t ic,
                                   double t, uint64_t seed, trixi_hip_stream stream);

/* One CarpenterKennedy2N54 step of size dt from time t (5 fused rhs + FP64
 * register-update launches): u <- u(t+dt), k is the 2N register.  With
 * use_graph != 0 the five stages are captured once into a hipGraph per
 * (u, k) pair and replayed; t and dt are passed through device memory.
 * Replaces the OrdinaryDiffEq stage loop around rhs_gpu! used by the
 * reference examples (e.g. examples/euler_ec_3d.jl:53-55). */
int trixi_hip_rk_step(trixi_hip_ctx* ctx, double* u, double* k, double t, double dt,
                      int32_t use_graph, trixi_hip_stream stream);

/* Stage s (0..4) of the CarpenterKennedy2N54 step that starts at time t:
 * k <- A_s k + dt rhs(u, t + c_s dt) (k is overwritten when s == 0),
 * u_out <- u + B_s k, in one fused launch (plus the indicator for shock
 * capturing).  u_out must not alias u: neighbouring elements read u's face
 * nodes while the stage writes.  For drivers that step stage by stage, as
 * OrdinaryDiffEq's 2N loop around rhs_gpu! does (solvers.jl:18-31). */
int trixi_hip_rk_stage(trixi_hip_ctx* ctx, const double* u, double* k, double* u_out, int32_t stage,
                       double t, double dt, trixi_hip_stream stream);

/* Block until the work queued on the stream has finished; reports an
 * asynchronous device error of earlier calls as TRIXI_HIP_EDEVICE.  The
 * counterpart of CUDA.synchronize() in the reference benchmarks
 * (benchmark/euler_ec_3d.jl:58-72). */
int trixi_hip_sync(trixi_hip_stream stream);

/* rhs! of an FP32-state context (precision TRIXI_HIP_FP32): as trixi_hip_rhs
 * with float state vectors.  Replaces rhs_gpu! with RealT = uEltype =
 * Float32 (solvers.jl:18-31; semidiscretization_hyperbolic.jl:60-66;
 * benchmark/euler_ec_3d.jl:6). */
int trixi_hip_rhs_f32(trixi_hip_ctx* ctx, float* du, const float* u, double t,
                      trixi_hip_stream stream);

/* One CarpenterKennedy2N54 step of an FP32-state context, as
 * trixi_hip_rk_step with float u and k (the update itself is evaluated in
 * FP64 and rounded to FP32).  Replaces the OrdinaryDiffEq step on Float32
 * CuArrays (examples/*.jl, e.g. euler_ec_3d.jl:53-55). */
int trixi_hip_rk_step_f32(trixi_hip_ctx* ctx, float* u, float* k, double t, double dt,
                          int32_t use_graph, trixi_hip_stream stream);

/* max_dt = 2 / (N * max_e invJ_e * sum_d lambda_d) on the device; blocks until
 * the scalar is on the host.  Replaces max_dt(u::CuArray, ...)
 * (stepsize_dg_{1,2,3}d.jl, e.g. stepsize_dg_3d.jl:20-45). */
int trixi_hip_max_dt(trixi_hip_ctx* ctx, const double* u, double* max_dt,
                     trixi_hip_stream stream);

/* trixi_hip_max_dt without any allocation: the block partials and the maximum
 * go to a static device scratch of 2048 doubles (16 KiB, held by every GPU
 * the library runs on until the process ends) instead of the context's
 * buffers; the result is the same, and like trixi_hip_max_dt it blocks until
 * the scalar is on the host.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (tight time-stepping loops,
 * harnesses that forbid hipMalloc).
 * Limit: at most 2047 block partials (the grid is at most 1024 blocks of 256
 * elements, so every mesh fits); larger grids are rejected with
 * TRIXI_HIP_EINVAL.  All calls on one GPU share the scratch: calls that
 * overlap on one GPU (several host threads or streams) race. */
int trixi_hip_max_dt_noalloc(trixi_hip_ctx* ctx, const double* u, double* max_dt,
                             trixi_hip_stream stream);

/* L2 and Linf errors against the initial condition at time t (2N-node
 * analysis basis); blocks.  Replaces calc_error_norms
 * (analysis_dg_{1,2,3}d.jl, e.g. analysis_dg_3d.jl:45-89). */
int trixi_hip_error_norms(trixi_hip_ctx* ctx, const double* u, int32_t ic, double t,
                          uint64_t seed, double* l2, double* linf,
                          trixi_hip_stream stream);

/* Quadrature integral of every conservative variable over the domain; blocks.
 * Replaces integrate / integrate_via_indices (analysis_dg_3d.jl:1-43). */
int trixi_hip_integrate(trixi_hip_ctx* ctx, const double* u, double* integral,
                        trixi_hip_stream stream);

/* Hennemann-Gassner blending factors (after smoothing) into the device array
 * alpha[nelements], in FP64.  Replaces the CPU indicator callable and its
 * D2H/H2D round trip (indicators.jl:7-42, dg_2d.jl:171-173).  Built for
 * compressible Euler (any indicator variable), 2D/3D GLM-MHD (density_pressure)
 * and Burgers (first); other equations return TRIXI_HIP_EUNSUPPORTED. */
int trixi_hip_indicator(trixi_hip_ctx* ctx, const double* u, double* alpha,
                        trixi_hip_stream stream);

/* Adjoint of rhs!: ubar = beta * ubar + (d rhs / d u)^T (alpha * vbar), FP64,
 * asynchronous.  No reference counterpart (TrixiCUDA.jl has no adjoint); it
 * backs a ChainRulesCore.rrule / torch.autograd.Function of rhs_gpu!
 * (solvers.jl:18-31).  Built for linear advection, compressible Euler (weak
 * form, flux differencing, HG shock capturing) and ideal GLM-MHD 2D and 3D
 * (flux differencing with the nonconservative Powell terms) on conforming
 * meshes; other configurations (and meshes with mortars) return
 * TRIXI_HIP_EUNSUPPORTED.  ubar must not alias u or vbar. */
int trixi_hip_rhs_vjp(trixi_hip_ctx* ctx, double* ubar, const double* u, const double* vbar,
                      double t, double alpha, double beta, trixi_hip_stream stream);

/* Allocate the adjoint's work buffers (7 state vectors).  Call once before
 * trixi_hip_rk_step_vjp.  Besides this call, only the first
 * trixi_hip_rhs_vjp of a shock-capturing context allocates after create
 * (two per-element arrays). */
int trixi_hip_enable_adjoint(trixi_hip_ctx* ctx);

/* Reverse of one CarpenterKennedy2N54 step from the step's initial state
 * u_n (a checkpoint, not modified): on entry ubar holds the cotangent of
 * u_{n+1}, on exit that of u_n.  The five stage inputs are recomputed in
 * FP64 from u_n, then the stages are reversed (register cotangent update +
 * rhs VJP per stage).  ubar must not alias u_n.  No reference counterpart
 * (multi-step gradients with RK-stage checkpointing). */
int trixi_hip_rk_step_vjp(trixi_hip_ctx* ctx, const double* u_n, double* ubar, double t,
                          double dt, trixi_hip_stream stream);

/* Positivity and finiteness check (FP64): writes a per-element flag
 * (bit 0: non-finite value, bit 1: rho <= 0, bit 2: p <= 0) into
 * flags[nelements] and the number of flagged elements to *nbad (blocks).
 * Returns TRIXI_HIP_ESTATE if any element is flagged. */
int trixi_hip_check_state(trixi_hip_ctx* ctx, const double* u, int32_t* flags,
                          int64_t* nbad, trixi_hip_stream stream);

#ifdef __cplusplus
}
#endif
#endif /* TRIXI_HIP_H */
