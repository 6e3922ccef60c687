"""trixi_hip -- MI355X-native DGSEM semidiscretization for Trixi.jl's TreeMesh.

Host package of libtrixihip.so (hand-written gfx950 HIP kernels behind a C
ABI, include/trixi_hip.h).  The public names are HIP-named twins of
TrixiCUDA.jl's exports (src/TrixiCUDA.jl:74-77): LobattoLegendreBasisHIP,
DGSEMHIP, SemidiscretizationHyperbolicHIP, semidiscretizeHIP, plus rhs_hip_
(rhs_gpu!) and the GPU-resident callbacks.

This directory is not a valid Python identifier; load it as ``trixi_hip``:

    spec = importlib.util.spec_from_file_location(
        "trixi_hip", "trixicuda.jl_amd/__init__.py",
        submodule_search_locations=["trixicuda.jl_amd"])
"""
from ._lib import LIBRARY_PATH, TrixiHipError, version
from .save_solution import (SaveSolutionCallback, cons2cons, cons2prim, save_mesh_file, save_solution_file, varnames,
                            write_solution_file)
from .callbacks import (AnalysisCallback, StepsizeCallback, blending_factors, calc_error_norms, check_state,
                        integrate, max_dt)
from .equations import (BoundaryConditionDirichlet, boundary_condition_periodic, source_terms_convergence_test,
                        CompressibleEulerEquations1D, CompressibleEulerEquations2D, CompressibleEulerEquations3D,
                        CompressibleEulerMulticomponentEquations1D, CompressibleEulerMulticomponentEquations2D,
                        CompressibleEulerEquationsQuasi1D, flux_chan_etal, flux_nonconservative_chan_etal, prim2cons,
                        flux_hlle, min_max_speed_einfeldt, ShallowWaterEquations2D, flux_wintermeyer_etal,
                        flux_fjordholm_etal, flux_nonconservative_wintermeyer_etal,
                        flux_nonconservative_fjordholm_etal, ShallowWaterEquations1D, FluxHydrostaticReconstruction,
                        hydrostatic_reconstruction_audusse_etal, flux_nonconservative_audusse_etal,
                        boundary_condition_slip_wall, waterheight_pressure,
                        FluxLaxFriedrichs, IdealGlmMhdEquations1D, IdealGlmMhdEquations2D, IdealGlmMhdEquations3D, LinearScalarAdvectionEquation1D,
                        LinearScalarAdvectionEquation2D, LinearScalarAdvectionEquation3D, density, density_pressure, pressure,
                        InviscidBurgersEquation1D, first, flux_ec, initial_condition_shock, FluxHLL, flux_hll,
                        min_max_speed_davis, initial_condition_eoc_test_coupled_euler_gravity,
                        source_terms_eoc_test_euler, HyperbolicDiffusionEquations1D, HyperbolicDiffusionEquations2D,
                        HyperbolicDiffusionEquations3D, initial_condition_poisson_nonperiodic,
                        initial_condition_harmonic_nonperiodic, source_terms_poisson_nonperiodic,
                        source_terms_harmonic, boundary_condition_poisson_nonperiodic,
                        initial_condition_rarefaction, boundary_condition_outflow,
                        flux_central, flux_hindenlang_gassner, flux_lax_friedrichs, flux_nonconservative_powell,
                        flux_ranocha, flux_shima_etal, initial_condition_constant,
                        initial_condition_convergence_test, initial_condition_seeded_random,
                        initial_condition_smooth_periodic, initial_condition_weak_blast_wave,
                        initial_condition_weak_blast_wave_seeded, max_abs_speed,
                        max_abs_speed_naive)
from .mesh import TreeMesh
from .semidiscretization import (ODEProblem, SemidiscretizationHyperbolicHIP, compute_coefficients, pack,
                                 rhs_hip_, semidiscretizeHIP, wrap_array, wrap_array_native)
from .solver import (DGSEMHIP, IndicatorHennemannGassner, LobattoLegendreBasisHIP, VolumeIntegralFluxDifferencing,
                     VolumeIntegralShockCapturingHG, VolumeIntegralWeakForm)
from .time_integration import CarpenterKennedy2N54, Solution, rk_stage, rk_step, solve, synchronize
from .adjoint import (backward, enable_adjoint, forward_with_checkpoints, l2_norm_and_gradient, quadrature_weights,
                      rhs, rhs_vjp, rk_step_vjp)

__all__ = [n for n in dir() if not n.startswith("_")]
