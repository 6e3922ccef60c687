"""Equations, numerical fluxes and initial conditions (host-side identities).

A Julia/Trixi function object cannot cross a C ABI, so each supported
Trixi.jl equation, flux and initial condition is an object carrying the
enum value the library understands (reference: the fluxes the tests pass to
DGSEMGPU, e.g. test/tree_dgsem_3d/euler_ec.jl:13-17,
test/tree_dgsem_3d/mhd_alfven_wave.jl:13-18, test/tree_dgsem_2d/euler_shock.jl:12-27).
"""
import math
from dataclasses import dataclass, field

from . import _lib as L


# ----------------------------------------------------------------------------
# equations
# ----------------------------------------------------------------------------
@dataclass(frozen=True)
class AbstractEquations:
    ndims: int
    equation_id: int

    @property
    def nvariables(self):
        if self.equation_id in (L.EQ_LINADV, L.EQ_BURGERS):
            return 1
        if self.equation_id == L.EQ_HYPDIFF:
            return self.ndims + 1
        if self.equation_id == L.EQ_EULER:
            return self.ndims + 2
        if self.equation_id == L.EQ_EULER_MULTI:
            return self.ndims + 1 + len(self.gammas)
        if self.equation_id == L.EQ_EULER_QUASI1D:
            return 4
        if self.equation_id == L.EQ_SHALLOW_WATER:
            return self.ndims + 2
        return 9 if self.ndims > 1 else 8  # ideal MHD: no cleaning field in 1D


def _linadv(ndims):
    @dataclass(frozen=True)
    class _LinAdv(AbstractEquations):
        advection_velocity: tuple = ()

        def __init__(self, advection_velocity):
            vel = tuple(float(a) for a in (advection_velocity if hasattr(advection_velocity, "__len__")
                                           else (advection_velocity,)))
            if len(vel) != ndims:
                raise ValueError(f"advection velocity must have {ndims} components")
            object.__setattr__(self, "ndims", ndims)
            object.__setattr__(self, "equation_id", L.EQ_LINADV)
            object.__setattr__(self, "advection_velocity", vel)

    _LinAdv.__name__ = _LinAdv.__qualname__ = f"LinearScalarAdvectionEquation{ndims}D"
    return _LinAdv


LinearScalarAdvectionEquation1D = _linadv(1)
LinearScalarAdvectionEquation2D = _linadv(2)
LinearScalarAdvectionEquation3D = _linadv(3)


def _euler(ndims):
    @dataclass(frozen=True)
    class _Euler(AbstractEquations):
        gamma: float = 1.4

        def __init__(self, gamma):
            object.__setattr__(self, "ndims", ndims)
            object.__setattr__(self, "equation_id", L.EQ_EULER)
            object.__setattr__(self, "gamma", float(gamma))

    _Euler.__name__ = _Euler.__qualname__ = f"CompressibleEulerEquations{ndims}D"
    return _Euler


CompressibleEulerEquations1D = _euler(1)
CompressibleEulerEquations2D = _euler(2)
CompressibleEulerEquations3D = _euler(3)


def _euler_multi(ndims):
    @dataclass(frozen=True)
    class _EulerMulti(AbstractEquations):
        """(rho v_1..d, rho e, rho_1..rho_n): n gas components with their own
        gamma_k and R_k (Trixi CompressibleEulerMulticomponentEquations{1,2}D;
        keyword arguments, a scalar meaning one component)."""
        gammas: tuple = ()
        gas_constants: tuple = ()

        def __init__(self, *, gammas, gas_constants):
            g = tuple(float(x) for x in (gammas if hasattr(gammas, "__len__") else (gammas,)))
            r = tuple(float(x) for x in (gas_constants if hasattr(gas_constants, "__len__") else (gas_constants,)))
            if len(g) != len(r) or not g:
                raise ValueError("gammas and gas_constants must have the same (non-zero) length")
            object.__setattr__(self, "ndims", ndims)
            object.__setattr__(self, "equation_id", L.EQ_EULER_MULTI)
            object.__setattr__(self, "gammas", g)
            object.__setattr__(self, "gas_constants", r)

        @property
        def ncomponents(self):
            return len(self.gammas)

    _EulerMulti.__name__ = _EulerMulti.__qualname__ = f"CompressibleEulerMulticomponentEquations{ndims}D"
    return _EulerMulti


CompressibleEulerMulticomponentEquations1D = _euler_multi(1)
CompressibleEulerMulticomponentEquations2D = _euler_multi(2)


@dataclass(frozen=True)
class CompressibleEulerEquationsQuasi1D(AbstractEquations):
    """(a rho, a rho v, a e, a) with the nozzle width a (Trixi
    CompressibleEulerEquationsQuasi1D); nonconservative pressure term a p_x."""
    gamma: float = 1.4

    def __init__(self, gamma):
        object.__setattr__(self, "ndims", 1)
        object.__setattr__(self, "equation_id", L.EQ_EULER_QUASI1D)
        object.__setattr__(self, "gamma", float(gamma))


def _shallow_water(ndims):
    @dataclass(frozen=True)
    class _ShallowWater(AbstractEquations):
        """(h, h v, b) with the bottom topography b (Trixi
        ShallowWaterEquations{1,2}D(; gravity_constant, H0))."""
        gravity: float = 9.81
        H0: float = 0.0

        def __init__(self, *, gravity_constant, H0=0.0):
            object.__setattr__(self, "ndims", ndims)
            object.__setattr__(self, "equation_id", L.EQ_SHALLOW_WATER)
            object.__setattr__(self, "gravity", float(gravity_constant))
            object.__setattr__(self, "H0", float(H0))

    _ShallowWater.__name__ = _ShallowWater.__qualname__ = f"ShallowWaterEquations{ndims}D"
    return _ShallowWater


ShallowWaterEquations1D = _shallow_water(1)
ShallowWaterEquations2D = _shallow_water(2)


def prim2cons(prim, equations):
    """Trixi's prim2cons for compressible Euler (rho, v.., p), quasi-1D Euler
    (rho, v, p, a), multicomponent Euler (v.., p, rho_1..rho_n) and shallow
    water (H, v1, v2, b), for
    initial conditions written as Python functions."""
    prim = [float(x) for x in prim]
    nd = equations.ndims
    if equations.equation_id == L.EQ_EULER:
        rho, v, p = prim[0], prim[1:1 + nd], prim[1 + nd]
        return [rho] + [rho * vi for vi in v] + [p / (equations.gamma - 1) + 0.5 * rho * sum(vi * vi for vi in v)]
    if equations.equation_id == L.EQ_SHALLOW_WATER:
        H, v, b = prim[0], prim[1:-1], prim[-1]
        return [H - b] + [(H - b) * vi for vi in v] + [b]
    if equations.equation_id == L.EQ_EULER_QUASI1D:
        rho, v, p, a = prim
        return [a * rho, a * rho * v, a * (p / (equations.gamma - 1) + 0.5 * rho * v * v), a]
    if equations.equation_id == L.EQ_EULER_MULTI:
        v, p, rk = prim[:nd], prim[nd], prim[nd + 1:]
        cv = [r / (g - 1) for g, r in zip(equations.gammas, equations.gas_constants)]
        gamma = sum(x * c * g for x, c, g in zip(rk, cv, equations.gammas)) / sum(x * c for x, c in zip(rk, cv))
        rho = sum(rk)
        return [rho * vi for vi in v] + [p / (gamma - 1) + 0.5 * rho * sum(vi * vi for vi in v)] + rk
    raise NotImplementedError(f"prim2cons is not provided for {type(equations).__name__}")


def _glm_mhd(ndims):
    @dataclass(frozen=True)
    class _GlmMhd(AbstractEquations):
        gamma: float = 5.0 / 3.0
        c_h: float = 1.0

        def __init__(self, gamma, initial_c_h=1.0):
            object.__setattr__(self, "ndims", ndims)
            object.__setattr__(self, "equation_id", L.EQ_GLM_MHD)
            object.__setattr__(self, "gamma", float(gamma))
            object.__setattr__(self, "c_h", float(initial_c_h))

    _GlmMhd.__name__ = _GlmMhd.__qualname__ = f"IdealGlmMhdEquations{ndims}D"
    return _GlmMhd


IdealGlmMhdEquations1D = _glm_mhd(1)  # (rho, rho v1..3, rho e, B1..3); no psi, no Powell terms
IdealGlmMhdEquations2D = _glm_mhd(2)
IdealGlmMhdEquations3D = _glm_mhd(3)


def _hypdiff(ndims):
    @dataclass(frozen=True)
    class _HypDiff(AbstractEquations):
        """(phi, q_1..d): phi_t - nu div q = f, q_t = (grad phi - q) / T_r,
        T_r = Lr^2 / nu (Trixi HyperbolicDiffusionEquations{1,2,3}D)."""
        nu: float = 1.0
        Lr: float = 1.0 / (2.0 * 3.141592653589793)

        def __init__(self, nu=1.0, Lr=None):
            object.__setattr__(self, "ndims", ndims)
            object.__setattr__(self, "equation_id", L.EQ_HYPDIFF)
            object.__setattr__(self, "nu", float(nu))
            object.__setattr__(self, "Lr", float(Lr) if Lr is not None else 1.0 / (2.0 * math.pi))

    _HypDiff.__name__ = _HypDiff.__qualname__ = f"HyperbolicDiffusionEquations{ndims}D"
    return _HypDiff


HyperbolicDiffusionEquations1D = _hypdiff(1)
HyperbolicDiffusionEquations2D = _hypdiff(2)
HyperbolicDiffusionEquations3D = _hypdiff(3)


@dataclass(frozen=True)
class InviscidBurgersEquation1D(AbstractEquations):
    def __init__(self):
        object.__setattr__(self, "ndims", 1)
        object.__setattr__(self, "equation_id", L.EQ_BURGERS)


# ----------------------------------------------------------------------------
# fluxes
# ----------------------------------------------------------------------------
@dataclass(frozen=True)
class NumericalFlux:
    name: str
    flux_id: int

    def __repr__(self):
        return self.name


def max_abs_speed(*args):  # identity marker of Trixi's max_abs_speed
    raise TypeError("wave-speed estimates are selected, not called, on the HIP path")


def max_abs_speed_naive(*args):  # identity marker of Trixi's max_abs_speed_naive
    raise TypeError("wave-speed estimates are selected, not called, on the HIP path")


def FluxLaxFriedrichs(max_abs_speed_fn=max_abs_speed):
    if max_abs_speed_fn is max_abs_speed_naive:
        return NumericalFlux("FluxLaxFriedrichs(max_abs_speed_naive)", L.FLUX_LLF_NAIVE)
    if max_abs_speed_fn is max_abs_speed:
        return NumericalFlux("FluxLaxFriedrichs(max_abs_speed)", L.FLUX_LLF)
    raise NotImplementedError("unsupported wave-speed estimate")


flux_central = NumericalFlux("flux_central", L.FLUX_CENTRAL)
flux_lax_friedrichs = FluxLaxFriedrichs(max_abs_speed)
flux_ranocha = NumericalFlux("flux_ranocha", L.FLUX_RANOCHA)
flux_shima_etal = NumericalFlux("flux_shima_etal", L.FLUX_SHIMA_ETAL)
flux_hindenlang_gassner = NumericalFlux("flux_hindenlang_gassner", L.FLUX_HINDENLANG_GASSNER)
flux_ec = NumericalFlux("flux_ec", L.FLUX_EC)  # entropy-conservative flux of the Burgers equation


def min_max_speed_davis(*args):  # identity marker of Trixi's min_max_speed_davis
    raise TypeError("wave-speed estimates are selected, not called, on the HIP path")


def min_max_speed_einfeldt(*args):  # identity marker of Trixi's min_max_speed_einfeldt
    raise TypeError("wave-speed estimates are selected, not called, on the HIP path")


def FluxHLL(min_max_speed=min_max_speed_davis):
    if min_max_speed is min_max_speed_einfeldt:
        return NumericalFlux("FluxHLL(min_max_speed_einfeldt)", L.FLUX_HLLE)
    if min_max_speed is not min_max_speed_davis:
        raise NotImplementedError("FluxHLL is built with min_max_speed_davis and min_max_speed_einfeldt")
    return NumericalFlux("FluxHLL(min_max_speed_davis)", L.FLUX_HLL)


flux_hll = FluxHLL()  # compressible Euler
flux_hlle = FluxHLL(min_max_speed_einfeldt)  # GLM-MHD
flux_nonconservative_powell = NumericalFlux("flux_nonconservative_powell", -1)
# quasi-1D Euler: Chan et al. (2023) entropy-conservative flux and its
# nonconservative pressure-gradient term (used as a pair)
flux_chan_etal = NumericalFlux("flux_chan_etal", L.FLUX_CHAN_ETAL)
flux_nonconservative_chan_etal = NumericalFlux("flux_nonconservative_chan_etal", -2)
# shallow water: the volume pair (flux_wintermeyer_etal, flux_nonconservative_wintermeyer_etal)
# and a surface flux with flux_nonconservative_fjordholm_etal
flux_wintermeyer_etal = NumericalFlux("flux_wintermeyer_etal", L.FLUX_WINTERMEYER_ETAL)
flux_fjordholm_etal = NumericalFlux("flux_fjordholm_etal", L.FLUX_FJORDHOLM_ETAL)
flux_nonconservative_wintermeyer_etal = NumericalFlux("flux_nonconservative_wintermeyer_etal", -3)
flux_nonconservative_fjordholm_etal = NumericalFlux("flux_nonconservative_fjordholm_etal", -4)
flux_nonconservative_audusse_etal = NumericalFlux("flux_nonconservative_audusse_etal", -5)


def hydrostatic_reconstruction_audusse_etal(*args):  # identity marker of Trixi's reconstruction
    raise TypeError("reconstructions are selected, not called, on the HIP path")


def FluxHydrostaticReconstruction(numerical_flux, hydrostatic_reconstruction):
    """LLF of Audusse's hydrostatically reconstructed states (shallow water),
    paired with flux_nonconservative_audusse_etal."""
    if hydrostatic_reconstruction is not hydrostatic_reconstruction_audusse_etal or \
            getattr(numerical_flux, "flux_id", None) != L.FLUX_LLF:
        raise NotImplementedError("FluxHydrostaticReconstruction is built for "
                                  "(flux_lax_friedrichs, hydrostatic_reconstruction_audusse_etal)")
    return NumericalFlux("FluxHydrostaticReconstruction(flux_lax_friedrichs, "
                         "hydrostatic_reconstruction_audusse_etal)", L.FLUX_HYDROSTATIC_LLF)


def density_pressure(*args):  # identity markers of the HG indicator variables
    raise TypeError("indicator variables are selected, not called, on the HIP path")


def pressure(*args):
    raise TypeError("indicator variables are selected, not called, on the HIP path")


def density(*args):
    raise TypeError("indicator variables are selected, not called, on the HIP path")


def first(*args):  # Julia's Base.first: the solution of a scalar equation as indicator variable
    raise TypeError("indicator variables are selected, not called, on the HIP path")


# ----------------------------------------------------------------------------
# initial conditions
# ----------------------------------------------------------------------------
@dataclass(frozen=True)
class InitialCondition:
    name: str
    ic_id: int
    seed: int = 0

    def __repr__(self):
        return self.name


initial_condition_constant = InitialCondition("initial_condition_constant", L.IC_CONSTANT)
initial_condition_convergence_test = InitialCondition("initial_condition_convergence_test", L.IC_CONVERGENCE_TEST)
initial_condition_weak_blast_wave = InitialCondition("initial_condition_weak_blast_wave", L.IC_WEAK_BLAST_WAVE)
# the Burgers step states of test/tree_dgsem_1d/burgers_shock.jl and burgers_rarefraction.jl
initial_condition_shock = InitialCondition("initial_condition_shock", L.IC_SHOCK)
initial_condition_rarefaction = InitialCondition("initial_condition_rarefaction", L.IC_RAREFACTION)
# compressible Euler 2D/3D: rho = 2 + 0.1 sin(pi (sum x - t)), v = 1, p = 2 rho^2 / (d pi) (gamma = 2)
initial_condition_eoc_test_coupled_euler_gravity = InitialCondition(
    "initial_condition_eoc_test_coupled_euler_gravity", L.IC_EOC_TEST_COUPLED_EULER_GRAVITY)
# hyperbolic diffusion: the initial guess phi = q = 1 at t = 0, else the steady
# solution of -lap phi = f (1D sin(pi x); 2D / 3D 2 cos(pi x) prod sin(2 pi x_k) + 2)
initial_condition_poisson_nonperiodic = InitialCondition("initial_condition_poisson_nonperiodic",
                                                         L.IC_POISSON_NONPERIODIC)
# test/tree_dgsem_1d/hypdiff_harmonic_nonperiodic.jl: phi = 5, q = 0 at t = 0, else 3 + e x, q = e
initial_condition_harmonic_nonperiodic = InitialCondition("initial_condition_harmonic_nonperiodic",
                                                          L.IC_HARMONIC_NONPERIODIC)


def initial_condition_seeded_random(seed=20261202):
    """rho = 1 + 0.1 U(-1,1), v = 0.1 N(0,1), p = 1 + 0.05 U(-1,1) per node,
    from a counter-based generator keyed by (seed, global node index)."""
    return InitialCondition(f"initial_condition_seeded_random(seed={seed})", L.IC_SEEDED_RANDOM, int(seed))


def initial_condition_weak_blast_wave_seeded(seed=20261202):
    """initial_condition_weak_blast_wave plus U(-0.01, 0.01) on rho, each v
    component and p per node (same generator), which removes the exact
    symmetries (zero velocity, identical neighbour states) at which the flux
    derivatives have kinks."""
    return InitialCondition(f"initial_condition_weak_blast_wave_seeded(seed={seed})", L.IC_WEAK_BLAST_WAVE_SEEDED,
                            int(seed))


# ----------------------------------------------------------------------------
# boundary conditions and source terms
# ----------------------------------------------------------------------------
class _Periodic:
    def __repr__(self):
        return "boundary_condition_periodic"


boundary_condition_periodic = _Periodic()


class _Outflow:
    """The outflow boundary of the reference's Burgers tests: the boundary flux
    is the physical flux of the inner state, flux(u_inner)."""

    def __repr__(self):
        return "boundary_condition_outflow"


boundary_condition_outflow = _Outflow()


class _PoissonNonperiodic:
    """Trixi's boundary_condition_poisson_nonperiodic: the steady Poisson state
    outside the face at any time (hyperbolic diffusion)."""

    def __repr__(self):
        return "boundary_condition_poisson_nonperiodic"


boundary_condition_poisson_nonperiodic = _PoissonNonperiodic()


class _SlipWall:
    """Trixi's boundary_condition_slip_wall (shallow water): the inner state
    with its normal momentum reversed outside the face."""

    def __repr__(self):
        return "boundary_condition_slip_wall"


boundary_condition_slip_wall = _SlipWall()


def waterheight_pressure(*args):  # identity marker of the shallow-water HG indicator variable
    raise TypeError("indicator variables are selected, not called, on the HIP path")


class BoundaryConditionDirichlet:
    """Trixi's BoundaryConditionDirichlet(boundary_value_function): the state
    boundary_value_function(x, t) outside the face; flux(u_inner, u_b) on +
    faces, flux(u_b, u_inner) on - faces.  Built for compressible Euler with
    initial_condition_convergence_test or initial_condition_constant."""

    def __init__(self, boundary_value_function, _steady=False):
        if not isinstance(boundary_value_function, InitialCondition):
            raise NotImplementedError("the boundary state must be one of the built initial conditions")
        self.boundary_value_function = boundary_value_function
        self._steady = _steady  # set by boundary_condition_poisson_nonperiodic

    def __repr__(self):
        return f"BoundaryConditionDirichlet({self.boundary_value_function})"


@dataclass(frozen=True)
class SourceTerms:
    name: str
    source_id: int

    def __repr__(self):
        return self.name


# Trixi's source_terms_convergence_test for compressible Euler: makes
# initial_condition_convergence_test an exact solution.
source_terms_convergence_test = SourceTerms("source_terms_convergence_test", L.SOURCE_CONVERGENCE_TEST)
# the manufactured source of initial_condition_eoc_test_coupled_euler_gravity (Euler 2D/3D)
source_terms_eoc_test_euler = SourceTerms("source_terms_eoc_test_euler", L.SOURCE_EOC_TEST_EULER)
# hyperbolic diffusion: f = -nu lap phi of the Poisson steady state, and -q / T_r
source_terms_poisson_nonperiodic = SourceTerms("source_terms_poisson_nonperiodic", L.SOURCE_POISSON_NONPERIODIC)
source_terms_harmonic = SourceTerms("source_terms_harmonic", L.SOURCE_HARMONIC)


def initial_condition_smooth_periodic(seed=20261202):
    """Seeded sums of three periodic sine modes per primitive variable."""
    return InitialCondition(f"initial_condition_smooth_periodic(seed={seed})", L.IC_SMOOTH_PERIODIC, int(seed))


def split_flux(flux):
    """Split a (conservative, nonconservative) tuple as Trixi accepts it."""
    if isinstance(flux, tuple):
        if len(flux) != 2 or flux[1] not in (flux_nonconservative_powell, flux_nonconservative_chan_etal,
                                             flux_nonconservative_wintermeyer_etal,
                                             flux_nonconservative_fjordholm_etal,
                                             flux_nonconservative_audusse_etal):
            raise NotImplementedError(f"unsupported flux tuple {flux}")
        return flux[0], flux[1]
    return flux, None
