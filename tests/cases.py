"""Shared case definitions: one description drives both the HIP path (via the
host package) and the oracle, so parity tests compare like with like."""
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Case:
    name: str
    ndims: int
    equation: str           # "linadv" | "euler" | "mhd" | "burgers" | "hypdiff" | "eulermulti" | "quasi1d" | "swe"
    level: int
    vi: str = "weak"        # "weak" | "fd" | "hg"
    vflux: str = "central"
    sflux: str = "llf"
    fvflux: str = "llf"
    polydeg: int = 3
    cmin: float = -2.0
    cmax: float = 2.0
    ic: str = "seeded_random"
    gamma: float = 1.4
    c_h: float = 1.0
    advection: tuple = (1.0, 0.5, 0.25)
    alpha_max: float = 0.5
    alpha_min: float = 0.001
    alpha_smooth: bool = True
    precision: str = "fp64"
    seed: int = 20261202
    periodic: bool = True
    bc_state: str = ""      # Dirichlet state on the boundary faces of a non-periodic mesh
    source: str = ""        # "convergence_test" for source_terms_convergence_test
    t: float = 0.0          # time of the rhs! evaluation
    patches: tuple = ()     # box refinement patches ((lo, hi), ...): L2 mortars
    indicator: str = "density_pressure"  # HG indicator variable: density_pressure | pressure | density | first
    outflow: tuple = ()     # boundary sides ("x_neg", ...) with the outflow condition
    nu: float = 1.0         # hyperbolic diffusion constant
    gammas: tuple = (1.4, 1.4)          # multicomponent Euler
    gas_constants: tuple = (0.4, 0.4)
    ic_fn: object = None    # initial condition as a Python function (x, t, equations) -> conservative state
    gravity: float = 9.81   # shallow water
    rk_cfl: float = 0.5     # RK step tests: dt = rk_cfl * max_dt


FLUX_NAMES = ("central", "llf", "llf_naive", "ranocha", "shima_etal", "hindenlang_gassner", "ec", "hll", "chan_etal",
              "hlle", "wintermeyer_etal", "fjordholm_etal", "hydrostatic_llf")
IC_NAMES = ("constant", "convergence_test", "weak_blast_wave", "seeded_random", "smooth_periodic",
            "weak_blast_wave_seeded", "shock", "rarefaction", "eoc_gravity", "poisson_nonperiodic",
            "harmonic_nonperiodic")
SOURCES = {"": 0, "convergence_test": 1, "eoc_test_euler": 2, "poisson_nonperiodic": 3, "harmonic": 4}
SIDES = ("x_neg", "x_pos", "y_neg", "y_pos", "z_neg", "z_pos")
INDICATORS = {"density_pressure": 0, "pressure": 1, "density": 2, "first": 2, "waterheight_pressure": 0}


def oracle_mesh(O, c: Case):
    if c.patches:
        return O.RefinedMesh(c.ndims, c.level, c.cmin, c.cmax, c.polydeg, c.patches, periodic=c.periodic)
    return O.Mesh(c.ndims, c.level, c.cmin, c.cmax, c.polydeg, periodic=c.periodic)


def oracle_cfg(O, c: Case):
    eq = {"linadv": O.EQ_LINADV, "euler": O.EQ_EULER, "mhd": O.EQ_GLM_MHD, "burgers": O.EQ_BURGERS,
          "hypdiff": O.EQ_HYPDIFF, "eulermulti": O.EQ_EULER_MULTI, "quasi1d": O.EQ_EULER_QUASI1D,
          "swe": O.EQ_SHALLOW_WATER}[c.equation]
    vi = {"weak": O.VI_WEAK, "fd": O.VI_FD, "hg": O.VI_HG}[c.vi]
    ext = dict(boundary_condition=(O.BC_SLIP_WALL if c.bc_state == "slip_wall" else
                                   O.BC_DIRICHLET if c.bc_state else O.BC_PERIODIC),
               boundary_ic=IC_NAMES.index(c.bc_state) if c.bc_state and c.bc_state != "slip_wall" else 0,
               source_terms=SOURCES[c.source],
               boundary_outflow_faces=sum(1 << SIDES.index(s) for s in c.outflow))
    return O.make_cfg(c.ndims, eq, c.polydeg, vi, FLUX_NAMES.index(c.vflux), FLUX_NAMES.index(c.sflux),
                      FLUX_NAMES.index(c.fvflux), gamma=c.gamma, c_h=c.c_h,
                      advection_velocity=c.advection[:c.ndims], alpha_max=c.alpha_max, alpha_min=c.alpha_min,
                      alpha_smooth=c.alpha_smooth, ic_length=c.cmax - c.cmin,
                      indicator_variable=INDICATORS[c.indicator], nu=c.nu,
                      gammas=c.gammas if c.equation == "eulermulti" else (),
                      gas_constants=c.gas_constants if c.equation == "eulermulti" else (), gravity=c.gravity, **ext)


def oracle_ic(O, c: Case):
    return IC_NAMES.index(c.ic)


def _flux(th, name, pair):
    f = {"central": th.flux_central, "llf": th.flux_lax_friedrichs,
         "llf_naive": th.FluxLaxFriedrichs(th.max_abs_speed_naive), "ranocha": th.flux_ranocha,
         "shima_etal": th.flux_shima_etal, "hindenlang_gassner": th.flux_hindenlang_gassner, "ec": th.flux_ec,
         "hll": th.flux_hll, "chan_etal": th.flux_chan_etal, "hlle": th.flux_hlle,
         "wintermeyer_etal": th.flux_wintermeyer_etal, "fjordholm_etal": th.flux_fjordholm_etal,
         "hydrostatic_llf": th.FluxHydrostaticReconstruction(th.flux_lax_friedrichs,
                                                             th.hydrostatic_reconstruction_audusse_etal)}[name]
    if pair == "mhd":
        return (f, th.flux_nonconservative_powell)
    if pair == "quasi1d":
        return (f, th.flux_nonconservative_chan_etal)
    if pair == "swe_volume":
        return (f, th.flux_nonconservative_wintermeyer_etal)
    if pair == "swe_surface":  # (the term belongs to the surface / subcell flux)
        return (f, th.flux_nonconservative_audusse_etal if name == "hydrostatic_llf"
                else th.flux_nonconservative_fjordholm_etal)
    return f


def product_semi(th, c: Case, precision=None):
    # (flux, nonconservative flux) pairs
    mhd = "mhd" if (c.equation == "mhd" and c.ndims > 1) else ("quasi1d" if c.equation == "quasi1d" else None)
    if c.equation == "linadv":
        eq = [None, th.LinearScalarAdvectionEquation1D, th.LinearScalarAdvectionEquation2D,
              th.LinearScalarAdvectionEquation3D][c.ndims](c.advection[:c.ndims])
    elif c.equation == "burgers":
        eq = th.InviscidBurgersEquation1D()
    elif c.equation == "hypdiff":
        eq = [None, th.HyperbolicDiffusionEquations1D, th.HyperbolicDiffusionEquations2D,
              th.HyperbolicDiffusionEquations3D][c.ndims](nu=c.nu)
    elif c.equation == "euler":
        eq = [None, th.CompressibleEulerEquations1D, th.CompressibleEulerEquations2D,
              th.CompressibleEulerEquations3D][c.ndims](c.gamma)
    elif c.equation == "quasi1d":
        eq = th.CompressibleEulerEquationsQuasi1D(c.gamma)
    elif c.equation == "swe":
        eq = [None, th.ShallowWaterEquations1D, th.ShallowWaterEquations2D][c.ndims](gravity_constant=c.gravity)
    elif c.equation == "eulermulti":
        eq = [None, th.CompressibleEulerMulticomponentEquations1D,
              th.CompressibleEulerMulticomponentEquations2D][c.ndims](gammas=c.gammas,
                                                                      gas_constants=c.gas_constants)
    else:
        eq = [None, th.IdealGlmMhdEquations1D, th.IdealGlmMhdEquations2D, th.IdealGlmMhdEquations3D][c.ndims](c.gamma,
                                                                                                        c.c_h)
    basis = th.LobattoLegendreBasisHIP(c.polydeg)
    if c.vi == "weak":
        vi = th.VolumeIntegralWeakForm()
    elif c.vi == "fd":
        vi = th.VolumeIntegralFluxDifferencing(_flux(th, c.vflux, "swe_volume" if c.equation == "swe" else mhd))
    else:
        ind = th.IndicatorHennemannGassner(eq, basis, alpha_max=c.alpha_max, alpha_min=c.alpha_min,
                                           alpha_smooth=c.alpha_smooth, variable=getattr(th, c.indicator))
        swe = c.equation == "swe"
        vi = th.VolumeIntegralShockCapturingHG(ind, volume_flux_dg=_flux(th, c.vflux, "swe_volume" if swe else mhd),
                                               volume_flux_fv=_flux(th, c.fvflux, "swe_surface" if swe else mhd))
    solver = th.DGSEMHIP(polydeg=c.polydeg, surface_flux=_flux(th, c.sflux, "swe_surface" if c.equation == "swe" else mhd),
                         volume_integral=vi,
                         precision=precision or c.precision)
    patches = tuple(dict(type="box", coordinates_min=lo, coordinates_max=hi) for lo, hi in c.patches)
    mesh = th.TreeMesh((c.cmin,) * c.ndims, (c.cmax,) * c.ndims, initial_refinement_level=c.level,
                       n_cells_max=10**8, periodicity=c.periodic, refinement_patches=patches)
    ics = {"constant": th.initial_condition_constant, "convergence_test": th.initial_condition_convergence_test,
           "weak_blast_wave": th.initial_condition_weak_blast_wave,
           "seeded_random": th.initial_condition_seeded_random(c.seed),
           "smooth_periodic": th.initial_condition_smooth_periodic(c.seed),
           "weak_blast_wave_seeded": th.initial_condition_weak_blast_wave_seeded(c.seed),
           "shock": th.initial_condition_shock, "rarefaction": th.initial_condition_rarefaction,
           "eoc_gravity": th.initial_condition_eoc_test_coupled_euler_gravity,
           "poisson_nonperiodic": th.initial_condition_poisson_nonperiodic,
           "harmonic_nonperiodic": th.initial_condition_harmonic_nonperiodic}
    kw = {}
    if c.bc_state == "slip_wall":
        kw["boundary_conditions"] = th.boundary_condition_slip_wall
    elif c.bc_state:
        bc = (th.boundary_condition_poisson_nonperiodic if c.bc_state == "poisson_nonperiodic"
              else th.BoundaryConditionDirichlet(ics[c.bc_state]))
        per = c.periodic if isinstance(c.periodic, tuple) else (c.periodic,) * c.ndims
        kw["boundary_conditions"] = {s: (th.boundary_condition_periodic if per[i // 2] else
                                         th.boundary_condition_outflow if s in c.outflow else bc)
                                     for i, s in enumerate(SIDES[:2 * c.ndims])}
    if c.source:
        kw["source_terms"] = {"convergence_test": th.source_terms_convergence_test,
                              "eoc_test_euler": th.source_terms_eoc_test_euler,
                              "poisson_nonperiodic": th.source_terms_poisson_nonperiodic,
                              "harmonic": th.source_terms_harmonic}[c.source]
    return th.SemidiscretizationHyperbolicHIP(mesh, eq, c.ic_fn if c.ic_fn is not None else ics[c.ic], solver, **kw)


def rel_max(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


# multicomponent Euler gases: equal components (the reference tests), one
# component, distinct components
MULTI_GASES = [((1.4, 1.4), (0.4, 0.4)), ((1.4,), (0.4,)), ((1.4, 1.648), (0.287, 1.578))]


def multi_entropy(u, nd, gammas, gas_constants):
    """U = -sum_k rho_k s_k, s_k = c_v,k log T - R_k log rho_k (complex-step safe)."""
    g, R = np.array(gammas), np.array(gas_constants)
    cv = R / (g - 1)
    rk = u[nd + 1:]
    rho = rk.sum()
    v = u[:nd] / rho
    T = (u[nd] - 0.5 * rho * np.sum(v * v)) / np.sum(rk * cv)
    return -np.sum(rk * (cv * np.log(T) - R * np.log(rk)))


def multi_entropy_vars(u, nd, gammas, gas_constants):
    h = 1e-30
    w = np.zeros(len(u))
    for j in range(len(u)):
        uc = u.astype(complex)
        uc[j] += 1j * h
        w[j] = multi_entropy(uc, nd, gammas, gas_constants).imag / h
    return w


# initial conditions written as functions in the reference's test files
def eulerquasi_ec_state(x, t, equations):
    """initial_condition_ec of test/tree_dgsem_1d/eulerquasi_ec.jl: v = 0.1,
    rho = 2 + 0.1 x, p = 3, a = 2 + x, as (a rho, a rho v, a e, a)."""
    v1, rho, p, a = 0.1, 2.0 + 0.1 * x[0], 3.0, 2.0 + x[0]
    return [a * rho, a * rho * v1, a * (p / (equations.gamma - 1) + 0.5 * rho * v1 ** 2), a]


def euler_blast_wave_1d_state(x, t, equations):
    """initial_condition_blast_wave of test/tree_dgsem_1d/euler_blast_wave.jl:
    inside |x| <= 0.5 rho = 1.1691, v = 0.1882 sign(x), p = 1.245; outside
    rho = 1, v = 0, p = 1e-3."""
    r = abs(x[0])
    cos_phi = 1.0 if x[0] > 0 else -1.0
    rho = 1.0 if r > 0.5 else 1.1691
    v1 = 0.0 if r > 0.5 else 0.1882 * cos_phi
    p = 1.0e-3 if r > 0.5 else 1.245
    return [rho, rho * v1, p / (equations.gamma - 1) + 0.5 * rho * v1 ** 2]


def oracle_state(O, c, mesh, t=0.0):
    """The case's initial state on the oracle mesh: the oracle's initial
    condition, or the case's function at the node coordinates (in 1D with
    the element end nodes one ulp inwards, as Trixi samples them)."""
    if c.ic_fn is None:
        return O.compute_coefficients(oracle_cfg(O, c), mesh, oracle_ic(O, c), t, c.seed)
    from types import SimpleNamespace
    eq = SimpleNamespace(gamma=c.gamma)
    x = mesh.node_coordinates.copy()
    if c.ndims == 1:
        x[:, 0, 0] = np.nextafter(x[:, 0, 0], np.inf)
        x[:, -1, 0] = np.nextafter(x[:, -1, 0], -np.inf)
    pts = x.reshape(-1, c.ndims)
    u = np.array([c.ic_fn(tuple(p), t, eq) for p in pts], dtype=np.float64)
    return u.reshape(x.shape[:-1] + (u.shape[-1],))


def stone_throw_state(x, t, equations):
    """initial_condition_stone_throw_discontinuous_bottom of
    test/tree_dgsem_1d/shallowwater_shock.jl: flat lake H = H0 = 1.75,
    v = -1 on [-0.75, 0] and 1 on [0, 0.75], bottom
    1.5 exp(-(x-1)^2/2) + 0.75 exp(-(x+1)^2/2), set to 0.5 on [-1.5, 0]."""
    x1 = x[0]
    H = 1.75
    v = 0.0
    if -0.75 <= x1 <= 0.0:
        v = -1.0
    elif 0.0 <= x1 <= 0.75:
        v = 1.0
    b = 1.5 / np.exp(0.5 * (x1 - 1.0) ** 2) + 0.75 / np.exp(0.5 * (x1 + 1.0) ** 2)
    if -1.5 <= x1 <= 0.0:
        b = 0.5
    return [H - b, (H - b) * v, b]
