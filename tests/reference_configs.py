"""The reference's own TreeMesh test configurations, as written.

One Case per test file that test/tree_dgsem_{1,2,3}d/tree_dgsem_{1,2,3}d.jl
includes (the shallow-water files are excluded there and here).  Every case
keeps the file's equation and constants, polydeg, domain, refinement level and
patches, volume integral, fluxes, indicator settings, initial condition,
boundary conditions and source terms; the reference tests compare u0 and one
rhs! at t = 0 (e.g. test/tree_dgsem_3d/euler_ec.jl:38-121), and so do the
tests that use this table.

The two benchmark configurations at their full size (3D GLM-MHD Alfven wave at
level 4, 2D shock capturing at level 6) follow the file cases.
"""
import math

from cases import Case, eulerquasi_ec_state, euler_blast_wave_1d_state


def euler_blob_state(x, t, equations):
    """initial_condition_blob of test/tree_dgsem_2d/euler_blob_mortar.jl:12-37:
    a dense blob (density contrast 10, radius 1, tanh edge of slope 2) at
    (-15, 0) in a Mach-2.7 background flow, at rest inside the blob."""
    R, dens0, chi = 1.0, 1.0, 10.0
    tau_cr = 1.0 / 1.6
    velx0 = 2.0 * R * math.sqrt(chi) / tau_cr
    c = velx0 / 2.7
    p0 = c * c * dens0 / equations.gamma
    r = math.sqrt((x[0] + 15.0) ** 2 + x[1] ** 2)
    bump = 0.5 * (1.0 + (math.tanh(2.0 * (r + R)) - (math.tanh(2.0 * (r - R)) + 1.0)))
    dens = dens0 + (chi - 1.0) * bump
    velx = velx0 - velx0 * bump
    return [dens, dens * velx, 0.0, p0 / (equations.gamma - 1.0) + 0.5 * dens * velx * velx]


# the x/y band refined twice: both patches of euler_blob_mortar.jl:56-61
BLOB_PATCHES = (((-40.0, -5.0), (40.0, 5.0)), ((-40.0, -5.0), (40.0, 5.0)))
SQRT2 = 2.0 ** 0.5

REFERENCE_CASES = [
    # ---- 1D (test/tree_dgsem_1d/) ----
    Case("1d/advection_basic", 1, "linadv", 4, cmin=-1.0, cmax=1.0, ic="convergence_test", advection=(1.0,)),
    Case("1d/advection_extended", 1, "linadv", 4, cmin=-1.0, cmax=1.0, ic="convergence_test", advection=(1.0,)),
    Case("1d/burgers_basic", 1, "burgers", 4, cmin=0.0, cmax=1.0, ic="convergence_test", source="convergence_test"),
    Case("1d/burgers_rarefraction", 1, "burgers", 6, "hg", "ec", "llf", "llf", cmin=0.0, cmax=1.0,
         ic="rarefaction", periodic=False, bc_state="rarefaction", outflow=("x_pos",), indicator="first",
         alpha_max=1.0),
    Case("1d/burgers_shock", 1, "burgers", 6, "hg", "ec", "llf", "llf", cmin=0.0, cmax=1.0, ic="shock",
         periodic=False, bc_state="shock", outflow=("x_pos",), indicator="first", alpha_max=1.0),
    Case("1d/euler_blast_wave", 1, "euler", 6, "hg", "ranocha", "llf", "llf", ic="constant",
         ic_fn=euler_blast_wave_1d_state),
    Case("1d/euler_ec", 1, "euler", 5, "fd", "ranocha", "ranocha", ic="weak_blast_wave"),
    Case("1d/euler_shock", 1, "euler", 5, "hg", "shima_etal", "llf", "llf", ic="weak_blast_wave"),
    Case("1d/euler_source_terms", 1, "euler", 4, polydeg=4, cmin=0.0, cmax=2.0, ic="convergence_test",
         source="convergence_test"),
    Case("1d/euler_source_terms_nonperiodic", 1, "euler", 4, cmin=0.0, cmax=2.0, ic="convergence_test",
         periodic=False, bc_state="convergence_test", source="convergence_test"),
    Case("1d/eulermulti_ec", 1, "eulermulti", 5, "fd", "ranocha", "ranocha", ic="weak_blast_wave"),
    Case("1d/eulermulti_es", 1, "eulermulti", 5, "fd", "ranocha", "llf", ic="weak_blast_wave"),
    Case("1d/eulerquasi_ec", 1, "quasi1d", 6, "fd", "chan_etal", "chan_etal", polydeg=4, cmin=-1.0, cmax=1.0,
         ic="constant", ic_fn=eulerquasi_ec_state),
    Case("1d/eulerquasi_source_terms", 1, "quasi1d", 4, "fd", "chan_etal", "chan_etal", polydeg=4, cmin=-1.0,
         cmax=1.0, ic="convergence_test", source="convergence_test"),
    Case("1d/hypdiff_harmonic_nonperiodic", 1, "hypdiff", 2, cmin=-1.0, cmax=2.0, ic="harmonic_nonperiodic",
         periodic=False, bc_state="harmonic_nonperiodic", source="harmonic", nu=1.25),
    Case("1d/hypdiff_nonperiodic", 1, "hypdiff", 3, polydeg=4, cmin=0.0, cmax=1.0, ic="poisson_nonperiodic",
         periodic=False, bc_state="poisson_nonperiodic", source="poisson_nonperiodic"),
    Case("1d/mhd_alfven_wave", 1, "mhd", 2, "fd", "hindenlang_gassner", polydeg=4, cmin=0.0, cmax=1.0,
         ic="convergence_test", gamma=5 / 3),
    Case("1d/mhd_ec", 1, "mhd", 4, "fd", "hindenlang_gassner", "hindenlang_gassner", cmin=0.0, cmax=1.0,
         ic="weak_blast_wave", gamma=2.0),
    # ---- 2D (test/tree_dgsem_2d/) ----
    Case("2d/advection_basic", 2, "linadv", 4, cmin=-1.0, cmax=1.0, ic="convergence_test", advection=(0.2, -0.7)),
    Case("2d/advection_mortar", 2, "linadv", 2, cmin=-1.0, cmax=1.0, ic="convergence_test", advection=(0.2, -0.7),
         patches=(((0.0, -1.0), (1.0, 1.0)),)),
    Case("2d/euler_blob_mortar", 2, "euler", 4, "hg", "ranocha", "llf", "llf", cmin=-32.0, cmax=32.0, gamma=5 / 3,
         ic="constant", ic_fn=euler_blob_state, alpha_max=0.05, alpha_min=0.0001, indicator="pressure",
         patches=BLOB_PATCHES),
    Case("2d/euler_ec", 2, "euler", 5, "fd", "ranocha", "ranocha", ic="weak_blast_wave"),
    Case("2d/euler_shock", 2, "euler", 5, "hg", "shima_etal", "llf", "llf", ic="weak_blast_wave"),
    Case("2d/euler_source_terms", 2, "euler", 4, cmin=0.0, cmax=2.0, ic="convergence_test",
         source="convergence_test"),
    Case("2d/euler_source_terms_nonperiodic", 2, "euler", 4, cmin=0.0, cmax=2.0, ic="convergence_test",
         periodic=False, bc_state="convergence_test", source="convergence_test"),
    Case("2d/eulermulti_ec", 2, "eulermulti", 5, "fd", "ranocha", "ranocha", ic="weak_blast_wave",
         gammas=(1.4,), gas_constants=(0.4,)),
    Case("2d/eulermulti_es", 2, "eulermulti", 5, "fd", "ranocha", "llf", ic="weak_blast_wave"),
    Case("2d/hypdiff_nonperiodic", 2, "hypdiff", 3, polydeg=4, cmin=0.0, cmax=1.0, ic="poisson_nonperiodic",
         periodic=(False, True), bc_state="poisson_nonperiodic", source="poisson_nonperiodic"),
    Case("2d/mhd_alfven_wave", 2, "mhd", 4, "fd", "central", "llf", cmin=0.0, cmax=SQRT2, ic="convergence_test",
         gamma=5 / 3),
    Case("2d/mhd_alfven_wave_mortar", 2, "mhd", 4, "fd", "hindenlang_gassner", "hlle", cmin=0.0, cmax=SQRT2,
         ic="convergence_test", gamma=5 / 3, patches=(((0.25 * SQRT2,) * 2, (0.75 * SQRT2,) * 2),)),
    Case("2d/mhd_ec", 2, "mhd", 4, "fd", "hindenlang_gassner", "hindenlang_gassner", ic="weak_blast_wave",
         gamma=1.4),
    Case("2d/mhd_shock", 2, "mhd", 3, "hg", "hindenlang_gassner", "hindenlang_gassner", "hindenlang_gassner",
         polydeg=4, ic="weak_blast_wave", gamma=1.4),
    # ---- 3D (test/tree_dgsem_3d/) ----
    Case("3d/advection_basic", 3, "linadv", 3, cmin=-1.0, cmax=1.0, ic="convergence_test",
         advection=(0.2, -0.7, 0.5)),
    Case("3d/advection_mortar", 3, "linadv", 2, cmin=-1.0, cmax=1.0, ic="convergence_test",
         advection=(0.2, -0.7, 0.5), patches=(((0.0, -1.0, -1.0), (1.0, 1.0, 1.0)),
                                               ((0.0, -0.5, -0.5), (0.5, 0.5, 0.5)))),
    Case("3d/euler_convergence", 3, "euler", 2, sflux="hll", cmin=0.0, cmax=2.0, ic="eoc_gravity",
         source="eoc_test_euler", gamma=2.0),
    Case("3d/euler_ec", 3, "euler", 3, "fd", "ranocha", "ranocha", ic="weak_blast_wave"),
    Case("3d/euler_mortar", 3, "euler", 2, cmin=0.0, cmax=2.0, ic="convergence_test", source="convergence_test",
         patches=(((0.5, 0.5, 0.5), (1.5, 1.5, 1.5)),)),
    Case("3d/euler_shock", 3, "euler", 3, "hg", "ranocha", "ranocha", "ranocha", ic="weak_blast_wave"),
    Case("3d/euler_source_terms", 3, "euler", 2, cmin=0.0, cmax=2.0, ic="convergence_test",
         source="convergence_test"),
    Case("3d/hypdiff_nonperiodic", 3, "hypdiff", 2, polydeg=4, cmin=0.0, cmax=1.0, ic="poisson_nonperiodic",
         periodic=(False, True, True), bc_state="poisson_nonperiodic", source="poisson_nonperiodic"),
    Case("3d/mhd_alfven_wave", 3, "mhd", 2, "fd", "hindenlang_gassner", cmin=-1.0, cmax=1.0, ic="convergence_test",
         gamma=5 / 3),
    Case("3d/mhd_alfven_wave_mortar", 3, "mhd", 2, "fd", "hindenlang_gassner", "hlle", cmin=-1.0, cmax=1.0,
         ic="convergence_test", gamma=5 / 3, patches=(((-0.5,) * 3, (0.5,) * 3),)),
    Case("3d/mhd_ec", 3, "mhd", 2, "fd", "hindenlang_gassner", "hindenlang_gassner", ic="weak_blast_wave",
         gamma=1.4),
    Case("3d/mhd_shock", 3, "mhd", 3, "hg", "hindenlang_gassner", "hindenlang_gassner", "hindenlang_gassner",
         polydeg=4, ic="weak_blast_wave", gamma=1.4),
]

# the benchmark's 3D GLM-MHD Alfven wave (level 4, 4096 elements) and 2D
# Hennemann-Gassner shock capturing (level 6, 4096 elements) at full size
BENCHMARK_CASES = [
    Case("bench/mhd3d_alfven_level4", 3, "mhd", 4, "fd", "hindenlang_gassner", cmin=-1.0, cmax=1.0,
         ic="convergence_test", gamma=5 / 3),
    Case("bench/euler2d_hg_level6", 2, "euler", 6, "hg", "ranocha", ic="weak_blast_wave_seeded"),
]
