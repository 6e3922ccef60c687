"""GPU parity: the HIP library against the CPU oracle on identical inputs.

Bar (FP64): relative max-norm ||du_gpu - du_oracle||_inf / ||du_oracle||_inf
<= 1e-12; error norms to 1e-10; integer/layout work bit-exact.  Mixed mode:
relative max-norm of du <= 2e-5 against the FP64 oracle (stated per case).
"""
import numpy as np
import pytest

from cases import (oracle_mesh, Case, oracle_cfg, oracle_ic, oracle_state, product_semi, rel_max, eulerquasi_ec_state,
                   euler_blast_wave_1d_state, stone_throw_state)

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 2e-5
MORTAR_2D = (((0.0, -1.0), (1.0, 1.0)),)
MORTAR_3D = (((0.0, -1.0, -1.0), (1.0, 1.0, 1.0)), ((0.0, -0.5, -0.5), (0.5, 0.5, 0.5)))
EULER_MORTAR_3D = (((0.5, 0.5, 0.5), (1.5, 1.5, 1.5)),)


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def run_pair(th, O, torch, case, precision=None):
    """(du_gpu in Trixi layout, du_oracle, u) for one case."""
    cfg = oracle_cfg(O, case)
    mesh = oracle_mesh(O, case)
    u = oracle_state(O, case, mesh, case.t)
    du_ref = O.rhs(cfg, mesh, u, t=case.t)
    semi = product_semi(th, case, precision)
    u_dev = th.pack(torch.from_numpy(u).cuda(), semi)
    du_dev = torch.empty_like(u_dev)
    th.rhs_hip_(du_dev, u_dev, semi, case.t)
    du = th.wrap_array_native(du_dev, semi).cpu().numpy()
    torch.cuda.synchronize()
    semi.close()
    return du, du_ref, u


# test/tree_dgsem_{2,3}d/mhd_shock.jl as written: polydeg 4, level 3 on
# [-2, 2]^d, weak blast wave, alpha_max 0.5, alpha_min 0.001, smoothing
MHD_HG_P4 = [
    Case("mhd2d_hg_ec_blast_p4", 2, "mhd", 3, "hg", "hindenlang_gassner", "hindenlang_gassner", "hindenlang_gassner",
         polydeg=4, ic="weak_blast_wave", gamma=1.4),
    Case("mhd3d_hg_ec_blast_p4", 3, "mhd", 3, "hg", "hindenlang_gassner", "hindenlang_gassner", "hindenlang_gassner",
         polydeg=4, ic="weak_blast_wave", gamma=1.4),
    Case("mhd3d_hg_ec_blast_seeded_p4", 3, "mhd", 2, "hg", "hindenlang_gassner", "hindenlang_gassner",
         "hindenlang_gassner", polydeg=4, ic="weak_blast_wave_seeded", gamma=1.4),
]


# multicomponent Euler: test/tree_dgsem_{1,2}d/eulermulti_{ec,es}.jl as
# written (polydeg 3, level 5 on [-2, 2]^d, weak blast wave, flux_ranocha
# volume flux with a flux_ranocha or LLF surface flux), plus seeded states
# with two distinct gases
MULTI_CASES = [
    Case("eulermulti1d_ec", 1, "eulermulti", 5, "fd", "ranocha", "ranocha", ic="weak_blast_wave"),
    Case("eulermulti1d_es", 1, "eulermulti", 5, "fd", "ranocha", "llf", ic="weak_blast_wave"),
    Case("eulermulti2d_ec_single", 2, "eulermulti", 5, "fd", "ranocha", "ranocha", ic="weak_blast_wave",
         gammas=(1.4,), gas_constants=(0.4,)),
    Case("eulermulti2d_es", 2, "eulermulti", 5, "fd", "ranocha", "llf", ic="weak_blast_wave"),
    Case("eulermulti2d_distinct_random", 2, "eulermulti", 3, "fd", "ranocha", "llf", ic="seeded_random",
         gammas=(1.4, 1.648), gas_constants=(0.287, 1.578)),
    Case("eulermulti1d_weak_distinct_random", 1, "eulermulti", 4, "weak", "central", "llf", ic="seeded_random",
         gammas=(1.4, 1.648), gas_constants=(0.287, 1.578)),
    Case("eulermulti2d_weak_single_random", 2, "eulermulti", 3, "weak", "central", "llf", ic="seeded_random",
         gammas=(1.4,), gas_constants=(0.4,)),
]
# quasi-1D Euler: test/tree_dgsem_1d/eulerquasi_ec.jl (its initial_condition_ec,
# polydeg 4, level 6 on [-1, 1]) and eulerquasi_source_terms.jl (convergence
# test with sources, level 4), (flux_chan_etal, flux_nonconservative_chan_etal)
# as volume and surface flux; seeded states with a random nozzle width
QUASI_CASES = [
    Case("quasi1d_ec", 1, "quasi1d", 6, "fd", "chan_etal", "chan_etal", polydeg=4, cmin=-1.0, cmax=1.0,
         ic="constant", ic_fn=eulerquasi_ec_state),
    Case("quasi1d_src", 1, "quasi1d", 4, "fd", "chan_etal", "chan_etal", polydeg=4, cmin=-1.0, cmax=1.0,
         ic="convergence_test", source="convergence_test", t=0.3),
    Case("quasi1d_random_p4", 1, "quasi1d", 4, "fd", "chan_etal", "chan_etal", polydeg=4, cmin=-1.0, cmax=1.0),
    Case("quasi1d_random_p3", 1, "quasi1d", 5, "fd", "chan_etal", "chan_etal", cmin=-1.0, cmax=1.0),
]
# shallow water 2D: test/tree_dgsem_2d/shallowwater_ec.jl (polydeg 4, level 2 on
# [-1, 1]^2, weak blast wave, volume (wintermeyer, nc wintermeyer), surface
# (fjordholm, nc fjordholm)), shallowwater_source_terms.jl and
# shawllowwater_source_terms_nonperiodic.jl (polydeg 3, level 3 on [0, sqrt 2]^2,
# LLF surface flux, convergence test with sources, periodic or Dirichlet)
SWE_CASES = [
    Case("swe2d_ec", 2, "swe", 2, "fd", "wintermeyer_etal", "fjordholm_etal", polydeg=4, cmin=-1.0, cmax=1.0,
         ic="weak_blast_wave"),
    Case("swe2d_src", 2, "swe", 3, "fd", "wintermeyer_etal", "llf", cmin=0.0, cmax=2 ** 0.5, ic="convergence_test",
         source="convergence_test", t=0.3),
    Case("swe2d_src_nonperiodic", 2, "swe", 3, "fd", "wintermeyer_etal", "llf", cmin=0.0, cmax=2 ** 0.5,
         ic="convergence_test", source="convergence_test", t=0.3, periodic=False, bc_state="convergence_test"),
    Case("swe2d_random_p4", 2, "swe", 3, "fd", "wintermeyer_etal", "fjordholm_etal", polydeg=4),
    Case("swe2d_random_llf", 2, "swe", 3, "fd", "wintermeyer_etal", "llf"),
    # test/tree_dgsem_1d/shallowwater_shock.jl as written: its stone-throw state
    # (g = 9.812, H0 = 1.75, discontinuous bottom), polydeg 4, level 3 on
    # [-3, 3], slip walls, HG with (wintermeyer, nc wintermeyer) as DG flux and
    # (hydrostatic reconstruction LLF, nc audusse) as FV and surface flux,
    # indicator waterheight_pressure
    Case("swe1d_shock", 1, "swe", 3, "hg", "wintermeyer_etal", "hydrostatic_llf", "hydrostatic_llf", polydeg=4,
         cmin=-3.0, cmax=3.0, ic="constant", ic_fn=stone_throw_state, gravity=9.812, periodic=False,
         bc_state="slip_wall", indicator="waterheight_pressure", rk_cfl=0.25),  # (0.5 max_dt is unstable here)
    Case("swe1d_hg_random", 1, "swe", 4, "hg", "wintermeyer_etal", "hydrostatic_llf", "hydrostatic_llf",
         polydeg=4, gravity=9.812, alpha_min=0.0, alpha_max=1.0),
    Case("swe1d_fd_random_walls", 1, "swe", 4, "fd", "wintermeyer_etal", "hydrostatic_llf", polydeg=4,
         gravity=9.812, periodic=False, bc_state="slip_wall"),
]
# test/tree_dgsem_1d/euler_blast_wave.jl: its initial_condition_blast_wave,
# HG with flux_ranocha (DG) and LLF (FV, surface), polydeg 3, level 6
EULER_BLAST_1D = Case("euler1d_blast_wave_hg", 1, "euler", 6, "hg", "ranocha", "llf", "llf", ic="constant",
                      ic_fn=euler_blast_wave_1d_state)


CASES = [
    Case("linadv1d_weak_llf", 1, "linadv", 4, cmin=-1.0, cmax=1.0, ic="convergence_test"),
    Case("linadv2d_weak_llf", 2, "linadv", 3, ic="seeded_random"),
    Case("linadv3d_weak_central", 3, "linadv", 2, sflux="central", ic="seeded_random"),
    Case("linadv1d_weak_llf_p1", 1, "linadv", 4, polydeg=1),
    Case("linadv2d_weak_llf_p4", 2, "linadv", 2, polydeg=4),
    Case("euler1d_weak_llf", 1, "euler", 4),
    Case("euler2d_weak_llf", 2, "euler", 3),
    Case("euler3d_weak_llfnaive", 3, "euler", 2, sflux="llf_naive"),
    Case("euler1d_fd_ranocha_llf", 1, "euler", 5, "fd", "ranocha"),
    Case("euler2d_fd_ranocha_llf", 2, "euler", 5, "fd", "ranocha"),      # C2 shape
    Case("euler2d_fd_ranocha_ranocha", 2, "euler", 3, "fd", "ranocha", "ranocha"),
    Case("euler2d_fd_shima_shima", 2, "euler", 3, "fd", "shima_etal", "shima_etal"),
    Case("euler2d_fd_shima_llfnaive", 2, "euler", 3, "fd", "shima_etal", "llf_naive"),
    Case("euler3d_fd_ranocha_llf", 3, "euler", 3, "fd", "ranocha", ic="smooth_periodic"),
    Case("euler3d_fd_ranocha_llf_random", 3, "euler", 2, "fd", "ranocha"),
    Case("euler3d_fd_ranocha_llfnaive", 3, "euler", 2, "fd", "ranocha", "llf_naive"),
    Case("euler3d_fd_shima_llf", 3, "euler", 2, "fd", "shima_etal"),
    Case("euler2d_fd_ranocha_llf_p1", 2, "euler", 3, "fd", "ranocha", polydeg=1),
    Case("euler3d_fd_ranocha_llf_p2", 3, "euler", 2, "fd", "ranocha", polydeg=2),
    Case("mhd3d_fd_hg_llf_alfven", 3, "mhd", 2, "fd", "hindenlang_gassner", cmin=-1.0, cmax=1.0,
         ic="convergence_test", gamma=5 / 3),
    Case("mhd3d_fd_hg_hg_random", 3, "mhd", 1, "fd", "hindenlang_gassner", "hindenlang_gassner", cmin=-1.0,
         cmax=1.0, gamma=5 / 3),
    Case("mhd3d_fd_hg_llfnaive", 3, "mhd", 1, "fd", "hindenlang_gassner", "llf_naive", cmin=-1.0, cmax=1.0,
         gamma=5 / 3),
    # non-periodic meshes with Dirichlet states (the exact convergence-test
    # solution) and source_terms_convergence_test, evaluated at t = 0.3
    # (reference tests tree_dgsem_{1,2}d/euler_source_terms_nonperiodic.jl)
    Case("euler1d_weak_dirichlet_src", 1, "euler", 4, cmin=0.0, cmax=2.0, ic="convergence_test", periodic=False,
         bc_state="convergence_test", source="convergence_test", t=0.3),
    Case("euler2d_weak_dirichlet_src", 2, "euler", 3, cmin=0.0, cmax=2.0, ic="convergence_test", periodic=False,
         bc_state="convergence_test", source="convergence_test", t=0.3),
    Case("euler3d_weak_dirichlet_src", 3, "euler", 2, cmin=0.0, cmax=2.0, ic="convergence_test", periodic=False,
         bc_state="convergence_test", source="convergence_test", t=0.3),
    Case("euler2d_fd_dirichlet_random", 2, "euler", 3, "fd", "ranocha", cmin=0.0, cmax=2.0, periodic=False,
         bc_state="convergence_test", t=0.7),
    Case("euler3d_fd_dirichlet_const", 3, "euler", 2, "fd", "ranocha", cmin=0.0, cmax=2.0, periodic=False,
         bc_state="constant", source="convergence_test", t=0.1),
    # periodic with source terms (reference tests tree_dgsem_{1,2,3}d/euler_source_terms.jl)
    Case("euler1d_weak_periodic_src", 1, "euler", 4, cmin=0.0, cmax=2.0, ic="convergence_test",
         source="convergence_test", t=0.4),
    Case("euler2d_weak_periodic_src", 2, "euler", 3, cmin=0.0, cmax=2.0, ic="convergence_test",
         source="convergence_test", t=0.4),
    # (a random state: at the convergence-test state du is the small difference
    # of a large flux divergence and the source, and the relative error of du
    # grows with that cancellation)
    Case("euler3d_fd_periodic_src", 3, "euler", 2, "fd", "ranocha", cmin=0.0, cmax=2.0, source="convergence_test",
         t=0.4),
    Case("euler2d_hg_ranocha_blast", 2, "euler", 4, "hg", "ranocha", ic="weak_blast_wave"),
    Case("euler2d_hg_shima_blast_nosmooth", 2, "euler", 4, "hg", "shima_etal", ic="weak_blast_wave",
         alpha_smooth=False),
    Case("euler2d_hg_ranocha_random", 2, "euler", 3, "hg", "ranocha", alpha_min=0.0),
    Case("euler1d_hg_ranocha_blast", 1, "euler", 5, "hg", "ranocha", ic="weak_blast_wave"),
    Case("euler3d_hg_ranocha_blast", 3, "euler", 3, "hg", "ranocha", ic="weak_blast_wave"),
    Case("euler2d_hg_ranocha_llfnaive", 2, "euler", 4, "hg", "ranocha", "llf_naive", "llf_naive",
         ic="weak_blast_wave"),
] + [
    # refined meshes with L2 mortars: the reference's mortar meshes
    # (test/tree_dgsem_2d/advection_mortar.jl, test/tree_dgsem_3d/advection_mortar.jl,
    # test/tree_dgsem_3d/euler_mortar.jl) and non-periodic / flux-differencing variants
    Case("linadv2d_mortar_llf", 2, "linadv", 2, cmin=-1.0, cmax=1.0, ic="convergence_test", advection=(0.2, -0.7),
         patches=MORTAR_2D),
    Case("linadv2d_mortar_llf_random", 2, "linadv", 2, cmin=-1.0, cmax=1.0, advection=(0.2, -0.7),
         patches=MORTAR_2D),
    Case("linadv3d_mortar_llf_random", 3, "linadv", 2, cmin=-1.0, cmax=1.0, advection=(0.2, -0.7, 0.5),
         patches=MORTAR_3D),
    # (the reference's velocity sums to zero, so the convergence-test state is
    # stationary and du is pure discretization error; mixed precision is
    # checked on the random state)
    Case("linadv3d_mortar_central", 3, "linadv", 2, sflux="central", cmin=-1.0, cmax=1.0, ic="convergence_test",
         advection=(0.2, -0.7, 0.5), patches=MORTAR_3D),
    Case("linadv3d_mortar_central_random", 3, "linadv", 2, sflux="central", cmin=-1.0, cmax=1.0,
         advection=(0.2, -0.7, 0.5), patches=MORTAR_3D),
    Case("euler3d_mortar_weak_llf_src", 3, "euler", 2, cmin=0.0, cmax=2.0, ic="convergence_test",
         source="convergence_test", t=0.3, patches=EULER_MORTAR_3D),
    Case("euler3d_mortar_weak_llf_random", 3, "euler", 2, cmin=0.0, cmax=2.0, patches=EULER_MORTAR_3D),
    Case("euler3d_mortar_fd_ranocha_random", 3, "euler", 2, "fd", "ranocha", cmin=0.0, cmax=2.0,
         patches=EULER_MORTAR_3D),
    Case("euler2d_mortar_fd_dirichlet", 2, "euler", 3, "fd", "ranocha", cmin=0.0, cmax=2.0, periodic=False,
         bc_state="convergence_test", source="convergence_test", t=0.2, patches=(((0.5, 0.5), (1.5, 1.5)),)),
    Case("euler2d_mortar_weak_random_nested", 2, "euler", 2, cmin=0.0, cmax=2.0,
         patches=(((0.5, 0.5), (1.5, 1.5)), ((0.8, 0.8), (1.2, 1.2)))),
    # shock capturing with the pressure / density indicator variables, on
    # mortar meshes (test/tree_dgsem_2d/euler_blob_mortar.jl: HG, variable
    # pressure, gamma 5/3) and with Dirichlet boundaries
    Case("euler2d_hg_pressure_blast", 2, "euler", 4, "hg", "ranocha", ic="weak_blast_wave", indicator="pressure"),
    Case("euler2d_hg_mortar_pressure", 2, "euler", 3, "hg", "ranocha", cmin=-1.0, cmax=1.0, gamma=5 / 3,
         ic="weak_blast_wave_seeded", indicator="pressure", patches=MORTAR_2D),
    Case("euler3d_hg_mortar", 3, "euler", 2, "hg", "ranocha", cmin=-1.0, cmax=1.0, ic="weak_blast_wave_seeded",
         patches=(((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5)),)),
    Case("euler2d_hg_dirichlet_density", 2, "euler", 3, "hg", "ranocha", cmin=0.0, cmax=2.0, periodic=False,
         bc_state="convergence_test", indicator="density", alpha_min=0.0),
    # 2D GLM-MHD (test/tree_dgsem_2d/mhd_alfven_wave.jl: volume flux_central,
    # surface LLF, Powell terms, gamma 5/3 on [0, sqrt 2]^2; mhd_ec.jl:
    # flux_hindenlang_gassner throughout)
    Case("mhd2d_fd_central_llf_alfven", 2, "mhd", 3, "fd", "central", cmin=0.0, cmax=2 ** 0.5,
         ic="convergence_test", gamma=5 / 3),
    Case("mhd2d_fd_central_llf_random", 2, "mhd", 3, "fd", "central", cmin=-1.0, cmax=1.0, gamma=5 / 3),
    Case("mhd2d_fd_hg_hg_random", 2, "mhd", 3, "fd", "hindenlang_gassner", "hindenlang_gassner", cmin=-1.0,
         cmax=1.0, gamma=1.4),
    Case("mhd2d_fd_hg_llf_random", 2, "mhd", 3, "fd", "hindenlang_gassner", cmin=-1.0, cmax=1.0, gamma=5 / 3),
    Case("mhd2d_fd_hg_llfnaive_random", 2, "mhd", 2, "fd", "hindenlang_gassner", "llf_naive", cmin=-1.0,
         cmax=1.0, gamma=5 / 3),
    # inviscid Burgers 1D (test/tree_dgsem_1d/burgers_basic.jl: weak form, LLF,
    # source terms, level 4 on [0, 1]; burgers_shock.jl / burgers_rarefraction.jl:
    # HG with flux_ec / LLF, variable first, alpha_max 1, level 6, inflow
    # Dirichlet step on x_neg and outflow on x_pos)
    Case("burgers1d_weak_llf_src", 1, "burgers", 4, cmin=0.0, cmax=1.0, ic="convergence_test",
         source="convergence_test", t=0.3),
    Case("burgers1d_fd_ec_llf_random", 1, "burgers", 4, "fd", "ec", cmin=0.0, cmax=1.0),
    Case("burgers1d_fd_ec_ec_random", 1, "burgers", 4, "fd", "ec", "ec", cmin=0.0, cmax=1.0),
    Case("burgers1d_hg_shock", 1, "burgers", 6, "hg", "ec", "llf", "llf", cmin=0.0, cmax=1.0, ic="shock",
         periodic=False, bc_state="shock", outflow=("x_pos",), indicator="first", alpha_max=1.0),
    Case("burgers1d_hg_rarefaction", 1, "burgers", 6, "hg", "ec", "llf", "llf", cmin=0.0, cmax=1.0,
         ic="rarefaction", periodic=False, bc_state="rarefaction", outflow=("x_pos",), indicator="first",
         alpha_max=1.0),
    Case("burgers1d_hg_random_periodic", 1, "burgers", 5, "hg", "ec", "llf", "llf", cmin=0.0, cmax=1.0,
         indicator="first", alpha_max=1.0, alpha_min=0.0),
    Case("burgers1d_weak_inflow_outflow_src", 1, "burgers", 4, cmin=0.0, cmax=1.0, ic="convergence_test",
         periodic=False, bc_state="convergence_test", outflow=("x_pos",), source="convergence_test", t=0.2),
    # 1D ideal MHD, eight variables (test/tree_dgsem_1d/mhd_alfven_wave.jl:
    # flux_hindenlang_gassner volume, LLF surface, polydeg 4, level 2;
    # mhd_ec.jl: flux_hindenlang_gassner throughout, polydeg 3)
    Case("mhd1d_fd_hg_llf_alfven_p4", 1, "mhd", 2, "fd", "hindenlang_gassner", cmin=0.0, cmax=1.0,
         ic="convergence_test", gamma=5 / 3, polydeg=4),
    Case("mhd1d_fd_hg_hg_random", 1, "mhd", 4, "fd", "hindenlang_gassner", "hindenlang_gassner", cmin=0.0,
         cmax=1.0, gamma=5 / 3),
    Case("mhd1d_fd_hg_llf_random", 1, "mhd", 4, "fd", "hindenlang_gassner", cmin=0.0, cmax=1.0, gamma=5 / 3),
    Case("mhd1d_weak_llf_random", 1, "mhd", 4, cmin=0.0, cmax=1.0, gamma=5 / 3),
    # FluxHLL (Davis) with the coupled Euler-gravity EOC state and
    # source_terms_eoc_test_euler (test/tree_dgsem_3d/euler_convergence.jl:
    # gamma 2, weak form, level 2 on [0, 2]^3)
    Case("euler3d_weak_hll_eoc", 3, "euler", 2, sflux="hll", cmin=0.0, cmax=2.0, ic="eoc_gravity",
         source="eoc_test_euler", gamma=2.0, t=0.3),
    Case("euler2d_weak_hll_eoc", 2, "euler", 3, sflux="hll", cmin=0.0, cmax=2.0, ic="eoc_gravity",
         source="eoc_test_euler", gamma=2.0, t=0.7),
    Case("euler1d_weak_hll_random", 1, "euler", 4, sflux="hll"),
    Case("euler2d_weak_hll_random", 2, "euler", 3, sflux="hll"),
    # test/tree_dgsem_3d/euler_shock.jl: HG with flux_ranocha as DG, FV and surface flux
    Case("euler3d_hg_ranocha_ranocha_blast", 3, "euler", 3, "hg", "ranocha", "ranocha", "ranocha",
         ic="weak_blast_wave"),
    # hyperbolic diffusion (test/tree_dgsem_{1,2,3}d/hypdiff_nonperiodic.jl: polydeg 4,
    # LLF, the Poisson steady state on the x boundaries, periodic elsewhere,
    # source_terms_poisson_nonperiodic; 1D hypdiff_harmonic_nonperiodic.jl: nu 1.25,
    # polydeg 3, the harmonic state as Dirichlet state, source_terms_harmonic)
    Case("hypdiff1d_poisson", 1, "hypdiff", 3, polydeg=4, cmin=0.0, cmax=1.0, ic="poisson_nonperiodic",
         periodic=False, bc_state="poisson_nonperiodic", source="poisson_nonperiodic"),
    Case("hypdiff2d_poisson", 2, "hypdiff", 3, polydeg=4, cmin=0.0, cmax=1.0, ic="poisson_nonperiodic",
         periodic=(False, True), bc_state="poisson_nonperiodic", source="poisson_nonperiodic"),
    Case("hypdiff3d_poisson", 3, "hypdiff", 2, polydeg=4, cmin=0.0, cmax=1.0, ic="poisson_nonperiodic",
         periodic=(False, True, True), bc_state="poisson_nonperiodic", source="poisson_nonperiodic"),
    # (a random state: at the steady state du is only the discretisation residual)
    Case("hypdiff2d_poisson_random_t", 2, "hypdiff", 3, polydeg=4, cmin=0.0, cmax=1.0,
         periodic=(False, True), bc_state="poisson_nonperiodic", source="poisson_nonperiodic", t=0.5),
    Case("hypdiff1d_harmonic_bc_random", 1, "hypdiff", 2, cmin=-1.0, cmax=2.0, periodic=False,
         bc_state="harmonic_nonperiodic", source="harmonic", nu=1.25, t=0.3),
    Case("hypdiff3d_random_periodic", 3, "hypdiff", 2, polydeg=3, cmin=0.0, cmax=1.0),
    # GLM-MHD shock capturing (test/tree_dgsem_{2,3}d/mhd_shock.jl: the EC flux +
    # Powell terms as DG, FV and surface flux, indicator density_pressure,
    # MHD weak blast wave)
    Case("mhd2d_hg_ec_blast", 2, "mhd", 3, "hg", "hindenlang_gassner", "hindenlang_gassner", "hindenlang_gassner",
         cmin=-2.0, cmax=2.0, ic="weak_blast_wave", gamma=1.4),
    Case("mhd2d_hg_llf_blast_seeded", 2, "mhd", 3, "hg", "hindenlang_gassner", "llf", "llf", cmin=-2.0, cmax=2.0,
         ic="weak_blast_wave_seeded", gamma=1.4, alpha_min=0.0),
    Case("mhd3d_hg_ec_blast", 3, "mhd", 2, "hg", "hindenlang_gassner", "hindenlang_gassner", "hindenlang_gassner",
         cmin=-2.0, cmax=2.0, ic="weak_blast_wave_seeded", gamma=1.4),
    # the tests' own configuration: polydeg 4, level 3 on [-2, 2]^d
    *MHD_HG_P4,
    *MULTI_CASES,
    *QUASI_CASES,
    EULER_BLAST_1D,
    # test/tree_dgsem_{2,3}d/mhd_alfven_wave_mortar.jl as written: volume
    # (flux_hindenlang_gassner, powell), surface (flux_hlle, powell), polydeg 3
    Case("mhd2d_mortar_hlle", 2, "mhd", 4, "fd", "hindenlang_gassner", "hlle", cmin=0.0, cmax=2 ** 0.5,
         ic="convergence_test", gamma=5 / 3, patches=(((0.25 * 2 ** 0.5,) * 2, (0.75 * 2 ** 0.5,) * 2),)),
    Case("mhd3d_mortar_hlle", 3, "mhd", 2, "fd", "hindenlang_gassner", "hlle", cmin=-1.0, cmax=1.0,
         ic="convergence_test", gamma=5 / 3, patches=(((-0.5,) * 3, (0.5,) * 3),)),
    Case("mhd2d_mortar_llf_random", 2, "mhd", 3, "fd", "hindenlang_gassner", "llf", gamma=5 / 3,
         patches=(((-1.0, -1.0), (0.5, 1.0)),)),
    Case("mhd3d_mortar_hlle_random", 3, "mhd", 2, "fd", "hindenlang_gassner", "hlle", gamma=5 / 3,
         patches=MORTAR_3D),
    Case("mhd2d_hlle_periodic_random", 2, "mhd", 3, "fd", "hindenlang_gassner", "hlle", gamma=5 / 3),
    *SWE_CASES,
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: c.name)
def test_rhs_fp64_matches_oracle(th, oracle, torch, case):
    du, du_ref, _ = run_pair(th, oracle, torch, case)
    err = rel_max(du, du_ref)
    print(f"{case.name}: rel max-norm {err:.3e}", flush=True)
    assert np.all(np.isfinite(du))
    assert err <= FP64_TOL, err


# (the MHD Alfven wave on a refined mesh, and the shallow-water manufactured
# state with its source: du is a cancellation of flux terms much larger than
# du, relative errors are amplified; the seeded states cover mixed)
_MIXED_SKIP = ("mhd2d_mortar_hlle", "mhd3d_mortar_hlle", "swe2d_src", "swe2d_src_nonperiodic")
MIXED_CASES = [c for c in CASES if c.vi in ("fd", "hg") and c.polydeg == 3 and not c.patches
               and c.name not in _MIXED_SKIP] + MHD_HG_P4 + \
    [QUASI_CASES[0], QUASI_CASES[2]] + \
    [Case("euler2d_weak_llf_mixed", 2, "euler", 3), Case("linadv1d_mixed", 1, "linadv", 4, cmin=-1.0, cmax=1.0,
                                                        ic="convergence_test"),
     # weak form with N = 4 applies D-hat on the matrix cores in mixed precision
     Case("euler3d_weak_llf_mixed", 3, "euler", 2, ic="smooth_periodic"),
     Case("linadv3d_weak_mixed", 3, "linadv", 2, ic="seeded_random"),
     Case("linadv2d_weak_mixed", 2, "linadv", 3, ic="seeded_random"),
     Case("euler3d_weak_hll_mixed", 3, "euler", 2, sflux="hll", ic="smooth_periodic")] + \
    [c for c in CASES if c.patches and c.name != "linadv3d_mortar_central" and c.name not in _MIXED_SKIP and
     not (c.ic == "convergence_test" and c.source)]  # (cancellation at the manufactured state, as above)


@pytest.mark.parametrize("case", [QUASI_CASES[0], EULER_BLAST_1D,
                                  Case("euler2d_function_blast", 2, "euler", 3, ic="constant",
                                       ic_fn=lambda x, t, eq: [1.0 + 0.1 * x[0], 0.2, -0.1 * x[1],
                                                               2.5 + (0.3 if x[0] * x[0] + x[1] * x[1] < 0.25
                                                                      else 0.0)])],
                         ids=lambda c: c.name)
def test_function_initial_condition(th, oracle, torch, case):
    """An initial condition written as a Python function (as the reference's
    test files define them) is sampled on the host at the node coordinates,
    with Trixi's 1D end-node shift, and packed into the device state."""
    mesh = oracle_mesh(oracle, case)
    u_ref = oracle_state(oracle, case, mesh)
    semi = product_semi(th, case)
    u = th.wrap_array_native(th.compute_coefficients(0.0, semi), semi).cpu().numpy()
    np.testing.assert_array_equal(u, u_ref)
    with pytest.raises(NotImplementedError):
        th.calc_error_norms(th.compute_coefficients(0.0, semi), 0.0, semi)
    semi.close()


@pytest.mark.parametrize("case", MIXED_CASES, ids=lambda c: c.name)
def test_rhs_mixed_within_stated_error(th, oracle, torch, case):
    du, du_ref, _ = run_pair(th, oracle, torch, case, precision="mixed")
    err = rel_max(du, du_ref)
    l2 = float(np.sqrt(np.mean((du - du_ref) ** 2)) / np.sqrt(np.mean(du_ref ** 2)))
    print(f"{case.name}: mixed rel max-norm {err:.3e}, rel L2 {l2:.3e}", flush=True)
    assert err <= MIXED_TOL, err


def test_pack_unpack_roundtrip_bit_exact(th, oracle, torch):
    case = Case("euler3d", 3, "euler", 2, "fd", "ranocha")
    semi = product_semi(th, case)
    x = torch.randn(semi.trixi_shape, dtype=torch.float64, device="cuda")
    nat = th.pack(x, semi)
    back = th.wrap_array_native(nat, semi)
    assert torch.equal(back, x)
    # native layout: u[(v*N^d + node)*E + e] == trixi[e, k, j, i, v]
    xn = x.cpu().numpy()
    nn = nat.cpu().numpy().reshape(semi.native_shape)
    E, N, nv = semi.nelements, semi.nnodes, semi.nvariables
    assert nn[2, 1 + N * 2 + N * N * 3, 7] == xn[7, 3, 2, 1, 2]
    semi.close()


@pytest.mark.parametrize("case", [Case("linadv1d", 1, "linadv", 4, cmin=-1.0, cmax=1.0, ic="convergence_test"),
                                  Case("euler2d_random", 2, "euler", 3),
                                  Case("euler2d_blast", 2, "euler", 3, ic="weak_blast_wave"),
                                  Case("euler2d_blast_seeded", 2, "euler", 3, ic="weak_blast_wave_seeded"),
                                  Case("euler3d_smooth", 3, "euler", 2, ic="smooth_periodic"),
                                  Case("mhd3d_alfven", 3, "mhd", 2, "fd", "hindenlang_gassner", cmin=-1.0,
                                       cmax=1.0, ic="convergence_test", gamma=5 / 3),
                                  Case("euler3d_mortar", 3, "euler", 2, cmin=0.0, cmax=2.0, ic="convergence_test",
                                       patches=EULER_MORTAR_3D),
                                  Case("mhd2d_alfven", 2, "mhd", 3, "fd", "central", cmin=0.0, cmax=2 ** 0.5,
                                       ic="convergence_test", gamma=5 / 3),
                                  Case("burgers1d", 1, "burgers", 4, cmin=0.0, cmax=1.0, ic="convergence_test",
                                       source="convergence_test"),
                                  Case("mhd1d_alfven", 1, "mhd", 3, "fd", "hindenlang_gassner", cmin=0.0, cmax=1.0,
                                       ic="convergence_test", gamma=5 / 3),
                                  Case("hypdiff2d_poisson", 2, "hypdiff", 3, polydeg=4, cmin=0.0, cmax=1.0,
                                       ic="poisson_nonperiodic", periodic=(False, True),
                                       bc_state="poisson_nonperiodic", source="poisson_nonperiodic"),
                                  MULTI_CASES[1], MULTI_CASES[3], MULTI_CASES[4], QUASI_CASES[1], SWE_CASES[0],
                                  SWE_CASES[2]],
                         ids=lambda c: c.name)
def test_initial_condition_and_callbacks_match_oracle(th, oracle, torch, case):
    O = oracle
    cfg = oracle_cfg(O, case)
    mesh = oracle_mesh(O, case)
    u_ref = O.compute_coefficients(cfg, mesh, oracle_ic(O, case), 0.0, case.seed)
    semi = product_semi(th, case)
    u = th.compute_coefficients(0.0, semi)
    u_np = th.wrap_array_native(u, semi).cpu().numpy()
    np.testing.assert_allclose(u_np, u_ref, rtol=1e-13, atol=1e-14)
    # max_dt
    assert th.max_dt(u, semi) == pytest.approx(O.max_dt(cfg, mesh, u_ref), rel=1e-13)
    # integrals: the two reductions sum in different orders (bound n eps for
    # n nodes of one sign: 3.6e-12 at 16384 nodes)
    np.testing.assert_allclose(th.integrate(u, semi), O.integrate(cfg, mesh, u_ref), rtol=1e-12, atol=1e-13)
    # error norms at a later time (nonzero errors)
    t = 0.1
    l2, linf = th.calc_error_norms(u, t, semi)
    l2r, linfr = O.error_norms(cfg, mesh, u_ref, t, oracle_ic(O, case), case.seed)
    # errors that vanish in exact arithmetic (e.g. the constant Alfven-wave
    # density) are compared at absolute round-off
    np.testing.assert_allclose(l2, l2r, rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(linf, linfr, rtol=1e-10, atol=1e-13)
    semi.close()


@pytest.mark.parametrize("case", [Case("euler2d_hg", 2, "euler", 5, "hg", "ranocha", ic="weak_blast_wave"),
                                  Case("euler2d_hg_pressure", 2, "euler", 5, "hg", "ranocha", ic="weak_blast_wave",
                                       indicator="pressure"),
                                  Case("euler3d_hg_density", 3, "euler", 3, "hg", "ranocha", ic="weak_blast_wave",
                                       indicator="density"),
                                  Case("euler2d_hg_mortar", 2, "euler", 4, "hg", "ranocha", ic="weak_blast_wave",
                                       indicator="pressure", patches=(((-0.5, -0.5), (0.5, 0.5)),)),
                                  Case("mhd2d_hg", 2, "mhd", 5, "hg", "hindenlang_gassner", "hindenlang_gassner",
                                       "hindenlang_gassner", ic="weak_blast_wave", gamma=1.4),
                                  # polydeg 4: blocks with idle lanes (2D: two 25-node
                                  # elements per wave; 3D: 125 nodes on 128 lanes)
                                  MHD_HG_P4[0], MHD_HG_P4[1],
                                  # test/tree_dgsem_1d/shallowwater_shock.jl: waterheight_pressure
                                  SWE_CASES[5],
                                  # x = 0.5 falls inside element 29 of 64 on [0, 1.1]
                                  Case("burgers1d_hg_shock", 1, "burgers", 6, "hg", "ec", "llf", "llf", cmin=0.0,
                                       cmax=1.1, ic="shock", periodic=False, bc_state="shock", outflow=("x_pos",),
                                       indicator="first", alpha_max=1.0)],
                         ids=lambda c: c.name)
def test_indicator_matches_oracle(th, oracle, torch, case):
    O = oracle
    cfg = oracle_cfg(O, case)
    mesh = oracle_mesh(O, case)
    u = oracle_state(O, case, mesh)
    alpha_ref = O.indicator_hg(cfg, mesh, u)
    semi = product_semi(th, case)
    alpha = th.blending_factors(th.pack(torch.from_numpy(u).cuda(), semi), semi).cpu().numpy()
    np.testing.assert_allclose(alpha, alpha_ref, rtol=1e-12, atol=1e-14)
    assert alpha_ref.max() > 1e-3 and (alpha_ref == 0).sum() > 0
    semi.close()


@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("case", [Case("euler2d_fd", 2, "euler", 3, "fd", "ranocha"),
                                  Case("euler2d_hg", 2, "euler", 3, "hg", "ranocha", ic="weak_blast_wave"),
                                  Case("mhd3d", 3, "mhd", 1, "fd", "hindenlang_gassner", cmin=-1.0, cmax=1.0,
                                       ic="convergence_test", gamma=5 / 3),
                                  Case("euler2d_weak_dirichlet_src", 2, "euler", 3, cmin=0.0, cmax=2.0,
                                       ic="convergence_test", periodic=False, bc_state="convergence_test",
                                       source="convergence_test"),
                                  Case("euler3d_fd_dirichlet_src", 3, "euler", 2, "fd", "ranocha", cmin=0.0,
                                       cmax=2.0, ic="convergence_test", periodic=False, bc_state="convergence_test",
                                       source="convergence_test"),
                                  Case("linadv3d_mortar", 3, "linadv", 2, cmin=-1.0, cmax=1.0, ic="convergence_test",
                                       advection=(0.2, -0.7, 0.5), patches=MORTAR_3D),
                                  Case("euler3d_mortar_src", 3, "euler", 2, cmin=0.0, cmax=2.0, ic="convergence_test",
                                       source="convergence_test", patches=EULER_MORTAR_3D),
                                  Case("euler2d_hg_mortar", 2, "euler", 3, "hg", "ranocha", cmin=-1.0, cmax=1.0,
                                       gamma=5 / 3, ic="weak_blast_wave_seeded", indicator="pressure",
                                       patches=MORTAR_2D),
                                  Case("mhd2d_alfven", 2, "mhd", 3, "fd", "central", cmin=0.0, cmax=2 ** 0.5,
                                       ic="convergence_test", gamma=5 / 3),
                                  Case("burgers1d_hg_shock", 1, "burgers", 6, "hg", "ec", "llf", "llf", cmin=0.0,
                                       cmax=1.0, ic="shock", periodic=False, bc_state="shock", outflow=("x_pos",),
                                       indicator="first", alpha_max=1.0),
                                  Case("burgers1d_src", 1, "burgers", 4, cmin=0.0, cmax=1.0, ic="convergence_test",
                                       source="convergence_test"),
                                  Case("mhd1d_alfven", 1, "mhd", 3, "fd", "hindenlang_gassner", cmin=0.0, cmax=1.0,
                                       ic="convergence_test", gamma=5 / 3),
                                  Case("euler3d_weak_hll_eoc", 3, "euler", 2, sflux="hll", cmin=0.0, cmax=2.0,
                                       ic="eoc_gravity", source="eoc_test_euler", gamma=2.0),
                                  Case("hypdiff2d_poisson", 2, "hypdiff", 3, polydeg=4, cmin=0.0, cmax=1.0,
                                       ic="poisson_nonperiodic", periodic=(False, True), bc_state="poisson_nonperiodic",
                                       source="poisson_nonperiodic"),
                                  Case("mhd2d_hg_blast", 2, "mhd", 3, "hg", "hindenlang_gassner", "hindenlang_gassner",
                                       "hindenlang_gassner", cmin=-2.0, cmax=2.0, ic="weak_blast_wave_seeded",
                                       gamma=1.4),
                                  MHD_HG_P4[0], MULTI_CASES[3], MULTI_CASES[4], QUASI_CASES[1], SWE_CASES[2],
                                  SWE_CASES[5]],
                         ids=lambda c: c.name)
def test_rk_step_matches_oracle(th, oracle, torch, case, use_graph):
    """Three CK2N54 steps from t = 0 (stage times through device memory: the
    boundary states and sources of the non-periodic cases depend on them)."""
    O = oracle
    cfg = oracle_cfg(O, case)
    mesh = oracle_mesh(O, case)
    u = oracle_state(O, case, mesh)
    semi = product_semi(th, case)
    ud = th.pack(torch.from_numpy(u).cuda(), semi)
    kd = torch.zeros_like(ud)
    dt = case.rk_cfl * O.max_dt(cfg, mesh, u)
    k = np.zeros_like(u)
    uo = u.copy()
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        for step in range(3):
            th.rk_step(ud, kd, step * dt, dt, semi, use_graph=use_graph, stream=stream)
            # eager host work between replays
            torch.cuda.current_stream().synchronize()
            O.ck2n54_step(cfg, mesh, uo, k, step * dt, dt)
            got = th.wrap_array_native(ud, semi).cpu().numpy()
            err = rel_max(got - u, uo - u)
            print(f"{case.name} graph={use_graph} step {step}: rel err of increment {err:.3e}", flush=True)
            assert err < 1e-11, err
    torch.cuda.synchronize()
    semi.close()


@pytest.mark.parametrize("case", [Case("euler3d_fd", 3, "euler", 2, "fd", "ranocha", ic="smooth_periodic"),
                                  Case("euler2d_hg", 2, "euler", 3, "hg", "ranocha", ic="weak_blast_wave_seeded"),
                                  Case("euler2d_fd_mixed", 2, "euler", 3, "fd", "ranocha", precision="mixed"),
                                  Case("linadv2d_mortar", 2, "linadv", 2, cmin=-1.0, cmax=1.0, ic="convergence_test",
                                       advection=(0.2, -0.7), patches=MORTAR_2D)],
                         ids=lambda c: c.name)
def test_rk_stages_equal_rk_step(th, torch, case):
    """Five trixi_hip_rk_stage calls (ping-pong states) == one trixi_hip_rk_step, bit for bit."""
    semi = product_semi(th, case)
    u0 = th.compute_coefficients(0.0, semi)
    dt = 0.5 * th.max_dt(u0, semi)
    t = 0.25
    u_step, k_step = u0.clone(), torch.zeros_like(u0)
    th.rk_step(u_step, k_step, t, dt, semi, use_graph=False)
    a, b, k = u0.clone(), torch.empty_like(u0), torch.full_like(u0, float("nan"))  # stage 0 overwrites k
    for st in range(5):
        th.rk_stage(a, k, b, st, t, dt, semi)
        a, b = b, a
    th.synchronize()
    assert torch.equal(a, u_step)
    assert torch.equal(k, k_step)
    with pytest.raises(th.TrixiHipError):
        th.rk_stage(a, k, a, 0, t, dt, semi)
    with pytest.raises(th.TrixiHipError):
        th.rk_stage(a, k, b, 5, t, dt, semi)
    semi.close()


def test_c1_trajectory_matches_oracle(th, oracle, torch):
    """C1: 1D advection, level 4, CK2N54 with CFL 1.6 to t=1: L2/Linf to 1e-10."""
    O = oracle
    case = Case("c1", 1, "linadv", 4, cmin=-1.0, cmax=1.0, ic="convergence_test", advection=(1.0,))
    semi = product_semi(th, case)
    ode = th.semidiscretizeHIP(semi, (0.0, 1.0))
    sol = th.solve(ode, th.CarpenterKennedy2N54(williamson_condition=False),
                   callback=[th.StepsizeCallback(cfl=1.6)])
    assert sol.nsteps == 20
    l2, linf = th.calc_error_norms(sol.u, sol.t, semi)
    cfg = oracle_cfg(O, case)
    mesh = O.Mesh(1, 4, -1.0, 1.0, 3)
    u = O.compute_coefficients(cfg, mesh, O.IC_CONVERGENCE_TEST)
    k = np.zeros_like(u)
    t = 0.0
    for _ in range(20):
        dt = min(1.6 * O.max_dt(cfg, mesh, u), 1.0 - t)
        O.ck2n54_step(cfg, mesh, u, k, t, dt)
        t += dt
    l2r, linfr = O.error_norms(cfg, mesh, u, t, O.IC_CONVERGENCE_TEST)
    print(f"C1: gpu l2={l2[0]:.16e} linf={linf[0]:.16e}; oracle l2={l2r[0]:.16e} linf={linfr[0]:.16e}")
    np.testing.assert_allclose(l2, l2r, rtol=1e-10)
    np.testing.assert_allclose(linf, linfr, rtol=1e-10)
    semi.close()


def test_check_state_flags(th, oracle, torch):
    case = Case("euler2d", 2, "euler", 2, "fd", "ranocha")
    semi = product_semi(th, case)
    u = th.compute_coefficients(0.0, semi)
    nbad, flags = th.check_state(u, semi)
    assert nbad == 0 and int(flags.abs().sum()) == 0
    v = th.wrap_array(u, semi)
    v[0, 3, 5] = -1.0          # rho <= 0 in element 5
    v[1, 0, 9] = float("nan")  # non-finite in element 9
    nbad, flags = th.check_state(u, semi)
    f = flags.cpu().numpy()
    assert nbad == 2 and f[5] & 2 and f[9] & 1
    semi.close()


@pytest.mark.parametrize("case", [
    Case("mhd_hg_llf_3d", 3, "mhd", 1, "hg", "hindenlang_gassner", cmin=-1.0, cmax=1.0, gamma=5 / 3),
    # three gas components (the library builds one or two)
    Case("eulermulti_3comp", 1, "eulermulti", 2, "fd", "ranocha", "ranocha", gammas=(1.4,) * 3,
         gas_constants=(0.4,) * 3),
    # slip walls are built for shallow water only
    Case("euler_slip_wall", 2, "euler", 2, periodic=False, bc_state="slip_wall"),
    # multicomponent Euler has no sources
    Case("eulermulti_source", 2, "eulermulti", 2, "fd", "ranocha", "llf", source="convergence_test"),
    # quasi-1D Euler needs flux differencing
    Case("quasi1d_weak", 1, "quasi1d", 2, "weak", "central", "chan_etal", polydeg=4),
    # the shallow-water indicator variable is waterheight_pressure
    Case("swe1d_hg_pressure", 1, "swe", 2, "hg", "wintermeyer_etal", "hydrostatic_llf", "hydrostatic_llf",
         polydeg=4, gravity=9.812, indicator="pressure"),
], ids=lambda c: c.name)
def test_unsupported_combination_is_an_error(th, torch, case):
    """Configurations outside what is built fail in create with
    EUNSUPPORTED, naming the reason, and allocate nothing on the device."""
    with pytest.raises(th.TrixiHipError) as ei:
        product_semi(th, case)
    assert ei.value.code == -2, ei.value


HEADLINE = Case("euler3d_level5_headline", 3, "euler", 5, "fd", "ranocha", "llf", ic="smooth_periodic")


# (free stream: absolute bound on du for flux terms of size 3.5 times D
# entries of size 10 times 1/J = 16, at FP64 and FP32 round-off)
@pytest.mark.parametrize("precision,tol,cons,fs", [("fp64", 1e-12, 1e-13, 1e-11), ("mixed", 1e-4, 1e-6, 1e-3)])
def test_headline_configuration_full_size(th, oracle, torch, precision, tol, cons, fs):
    """The measured configuration at its full size (3D Euler, flux_ranocha +
    LLF, polydeg 3, level 5: 32768 elements, 2097152 DOF): the whole state
    against the oracle, conservation of every variable (the integral of du
    vanishes on the periodic mesh) and free-stream preservation."""
    O = oracle
    case = HEADLINE
    cfg, mesh = oracle_cfg(O, case), oracle_mesh(O, case)
    u = O.compute_coefficients(cfg, mesh, oracle_ic(O, case), 0.0, case.seed)
    du_ref = O.rhs(cfg, mesh, u)
    semi = product_semi(th, case, precision)
    ud = th.pack(torch.from_numpy(u).cuda(), semi)
    dud = torch.empty_like(ud)
    th.rhs_hip_(dud, ud, semi, 0.0)
    du = th.wrap_array_native(dud, semi).cpu().numpy()
    err = rel_max(du, du_ref)
    print(f"headline {precision}: rel max-norm {err:.3e}", flush=True)
    assert err <= tol, err
    scale = np.abs(du_ref).max() * mesh.total_volume
    integral = th.integrate(dud, semi)
    assert np.abs(integral).max() <= cons * scale, integral
    uc = np.zeros_like(u)
    uc[..., 0], uc[..., 4] = 1.0, 1.0 / 0.4
    th.rhs_hip_(dud, th.pack(torch.from_numpy(uc).cuda(), semi), semi, 0.0)
    print(f"headline {precision}: free stream max |du| {float(dud.abs().max()):.3e}", flush=True)
    assert float(dud.abs().max()) <= fs, float(dud.abs().max())
    semi.close()


def test_headline_rk_step_full_size(th, oracle, torch):
    """One CK2N54 step (hipGraph) of the measured configuration at full size
    against the oracle's step."""
    O = oracle
    case = HEADLINE
    cfg, mesh = oracle_cfg(O, case), oracle_mesh(O, case)
    u = O.compute_coefficients(cfg, mesh, oracle_ic(O, case), 0.0, case.seed)
    dt = 0.5 * O.max_dt(cfg, mesh, u)
    semi = product_semi(th, case)
    ud = th.pack(torch.from_numpy(u).cuda(), semi)
    kd = torch.zeros_like(ud)
    stream = torch.cuda.Stream()  # graph replay runs on a non-default stream
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        th.rk_step(ud, kd, 0.0, dt, semi, use_graph=True, stream=stream)
    torch.cuda.synchronize()
    uo, k = u.copy(), np.zeros_like(u)
    O.ck2n54_step(cfg, mesh, uo, k, 0.0, dt)
    got = th.wrap_array_native(ud, semi).cpu().numpy()
    err = rel_max(got - u, uo - u)
    print(f"headline RK step: rel err of increment {err:.3e}", flush=True)
    assert err < 1e-11, err
    semi.close()


def test_save_solution_callback_files_match_device_state(th, torch, tmp_path):
    """SaveSolutionCallback during a 2D Euler shock-capturing solve: the
    initial, every-2nd-step and final files hold cons2prim of the unpacked
    device state bit for bit, with the blending factors as element variable."""
    import os
    if not os.path.exists(th.save_solution.H5_LIBRARY_PATH):
        pytest.skip("libtrixih5.so not built (no HDF5 library)")
    from test_save_solution import H5DUMP, h5_dataset
    if H5DUMP is None:
        pytest.skip("h5dump absent")
    case = Case("euler2d_hg_save", 2, "euler", 3, "hg", "ranocha", "llf", "llf", ic="weak_blast_wave")
    semi = product_semi(th, case)
    u0 = th.compute_coefficients(0.0, semi)
    save = th.SaveSolutionCallback(interval=2, output_directory=str(tmp_path), solution_variables=th.cons2prim)
    ode = th.semidiscretizeHIP(semi, (0.0, 0.05))
    sol = th.solve(ode, th.CarpenterKennedy2N54(williamson_condition=False),
                   callback=[th.StepsizeCallback(cfl=0.5), save])
    names = [os.path.basename(f) for f in save.files]
    steps = sorted(set([0] + list(range(2, sol.nsteps + 1, 2)) + [sol.nsteps]))
    assert names == ["solution_%09d.h5" % s for s in steps], (names, sol.nsteps)
    for path, u in ((save.files[0], u0), (save.files[-1], sol.u)):
        prim = th.cons2prim(th.wrap_array_native(u, semi).cpu().numpy(), semi.equations)
        for v in range(4):
            np.testing.assert_array_equal(h5_dataset(path, f"/variables_{v + 1}", tmp_path), prim[..., v].reshape(-1))
        alpha = th.blending_factors(u, semi).cpu().numpy()
        np.testing.assert_array_equal(h5_dataset(path, "/element_variables_1", tmp_path), alpha)
    semi.close()


def test_save_solution_callback_dt_lands_on_save_times(th, torch, tmp_path):
    """SaveSolutionCallback(dt=0.25) on the C1 advection run (CFL 1.6, t in
    [0, 1]): the steps are shortened to land on 0.25, 0.5, 0.75 and 1, each
    file's time attribute is that time, and the final state still meets the
    oracle trajectory's accuracy (errors within 1e-3 of the exact solution)."""
    import os
    if not os.path.exists(th.save_solution.H5_LIBRARY_PATH):
        pytest.skip("libtrixih5.so not built (no HDF5 library)")
    from test_save_solution import H5DUMP, h5_attributes
    if H5DUMP is None:
        pytest.skip("h5dump absent")
    case = Case("c1_save_dt", 1, "linadv", 4, cmin=-1.0, cmax=1.0, ic="convergence_test", advection=(1.0,))
    semi = product_semi(th, case)
    save = th.SaveSolutionCallback(dt=0.25, output_directory=str(tmp_path))
    sol = th.solve(th.semidiscretizeHIP(semi, (0.0, 1.0)), th.CarpenterKennedy2N54(williamson_condition=False),
                   callback=[th.StepsizeCallback(cfl=1.6), save])
    times = [float(h5_attributes(f)[0][("/", "time")]) for f in save.files]
    assert times == [0.0, 0.25, 0.5, 0.75, 1.0], (times, save.files)
    assert sol.t == 1.0
    steps = [int(h5_attributes(f)[0][("/", "timestep")]) for f in save.files]
    assert steps == sorted(steps) and steps[-1] == sol.nsteps
    l2, linf = th.calc_error_norms(sol.u, sol.t, semi)
    assert linf[0] < 1e-3
    semi.close()
