"""GPU adjoint: the hand-written VJP kernels against exact derivatives.

* rhs! VJP == J^T vbar with J assembled column by column from the oracle's
  dual-number JVP (tiny meshes), relative max-norm <= 1e-11;
* dot-product test <vbar, J w> (oracle JVP) == <J^T vbar, w> (GPU VJP) on a
  larger mesh, relative 1e-11;
* reverse CK2N54 step against the oracle's dual-number step JVP;
* 20-step gradient of the L2 norm with step checkpointing against
  fourth-order central differences of the GPU forward map (smooth fluxes,
  relative 1e-8), and, with LLF, against the oracle's chained step JVPs
  (relative 1e-10);
* the autograd function backward == VJP.
"""
import numpy as np
import pytest

from cases import Case, oracle_cfg, oracle_ic, oracle_mesh, product_semi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def setup(th, O, torch, case):
    cfg = oracle_cfg(O, case)
    mesh = oracle_mesh(O, case)
    u = O.compute_coefficients(cfg, mesh, oracle_ic(O, case), case.t, case.seed)
    semi = product_semi(th, case)
    ud = th.pack(torch.from_numpy(u).cuda(), semi)
    return cfg, mesh, u, semi, ud


TINY = [
    Case("linadv1d_weak_llf", 1, "linadv", 2, cmin=-1.0, cmax=1.0, ic="convergence_test", advection=(0.7,)),
    Case("euler1d_weak_llf", 1, "euler", 2),
    Case("euler1d_fd_ranocha_llf", 1, "euler", 2, "fd", "ranocha"),
    Case("euler2d_fd_ranocha_llf", 2, "euler", 1, "fd", "ranocha"),
    Case("euler2d_fd_ranocha_llfnaive", 2, "euler", 1, "fd", "ranocha", "llf_naive"),
    Case("euler2d_fd_shima_shima", 2, "euler", 1, "fd", "shima_etal", "shima_etal"),
    Case("euler2d_weak_central", 2, "euler", 1, sflux="central"),
    Case("euler3d_fd_ranocha_ranocha", 3, "euler", 1, "fd", "ranocha", "ranocha"),
    Case("mhd3d_fd_hg_llf", 3, "mhd", 1, "fd", "hindenlang_gassner", cmin=-1.0, cmax=1.0, gamma=5 / 3),
    # Hennemann-Gassner shock capturing: blending factors inside the logistic
    # range (weak blast, level 1: 0.0065), smoothed from a neighbour and
    # clipped (seeded random, level 2), and alpha_max = 1
    Case("euler2d_hg_ranocha_blast", 2, "euler", 1, "hg", "ranocha", ic="weak_blast_wave_seeded"),
    Case("euler2d_hg_ranocha_random", 2, "euler", 2, "hg", "ranocha"),
    Case("euler2d_hg_shima_llfnaive_amax1", 2, "euler", 2, "hg", "shima_etal", "llf_naive", "llf_naive",
         alpha_max=1.0),
    Case("euler1d_hg_ranocha_blast", 1, "euler", 2, "hg", "ranocha", ic="weak_blast_wave_seeded"),
    # indicator variables pressure and density
    Case("euler2d_hg_pressure_blast", 2, "euler", 1, "hg", "ranocha", ic="weak_blast_wave_seeded", indicator="pressure"),
    Case("euler2d_hg_density_random", 2, "euler", 2, "hg", "ranocha", indicator="density"),
    # non-periodic: Dirichlet states do not depend on u, the source depends on x and t only
    Case("euler2d_weak_dirichlet_src", 2, "euler", 1, cmin=0.0, cmax=2.0, periodic=False, bc_state="convergence_test",
         source="convergence_test", t=0.3),
    Case("euler1d_fd_dirichlet", 1, "euler", 2, "fd", "ranocha", cmin=0.0, cmax=2.0, periodic=False,
         bc_state="convergence_test", t=0.6),
    Case("euler2d_fd_dirichlet", 2, "euler", 2, "fd", "ranocha", cmin=0.0, cmax=2.0, periodic=False,
         bc_state="convergence_test", t=0.6),
    Case("mhd3d_fd_hg_llfnaive", 3, "mhd", 1, "fd", "hindenlang_gassner", "llf_naive", cmin=-1.0, cmax=1.0,
         gamma=5 / 3),
    Case("mhd3d_fd_hg_hg", 3, "mhd", 1, "fd", "hindenlang_gassner", "hindenlang_gassner", cmin=-1.0, cmax=1.0,
         gamma=5 / 3),
    Case("mhd2d_fd_central_llf", 2, "mhd", 2, "fd", "central", cmin=-1.0, cmax=1.0, gamma=5 / 3),
    Case("mhd2d_fd_hg_hg", 2, "mhd", 2, "fd", "hindenlang_gassner", "hindenlang_gassner", cmin=-1.0, cmax=1.0),
    Case("mhd2d_fd_hg_llfnaive", 2, "mhd", 1, "fd", "hindenlang_gassner", "llf_naive", cmin=-1.0, cmax=1.0),
]


@pytest.mark.parametrize("case", TINY, ids=lambda c: c.name)
def test_vjp_equals_transposed_oracle_jacobian(th, oracle, torch, case):
    O = oracle
    cfg, mesh, u, semi, ud = setup(th, O, torch, case)
    J = O.jacobian(cfg, mesh, u, case.t)
    rng = np.random.default_rng(7)
    vbar = rng.standard_normal(u.shape)
    expect = (J.T @ vbar.ravel()).reshape(u.shape)
    vd = th.pack(torch.from_numpy(vbar).cuda(), semi)
    ubar = torch.empty_like(vd)
    th.rhs_vjp(ubar, ud, vd, semi, t=case.t)
    got = th.wrap_array_native(ubar, semi).cpu().numpy()
    err = float(np.abs(got - expect).max() / np.abs(expect).max())
    print(f"{case.name}: VJP vs J^T vbar rel max {err:.3e}", flush=True)
    assert err < 1e-11, err
    # beta accumulation: ubar = 2 ubar + J^T (3 vbar)
    th.rhs_vjp(ubar, ud, vd, semi, t=case.t, alpha=3.0, beta=2.0)
    got2 = th.wrap_array_native(ubar, semi).cpu().numpy()
    assert float(np.abs(got2 - 5 * expect).max() / np.abs(expect).max()) < 1e-11
    semi.close()


@pytest.mark.parametrize("case", [Case("euler3d_fd_ranocha_llf", 3, "euler", 2, "fd", "ranocha", ic="smooth_periodic"),
                                  Case("euler2d_fd_shima_llf", 2, "euler", 3, "fd", "shima_etal"),
                                  Case("linadv3d_weak_llf", 3, "linadv", 2),
                                  Case("mhd3d_fd_hg_llf_alfven", 3, "mhd", 2, "fd", "hindenlang_gassner",
                                       cmin=-1.0, cmax=1.0, ic="convergence_test", gamma=5 / 3),
                                  Case("euler3d_hg_ranocha_blast", 3, "euler", 2, "hg", "ranocha",
                                       ic="weak_blast_wave_seeded"),
                                  Case("euler3d_hg_ranocha_fv_ranocha", 3, "euler", 2, "hg", "ranocha", "ranocha",
                                       "ranocha", ic="weak_blast_wave_seeded")],
                         ids=lambda c: c.name)
def test_vjp_dot_product(th, oracle, torch, case):
    O = oracle
    cfg, mesh, u, semi, ud = setup(th, O, torch, case)
    rng = np.random.default_rng(8)
    w = rng.standard_normal(u.shape)
    vbar = rng.standard_normal(u.shape)
    _, jw = O.rhs_jvp(cfg, mesh, u, w)
    ubar = torch.empty_like(ud)
    th.rhs_vjp(ubar, ud, th.pack(torch.from_numpy(vbar).cuda(), semi), semi)
    lhs = float((vbar * jw).sum())
    rhs = float((th.wrap_array_native(ubar, semi).cpu().numpy() * w).sum())
    print(f"{case.name}: <v,Jw>={lhs:.15e} <J^T v,w>={rhs:.15e}", flush=True)
    assert abs(lhs - rhs) <= 1e-11 * max(abs(lhs), np.abs(vbar).sum() * np.abs(jw).max() * 1e-3)
    semi.close()


def test_rk_step_vjp_dot_product(th, oracle, torch):
    O = oracle
    case = Case("euler2d_fd", 2, "euler", 2, "fd", "ranocha")
    cfg, mesh, u, semi, ud = setup(th, O, torch, case)
    dt = 0.5 * O.max_dt(cfg, mesh, u)
    rng = np.random.default_rng(9)
    w = rng.standard_normal(u.shape)
    vbar = rng.standard_normal(u.shape)
    _, dphi = O.ck2n54_step_jvp(cfg, mesh, u, w, dt)
    th.enable_adjoint(semi)
    ubar = th.pack(torch.from_numpy(vbar).cuda(), semi)
    th.rk_step_vjp(ud, ubar, 0.0, dt, semi)
    lhs = float((vbar * dphi).sum())
    rhs = float((th.wrap_array_native(ubar, semi).cpu().numpy() * w).sum())
    print(f"rk step: <v,dPhi w>={lhs:.15e} <dPhi^T v,w>={rhs:.15e}", flush=True)
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    semi.close()


def test_multistep_gradient_matches_finite_differences(th, oracle, torch):
    """Gradient of ||u(t_20)||_L2 w.r.t. u0 over 20 CK2N54 steps, checkpointed,
    against fourth-order central differences of the GPU forward map.  The
    surface flux is flux_ranocha, so the map is smooth (LLF's max wave speed
    has kinks that finite differences step across: covered by the next test)."""
    case = Case("euler2d_fd", 2, "euler", 2, "fd", "ranocha", "ranocha", ic="smooth_periodic")
    semi = product_semi(th, case)
    u0 = th.compute_coefficients(0.0, semi)
    dt = 0.5 * th.max_dt(u0, semi)
    nsteps = 20
    traj = th.forward_with_checkpoints(semi, u0, dt, nsteps)
    loss, g_final = th.l2_norm_and_gradient(traj.u_final, semi)
    grad = th.backward(traj, g_final)

    def loss_of(u_init):
        tr = th.forward_with_checkpoints(semi, u_init, dt, nsteps)
        return float(th.l2_norm_and_gradient(tr.u_final, semi)[0])

    gen = torch.Generator(device="cuda").manual_seed(10)
    for i in range(3):
        d = torch.randn(u0.shape, dtype=torch.float64, device="cuda", generator=gen)
        h = 1e-4
        fd = (-loss_of(u0 + 2 * h * d) + 8 * loss_of(u0 + h * d) - 8 * loss_of(u0 - h * d)
              + loss_of(u0 - 2 * h * d)) / (12 * h)
        ad = float((grad * d).sum())
        print(f"direction {i}: adjoint {ad:.12e} fourth-order FD {fd:.12e}", flush=True)
        assert abs(ad - fd) <= 1e-8 * abs(fd) + 1e-12
    semi.close()


@pytest.mark.parametrize("sflux", ["llf", "llf_naive"])
def test_multistep_gradient_matches_oracle_tangent(th, oracle, torch, sflux):
    """<d u0, adjoint of g> over 20 checkpointed CK2N54 steps with an LLF
    surface flux == <g, d u_20> from the oracle's chained dual-number step JVPs."""
    O = oracle
    case = Case("euler2d_fd", 2, "euler", 2, "fd", "ranocha", sflux, ic="smooth_periodic")
    cfg, mesh, u, semi, ud = setup(th, O, torch, case)
    dt = 0.5 * O.max_dt(cfg, mesh, u)
    nsteps = 20
    rng = np.random.default_rng(11)
    d = rng.standard_normal(u.shape)
    g = rng.standard_normal(u.shape)
    ut, dt_ = u.copy(), d.copy()
    for _ in range(nsteps):
        ut, dt_ = O.ck2n54_step_jvp(cfg, mesh, ut, dt_, dt)
    exact = float((g * dt_).sum())
    traj = th.forward_with_checkpoints(semi, ud, dt, nsteps)
    grad = th.backward(traj, th.pack(torch.from_numpy(g).cuda(), semi))
    ad = float((th.wrap_array_native(grad, semi).cpu().numpy() * d).sum())
    print(f"{sflux}: adjoint {ad:.15e} oracle tangent {exact:.15e}", flush=True)
    assert abs(ad - exact) <= 1e-10 * abs(exact)
    semi.close()


MHD_ALFVEN = dict(cmin=-1.0, cmax=1.0, ic="convergence_test", gamma=5 / 3)


def test_mhd_multistep_gradient_matches_oracle_tangent(th, oracle, torch):
    """GLM-MHD Alfven wave (flux_hindenlang_gassner volume flux, LLF surface
    flux, Powell terms): <d u0, adjoint of g> over 20 checkpointed CK2N54
    steps == <g, d u_20> from the oracle's chained dual-number step JVPs."""
    O = oracle
    case = Case("mhd3d_alfven", 3, "mhd", 2, "fd", "hindenlang_gassner", "llf", **MHD_ALFVEN)
    cfg, mesh, u, semi, ud = setup(th, O, torch, case)
    dt = 0.5 * O.max_dt(cfg, mesh, u)
    nsteps = 20
    rng = np.random.default_rng(12)
    d = rng.standard_normal(u.shape)
    g = rng.standard_normal(u.shape)
    ut, dt_ = u.copy(), d.copy()
    for _ in range(nsteps):
        ut, dt_ = O.ck2n54_step_jvp(cfg, mesh, ut, dt_, dt)
    exact = float((g * dt_).sum())
    traj = th.forward_with_checkpoints(semi, ud, dt, nsteps)
    grad = th.backward(traj, th.pack(torch.from_numpy(g).cuda(), semi))
    ad = float((th.wrap_array_native(grad, semi).cpu().numpy() * d).sum())
    print(f"mhd llf: adjoint {ad:.15e} oracle tangent {exact:.15e}", flush=True)
    assert abs(ad - exact) <= 1e-10 * abs(exact)
    semi.close()


@pytest.mark.parametrize("sflux", ["hindenlang_gassner", "llf"])
def test_mhd_l2_gradient_matches_finite_differences(th, oracle, torch, sflux):
    """Gradient of ||u(t_20)||_L2 with respect to u0 for the GLM-MHD Alfven
    wave, against fourth-order central differences of the GPU forward map.
    With flux_hindenlang_gassner at the surfaces the map is smooth (relative
    1e-8); LLF's max wave speed has kinks that the differences step across,
    so there the bound is 1e-4 (its exact check is the oracle tangent test)."""
    case = Case("mhd3d_alfven", 3, "mhd", 2, "fd", "hindenlang_gassner", sflux, **MHD_ALFVEN)
    semi = product_semi(th, case)
    u0 = th.compute_coefficients(0.0, semi)
    dt = 0.5 * th.max_dt(u0, semi)
    nsteps = 20
    traj = th.forward_with_checkpoints(semi, u0, dt, nsteps)
    loss, g_final = th.l2_norm_and_gradient(traj.u_final, semi)
    grad = th.backward(traj, g_final)

    def loss_of(u_init):
        tr = th.forward_with_checkpoints(semi, u_init, dt, nsteps)
        return float(th.l2_norm_and_gradient(tr.u_final, semi)[0])

    gen = torch.Generator(device="cuda").manual_seed(13)
    tol = 1e-8 if sflux == "hindenlang_gassner" else 1e-4
    for i in range(2):
        d = torch.randn(u0.shape, dtype=torch.float64, device="cuda", generator=gen)
        h = 1e-4
        fd = (-loss_of(u0 + 2 * h * d) + 8 * loss_of(u0 + h * d) - 8 * loss_of(u0 - h * d)
              + loss_of(u0 - 2 * h * d)) / (12 * h)
        ad = float((grad * d).sum())
        print(f"mhd {sflux} direction {i}: adjoint {ad:.12e} fourth-order FD {fd:.12e}", flush=True)
        assert abs(ad - fd) <= tol * abs(fd) + 1e-12
    semi.close()


def test_c4_l2_gradient_full_config(th, torch):
    """C4 as stated: GLM-MHD Alfven wave on [-1,1]^3 at level 4 (4096
    elements), flux_hindenlang_gassner + Powell volume terms, LLF + Powell
    surface terms, c_h = 1, gamma = 5/3; gradient of ||u(t_20)||_L2 with
    respect to u0 over 20 CK2N54 steps of dt = max_dt (CFL 1.0), against
    fourth-order central differences on 16 random directions (relative 1e-5:
    LLF's wave-speed max is not differentiable everywhere)."""
    case = Case("c4", 3, "mhd", 4, "fd", "hindenlang_gassner", "llf", **MHD_ALFVEN)
    semi = product_semi(th, case)
    u0 = th.compute_coefficients(0.0, semi)
    dt = 1.0 * th.max_dt(u0, semi)
    nsteps = 20
    traj = th.forward_with_checkpoints(semi, u0, dt, nsteps)
    _, g_final = th.l2_norm_and_gradient(traj.u_final, semi)
    grad = th.backward(traj, g_final)

    def loss_of(u_init):
        tr = th.forward_with_checkpoints(semi, u_init, dt, nsteps)
        return float(th.l2_norm_and_gradient(tr.u_final, semi)[0])

    gen = torch.Generator(device="cuda").manual_seed(20261202)
    worst = 0.0
    for i in range(16):
        d = torch.randn(u0.shape, dtype=torch.float64, device="cuda", generator=gen)
        h = 1e-4
        fd = (-loss_of(u0 + 2 * h * d) + 8 * loss_of(u0 + h * d) - 8 * loss_of(u0 - h * d)
              + loss_of(u0 - 2 * h * d)) / (12 * h)
        ad = float((grad * d).sum())
        worst = max(worst, abs(ad - fd) / abs(fd))
        assert abs(ad - fd) <= 1e-5 * abs(fd), (i, ad, fd)
    print(f"c4: 16 directions, worst relative difference adjoint vs FD {worst:.3e}", flush=True)
    semi.close()


def test_nonperiodic_multistep_gradient_matches_oracle_tangent(th, oracle, torch):
    """2D Euler on a non-periodic mesh with the exact convergence-test state on
    the boundary and source_terms_convergence_test: 10 checkpointed CK2N54
    steps from t = 0 (time-dependent boundary states and sources at every
    stage) against the oracle's chained step JVPs."""
    O = oracle
    case = Case("euler2d_dirichlet_src", 2, "euler", 2, cmin=0.0, cmax=2.0, ic="convergence_test", periodic=False,
                bc_state="convergence_test", source="convergence_test")
    cfg, mesh, u, semi, ud = setup(th, O, torch, case)
    dt = 0.5 * O.max_dt(cfg, mesh, u)
    nsteps = 10
    rng = np.random.default_rng(21)
    d = rng.standard_normal(u.shape)
    g = rng.standard_normal(u.shape)
    ut, dt_ = u.copy(), d.copy()
    for n in range(nsteps):
        ut, dt_ = O.ck2n54_step_jvp(cfg, mesh, ut, dt_, dt, t=n * dt)
    exact = float((g * dt_).sum())
    traj = th.forward_with_checkpoints(semi, ud, dt, nsteps)
    grad = th.backward(traj, th.pack(torch.from_numpy(g).cuda(), semi))
    ad = float((th.wrap_array_native(grad, semi).cpu().numpy() * d).sum())
    print(f"non-periodic: adjoint {ad:.15e} oracle tangent {exact:.15e}", flush=True)
    assert abs(ad - exact) <= 1e-10 * abs(exact)
    semi.close()


HG_BLAST = dict(ic="weak_blast_wave_seeded")


def test_hg_multistep_gradient_matches_oracle_tangent(th, oracle, torch):
    """2D Euler with Hennemann-Gassner shock capturing (weak blast, level 3,
    alpha_max 0.5, flux_ranocha, FV subcells with LLF): <d u0, adjoint of g>
    over 100 checkpointed CK2N54 steps == <g, d u_100> from the oracle's
    chained dual-number step JVPs, which differentiate through the blending
    factor, its smoothing and clipping."""
    O = oracle
    case = Case("euler2d_hg_blast", 2, "euler", 3, "hg", "ranocha", **HG_BLAST)
    cfg, mesh, u, semi, ud = setup(th, O, torch, case)
    dt = 0.5 * O.max_dt(cfg, mesh, u)
    nsteps = 100
    rng = np.random.default_rng(14)
    d = rng.standard_normal(u.shape)
    g = rng.standard_normal(u.shape)
    ut, dt_ = u.copy(), d.copy()
    for _ in range(nsteps):
        ut, dt_ = O.ck2n54_step_jvp(cfg, mesh, ut, dt_, dt)
    exact = float((g * dt_).sum())
    traj = th.forward_with_checkpoints(semi, ud, dt, nsteps)
    grad = th.backward(traj, th.pack(torch.from_numpy(g).cuda(), semi))
    ad = float((th.wrap_array_native(grad, semi).cpu().numpy() * d).sum())
    print(f"hg: adjoint {ad:.15e} oracle tangent {exact:.15e}", flush=True)
    assert abs(ad - exact) <= 1e-9 * abs(exact)
    semi.close()


def test_hg_mixed_forward_checkpointed_adjoint(th, oracle, torch):
    """The mixed-precision forward trajectory (FP32 fluxes, FP64 state and RK
    update) with the FP64 adjoint evaluated from its checkpoints, against the
    all-FP64 gradient of the same loss over 100 steps: the difference is the
    forward's FP32 rounding carried through the gradient (stated, bounded
    by 1e-3 relative in the L2 norm)."""
    case = Case("euler2d_hg_blast", 2, "euler", 3, "hg", "ranocha", **HG_BLAST)
    s64 = product_semi(th, case)
    smx = product_semi(th, case, precision="mixed")
    u0 = th.compute_coefficients(0.0, s64)
    dt = 0.5 * th.max_dt(u0, s64)
    nsteps = 100
    grads = {}
    for name, semi in (("fp64", s64), ("mixed", smx)):
        traj = th.forward_with_checkpoints(semi, u0, dt, nsteps)
        _, g_final = th.l2_norm_and_gradient(traj.u_final, semi)
        grads[name] = th.backward(traj, g_final)
    a, b = grads["mixed"], grads["fp64"]
    rel = float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
    print(f"hg gradient, mixed forward vs FP64 forward: relative L2 difference {rel:.3e}", flush=True)
    assert rel < 1e-3
    s64.close()
    smx.close()


def test_autograd_function_backward(th, oracle, torch):
    case = Case("euler3d_fd", 3, "euler", 1, "fd", "ranocha")
    semi = product_semi(th, case)
    u = th.compute_coefficients(0.0, semi).requires_grad_(True)
    c = torch.randn_like(u)
    (th.rhs(u, semi) * c).sum().backward()
    ubar = torch.empty_like(c)
    th.rhs_vjp(ubar, u.detach(), c, semi)
    assert torch.allclose(u.grad, ubar, rtol=0, atol=0)
    semi.close()


@pytest.mark.parametrize("c", [Case("linadv2d_p2", 2, "linadv", 1, polydeg=2),
                               Case("linadv2d_mortar", 2, "linadv", 2, cmin=-1.0, cmax=1.0,
                                    patches=(((0.0, -1.0), (1.0, 1.0)),))], ids=lambda c: c.name)
def test_adjoint_unsupported_is_an_error(th, torch, c):
    semi = product_semi(th, c)
    u = th.compute_coefficients(0.0, semi)
    with pytest.raises(th.TrixiHipError) as ei:
        th.rhs_vjp(torch.empty_like(u), u, torch.ones_like(u), semi)
    assert ei.value.code == -2
    semi.close()


@pytest.mark.parametrize("case", [c for c in TINY if c.ndims > 1 and not c.patches], ids=lambda c: c.name)
def test_vjp_direction_groups_equal_one_group(th, torch, case):
    """The direction-group form of the VJP kernel (small meshes: one lane
    group and LDS accumulator per direction, summed in direction order; for
    GLM-MHD and 2D shock capturing one more group with all the interface
    fluxes) against the one-group form (directions one after the other in
    every lane): the same terms, summed in another order, so they agree to
    round-off (relative max-norm 1e-13)."""
    import os
    out = {}
    for grouped in (True, False):
        old = os.environ.get("TRIXI_HIP_VJP_PD_LANES")
        os.environ["TRIXI_HIP_VJP_PD_LANES"] = str(1 << 40) if grouped else "0"
        try:
            semi = product_semi(th, case)
        finally:
            if old is None:
                del os.environ["TRIXI_HIP_VJP_PD_LANES"]
            else:
                os.environ["TRIXI_HIP_VJP_PD_LANES"] = old
        try:
            u = th.compute_coefficients(0.0, semi)
            gen = torch.Generator(device="cuda").manual_seed(20261202)
            vb = torch.randn(u.numel(), dtype=torch.float64, device="cuda", generator=gen)
            ub = torch.empty_like(u)
            th.rhs_vjp(ub, u, vb, semi)
            torch.cuda.synchronize()
            out[grouped] = ub.cpu().numpy()
        finally:
            semi.close()
    a, b = out[True], out[False]
    err = float(np.abs(a - b).max() / np.abs(b).max())
    print(f"{case.name}: grouped vs one-group VJP rel max-norm {err:.3e}", flush=True)
    assert np.all(np.isfinite(a))
    assert err <= 1e-13, err
