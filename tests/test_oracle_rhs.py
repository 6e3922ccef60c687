"""Pins the oracle's full semidiscretization by invariants of the scheme.

* free-stream preservation, conservation, entropy conservation (EC volume +
  EC surface flux, Euler and GLM-MHD with the Powell term), entropy
  dissipation with LLF;
* pressure-equilibrium preservation of flux_shima_etal / flux_ranocha;
* exactness for polynomial data of linear advection away from the periodic seam;
* CK2N54 order of accuracy against the matrix exponential of the (linear)
  semidiscretization, and the C1 trajectory against Trixi's published test
  values (a recalled sanity check only, not part of the reference tree);
* the HG indicator's special cases and smoothing;
* the dual-number JVP against central finite differences.
"""
import numpy as np
import pytest
import scipy.linalg


def entropy_rate(O, cfg, mesh, u, du):
    nv = O.nvars(cfg)
    w = np.apply_along_axis(lambda x: O.cons2entropy(cfg, x), -1, u.reshape(-1, nv)).reshape(u.shape)
    wd = (w * du).sum(-1, keepdims=True)
    return O.integrate(cfg, mesh, np.repeat(wd, nv, axis=-1))[0]


@pytest.mark.parametrize("nd,level", [(1, 3), (2, 2), (3, 1)])
@pytest.mark.parametrize("vi,vf", [("WEAK", "CENTRAL"), ("FD", "RANOCHA"), ("FD", "SHIMA_ETAL"), ("HG", "RANOCHA")])
def test_euler_free_stream_and_conservation(oracle, nd, level, vi, vf):
    O = oracle
    cfg = O.make_cfg(nd, O.EQ_EULER, 3, getattr(O, "VI_" + vi), getattr(O, "FLUX_" + vf), O.FLUX_LLF)
    mesh = O.Mesh(nd, level, -2.0, 2.0, 3)
    u = O.compute_coefficients(cfg, mesh, O.IC_CONSTANT)
    assert np.abs(O.rhs(cfg, mesh, u)).max() < 1e-13
    u = O.compute_coefficients(cfg, mesh, O.IC_SEEDED_RANDOM)
    du = O.rhs(cfg, mesh, u)
    scale = np.abs(du).max() * mesh.total_volume
    assert np.abs(O.integrate(cfg, mesh, du)).max() < 1e-13 * scale


@pytest.mark.parametrize("nd,level", [(1, 3), (2, 2), (3, 1)])
def test_euler_entropy_conservation_and_dissipation(oracle, nd, level):
    O = oracle
    mesh = O.Mesh(nd, level, -2.0, 2.0, 3)
    cfg = O.make_cfg(nd, O.EQ_EULER, 3, O.VI_FD, O.FLUX_RANOCHA, O.FLUX_RANOCHA)
    u = O.compute_coefficients(cfg, mesh, O.IC_SEEDED_RANDOM)
    du = O.rhs(cfg, mesh, u)
    assert abs(entropy_rate(O, cfg, mesh, u, du)) < 1e-13 * np.abs(du).max() * mesh.total_volume
    cfg = O.make_cfg(nd, O.EQ_EULER, 3, O.VI_FD, O.FLUX_RANOCHA, O.FLUX_LLF)
    du = O.rhs(cfg, mesh, u)
    assert entropy_rate(O, cfg, mesh, u, du) < -1e-6


def test_mhd_entropy_conservation_with_powell(oracle):
    O = oracle
    mesh = O.Mesh(3, 1, -1.0, 1.0, 3)
    cfg = O.make_cfg(3, O.EQ_GLM_MHD, 3, O.VI_FD, O.FLUX_HINDENLANG_GASSNER, O.FLUX_HINDENLANG_GASSNER,
                     gamma=5.0 / 3.0, c_h=1.0)
    u = O.compute_coefficients(cfg, mesh, O.IC_SEEDED_RANDOM)
    du = O.rhs(cfg, mesh, u)
    assert abs(entropy_rate(O, cfg, mesh, u, du)) < 1e-12 * np.abs(du).max() * mesh.total_volume
    cfg = O.make_cfg(3, O.EQ_GLM_MHD, 3, O.VI_FD, O.FLUX_HINDENLANG_GASSNER, O.FLUX_LLF,
                     gamma=5.0 / 3.0, c_h=1.0)
    du = O.rhs(cfg, mesh, u)
    assert entropy_rate(O, cfg, mesh, u, du) < -1e-6
    # free stream (constant state, constant field) is preserved
    u = O.compute_coefficients(cfg, mesh, O.IC_CONSTANT)
    assert np.abs(O.rhs(cfg, mesh, u)).max() < 1e-13


@pytest.mark.parametrize("vf", ["RANOCHA", "SHIMA_ETAL"])
def test_pressure_equilibrium_preservation(oracle, vf):
    O = oracle
    nd = 2
    mesh = O.Mesh(nd, 2, -2.0, 2.0, 3)
    k = getattr(O, "FLUX_" + vf)
    cfg = O.make_cfg(nd, O.EQ_EULER, 3, O.VI_FD, k, k)
    x = mesh.node_coordinates
    rho = 1.0 + 0.3 * np.sin(np.pi * x[..., 0] / 2) * np.cos(np.pi * x[..., 1] / 2)
    v = np.array([0.3, -0.2])
    p = 1.0
    u = np.stack([rho, rho * v[0], rho * v[1], p / 0.4 + 0.5 * rho * (v @ v)], axis=-1)
    du = O.rhs(cfg, mesh, u)
    # dp/dt = (gamma-1) (dE - v.dm + |v|^2/2 drho) must vanish
    dp = 0.4 * (du[..., 3] - v[0] * du[..., 1] - v[1] * du[..., 2] + 0.5 * (v @ v) * du[..., 0])
    assert np.abs(dp).max() < 1e-12 * np.abs(du).max()


@pytest.mark.parametrize("nd,level", [(1, 3), (2, 2), (3, 2)])
def test_linadv_polynomial_exactness(oracle, nd, level):
    O = oracle
    a = (1.0, -0.5, 0.25)[:nd]
    cfg = O.make_cfg(nd, O.EQ_LINADV, 3, O.VI_WEAK, O.FLUX_CENTRAL, O.FLUX_LLF, advection_velocity=a)
    mesh = O.Mesh(nd, level, -1.0, 1.0, 3)
    x = mesh.node_coordinates
    coef = [0.3, -0.7, 0.2]
    u = 1.0 + sum(coef[d] * x[..., d] ** 3 + 0.1 * x[..., d] ** 2 for d in range(nd))
    du = O.rhs(cfg, mesh, u[..., None])[..., 0]
    exact = -sum(a[d] * (3 * coef[d] * x[..., d] ** 2 + 0.2 * x[..., d]) for d in range(nd))
    n1 = 2**level
    interior = np.all((mesh.cell_index > 0) & (mesh.cell_index < n1 - 1), axis=1)
    assert interior.any()
    np.testing.assert_allclose(du[interior], exact[interior], atol=1e-12)


def test_c1_trajectory_matches_published_trixi_values(oracle):
    # Trixi.jl's elixir_advection_basic.jl (1D, polydeg 3, level 4, CFL 1.6,
    # t in [0, 1]) test values; recalled, used as a sanity check only.
    O = oracle
    cfg = O.make_cfg(1, O.EQ_LINADV, 3, O.VI_WEAK, surface_flux=O.FLUX_LLF, advection_velocity=(1.0,))
    mesh = O.Mesh(1, 4, -1.0, 1.0, 3)
    u = O.compute_coefficients(cfg, mesh, O.IC_CONVERGENCE_TEST)
    dt = 1.6 * O.max_dt(cfg, mesh, u)
    assert dt == pytest.approx(0.05, rel=1e-15)
    k = np.zeros_like(u)
    t = 0.0
    for _ in range(20):
        O.ck2n54_step(cfg, mesh, u, k, t, dt)
        t += dt
    l2, linf = O.error_norms(cfg, mesh, u, t, O.IC_CONVERGENCE_TEST)
    assert l2[0] == pytest.approx(6.0388296447998465e-6, rel=1e-8)
    assert linf[0] == pytest.approx(3.217887726258972e-5, rel=1e-8)


def test_ck2n54_is_fourth_order(oracle):
    O = oracle
    cfg = O.make_cfg(1, O.EQ_LINADV, 3, O.VI_WEAK, surface_flux=O.FLUX_LLF, advection_velocity=(1.0,))
    mesh = O.Mesh(1, 2, -1.0, 1.0, 3)
    u0 = O.compute_coefficients(cfg, mesh, O.IC_CONVERGENCE_TEST)
    n = u0.size
    A = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        A[:, j] = O.rhs(cfg, mesh, e.reshape(u0.shape)).ravel()
    T = 0.5
    exact = scipy.linalg.expm(A * T) @ u0.ravel()
    errs = []
    for steps in (20, 40, 80):
        u = u0.copy()
        k = np.zeros_like(u)
        dt = T / steps
        for s in range(steps):
            O.ck2n54_step(cfg, mesh, u, k, s * dt, dt)
        errs.append(np.abs(u.ravel() - exact).max())
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates > 3.8), rates


def test_hg_indicator_special_cases(oracle):
    O = oracle
    cfg = O.make_cfg(2, O.EQ_EULER, 3, O.VI_HG, O.FLUX_RANOCHA, O.FLUX_LLF, O.FLUX_LLF,
                     alpha_max=0.5, alpha_min=0.001, alpha_smooth=False)
    mesh = O.Mesh(2, 2, -2.0, 2.0, 3)
    u = O.compute_coefficients(cfg, mesh, O.IC_CONSTANT)
    assert np.all(O.indicator_hg(cfg, mesh, u) == 0.0)
    # put the highest Legendre mode (in x) into rho*p on element 5
    b = O.basis(4)
    V = np.linalg.inv(b["Vinv"])
    mode = V[:, 3]
    rho = u[..., 0].copy()
    p = 1.0 + 0.0 * rho
    rho[5] = 1.0 + 0.2 * mode[:, None] * np.ones((1, 4))  # varies in y only (axis 1 is j)
    u2 = np.stack([rho, 0 * rho, 0 * rho, p / 0.4], axis=-1)
    alpha = O.indicator_hg(cfg, mesh, u2)
    assert alpha[5] == pytest.approx(0.5)
    assert np.all(np.delete(alpha, 5) == 0.0)
    # smoothing: neighbours get max(own, alpha/2)
    cfg.alpha_smooth = 1
    alpha_s = O.indicator_hg(cfg, mesh, u2)
    nbrs = set()
    for (l, r) in mesh.neighbor_ids:
        if l == 5:
            nbrs.add(r)
        if r == 5:
            nbrs.add(l)
    for e in range(mesh.nelements):
        expect = 0.5 if e == 5 else (0.25 if e in nbrs else 0.0)
        assert alpha_s[e] == pytest.approx(expect)


def test_hg_rhs_reduces_to_fd_and_fv(oracle):
    O = oracle
    mesh = O.Mesh(2, 2, -2.0, 2.0, 3)
    cfg_fd = O.make_cfg(2, O.EQ_EULER, 3, O.VI_FD, O.FLUX_RANOCHA, O.FLUX_LLF)
    cfg_hg = O.make_cfg(2, O.EQ_EULER, 3, O.VI_HG, O.FLUX_RANOCHA, O.FLUX_LLF, O.FLUX_LLF,
                        alpha_max=0.0, alpha_min=0.001)
    u = O.compute_coefficients(cfg_fd, mesh, O.IC_SEEDED_RANDOM)
    # alpha_max = 0 -> pure DG
    np.testing.assert_allclose(O.rhs(cfg_hg, mesh, u), O.rhs(cfg_fd, mesh, u), rtol=0, atol=1e-14)
    # blend is affine in alpha: rhs(alpha) = (1-alpha) DG + alpha FV
    cfg_hg.alpha_max = 1.0
    cfg_hg.alpha_min = 0.0
    u2 = O.compute_coefficients(cfg_fd, mesh, O.IC_WEAK_BLAST_WAVE)
    du, alpha = O.rhs(cfg_hg, mesh, u2, return_alpha=True)
    assert alpha.max() > 0
    # a first-order FV volume term is still conservative
    assert np.abs(O.integrate(cfg_hg, mesh, du)).max() < 1e-12 * np.abs(du).max() * mesh.total_volume


def test_jvp_matches_central_differences(oracle):
    O = oracle
    mesh = O.Mesh(2, 2, -2.0, 2.0, 3)
    cfg = O.make_cfg(2, O.EQ_EULER, 3, O.VI_FD, O.FLUX_RANOCHA, O.FLUX_LLF)
    u = O.compute_coefficients(cfg, mesh, O.IC_SEEDED_RANDOM)
    rng = np.random.default_rng(6)
    w = rng.standard_normal(u.shape)
    du, jv = O.rhs_jvp(cfg, mesh, u, w)
    np.testing.assert_array_equal(du, O.rhs(cfg, mesh, u))
    h = 1e-6
    fd = (O.rhs(cfg, mesh, u + h * w) - O.rhs(cfg, mesh, u - h * w)) / (2 * h)
    assert np.abs(jv - fd).max() < 1e-6 * np.abs(jv).max()


def test_shock_capturing_jvp_matches_central_differences(oracle):
    """Dual-number derivative through the Hennemann-Gassner blending factor
    (modal energy, logistic, clipping, neighbour smoothing) and the DG/FV
    blend.  The seeded weak blast has no exact ties or zero velocities, where
    LLF's wave-speed max and |v| have kinks that central differences average."""
    O = oracle
    mesh = O.Mesh(2, 3, -2.0, 2.0, 3)
    cfg = O.make_cfg(2, O.EQ_EULER, 3, O.VI_HG, O.FLUX_RANOCHA, O.FLUX_LLF, O.FLUX_LLF)
    u = O.compute_coefficients(cfg, mesh, O.IC_WEAK_BLAST_WAVE_SEEDED)
    alpha = O.indicator_hg(cfg, mesh, u)
    assert ((alpha > 1e-3) & (alpha < 0.5)).sum() >= 4  # blending factors on the logistic branch
    rng = np.random.default_rng(16)
    w = rng.standard_normal(u.shape)
    _, jv = O.rhs_jvp(cfg, mesh, u, w)
    h = 1e-6
    fd = (O.rhs(cfg, mesh, u + h * w) - O.rhs(cfg, mesh, u - h * w)) / (2 * h)
    assert np.abs(jv - fd).max() < 1e-6 * np.abs(jv).max()
    # and the blending factor alone
    ja = (O.indicator_hg(cfg, mesh, u + h * w) - O.indicator_hg(cfg, mesh, u - h * w)) / (2 * h)
    assert np.abs(ja).max() > 1e-2


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_source_terms_manufacture_the_convergence_test(oracle, nd):
    """S = u_t + div f(u) at the convergence-test state (central differences
    in x and t of the initial condition and the physical flux)."""
    O = oracle
    cfg = O.make_cfg(nd, O.EQ_EULER, 3, O.VI_WEAK)
    rng = np.random.default_rng(20 + nd)
    h = 1e-5
    for _ in range(6):
        x, t = rng.uniform(0, 2, nd), rng.uniform(0, 1)
        ic = lambda xx, tt: O.initial_condition(cfg, O.IC_CONVERGENCE_TEST, xx, tt)
        ut = (ic(x, t + h) - ic(x, t - h)) / (2 * h)
        div = sum((O.flux(cfg, ic(x + h * np.eye(nd)[d], t), d + 1) - O.flux(cfg, ic(x - h * np.eye(nd)[d], t), d + 1))
                  / (2 * h) for d in range(nd))
        S = O.source_convergence_test(cfg, x, t)
        assert np.abs(S - (ut + div)).max() < 1e-8 * max(1.0, np.abs(S).max())


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_dirichlet_constant_state_is_preserved(oracle, nd):
    """Constant state inside and on the boundary: du = 0 (free stream through
    the boundary fluxes)."""
    O = oracle
    cfg = O.make_cfg(nd, O.EQ_EULER, 3, O.VI_FD, O.FLUX_RANOCHA, O.FLUX_LLF, boundary_condition=O.BC_DIRICHLET,
                     boundary_ic=O.IC_CONSTANT)
    mesh = O.Mesh(nd, 2, 0.0, 2.0, 3, periodic=False)
    u = O.compute_coefficients(cfg, mesh, O.IC_CONSTANT)
    assert np.abs(O.rhs(cfg, mesh, u)).max() < 1e-13


@pytest.mark.parametrize("nd,vi", [(1, "weak"), (2, "weak"), (2, "fd"), (3, "weak")])
def test_dirichlet_and_sources_converge_to_the_time_derivative(oracle, nd, vi):
    """With Dirichlet states and source_terms_convergence_test, rhs!(u_exact(t))
    approximates d/dt u_exact; the max-norm residual of polydeg 3 drops like
    h^3 on the non-periodic mesh (a wrong boundary flux or source leaves an
    O(1) residual that does not converge)."""
    O = oracle
    kind = O.VI_WEAK if vi == "weak" else O.VI_FD
    cfg = O.make_cfg(nd, O.EQ_EULER, 3, kind, O.FLUX_RANOCHA, O.FLUX_LLF, boundary_condition=O.BC_DIRICHLET,
                     boundary_ic=O.IC_CONVERGENCE_TEST, source_terms=O.SOURCE_CONVERGENCE_TEST, ic_length=2.0)
    errs = []
    t = 0.3
    for lvl in {1: (3, 4), 2: (2, 3), 3: (2, 3)}[nd]:
        mesh = O.Mesh(nd, lvl, 0.0, 2.0, 3, periodic=False)
        u = O.compute_coefficients(cfg, mesh, O.IC_CONVERGENCE_TEST, t)
        h = 1e-6
        ut = (O.compute_coefficients(cfg, mesh, O.IC_CONVERGENCE_TEST, t + h)
              - O.compute_coefficients(cfg, mesh, O.IC_CONVERGENCE_TEST, t - h)) / (2 * h)
        errs.append(np.abs(O.rhs(cfg, mesh, u, t=t) - ut).max())
    rate = np.log2(errs[0] / errs[1])
    assert rate > 2.5, (errs, rate)


def test_mesh_morton_order_and_connectivity(oracle):
    O = oracle
    for nd in (1, 2, 3):
        mesh = O.Mesh(nd, 3, -2.0, 2.0, 3)
        n1 = 8
        E = mesh.nelements
        ci = mesh.cell_index
        # Morton: element id equals the bit interleave of the cell coordinates
        code = np.zeros(E, dtype=np.int64)
        for lvl in range(3):
            for d in range(nd):
                code |= ((ci[:, d] >> lvl) & 1).astype(np.int64) << (lvl * nd + d)
        np.testing.assert_array_equal(code, np.arange(E))
        # interfaces: 3 per element in +x, +y, +z order; right = left + e_d (periodic)
        assert mesh.ninterfaces == nd * E
        for i, ((l, r), o) in enumerate(zip(mesh.neighbor_ids, mesh.orientations)):
            assert l == i // nd and o == i % nd + 1
            expect = ci[l].copy()
            expect[o - 1] = (expect[o - 1] + 1) % n1
            np.testing.assert_array_equal(ci[r], expect)
        dx = 4.0 / n1
        np.testing.assert_allclose(mesh.inverse_jacobian, 2.0 / dx)


@pytest.mark.parametrize("polydeg", [1, 2, 3, 4, 5, 7])
def test_mortar_operators_are_interpolation_and_l2_projection(oracle, polydeg):
    """forward_{lower,upper} interpolate the Lobatto polynomial to the lower /
    upper half (exact on monomials), reverse_{lower,upper} are the L2
    projection of the two halves back to one polynomial: Galerkin
    orthogonality against every Legendre polynomial of degree < N, integrated
    with a 2N-point Gauss rule of numpy (independent of the oracle)."""
    from numpy.polynomial import legendre as Lg
    O = oracle
    N = polydeg + 1
    fl, fu, rl, ru = O.mortar_matrices(polydeg)
    x = O.basis(N)["nodes"]
    V = np.vander(x, N, increasing=True)
    np.testing.assert_allclose(fl @ V, np.vander((x - 1) / 2, N, increasing=True), atol=1e-14)
    np.testing.assert_allclose(fu @ V, np.vander((x + 1) / 2, N, increasing=True), atol=1e-14)
    g, wg = Lg.leggauss(2 * N)

    def interp(vals, s):
        return np.vander(s, N, increasing=True) @ np.linalg.solve(V, vals)

    rng = np.random.default_rng(polydeg)
    for _ in range(3):
        f1, f2 = rng.standard_normal(N), rng.standard_normal(N)
        q = rl @ f1 + ru @ f2
        for k in range(N):
            phi = Lg.Legendre.basis(k)
            lhs = np.sum(wg * interp(q, g) * phi(g))
            rhs = 0.5 * np.sum(wg * interp(f1, g) * phi((g - 1) / 2)) + 0.5 * np.sum(wg * interp(f2, g) * phi((g + 1) / 2))
            assert abs(lhs - rhs) < 1e-13 * max(1.0, abs(rhs))
    # projection of an interpolated polynomial reproduces it
    np.testing.assert_allclose(rl @ fl + ru @ fu, np.eye(N), atol=1e-13)


# the reference's mortar meshes (test/tree_dgsem_2d/advection_mortar.jl:16-23,
# test/tree_dgsem_3d/advection_mortar.jl:16-25, test/tree_dgsem_3d/euler_mortar.jl:15-22)
MORTAR_MESHES = {
    "advection_mortar_2d": (2, 2, -1.0, 1.0, [((0.0, -1.0), (1.0, 1.0))]),
    "advection_mortar_3d": (3, 2, -1.0, 1.0, [((0.0, -1.0, -1.0), (1.0, 1.0, 1.0)), ((0.0, -0.5, -0.5), (0.5, 0.5, 0.5))]),
    "euler_mortar_3d": (3, 2, 0.0, 2.0, [((0.5,) * 3, (1.5,) * 3)]),
}


def test_refined_mesh_mortar_geometry(oracle):
    """Every mortar joins 2^(d-1) leaves of half the size that tile the large
    element's face, each small face touches the large face, and the face
    count balances: 2d faces per element = interfaces both sides + mortar
    faces + boundary faces."""
    O = oracle
    for name, (nd, lvl, a, b, patches) in MORTAR_MESHES.items():
        for per in (True, False):
            m = O.RefinedMesh(nd, lvl, a, b, 3, patches, periodic=per)
            K = 1 << (nd - 1)
            h = (b - a) / (1 << m.levels.astype(np.int64))
            ctr = m.node_coordinates.reshape(m.nelements, -1, nd).mean(axis=1)
            for ids, o, ls in zip(m.mortar_neighbor_ids, m.mortar_orientations, m.mortar_large_sides):
                large, small = ids[K], ids[:K]
                assert np.all(h[small] == h[large] / 2)
                d = o - 1
                sgn = 1 if ls == 1 else -1   # small elements on the + side of the large one
                for s in small:
                    gap = (ctr[s, d] - ctr[large, d]) * sgn
                    if per:
                        gap = (gap + (b - a) / 2) % (b - a) - (b - a) / 2
                    assert abs(gap - 0.75 * h[large]) < 1e-12
                    for t in range(nd):
                        if t != d:
                            assert abs(ctr[s, t] - ctr[large, t]) < h[large] / 2
                tc = np.delete(ctr[small], d, axis=1)
                assert len({tuple(np.round(r, 12)) for r in tc}) == K
            faces = 2 * m.ninterfaces + m.nmortars * (K + 1) + len(m.boundary_faces)
            assert faces == 2 * nd * m.nelements, name


@pytest.mark.parametrize("name", list(MORTAR_MESHES))
def test_mortar_free_stream_and_conservation(oracle, name):
    """Constant states stay constant across mortars, and the rhs integrates to
    zero over a periodic refined mesh (the L2 projection preserves the face
    integral of the small-face fluxes)."""
    O = oracle
    nd, lvl, a, b, patches = MORTAR_MESHES[name]
    mesh = O.RefinedMesh(nd, lvl, a, b, 3, patches)
    for eq, vi, vf in ((O.EQ_LINADV, O.VI_WEAK, O.FLUX_CENTRAL), (O.EQ_EULER, O.VI_WEAK, O.FLUX_CENTRAL),
                       (O.EQ_EULER, O.VI_FD, O.FLUX_RANOCHA)):
        cfg = O.make_cfg(nd, eq, 3, vi, vf, O.FLUX_LLF, advection_velocity=(0.2, -0.7, 0.5)[:nd])
        u = O.compute_coefficients(cfg, mesh, O.IC_CONSTANT)
        assert np.abs(O.rhs(cfg, mesh, u)).max() < 1e-12
        u = O.compute_coefficients(cfg, mesh, O.IC_SEEDED_RANDOM)
        du = O.rhs(cfg, mesh, u)
        assert np.abs(O.integrate(cfg, mesh, du)).max() < 1e-13 * np.abs(du).max() * mesh.total_volume


@pytest.mark.parametrize("nd,eq,vi,per,src", [(2, "linadv", "weak", True, False), (3, "linadv", "weak", True, False),
                                              (2, "euler", "fd", True, True), (2, "euler", "weak", False, True),
                                              (3, "euler", "fd", True, True)])
def test_mortar_meshes_converge_to_the_time_derivative(oracle, nd, eq, vi, per, src):
    """rhs!(u_exact(t)) on a mesh refined in [0.5, 1.5]^d approaches
    d/dt u_exact at the DG rate (polydeg 3: about h^3 in the max norm) when
    both levels are refined; a wrong mortar interpolation or projection (a
    half swapped, an operator transposed) leaves an O(1) residual at the
    mortars that does not converge."""
    O = oracle
    E = {"linadv": O.EQ_LINADV, "euler": O.EQ_EULER}[eq]
    kw = dict(ic_length=2.0, advection_velocity=(1.0, 0.5, 0.25)[:nd])
    if not per:
        kw.update(boundary_condition=O.BC_DIRICHLET, boundary_ic=O.IC_CONVERGENCE_TEST)
    if src:
        kw.update(source_terms=O.SOURCE_CONVERGENCE_TEST)
    cfg = O.make_cfg(nd, E, 3, O.VI_WEAK if vi == "weak" else O.VI_FD, O.FLUX_RANOCHA if vi == "fd" else O.FLUX_CENTRAL,
                     O.FLUX_LLF, **kw)
    errs = []
    t, h = 0.3, 1e-6
    for lvl in ((3, 4) if (nd == 2 and eq == "linadv") else (2, 3)):
        mesh = O.RefinedMesh(nd, lvl, 0.0, 2.0, 3, [((0.5,) * nd, (1.5,) * nd)], periodic=per)
        assert mesh.nmortars > 0
        u = O.compute_coefficients(cfg, mesh, O.IC_CONVERGENCE_TEST, t)
        ut = (O.compute_coefficients(cfg, mesh, O.IC_CONVERGENCE_TEST, t + h)
              - O.compute_coefficients(cfg, mesh, O.IC_CONVERGENCE_TEST, t - h)) / (2 * h)
        errs.append(np.abs(O.rhs(cfg, mesh, u, t=t) - ut).max())
    rate = np.log2(errs[0] / errs[1])
    assert rate > 2.5, (errs, rate)


def _prim2cons(rho, v, p, gamma):
    return np.concatenate([rho[..., None], rho[..., None] * v, (p / (gamma - 1) + 0.5 * rho * (v ** 2).sum(-1))[..., None]],
                          axis=-1)


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_indicator_variables_reduce_to_density_pressure(oracle, nd):
    """pressure and density as HG indicator variables: with rho = 1 the
    pressure indicator equals the density_pressure one, with p = 1 the
    density indicator does (rho p is then p, resp. rho); and on a state whose
    density carries the top mode while the pressure is smooth, the density
    indicator fires where the pressure one does not."""
    O = oracle
    gamma = 1.4
    mesh = O.Mesh(nd, {1: 4, 2: 2, 3: 1}[nd], -1.0, 1.0, 3)
    shape = mesh.node_coordinates.shape[:-1]
    rng = np.random.default_rng(nd)
    mk = lambda var: O.make_cfg(nd, O.EQ_EULER, 3, O.VI_HG, O.FLUX_RANOCHA, O.FLUX_LLF, O.FLUX_LLF, gamma=gamma,
                                alpha_smooth=False, indicator_variable=var)
    v = 0.3 * rng.standard_normal(shape + (nd,))
    field = 1.0 + 0.3 * rng.uniform(-1, 1, shape)
    one = np.ones(shape)
    u = _prim2cons(one, v, field, gamma)      # rho = 1
    a_dp = O.indicator_hg(mk(O.IND_DENSITY_PRESSURE), mesh, u)
    assert a_dp.max() > 0.1
    np.testing.assert_allclose(O.indicator_hg(mk(O.IND_PRESSURE), mesh, u), a_dp, rtol=1e-12, atol=1e-14)
    u = _prim2cons(field, v, one, gamma)      # p = 1
    a_dp = O.indicator_hg(mk(O.IND_DENSITY_PRESSURE), mesh, u)
    np.testing.assert_allclose(O.indicator_hg(mk(O.IND_DENSITY), mesh, u), a_dp, rtol=1e-12, atol=1e-14)
    # top Legendre mode in rho only (element 0), p constant
    V = np.linalg.inv(O.basis(4)["Vinv"])
    rho = np.ones(shape)
    rho[0] = 1.0 + 0.2 * V[:, 3].reshape((1,) * (nd - 1) + (4,))
    u = _prim2cons(rho, 0 * v, one, gamma)
    assert O.indicator_hg(mk(O.IND_DENSITY), mesh, u)[0] > 0.1
    assert np.all(O.indicator_hg(mk(O.IND_PRESSURE), mesh, u) == 0.0)


def test_hg_smoothing_crosses_mortars(oracle):
    """Blending-factor smoothing over interfaces and mortars: a small element
    takes half of its large neighbour's factor and the large one half of each
    small neighbour's, as across an interface."""
    O = oracle
    nd, lvl, a, b, patches = MORTAR_MESHES["advection_mortar_2d"]
    mesh = O.RefinedMesh(nd, lvl, a, b, 3, patches)
    cfg = O.make_cfg(2, O.EQ_EULER, 3, O.VI_HG, O.FLUX_RANOCHA, O.FLUX_LLF, O.FLUX_LLF, alpha_smooth=False)
    u = O.compute_coefficients(cfg, mesh, O.IC_CONSTANT)
    V = np.linalg.inv(O.basis(4)["Vinv"])
    m = 0
    ids = mesh.mortar_neighbor_ids[m]
    large, small = ids[2], ids[:2]
    rho = u[..., 0].copy()
    rho[large] = 1.0 + 0.2 * V[:, 3][:, None] * np.ones((1, 4))
    u2 = u.copy()
    u2[..., 0] = rho
    raw = O.indicator_hg(cfg, mesh, u2)
    assert raw[large] > 0.1 and np.count_nonzero(raw) == 1
    cfg.alpha_smooth = 1
    sm = O.indicator_hg(cfg, mesh, u2)
    for s in small:
        assert sm[s] == pytest.approx(0.5 * raw[large])
    # the other direction: the small element's factor reaches the large one
    u3 = u.copy()
    rho = u[..., 0].copy()
    rho[small[0]] = 1.0 + 0.2 * V[:, 3][:, None] * np.ones((1, 4))
    u3[..., 0] = rho
    cfg.alpha_smooth = 0
    raw = O.indicator_hg(cfg, mesh, u3)
    cfg.alpha_smooth = 1
    assert O.indicator_hg(cfg, mesh, u3)[large] == pytest.approx(0.5 * raw[small[0]])


def test_hg_on_mortar_meshes_is_conservative(oracle):
    O = oracle
    for name in ("advection_mortar_2d", "euler_mortar_3d"):
        nd, lvl, a, b, patches = MORTAR_MESHES[name]
        mesh = O.RefinedMesh(nd, lvl, a, b, 3, patches)
        cfg = O.make_cfg(nd, O.EQ_EULER, 3, O.VI_HG, O.FLUX_RANOCHA, O.FLUX_LLF, O.FLUX_LLF, alpha_min=0.0,
                         indicator_variable=O.IND_PRESSURE)
        u = O.compute_coefficients(cfg, mesh, O.IC_CONSTANT)
        assert np.abs(O.rhs(cfg, mesh, u)).max() < 1e-12
        u = O.compute_coefficients(cfg, mesh, O.IC_WEAK_BLAST_WAVE_SEEDED)
        du, alpha = O.rhs(cfg, mesh, u, return_alpha=True)
        assert alpha.max() > 0.05
        assert np.abs(O.integrate(cfg, mesh, du)).max() < 1e-13 * np.abs(du).max() * mesh.total_volume


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_mhd_alfven_wave_is_an_exact_solution(oracle, nd):
    """The MHD convergence-test states (1D: circularly polarised Alfven wave
    on [0, 1], no cleaning field; 2D: Derigs et al.'s Alfven wave along the
    diagonal of [0, sqrt 2]^2; 3D: Altmann's wave on [-1, 1]^3)
    solve the equations with the Powell terms: rhs!(u_exact(t)) - u_t drops
    at the DG rate.  A wrong sign of t in the wave or a flux term that breaks
    the 2D reduction leaves an O(1) residual."""
    O = oracle
    lo, hi = {1: (0.0, 1.0), 2: (0.0, np.sqrt(2.0)), 3: (-1.0, 1.0)}[nd]
    cfg = O.make_cfg(nd, O.EQ_GLM_MHD, 3, O.VI_FD, O.FLUX_HINDENLANG_GASSNER, O.FLUX_LLF, gamma=5 / 3, c_h=1.0)
    errs = []
    t, h = 0.1, 1e-6
    for lvl in ((3, 4) if nd < 3 else (2, 3)):
        mesh = O.Mesh(nd, lvl, lo, hi, 3)
        u = O.compute_coefficients(cfg, mesh, O.IC_CONVERGENCE_TEST, t)
        ut = (O.compute_coefficients(cfg, mesh, O.IC_CONVERGENCE_TEST, t + h)
              - O.compute_coefficients(cfg, mesh, O.IC_CONVERGENCE_TEST, t - h)) / (2 * h)
        assert np.abs(ut).max() > 0.5
        errs.append(np.abs(O.rhs(cfg, mesh, u, t=t) - ut).max())
    rate = np.log2(errs[0] / errs[1])
    assert rate > 2.5, (errs, rate)


def test_mhd2d_entropy_conservation_with_powell(oracle):
    O = oracle
    mesh = O.Mesh(2, 2, -1.0, 1.0, 3)
    cfg = O.make_cfg(2, O.EQ_GLM_MHD, 3, O.VI_FD, O.FLUX_HINDENLANG_GASSNER, O.FLUX_HINDENLANG_GASSNER,
                     gamma=1.4, c_h=1.0)
    u = O.compute_coefficients(cfg, mesh, O.IC_SEEDED_RANDOM)
    du = O.rhs(cfg, mesh, u)
    assert abs(entropy_rate(O, cfg, mesh, u, du)) < 1e-12 * np.abs(du).max() * mesh.total_volume
    cfg = O.make_cfg(2, O.EQ_GLM_MHD, 3, O.VI_FD, O.FLUX_HINDENLANG_GASSNER, O.FLUX_LLF, gamma=1.4, c_h=1.0)
    du = O.rhs(cfg, mesh, u)
    assert entropy_rate(O, cfg, mesh, u, du) < -1e-6
    # a constant state with constant B stays constant
    u = O.compute_coefficients(cfg, mesh, O.IC_CONSTANT)
    assert np.abs(O.rhs(cfg, mesh, u)).max() < 1e-13


def test_burgers_source_terms_manufacture_the_convergence_test(oracle):
    O = oracle
    cfg = O.make_cfg(1, O.EQ_BURGERS, 3)
    rng = np.random.default_rng(5)
    h = 1e-5
    ic = lambda xx, tt: O.initial_condition(cfg, O.IC_CONVERGENCE_TEST, np.array([xx]), tt)
    for _ in range(8):
        x, t = rng.uniform(0, 1), rng.uniform(0, 1)
        ut = (ic(x, t + h) - ic(x, t - h)) / (2 * h)
        fx = (O.flux(cfg, ic(x + h, t), 1) - O.flux(cfg, ic(x - h, t), 1)) / (2 * h)
        S = O.source_convergence_test(cfg, np.array([x]), t)
        assert abs(S[0] - (ut + fx)[0]) < 1e-7 * max(1.0, abs(S[0]))


@pytest.mark.parametrize("periodic,vi", [(True, "weak"), (True, "fd"), (False, "weak"), (False, "fd")])
def test_burgers_converges_to_the_time_derivative(oracle, periodic, vi):
    """rhs!(u_exact(t)) - u_t -> 0 at the DG rate for the Burgers convergence
    test with its source, periodic or with the exact state as inflow
    (Dirichlet, x_neg) and the outflow condition flux(u_inner) on x_pos (the
    wave moves right: u >= 1)."""
    O = oracle
    kw = dict(source_terms=O.SOURCE_CONVERGENCE_TEST)
    if not periodic:
        kw.update(boundary_condition=O.BC_DIRICHLET, boundary_ic=O.IC_CONVERGENCE_TEST, boundary_outflow_faces=2)
    cfg = O.make_cfg(1, O.EQ_BURGERS, 3, O.VI_WEAK if vi == "weak" else O.VI_FD, O.FLUX_EC, O.FLUX_LLF, **kw)
    errs = []
    t, h = 0.3, 1e-6
    for lvl in (4, 5):
        mesh = O.Mesh(1, lvl, 0.0, 1.0, 3, periodic=periodic)
        u = O.compute_coefficients(cfg, mesh, O.IC_CONVERGENCE_TEST, t)
        ut = (O.compute_coefficients(cfg, mesh, O.IC_CONVERGENCE_TEST, t + h)
              - O.compute_coefficients(cfg, mesh, O.IC_CONVERGENCE_TEST, t - h)) / (2 * h)
        errs.append(np.abs(O.rhs(cfg, mesh, u, t=t) - ut).max())
    rate = np.log2(errs[0] / errs[1])
    assert rate > 2.5, (errs, rate)


def test_burgers_entropy_conservation_and_outflow(oracle):
    O = oracle
    mesh = O.Mesh(1, 4, 0.0, 1.0, 3)
    cfg = O.make_cfg(1, O.EQ_BURGERS, 3, O.VI_FD, O.FLUX_EC, O.FLUX_EC)
    u = O.compute_coefficients(cfg, mesh, O.IC_SEEDED_RANDOM)
    du = O.rhs(cfg, mesh, u)
    assert abs(entropy_rate(O, cfg, mesh, u, du)) < 1e-13 * np.abs(du).max()
    cfg = O.make_cfg(1, O.EQ_BURGERS, 3, O.VI_FD, O.FLUX_EC, O.FLUX_LLF)
    du = O.rhs(cfg, mesh, u)
    assert entropy_rate(O, cfg, mesh, u, du) < -1e-6
    # outflow on both sides: the total changes only by the boundary fluxes
    # f(u) = u^2/2 of the end nodes, d/dt int u = u(0)^2/2 - u(1)^2/2
    mesh = O.Mesh(1, 4, 0.0, 1.0, 3, periodic=False)
    cfg = O.make_cfg(1, O.EQ_BURGERS, 3, O.VI_FD, O.FLUX_EC, O.FLUX_LLF, boundary_condition=O.BC_DIRICHLET,
                     boundary_outflow_faces=3)
    du = O.rhs(cfg, mesh, u)
    expect = 0.5 * u[0, 0, 0] ** 2 - 0.5 * u[-1, -1, 0] ** 2
    assert O.integrate(cfg, mesh, du)[0] == pytest.approx(expect, rel=1e-12)


def test_mhd1d_entropy_conservation(oracle):
    """1D ideal MHD (eight variables, no Powell terms): with B1 constant the
    EC flux conserves entropy and LLF dissipates it; constant states stay."""
    O = oracle
    mesh = O.Mesh(1, 4, 0.0, 1.0, 3)
    cfg = O.make_cfg(1, O.EQ_GLM_MHD, 3, O.VI_FD, O.FLUX_HINDENLANG_GASSNER, O.FLUX_HINDENLANG_GASSNER, gamma=5 / 3)
    assert O.nvars(cfg) == 8
    u = O.compute_coefficients(cfg, mesh, O.IC_SEEDED_RANDOM)
    u[..., 4] += 0.5 * (1.0 - u[..., 5] ** 2)  # B1 = 1 at the same pressure
    u[..., 5] = 1.0
    du = O.rhs(cfg, mesh, u)
    assert abs(entropy_rate(O, cfg, mesh, u, du)) < 1e-12 * np.abs(du).max()
    cfg = O.make_cfg(1, O.EQ_GLM_MHD, 3, O.VI_FD, O.FLUX_HINDENLANG_GASSNER, O.FLUX_LLF, gamma=5 / 3)
    du = O.rhs(cfg, mesh, u)
    assert entropy_rate(O, cfg, mesh, u, du) < -1e-6
    u = O.compute_coefficients(cfg, mesh, O.IC_CONSTANT)
    assert np.abs(O.rhs(cfg, mesh, u)).max() < 1e-13


@pytest.mark.parametrize("nd", [2, 3])
def test_eoc_test_euler_source_is_manufactured(oracle, nd):
    """source_terms_eoc_test_euler = u_t + div f(u) at the coupled
    Euler-gravity EOC state (gamma = 2; central differences), and with it
    rhs!(u_exact) - u_t drops at the DG rate on the weak form with flux_hll
    (test/tree_dgsem_3d/euler_convergence.jl)."""
    O = oracle
    cfg = O.make_cfg(nd, O.EQ_EULER, 3, O.VI_WEAK, O.FLUX_CENTRAL, O.FLUX_HLL, gamma=2.0,
                     source_terms=O.SOURCE_EOC_TEST_EULER)
    rng = np.random.default_rng(60 + nd)
    h = 1e-5
    ic = lambda xx, tt: O.initial_condition(cfg, O.IC_EOC_GRAVITY, xx, tt)
    for _ in range(6):
        x, t = rng.uniform(0, 2, nd), rng.uniform(0, 1)
        ut = (ic(x, t + h) - ic(x, t - h)) / (2 * h)
        div = sum((O.flux(cfg, ic(x + h * np.eye(nd)[d], t), d + 1) - O.flux(cfg, ic(x - h * np.eye(nd)[d], t), d + 1))
                  / (2 * h) for d in range(nd))
        S = O.source_convergence_test(cfg, x, t)
        assert np.abs(S - (ut + div)).max() < 1e-8 * max(1.0, np.abs(S).max())
    errs = []
    t, hh = 0.3, 1e-6
    for lvl in (2, 3):
        mesh = O.Mesh(nd, lvl, 0.0, 2.0, 3)
        u = O.compute_coefficients(cfg, mesh, O.IC_EOC_GRAVITY, t)
        ut = (O.compute_coefficients(cfg, mesh, O.IC_EOC_GRAVITY, t + hh)
              - O.compute_coefficients(cfg, mesh, O.IC_EOC_GRAVITY, t - hh)) / (2 * hh)
        errs.append(np.abs(O.rhs(cfg, mesh, u, t=t) - ut).max())
    rate = np.log2(errs[0] / errs[1])
    assert rate > 2.5, (errs, rate)


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_hypdiff_poisson_steady_state_is_an_equilibrium(oracle, nd):
    """Hyperbolic diffusion with source_terms_poisson_nonperiodic and the
    steady solution on the x boundaries (periodic in the other directions):
    rhs!(u_steady) -> 0 at the DG rate (polydeg 4: about h^4); a wrong sign
    in the flux, the relaxation or f leaves an O(1) residual."""
    O = oracle
    per = (False,) + (True,) * (nd - 1)
    cfg = O.make_cfg(nd, O.EQ_HYPDIFF, 4, O.VI_WEAK, O.FLUX_CENTRAL, O.FLUX_LLF,
                     boundary_condition=O.BC_DIRICHLET, boundary_ic=O.IC_POISSON_NONPERIODIC,
                     source_terms=O.SOURCE_POISSON_NONPERIODIC)
    errs = []
    for lvl in ((3, 4) if nd < 3 else (2, 3)):
        mesh = O.Mesh(nd, lvl, 0.0, 1.0, 4, periodic=per)
        u = O.compute_coefficients(cfg, mesh, O.IC_POISSON_NONPERIODIC, 1.0)  # the steady state (t != 0)
        errs.append(np.abs(O.rhs(cfg, mesh, u, t=0.0)).max())
    rate = np.log2(errs[0] / errs[1])
    assert rate > 3.3, (errs, rate)


def test_hypdiff_harmonic_and_llf(oracle):
    """The harmonic state phi = 3 + e x, q = e is an exact equilibrium of
    hyperbolic diffusion (nu = 1.25) with source_terms_harmonic and itself as
    the Dirichlet state; LLF's wave speed is the constant sqrt(nu / T_r) = nu / Lr."""
    O = oracle
    cfg = O.make_cfg(1, O.EQ_HYPDIFF, 3, O.VI_WEAK, O.FLUX_CENTRAL, O.FLUX_LLF, nu=1.25,
                     boundary_condition=O.BC_DIRICHLET, boundary_ic=O.IC_HARMONIC_NONPERIODIC,
                     source_terms=O.SOURCE_HARMONIC)
    mesh = O.Mesh(1, 2, -1.0, 2.0, 3, periodic=False)
    u = O.compute_coefficients(cfg, mesh, O.IC_HARMONIC_NONPERIODIC, 0.5)
    assert np.abs(O.rhs(cfg, mesh, u, t=0.5)).max() < 1e-10  # terms of size phi / T_r * invJ ~ 1e3
    ul, ur = np.array([0.3, -0.2]), np.array([1.1, 0.4])
    lam = 1.25 * 2 * np.pi
    fl, fr = O.flux(cfg, ul, 1), O.flux(cfg, ur, 1)
    np.testing.assert_allclose(fl, [-1.25 * -0.2, -0.3 * lam ** 2 / 1.25], rtol=1e-14)
    np.testing.assert_allclose(O.two_point_flux(cfg, O.FLUX_LLF, ul, ur, 1), 0.5 * (fl + fr) - 0.5 * lam * (ur - ul),
                               rtol=1e-14)


@pytest.mark.parametrize("nd", [2, 3])
def test_mhd_shock_capturing_entropy_and_limits(oracle, nd):
    """GLM-MHD shock capturing (test/tree_dgsem_{2,3}d/mhd_shock.jl): with the EC
    flux + Powell terms as DG, FV (subcell interfaces f + nc(u_l, u_r)/2 on the
    left, f + nc(u_r, u_l)/2 on the right) and surface flux, the blended scheme
    conserves entropy at any blending factor; alpha_max = 0 reduces it to flux
    differencing; LLF dissipates."""
    O = oracle
    mesh = O.Mesh(nd, 2 if nd == 2 else 1, -1.0, 1.0, 3)
    HG = O.FLUX_HINDENLANG_GASSNER
    cfg = O.make_cfg(nd, O.EQ_GLM_MHD, 3, O.VI_HG, HG, HG, HG, gamma=1.4, c_h=1.0, alpha_min=0.0)
    u = O.compute_coefficients(cfg, mesh, O.IC_WEAK_BLAST_WAVE_SEEDED)
    du, alpha = O.rhs(cfg, mesh, u, return_alpha=True)
    assert alpha.max() > 0.01
    assert abs(entropy_rate(O, cfg, mesh, u, du)) < 1e-12 * np.abs(du).max() * mesh.total_volume
    cfg0 = O.make_cfg(nd, O.EQ_GLM_MHD, 3, O.VI_HG, HG, HG, HG, gamma=1.4, c_h=1.0, alpha_max=0.0)
    cfg_fd = O.make_cfg(nd, O.EQ_GLM_MHD, 3, O.VI_FD, HG, HG, gamma=1.4, c_h=1.0)
    np.testing.assert_allclose(O.rhs(cfg0, mesh, u), O.rhs(cfg_fd, mesh, u), rtol=0, atol=1e-12 * np.abs(du).max())
    cfg = O.make_cfg(nd, O.EQ_GLM_MHD, 3, O.VI_HG, HG, O.FLUX_LLF, O.FLUX_LLF, gamma=1.4, c_h=1.0, alpha_min=0.0)
    du = O.rhs(cfg, mesh, u)
    assert entropy_rate(O, cfg, mesh, u, du) < -1e-6


@pytest.mark.parametrize("nd,level", [(1, 3), (2, 2)])
def test_multicomponent_euler_invariants(oracle, nd, level):
    """Multicomponent Euler (test/tree_dgsem_{1,2}d/eulermulti_{ec,es}.jl):
    free stream, conservation of each partial mass, momentum and energy,
    entropy conservation with flux_ranocha throughout (equal gases, entropy
    -sum rho_k (c_v log T - R log rho_k) with complex-step entropy variables)
    and dissipation with an LLF surface flux."""
    from cases import multi_entropy_vars
    O = oracle
    mesh = O.Mesh(nd, level, -2.0, 2.0, 3)
    gas = ((1.4, 1.4), (0.4, 0.4))
    kw = dict(gammas=gas[0], gas_constants=gas[1])
    cfg = O.make_cfg(nd, O.EQ_EULER_MULTI, 3, O.VI_FD, O.FLUX_RANOCHA, O.FLUX_RANOCHA, **kw)
    u = O.compute_coefficients(cfg, mesh, O.IC_CONSTANT)
    assert np.abs(O.rhs(cfg, mesh, u)).max() < 1e-13
    u = O.compute_coefficients(cfg, mesh, O.IC_SEEDED_RANDOM)
    nv = nd + 3

    def entropy_rate(du):
        w = np.apply_along_axis(lambda x: multi_entropy_vars(x, nd, *gas), -1, u.reshape(-1, nv)).reshape(u.shape)
        return O.integrate(cfg, mesh, np.repeat((w * du).sum(-1, keepdims=True), nv, axis=-1))[0]

    for sflux, vi in ((O.FLUX_RANOCHA, O.VI_FD), (O.FLUX_LLF, O.VI_FD), (O.FLUX_LLF, O.VI_WEAK)):
        c = O.make_cfg(nd, O.EQ_EULER_MULTI, 3, vi, O.FLUX_RANOCHA, sflux, **kw)
        du = O.rhs(c, mesh, u)
        scale = np.abs(du).max() * mesh.total_volume
        assert np.abs(O.integrate(c, mesh, du)).max() < 1e-13 * scale
        if vi == O.VI_FD and sflux == O.FLUX_RANOCHA:
            assert abs(entropy_rate(du)) < 1e-13 * scale
        elif vi == O.VI_FD:
            assert entropy_rate(du) < -1e-6


@pytest.mark.parametrize("nd", [1, 2])
def test_multicomponent_weak_blast_state(oracle, nd):
    """Trixi's multicomponent weak blast: partial densities 2^(k-1) / (2^n - 1)
    of the Euler blast's density, the same velocity and pressure (checked by
    the Euler flux of the summed state), and the single-gas case equal to the
    Euler state."""
    O = oracle
    mesh = O.Mesh(nd, 3, -2.0, 2.0, 3)
    ce = O.make_cfg(nd, O.EQ_EULER, gamma=1.4)
    ue = O.compute_coefficients(ce, mesh, O.IC_WEAK_BLAST_WAVE)
    for gas in (((1.4, 1.4), (0.4, 0.4)), ((1.4,), (0.4,))):
        cm = O.make_cfg(nd, O.EQ_EULER_MULTI, gammas=gas[0], gas_constants=gas[1])
        um = O.compute_coefficients(cm, mesh, O.IC_WEAK_BLAST_WAVE)
        n = len(gas[0])
        frac = np.array([2.0 ** k / (2.0 ** n - 1) for k in range(n)])
        np.testing.assert_allclose(um[..., nd + 1:], ue[..., :1] * frac, rtol=1e-15)
        np.testing.assert_allclose(um[..., :nd + 1], ue[..., 1:], rtol=1e-14, atol=1e-15)


# ---------------------------------------------------------------------------
# quasi-1D compressible Euler: (a rho, a rho v, a e, a)
# ---------------------------------------------------------------------------
def quasi_entropy(u, gamma=1.4):
    """a U_Euler(u / a) with U_Euler = -rho s / (gamma - 1), s = log p - gamma log rho."""
    a = u[3]
    rho, rv, re = u[0] / a, u[1] / a, u[2] / a
    p = (gamma - 1) * (re - 0.5 * rv * rv / rho)
    return -a * rho * (np.log(p) - gamma * np.log(rho)) / (gamma - 1)


@pytest.mark.parametrize("N", [4, 5])
def test_quasi1d_reduces_to_euler_at_unit_width(oracle, N):
    """With a = 1 the quasi-1D scheme (flux_chan_etal + a_ll p_rr in the
    nonconservative form with Trixi's factor 1/2) is the 1D Euler flux
    differencing scheme with flux_ranocha, node for node; a stays constant.
    A nonconservative term with a mean pressure, or without the factor 1/2,
    fails this."""
    O = oracle
    mesh = O.Mesh(1, 3, -1.0, 1.0, N - 1)
    cq = O.make_cfg(1, O.EQ_EULER_QUASI1D, N - 1, O.VI_FD, O.FLUX_CHAN_ETAL, O.FLUX_CHAN_ETAL)
    ce = O.make_cfg(1, O.EQ_EULER, N - 1, O.VI_FD, O.FLUX_RANOCHA, O.FLUX_RANOCHA)
    ue = O.compute_coefficients(ce, mesh, O.IC_SEEDED_RANDOM)
    uq = np.concatenate([ue, np.ones(ue.shape[:-1] + (1,))], axis=-1)
    dq, de = O.rhs(cq, mesh, uq), O.rhs(ce, mesh, ue)
    assert np.abs(dq[..., :3] - de).max() < 1e-13 * np.abs(de).max()
    assert np.abs(dq[..., 3]).max() == 0.0


def test_quasi1d_flux_consistency_and_entropy_conservation(oracle):
    """flux_chan_etal is consistent (f*(u, u) = (a rho v, a rho v^2, a v (e + p), 0))
    and symmetric, and the semidiscretisation with it conserves the entropy
    a U_Euler(u / a) on a periodic mesh with a varying width (entropy
    variables by complex step)."""
    O = oracle
    cfg = O.make_cfg(1, O.EQ_EULER_QUASI1D, 4, O.VI_FD, O.FLUX_CHAN_ETAL, O.FLUX_CHAN_ETAL)
    rng = np.random.default_rng(30)
    for _ in range(20):
        ul = np.array([1.2, 0.3, 3.0, 1.5]) * (1 + 0.2 * rng.random(4))
        ur = np.array([1.0, -0.2, 2.5, 2.0]) * (1 + 0.2 * rng.random(4))
        np.testing.assert_allclose(O.two_point_flux(cfg, O.FLUX_CHAN_ETAL, ul, ul, 1), O.flux(cfg, ul, 1),
                                   rtol=1e-14, atol=1e-15)
        fs = O.two_point_flux(cfg, O.FLUX_CHAN_ETAL, ul, ur, 1)
        # (the energy component cancels terms of size one: absolute round-off)
        np.testing.assert_allclose(fs, O.two_point_flux(cfg, O.FLUX_CHAN_ETAL, ur, ul, 1), rtol=1e-13, atol=1e-15)
    mesh = O.Mesh(1, 3, -1.0, 1.0, 4)
    u = O.compute_coefficients(cfg, mesh, O.IC_SEEDED_RANDOM)
    du = O.rhs(cfg, mesh, u)
    assert np.abs(du[..., 3]).max() == 0.0
    h = 1e-30
    w = np.zeros_like(u)
    for j in range(4):
        uc = u.astype(complex)
        uc[..., j] += 1j * h
        w[..., j] = quasi_entropy(np.moveaxis(uc, -1, 0)).imag / h
    rate = O.integrate(cfg, mesh, np.repeat((w * du).sum(-1, keepdims=True), 4, axis=-1))[0]
    assert abs(rate) < 1e-13 * np.abs(du).max() * mesh.total_volume, rate


def test_quasi1d_source_and_convergence(oracle):
    """source_terms_convergence_test equals the residual of the manufactured
    state (sympy derivatives of the quasi-1D equations), and with it
    rhs!(u_exact(t)) approximates d/dt u_exact at the DG rate (polydeg 4)."""
    import sympy as sp
    O = oracle
    gamma = 1.4
    cfg = O.make_cfg(1, O.EQ_EULER_QUASI1D, 4, O.VI_FD, O.FLUX_CHAN_ETAL, O.FLUX_CHAN_ETAL,
                     source_terms=O.SOURCE_CONVERGENCE_TEST, gamma=gamma)
    X, T = sp.symbols("x t")
    ini = 2 + sp.Rational(1, 10) * sp.sin(sp.pi * (X - T))
    rho, v, a = ini, 1, sp.Rational(3, 2) - sp.cos(sp.pi * X) / 2
    e = ini ** 2 / rho
    p = (gamma - 1) * (e - rho * v ** 2 / 2)
    S = [sp.diff(a * rho, T) + sp.diff(a * rho * v, X),
         sp.diff(a * rho * v, T) + sp.diff(a * rho * v * v, X) + a * sp.diff(p, X),
         sp.diff(a * e, T) + sp.diff(a * v * (e + p), X)]
    Sf = sp.lambdify((X, T), S, "numpy")
    rng = np.random.default_rng(31)
    for _ in range(8):
        x, t = rng.uniform(-1, 1), rng.uniform(0, 1)
        s = O.source_convergence_test(cfg, np.array([x]), t)
        np.testing.assert_allclose(s[:3], np.array(Sf(x, t), dtype=float), rtol=1e-13, atol=1e-14)
        assert s[3] == 0.0
    errs = []
    t = 0.3
    for lvl in (3, 4):
        mesh = O.Mesh(1, lvl, -1.0, 1.0, 4)
        u = O.compute_coefficients(cfg, mesh, O.IC_CONVERGENCE_TEST, t)
        h = 1e-6
        ut = (O.compute_coefficients(cfg, mesh, O.IC_CONVERGENCE_TEST, t + h)
              - O.compute_coefficients(cfg, mesh, O.IC_CONVERGENCE_TEST, t - h)) / (2 * h)
        errs.append(np.abs(O.rhs(cfg, mesh, u, t=t) - ut).max())
    assert np.log2(errs[0] / errs[1]) > 3.5, errs


@pytest.mark.parametrize("nd", [2, 3])
def test_mhd_mortar_free_stream_and_side_order(oracle, nd):
    """GLM-MHD on refined meshes (test/tree_dgsem_{2,3}d/mhd_alfven_wave_mortar.jl):
    a constant state stays constant through the mortars (Powell terms
    included); mass, momentum-free components and B stay conservative where
    the nonconservative terms vanish (rho); and the Alfven wave is
    approximated at the DG rate across the refinement.  The side order of
    the nonconservative mortar terms (small faces 1/2 nc(u_large, u_small))
    follows the reference's mortar_flux_kernel! text."""
    O = oracle
    patches = (((0.25 * 2 ** 0.5,) * nd, (0.75 * 2 ** 0.5,) * nd),)
    for sflux in (O.FLUX_HLLE, O.FLUX_LLF):
        cfg = O.make_cfg(nd, O.EQ_GLM_MHD, 3, O.VI_FD, O.FLUX_HINDENLANG_GASSNER, sflux, gamma=5 / 3)
        mesh = O.RefinedMesh(nd, 2, 0.0, 2 ** 0.5, 3, patches)
        assert mesh.nmortars > 0
        u = O.compute_coefficients(cfg, mesh, O.IC_CONSTANT)
        u[..., 5:8] = (0.3, -0.2, 0.1)
        u[..., 4] += 0.5 * (0.09 + 0.04 + 0.01)
        assert np.abs(O.rhs(cfg, mesh, u)).max() < 1e-13
        u = O.compute_coefficients(cfg, mesh, O.IC_SEEDED_RANDOM)
        du = O.rhs(cfg, mesh, u)
        # the density equation has no nonconservative term
        assert abs(O.integrate(cfg, mesh, du)[0]) < 1e-13 * np.abs(du).max() * mesh.total_volume
        if nd == 2:
            errs = []
            t = 0.1
            for lvl in (2, 3):
                mesh = O.RefinedMesh(nd, lvl, 0.0, 2 ** 0.5, 3, patches)
                u = O.compute_coefficients(cfg, mesh, O.IC_CONVERGENCE_TEST, t)
                h = 1e-6
                ut = (O.compute_coefficients(cfg, mesh, O.IC_CONVERGENCE_TEST, t + h)
                      - O.compute_coefficients(cfg, mesh, O.IC_CONVERGENCE_TEST, t - h)) / (2 * h)
                errs.append(np.abs(O.rhs(cfg, mesh, u, t=t) - ut).max())
            assert np.log2(errs[0] / errs[1]) > 2.5, errs


# ---------------------------------------------------------------------------
# shallow water 2D: (h, h v1, h v2, b)
# ---------------------------------------------------------------------------
def _swe_cfg(O, vflux, sflux, polydeg=4, **kw):
    return O.make_cfg(2, O.EQ_SHALLOW_WATER, polydeg, O.VI_FD, vflux, sflux, gravity=9.81, **kw)


def test_shallow_water_fluxes_and_lake_at_rest(oracle):
    """flux_wintermeyer_etal and flux_fjordholm_etal are consistent and
    symmetric; the lake at rest (H = h + b constant, v = 0) is a discrete
    steady state with the volume pair (wintermeyer, g h_l b_r) and the
    surface pair (fjordholm, Fjordholm's term) even for a bottom that jumps
    at every node, and with LLF at the surface for a continuous bottom (LLF
    leaves b undissipated)."""
    O = oracle
    rng = np.random.default_rng(50)
    cfg = _swe_cfg(O, O.FLUX_WINTERMEYER_ETAL, O.FLUX_FJORDHOLM_ETAL)
    for _ in range(20):
        ul = np.array([1.0 + rng.random(), 0.3 * rng.standard_normal(), 0.3 * rng.standard_normal(), rng.random()])
        ur = np.array([1.0 + rng.random(), 0.3 * rng.standard_normal(), 0.3 * rng.standard_normal(), rng.random()])
        for o in (1, 2):
            for k in (O.FLUX_WINTERMEYER_ETAL, O.FLUX_FJORDHOLM_ETAL, O.FLUX_LLF):
                np.testing.assert_allclose(O.two_point_flux(cfg, k, ul, ul, o), O.flux(cfg, ul, o), rtol=1e-14,
                                           atol=1e-14)
            for k in (O.FLUX_WINTERMEYER_ETAL, O.FLUX_FJORDHOLM_ETAL):
                np.testing.assert_allclose(O.two_point_flux(cfg, k, ul, ur, o), O.two_point_flux(cfg, k, ur, ul, o),
                                           rtol=1e-14, atol=1e-14)
    mesh = O.Mesh(2, 2, 0.0, 2 ** 0.5, 4)
    u = np.zeros((mesh.nelements, 5, 5, 4))
    b = 0.5 * rng.random(u.shape[:-1])
    u[..., 0], u[..., 3] = 3.0 - b, b
    # the cancelling terms are of size g H^2 N^2 invJ (about 1e4): round-off 1e-16 of that
    tol = 1e-16 * 9.81 * 3.0 ** 2 * 25 * mesh.inverse_jacobian.max() * 10
    assert np.abs(O.rhs(cfg, mesh, u)).max() < tol
    x = mesh.node_coordinates
    b = 0.5 + 0.2 * np.sin(2 * np.pi * x[..., 0] / 2 ** 0.5) * np.cos(2 * np.pi * x[..., 1] / 2 ** 0.5)
    u[..., 0], u[..., 3] = 3.0 - b, b
    for sflux in (O.FLUX_FJORDHOLM_ETAL, O.FLUX_LLF):
        assert np.abs(O.rhs(_swe_cfg(O, O.FLUX_WINTERMEYER_ETAL, sflux), mesh, u)).max() < tol


def test_shallow_water_energy_conservation(oracle):
    """The total energy E = h |v|^2 / 2 + g h^2 / 2 + g h b is conserved by the
    semidiscretisation with the reference's EC pairs (entropy variables by
    complex step); LLF at the surface dissipates it."""
    O = oracle
    g = 9.81
    mesh = O.Mesh(2, 2, 0.0, 2 ** 0.5, 4)
    cfg = _swe_cfg(O, O.FLUX_WINTERMEYER_ETAL, O.FLUX_FJORDHOLM_ETAL)
    u = O.compute_coefficients(cfg, mesh, O.IC_SEEDED_RANDOM)

    def energy(q):
        h, hv1, hv2, b = q
        return 0.5 * (hv1 * hv1 + hv2 * hv2) / h + 0.5 * g * h * h + g * h * b

    w = np.zeros_like(u)
    for j in range(4):
        uc = u.astype(complex)
        uc[..., j] += 1j * 1e-30
        w[..., j] = energy(np.moveaxis(uc, -1, 0)).imag / 1e-30
    for sflux, ec in ((O.FLUX_FJORDHOLM_ETAL, True), (O.FLUX_LLF, False)):
        du = O.rhs(_swe_cfg(O, O.FLUX_WINTERMEYER_ETAL, sflux), mesh, u)
        assert np.abs(du[..., 3]).max() == 0.0
        rate = O.integrate(cfg, mesh, np.repeat((w * du).sum(-1, keepdims=True), 4, axis=-1))[0]
        scale = np.abs(du).max() * mesh.total_volume
        if ec:
            assert abs(rate) < 1e-12 * scale, rate
        else:
            assert rate < -1e-6 * scale, rate


def test_shallow_water_source_and_convergence(oracle):
    """source_terms_convergence_test equals the sympy residual of the
    manufactured state (h = H - b, v constant, with g h grad b), and with it
    rhs!(u_exact) - u_t converges at the DG rate, periodic and with the
    Dirichlet state on every side."""
    import sympy as sp
    O = oracle
    g = 9.81
    X, Y, T = sp.symbols("x y t")
    w, wt, wb = 2 * sp.pi * sp.sqrt(2), 2 * sp.pi, sp.sqrt(2) * sp.pi
    H = 7 + sp.cos(w * X) * sp.sin(w * Y) * sp.cos(wt * T)
    b = 2 + sp.sin(wb * X) / 2 + sp.sin(wb * Y) / 2
    h, v1, v2 = H - b, sp.Rational(1, 2), sp.Rational(3, 2)
    S = [sp.diff(h, T) + sp.diff(h * v1, X) + sp.diff(h * v2, Y),
         sp.diff(h * v1, T) + sp.diff(h * v1 * v1 + g * h * h / 2, X) + sp.diff(h * v1 * v2, Y) + g * h * sp.diff(b, X),
         sp.diff(h * v2, T) + sp.diff(h * v1 * v2, X) + sp.diff(h * v2 * v2 + g * h * h / 2, Y) + g * h * sp.diff(b, Y)]
    Sf = sp.lambdify((X, Y, T), S, "numpy")
    cfg = _swe_cfg(O, O.FLUX_WINTERMEYER_ETAL, O.FLUX_LLF, 3, source_terms=O.SOURCE_CONVERGENCE_TEST)
    rng = np.random.default_rng(51)
    for _ in range(8):
        x, t = rng.uniform(0, 2 ** 0.5, 2), rng.uniform(0, 1)
        s = O.source_convergence_test(cfg, x, t)
        np.testing.assert_allclose(s[:3], np.array(Sf(x[0], x[1], t), dtype=float), rtol=1e-12, atol=1e-11)
        assert s[3] == 0.0
    for periodic in (True, False):
        kw = {} if periodic else dict(boundary_condition=O.BC_DIRICHLET, boundary_ic=O.IC_CONVERGENCE_TEST)
        c = _swe_cfg(O, O.FLUX_WINTERMEYER_ETAL, O.FLUX_LLF, 3, source_terms=O.SOURCE_CONVERGENCE_TEST, **kw)
        errs, t = [], 0.1
        for lvl in (2, 3):
            mesh = O.Mesh(2, lvl, 0.0, 2 ** 0.5, 3, periodic=periodic)
            u = O.compute_coefficients(c, mesh, O.IC_CONVERGENCE_TEST, t)
            hh = 1e-6
            ut = (O.compute_coefficients(c, mesh, O.IC_CONVERGENCE_TEST, t + hh)
                  - O.compute_coefficients(c, mesh, O.IC_CONVERGENCE_TEST, t - hh)) / (2 * hh)
            errs.append(np.abs(O.rhs(c, mesh, u, t=t) - ut).max())
        assert np.log2(errs[0] / errs[1]) > 2.5, (periodic, errs)


def test_shallow_water_1d_hydrostatic_reconstruction(oracle):
    """1D shallow water with the hydrostatic reconstruction
    (test/tree_dgsem_1d/shallowwater_shock.jl): the reconstructed LLF flux is
    consistent; the lake at rest stays at rest with a bottom that jumps at
    every node, for flux differencing and for shock capturing at any blending
    (Audusse's term; its sign and factor are fixed by this); the total
    energy does not grow with the reconstruction at the surface; slip walls
    keep a lake at rest and reflect: with v = 0 at the wall the mass flux
    through it vanishes (total mass conserved)."""
    O = oracle
    g = 9.812
    rng = np.random.default_rng(60)
    kw = dict(gravity=g)
    cfg = O.make_cfg(1, O.EQ_SHALLOW_WATER, 4, O.VI_FD, O.FLUX_WINTERMEYER_ETAL, O.FLUX_HYDROSTATIC_LLF, **kw)
    for _ in range(10):
        u = np.array([1.0 + rng.random(), 0.3 * rng.standard_normal(), rng.random()])
        np.testing.assert_allclose(O.two_point_flux(cfg, O.FLUX_HYDROSTATIC_LLF, u, u, 1), O.flux(cfg, u, 1),
                                   rtol=1e-14, atol=1e-14)
    for periodic, bc in ((True, O.BC_PERIODIC), (False, O.BC_SLIP_WALL)):
        mesh = O.Mesh(1, 3, -3.0, 3.0, 4, periodic=periodic)
        u = np.zeros((mesh.nelements, 5, 3))
        b = 0.8 * rng.random(u.shape[:-1])
        u[..., 0], u[..., 2] = 1.75 - b, b
        tol = 1e-15 * g * 1.75 ** 2 * 25 * mesh.inverse_jacobian.max() * 10
        for vi, fv in ((O.VI_FD, O.FLUX_CENTRAL), (O.VI_HG, O.FLUX_HYDROSTATIC_LLF)):
            for amax in (0.5, 1.0):
                c = O.make_cfg(1, O.EQ_SHALLOW_WATER, 4, vi, O.FLUX_WINTERMEYER_ETAL, O.FLUX_HYDROSTATIC_LLF, fv,
                               alpha_max=amax, alpha_min=0.0, boundary_condition=bc, **kw)
                assert np.abs(O.rhs(c, mesh, u)).max() < tol, (periodic, vi, amax)
    mesh = O.Mesh(1, 3, -3.0, 3.0, 4, periodic=False)
    c = O.make_cfg(1, O.EQ_SHALLOW_WATER, 4, O.VI_FD, O.FLUX_WINTERMEYER_ETAL, O.FLUX_HYDROSTATIC_LLF,
                   boundary_condition=O.BC_SLIP_WALL, **kw)
    u = O.compute_coefficients(c, mesh, O.IC_SEEDED_RANDOM)
    du = O.rhs(c, mesh, u)
    assert abs(O.integrate(c, mesh, du)[0]) < 1e-13 * np.abs(du).max() * mesh.total_volume

    def energy(q):
        h, hv, bb = q
        return 0.5 * hv * hv / h + 0.5 * g * h * h + g * h * bb

    w = np.zeros_like(u)
    for j in range(3):
        uc = u.astype(complex)
        uc[..., j] += 1j * 1e-30
        w[..., j] = energy(np.moveaxis(uc, -1, 0)).imag / 1e-30
    rate = O.integrate(c, mesh, np.repeat((w * du).sum(-1, keepdims=True), 3, axis=-1))[0]
    assert rate <= 1e-12 * np.abs(du).max() * mesh.total_volume, rate
