"""Closed-form pins of the oracle functions that the invariant tests leave
open: the Hennemann-Gassner indicator between its saturation ends (the
logistic's interior and the second-highest-mode energy fraction), the
AnalysisCallback error norms, the quadrature integral and the weak-blast
initial state.

Every expected value comes from an independent construction: Legendre modes
from numpy.polynomial.legendre (not the oracle's Vandermonde), integrals of
polynomials in closed form, and the constants the reference and the
benchmark definition state.  Each check is chosen so that a plausible mistake
in the oracle (a wrong denominator, a mis-scaled logistic, a missing Jacobian
or volume normalisation, a swapped state constant) changes the result.
"""
import numpy as np
import pytest
from numpy.polynomial import legendre as npleg

# Reference constants: src/solvers/indicators.jl:22-24
#   threshold = 0.5 * 10^(-1.8 * nnodes^0.25), parameter_s = log((1 - 0.0001) / 0.0001)
N4 = 4
THRESHOLD = 0.5 * 10.0 ** (-1.8 * N4 ** 0.25)
PARAMETER_S = np.log((1.0 - 0.0001) / 0.0001)


def _lgl_nodes(n):
    """Legendre-Gauss-Lobatto nodes: +-1 and the roots of P'_{n-1}."""
    c = np.zeros(n)
    c[-1] = 1.0
    inner = npleg.legroots(npleg.legder(c))
    return np.concatenate(([-1.0], np.sort(inner), [1.0]))


def _mode(m, x):
    """Normalised Legendre polynomial sqrt(m + 1/2) P_m(x) (the modal basis of
    the indicator's inverse Vandermonde)."""
    c = np.zeros(m + 1)
    c[m] = 1.0
    return np.sqrt(m + 0.5) * npleg.legval(x, c)


def _logistic(energy):
    return 1.0 / (1.0 + np.exp(-PARAMETER_S / THRESHOLD * (energy - THRESHOLD)))


def _euler2d_state_with_indicator(values):
    """A 2D Euler element state (v = 0, rho = 1) whose density_pressure
    (gamma - 1)(rho rho_e - |rho v|^2 / 2) equals `values` [i, j] at the nodes."""
    g = 1.4
    u = np.zeros(values.shape + (4,))
    u[..., 0] = 1.0
    u[..., 3] = values / (g - 1.0)
    return u


def _mesh_and_cfg(O, alpha_max=1.0, alpha_min=0.001):
    cfg = O.make_cfg(2, O.EQ_EULER, 3, O.VI_HG, O.FLUX_RANOCHA, O.FLUX_LLF, O.FLUX_LLF,
                     alpha_max=alpha_max, alpha_min=alpha_min, alpha_smooth=False)
    mesh = O.Mesh(2, 2, -2.0, 2.0, 3)
    return cfg, mesh


def _indicator_of_modes(O, modes, alpha_max=1.0, alpha_min=0.001, element=5):
    """alpha of `element` whose density_pressure is sum c phi_i(x) phi_j(y)
    over modes {(i, j): c}; every other element is a constant state."""
    cfg, mesh = _mesh_and_cfg(O, alpha_max, alpha_min)
    x = _lgl_nodes(N4)
    vals = np.zeros((N4, N4))
    for (i, j), c in modes.items():
        vals += c * _mode(i, x)[:, None] * _mode(j, x)[None, :]
    u = O.compute_coefficients(cfg, mesh, O.IC_CONSTANT)
    # state layout [element, i, j, variable], i the x index
    u[element] = _euler2d_state_with_indicator(vals)
    alpha = O.indicator_hg(cfg, mesh, u)
    assert np.all(np.delete(alpha, element) == 0.0)
    return alpha[element]


@pytest.mark.parametrize("scale", [0.7, 0.9, 1.0, 1.15, 1.4])
def test_indicator_logistic_interior_highest_mode(oracle, scale):
    """Energy in the constant mode and the highest mode: e = frac1 =
    c3^2 / (c0^2 + c3^2), set to `scale` x threshold, so that alpha is the
    logistic strictly between alpha_min and alpha_max."""
    e = scale * THRESHOLD
    c0 = 1.0
    c3 = c0 * np.sqrt(e / (1.0 - e))
    expected = _logistic(e)
    assert 0.001 < expected < 1.0 - 0.001
    got = _indicator_of_modes(oracle, {(0, 0): c0, (3, 0): c3})
    assert got == pytest.approx(expected, rel=1e-10)
    # the same fraction in the y direction
    got_y = _indicator_of_modes(oracle, {(0, 0): c0, (0, 3): c3})
    assert got_y == pytest.approx(expected, rel=1e-10)


@pytest.mark.parametrize("scale", [0.8, 1.0, 1.25])
def test_indicator_second_highest_mode_fraction(oracle, scale):
    """Energy in modes of degree 0, 2 and 3: frac1 = c3^2 / total
    stays below frac2 = c2^2 / (c0^2 + c2^2), whose denominator is the energy
    without the highest modes (not the total energy), so alpha = logistic(frac2).
    Mode (2, 1) has max degree 2: it counts in frac2's numerator."""
    e2 = scale * THRESHOLD
    c0 = 1.0
    c2 = c0 * np.sqrt(e2 / (1.0 - e2))
    c3 = 0.9 * c2
    # split the degree-2 energy between (2, 0) and (2, 1)
    a, b = c2 * np.sqrt(0.6), c2 * np.sqrt(0.4)
    total = c0 ** 2 + c2 ** 2 + c3 ** 2
    frac1 = c3 ** 2 / total
    frac2 = c2 ** 2 / (c0 ** 2 + c2 ** 2)
    assert frac1 < frac2
    expected = _logistic(frac2)
    assert 0.001 < expected < 1.0 - 0.001
    got = _indicator_of_modes(oracle, {(0, 0): c0, (2, 0): a, (2, 1): b, (0, 3): c3})
    assert got == pytest.approx(expected, rel=1e-10)
    # a denominator of the total energy would give a factor that differs by
    # far more than the tolerance above
    assert abs(_logistic(c2 ** 2 / total) - expected) > 1e-5 * expected


def test_indicator_clipping_ends(oracle):
    """alpha below alpha_min is 0, above 1 - alpha_min is 1, then capped at
    alpha_max (Trixi's clipping order)."""
    e_low = 0.2 * THRESHOLD  # logistic below alpha_min = 0.001
    assert _logistic(e_low) < 0.001
    c3 = np.sqrt(e_low / (1.0 - e_low))
    assert _indicator_of_modes(oracle, {(0, 0): 1.0, (3, 0): c3}) == 0.0
    e_high = 2.0 * THRESHOLD
    assert _logistic(e_high) > 1.0 - 0.001
    c3 = np.sqrt(e_high / (1.0 - e_high))
    assert _indicator_of_modes(oracle, {(0, 0): 1.0, (3, 0): c3}, alpha_max=0.5) == 0.5
    assert _indicator_of_modes(oracle, {(0, 0): 1.0, (3, 0): c3}, alpha_max=1.0) == 1.0
    # a value inside (alpha_min, 1 - alpha_min) but above alpha_max is capped
    e_mid = 1.2 * THRESHOLD
    assert 0.3 < _logistic(e_mid) < 0.999
    c3 = np.sqrt(e_mid / (1.0 - e_mid))
    assert _indicator_of_modes(oracle, {(0, 0): 1.0, (3, 0): c3}, alpha_max=0.3) == 0.3


def _linadv_cfg(O, nd):
    return O.make_cfg(nd, O.EQ_LINADV, 3, O.VI_WEAK, O.FLUX_CENTRAL, O.FLUX_LLF,
                      advection_velocity=(1.0, 0.5, 0.25)[:nd])


@pytest.mark.parametrize("nd,level,lo,hi", [(1, 3, -1.0, 1.0), (2, 2, -1.0, 3.0), (3, 1, 0.0, 2.0)])
def test_error_norms_of_a_constant_offset(oracle, nd, level, lo, hi):
    """u = u_exact + c (u_exact the constant state 2): L2 = Linf = |c| on
    any mesh (the weights, the Jacobian and the volume normalisation)."""
    O = oracle
    cfg = _linadv_cfg(O, nd)
    mesh = O.Mesh(nd, level, lo, hi, 3)
    u = O.compute_coefficients(cfg, mesh, O.IC_CONSTANT)
    assert np.all(u == 2.0)
    for c in (0.25, -1.5e-3):
        l2, linf = O.error_norms(cfg, mesh, u + c, 0.0, O.IC_CONSTANT)
        assert l2[0] == pytest.approx(abs(c), rel=1e-13)
        assert linf[0] == pytest.approx(abs(c), rel=1e-13)


@pytest.mark.parametrize("nd,level", [(1, 2), (2, 1), (3, 1)])
def test_error_norms_of_a_polynomial_error(oracle, nd, level):
    """u = u_exact + g with g = prod_d x_d (degree 1 per direction, exactly
    interpolated to the analysis nodes) on [-1, 1]^d:
    L2 = sqrt(int g^2 / |Omega|) = 3^(-d/2), Linf = max |g| = 1 (a corner)."""
    O = oracle
    cfg = _linadv_cfg(O, nd)
    mesh = O.Mesh(nd, level, -1.0, 1.0, 3)
    x = mesh.node_coordinates
    g = np.prod(x, axis=-1)[..., None]
    u = O.compute_coefficients(cfg, mesh, O.IC_CONSTANT) + g
    l2, linf = O.error_norms(cfg, mesh, u, 0.0, O.IC_CONSTANT)
    assert l2[0] == pytest.approx(3.0 ** (-nd / 2), rel=1e-13)
    assert linf[0] == pytest.approx(1.0, rel=1e-13)
    # a quadratic error: g = x_0^2 - 1/3 on [-1, 1]^d, int g^2 / |Omega| = 4/45
    g2 = (x[..., 0] ** 2 - 1.0 / 3.0)[..., None]
    l2, linf = O.error_norms(cfg, mesh, O.compute_coefficients(cfg, mesh, O.IC_CONSTANT) + g2, 0.0,
                             O.IC_CONSTANT)
    assert l2[0] == pytest.approx(np.sqrt(4.0 / 45.0), rel=1e-12)
    assert linf[0] == pytest.approx(2.0 / 3.0, rel=1e-13)


@pytest.mark.parametrize("nd,level,lo,hi", [(1, 3, -1.0, 1.0), (2, 2, -1.0, 3.0), (3, 1, -1.0, 1.0)])
def test_integrate_constants_and_polynomials(oracle, nd, level, lo, hi):
    """The LGL quadrature of a constant is c |Omega|; of x_0^4 x_1^2 (degree
    <= 2N - 3 = 5 per direction, integrated exactly by 4 LGL nodes) it is the
    closed form prod_d (hi^(k+1) - lo^(k+1)) / (k + 1)."""
    O = oracle
    cfg = _linadv_cfg(O, nd)
    mesh = O.Mesh(nd, level, lo, hi, 3)
    x = mesh.node_coordinates
    assert O.integrate(cfg, mesh, np.full(x.shape[:-1] + (1,), 3.5))[0] == pytest.approx(
        3.5 * (hi - lo) ** nd, rel=1e-14)
    assert mesh.total_volume == pytest.approx((hi - lo) ** nd, rel=1e-15)
    powers = (4, 2, 1)[:nd]
    f = np.ones(x.shape[:-1])
    exact = 1.0
    for d, k in enumerate(powers):
        f = f * x[..., d] ** k
        exact *= (hi ** (k + 1) - lo ** (k + 1)) / (k + 1)
    got = O.integrate(cfg, mesh, f[..., None])[0]
    assert got == pytest.approx(exact, rel=1e-12, abs=1e-13)


def test_weak_blast_wave_state(oracle):
    """initial_condition_weak_blast_wave (the shock-capturing benchmark's
    state): inside r <= 0.5 rho = 1.1691, p = 1.245, |v| = 0.1882 radially
    outwards; outside rho = 1, v = 0, p = 1 (conservative variables with
    gamma = 1.4)."""
    O = oracle
    g = 1.4
    cfg2 = O.make_cfg(2, O.EQ_EULER, 3, O.VI_HG, O.FLUX_RANOCHA, O.FLUX_LLF, O.FLUX_LLF, gamma=g)
    for x in ([0.1, 0.2], [-0.3, 0.35], [0.0, -0.5]):
        u = O.initial_condition(cfg2, O.IC_WEAK_BLAST_WAVE, x)
        r = np.hypot(*x)
        rho, p = 1.1691, 1.245
        v = 0.1882 * np.array(x) / r
        assert u == pytest.approx([rho, rho * v[0], rho * v[1], p / (g - 1) + 0.5 * rho * v @ v], rel=1e-14,
                                  abs=1e-15)
    for x in ([0.6, 0.0], [1.0, 1.0], [-2.0, 1.5]):
        u = O.initial_condition(cfg2, O.IC_WEAK_BLAST_WAVE, x)
        assert u == pytest.approx([1.0, 0.0, 0.0, 1.0 / (g - 1)], rel=1e-15, abs=0.0)
    cfg3 = O.make_cfg(3, O.EQ_EULER, 3, O.VI_FD, O.FLUX_RANOCHA, O.FLUX_RANOCHA, gamma=g)
    x = np.array([0.1, -0.2, 0.15])
    u = O.initial_condition(cfg3, O.IC_WEAK_BLAST_WAVE, x)
    v = 0.1882 * x / np.linalg.norm(x)
    assert u == pytest.approx([1.1691, *(1.1691 * v), 1.245 / (g - 1) + 0.5 * 1.1691 * v @ v], rel=1e-14)
    cfg1 = O.make_cfg(1, O.EQ_EULER, 3, O.VI_FD, O.FLUX_RANOCHA, O.FLUX_RANOCHA, gamma=g)
    for xx, sgn in ((0.2, 1.0), (-0.2, -1.0)):
        u = O.initial_condition(cfg1, O.IC_WEAK_BLAST_WAVE, [xx])
        assert u == pytest.approx([1.1691, 1.1691 * 0.1882 * sgn, 1.245 / (g - 1) + 0.5 * 1.1691 * 0.1882 ** 2],
                                  rel=1e-14)
