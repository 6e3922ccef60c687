"""Pins the oracle's point fluxes by properties the mathematics fixes.

* ln_mean / inv_ln_mean against 50-digit mpmath values (tests/golden/ln_mean.json);
* consistency f*(u,u) = f(u) and symmetry of every two-point flux;
* Tadmor's entropy-conservation condition [[w]].f* = [[psi]] for flux_ranocha;
* wave-speed estimates against eigenvalues of the primitive-variable
  quasi-linear systems (Euler and ideal MHD), computed by numpy;
* LLF = central - lambda/2 [[u]] with lambda at least the spectral radius.
"""
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rng_state_euler(O, cfg, rng):
    nd = cfg.ndims
    rho = 1.0 + 0.5 * rng.random()
    v = 0.5 * rng.standard_normal(nd)
    p = 1.0 + 0.5 * rng.random()
    u = np.zeros(nd + 2)
    u[0] = rho
    u[1:1 + nd] = rho * v
    u[-1] = p / (cfg.gamma - 1) + 0.5 * rho * v @ v
    return u


def rng_state_mhd(O, cfg, rng):
    rho = 1.0 + 0.5 * rng.random()
    v = 0.5 * rng.standard_normal(3)
    p = 1.0 + 0.5 * rng.random()
    B = 0.7 * rng.standard_normal(3)
    psi = 0.1 * rng.standard_normal()
    u = np.zeros(9)
    u[0] = rho
    u[1:4] = rho * v
    u[4] = p / (cfg.gamma - 1) + 0.5 * rho * v @ v + 0.5 * B @ B + 0.5 * psi * psi
    u[5:8] = B
    u[8] = psi
    return u


def test_ln_mean_against_mpmath(oracle):
    with open(os.path.join(GOLDEN, "ln_mean.json")) as fh:
        cases = json.load(fh)
    for c in cases:
        lm = oracle.ln_mean(c["x"], c["y"])
        ilm = oracle.inv_ln_mean(c["x"], c["y"])
        ref, iref = float(c["ln_mean"]), float(c["inv_ln_mean"])
        # the closed-form branch loses digits as f^2 -> 1e-4 from above:
        # |d log| ~ eps / (y/x - 1) ~ 1e-14 at the threshold
        assert abs(lm - ref) <= 5e-14 * abs(ref), (c, lm)
        assert abs(ilm - iref) <= 5e-14 * abs(iref), (c, ilm)
        assert oracle.ln_mean(c["y"], c["x"]) == pytest.approx(ref, rel=5e-14)


@pytest.mark.parametrize("nd", [1, 2, 3])
@pytest.mark.parametrize("kind", ["RANOCHA", "SHIMA_ETAL", "LLF", "LLF_NAIVE", "CENTRAL"])
def test_euler_two_point_consistency_and_symmetry(oracle, nd, kind):
    O = oracle
    cfg = O.make_cfg(nd, O.EQ_EULER, gamma=1.4)
    k = getattr(O, "FLUX_" + kind)
    rng = np.random.default_rng(1)
    for _ in range(20):
        ul, ur = rng_state_euler(O, cfg, rng), rng_state_euler(O, cfg, rng)
        for o in range(1, nd + 1):
            f = O.flux(cfg, ul, o)
            np.testing.assert_allclose(O.two_point_flux(cfg, k, ul, ul, o), f, rtol=1e-14, atol=1e-14)
            if kind not in ("LLF", "LLF_NAIVE"):
                np.testing.assert_allclose(O.two_point_flux(cfg, k, ul, ur, o),
                                           O.two_point_flux(cfg, k, ur, ul, o), rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_ranocha_is_entropy_conservative(oracle, nd):
    O = oracle
    cfg = O.make_cfg(nd, O.EQ_EULER, gamma=1.4)
    rng = np.random.default_rng(2)
    for _ in range(50):
        ul, ur = rng_state_euler(O, cfg, rng), rng_state_euler(O, cfg, rng)
        # also states close together (series branch of ln_mean)
        if _ % 2:
            ur = ul * (1 + 1e-4 * rng.standard_normal(len(ul)))
        wl, wr = O.cons2entropy(cfg, ul), O.cons2entropy(cfg, ur)
        for o in range(1, nd + 1):
            f = O.two_point_flux(cfg, O.FLUX_RANOCHA, ul, ur, o)
            lhs = (wr - wl) @ f
            rhs = O.entropy_potential(cfg, ur, o) - O.entropy_potential(cfg, ul, o)
            assert abs(lhs - rhs) <= 1e-13 * (1 + np.abs(wr - wl) @ np.abs(f)), (lhs, rhs)
    # shima_etal is kinetic-energy preserving but not entropy conservative
    ul, ur = rng_state_euler(O, cfg, rng), rng_state_euler(O, cfg, rng)
    wl, wr = O.cons2entropy(cfg, ul), O.cons2entropy(cfg, ur)
    f = O.two_point_flux(cfg, O.FLUX_SHIMA_ETAL, ul, ur, 1)
    assert abs((wr - wl) @ f - (O.entropy_potential(cfg, ur, 1) - O.entropy_potential(cfg, ul, 1))) > 1e-8


def euler_prim_matrix(rho, vn, p, gamma):
    return np.array([[vn, rho, 0.0], [0.0, vn, 1.0 / rho], [0.0, gamma * p, vn]])


def mhd_prim_matrix(rho, v, p, B, gamma, o):
    # 1D ideal MHD in primitive variables (rho, v_n, v_t1, v_t2, B_t1, B_t2, p)
    # along direction o with the normal field as a parameter.
    t1, t2 = [d for d in range(3) if d != o]
    vn, Bn = v[o], B[o]
    A = np.zeros((7, 7))
    A[0, 0], A[0, 1] = vn, rho
    A[1, 1], A[1, 4], A[1, 5], A[1, 6] = vn, B[t1] / rho, B[t2] / rho, 1.0 / rho
    A[2, 2], A[2, 4] = vn, -Bn / rho
    A[3, 3], A[3, 5] = vn, -Bn / rho
    A[4, 1], A[4, 2], A[4, 4] = B[t1], -Bn, vn
    A[5, 1], A[5, 3], A[5, 5] = B[t2], -Bn, vn
    A[6, 1], A[6, 6] = gamma * p, vn
    return A


def test_wave_speeds_match_eigenvalues(oracle):
    O = oracle
    rng = np.random.default_rng(3)
    cfg = O.make_cfg(3, O.EQ_EULER, gamma=1.4)
    for _ in range(20):
        u = rng_state_euler(O, cfg, rng)
        rho, v = u[0], u[1:4] / u[0]
        p = 0.4 * (u[4] - 0.5 * rho * v @ v)
        lam = O.max_abs_speeds(cfg, u)
        for o in range(3):
            ev = np.linalg.eigvals(euler_prim_matrix(rho, v[o], p, 1.4))
            assert lam[o] == pytest.approx(np.abs(ev).max(), rel=1e-12)
    cfgm = O.make_cfg(3, O.EQ_GLM_MHD, gamma=5.0 / 3.0, c_h=1.0)
    for _ in range(20):
        u = rng_state_mhd(O, cfgm, rng)
        rho, v, B, psi = u[0], u[1:4] / u[0], u[5:8], u[8]
        p = (2.0 / 3.0) * (u[4] - 0.5 * rho * v @ v - 0.5 * B @ B - 0.5 * psi**2)
        lam = O.max_abs_speeds(cfgm, u)
        for o in range(3):
            ev = np.linalg.eigvals(mhd_prim_matrix(rho, v, p, B, 5.0 / 3.0, o))
            assert lam[o] == pytest.approx(np.abs(ev).max(), rel=1e-9)


@pytest.mark.parametrize("eq", ["EULER", "GLM_MHD"])
def test_llf_structure(oracle, eq):
    O = oracle
    rng = np.random.default_rng(4)
    if eq == "EULER":
        cfg = O.make_cfg(3, O.EQ_EULER, gamma=1.4)
        gen = rng_state_euler
    else:
        cfg = O.make_cfg(3, O.EQ_GLM_MHD, gamma=5.0 / 3.0, c_h=1.0)
        gen = rng_state_mhd
    for _ in range(20):
        ul, ur = gen(O, cfg, rng), gen(O, cfg, rng)
        for o in (1, 2, 3):
            fc = O.two_point_flux(cfg, O.FLUX_CENTRAL, ul, ur, o)
            sl, sr = O.max_abs_speeds(cfg, ul)[o - 1], O.max_abs_speeds(cfg, ur)[o - 1]
            for kind, lam_expect in ((O.FLUX_LLF, max(sl, sr)), (O.FLUX_LLF_NAIVE, None)):
                f = O.two_point_flux(cfg, kind, ul, ur, o)
                d = (fc - f)
                # dissipation is a multiple of the jump with lambda >= local spectral radius
                lam = d @ (ur - ul) / ((ur - ul) @ (ur - ul)) * 2
                np.testing.assert_allclose(d, 0.5 * lam * (ur - ul), rtol=1e-12, atol=1e-13)
                assert lam >= max(sl, sr) * (1 - 1e-14)
                if lam_expect is not None:
                    assert lam == pytest.approx(lam_expect, rel=1e-13)


def test_mhd_two_point_consistency_symmetry(oracle):
    O = oracle
    cfg = O.make_cfg(3, O.EQ_GLM_MHD, gamma=5.0 / 3.0, c_h=1.0)
    rng = np.random.default_rng(5)
    for _ in range(20):
        ul, ur = rng_state_mhd(O, cfg, rng), rng_state_mhd(O, cfg, rng)
        for o in (1, 2, 3):
            np.testing.assert_allclose(O.two_point_flux(cfg, O.FLUX_HINDENLANG_GASSNER, ul, ul, o),
                                       O.flux(cfg, ul, o), rtol=1e-13, atol=1e-13)
            np.testing.assert_allclose(O.two_point_flux(cfg, O.FLUX_HINDENLANG_GASSNER, ul, ur, o),
                                       O.two_point_flux(cfg, O.FLUX_HINDENLANG_GASSNER, ur, ul, o),
                                       rtol=1e-13, atol=1e-13)
            # Powell/GLM term is linear in the right state's normal field and psi
            z = ur.copy()
            z[5 + o - 1] = 0.0
            z[8] = 0.0
            np.testing.assert_allclose(O.noncons_flux(cfg, ul, z, o), 0.0, atol=0)


def test_burgers_flux_ec_and_llf(oracle):
    """flux_ec of the inviscid Burgers equation: consistent (f*(u, u) = u^2/2),
    symmetric, and entropy conservative for U = u^2/2, w = u, F = u^3/3:
    (u_r - u_l) f* = psi_r - psi_l with psi = w f - F = u^3/6.  LLF: the
    central flux minus max(|u_l|, |u_r|) / 2 [[u]], both wave-speed variants."""
    O = oracle
    cfg = O.make_cfg(1, O.EQ_BURGERS, 3)
    rng = np.random.default_rng(3)
    for _ in range(50):
        ul, ur = rng.uniform(-2, 2, 1), rng.uniform(-2, 2, 1)
        f = O.two_point_flux(cfg, O.FLUX_EC, ul, ur, 1)
        assert f[0] == pytest.approx(O.two_point_flux(cfg, O.FLUX_EC, ur, ul, 1)[0], rel=1e-15)
        assert (ur - ul)[0] * f[0] == pytest.approx((ur[0] ** 3 - ul[0] ** 3) / 6, rel=1e-12, abs=1e-14)
        assert O.two_point_flux(cfg, O.FLUX_EC, ul, ul, 1)[0] == pytest.approx(0.5 * ul[0] ** 2, rel=1e-15)
        lam = max(abs(ul[0]), abs(ur[0]))
        expect = 0.25 * (ul[0] ** 2 + ur[0] ** 2) - 0.5 * lam * (ur[0] - ul[0])
        for kind in (O.FLUX_LLF, O.FLUX_LLF_NAIVE):
            assert O.two_point_flux(cfg, kind, ul, ur, 1)[0] == pytest.approx(expect, rel=1e-14, abs=1e-15)


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_hll_flux_structure(oracle, nd):
    """FluxHLL(min_max_speed_davis) for Euler: consistent; the left (right)
    physical flux when every wave moves right (left); otherwise
    (l_max f_l - l_min f_r + l_min l_max [[u]]) / (l_max - l_min) with the
    Davis estimates min / max of v -+ c over both states; and its
    dissipation never exceeds LLF's."""
    O = oracle
    cfg = O.make_cfg(nd, O.EQ_EULER, 3, gamma=1.4)
    rng = np.random.default_rng(40 + nd)

    def prim2cons(rho, v, p):
        return np.concatenate([[rho], rho * v, [p / 0.4 + 0.5 * rho * v @ v]])

    for shift in (0.0, 3.0, -3.0):  # subsonic, supersonic right, supersonic left
        for _ in range(20):
            rl, rr = rng.uniform(0.5, 2, 2)
            pl, pr = rng.uniform(0.5, 2, 2)
            vl, vr = rng.normal(0, 0.3, nd), rng.normal(0, 0.3, nd)
            vl[0] += shift
            vr[0] += shift
            ul, ur = prim2cons(rl, vl, pl), prim2cons(rr, vr, pr)
            f = O.two_point_flux(cfg, O.FLUX_HLL, ul, ur, 1)
            fl, fr = O.flux(cfg, ul, 1), O.flux(cfg, ur, 1)
            cl, cr = np.sqrt(1.4 * pl / rl), np.sqrt(1.4 * pr / rr)
            lmin, lmax = min(vl[0] - cl, vr[0] - cr), max(vl[0] + cl, vr[0] + cr)
            if lmin >= 0:
                np.testing.assert_allclose(f, fl, rtol=1e-14, atol=1e-14)
            elif lmax <= 0:
                np.testing.assert_allclose(f, fr, rtol=1e-14, atol=1e-14)
            else:
                expect = (lmax * fl - lmin * fr + lmin * lmax * (ur - ul)) / (lmax - lmin)
                np.testing.assert_allclose(f, expect, rtol=1e-12, atol=1e-13)
        np.testing.assert_allclose(O.two_point_flux(cfg, O.FLUX_HLL, ul, ul, 1), O.flux(cfg, ul, 1), rtol=1e-14,
                                   atol=1e-14)


# ---------------------------------------------------------------------------
# multicomponent compressible Euler: (rho v, rho e, rho_1..rho_nc)
# ---------------------------------------------------------------------------
from cases import MULTI_GASES, multi_entropy, multi_entropy_vars  # noqa: E402


def multi_state(nd, gammas, gas_constants, rng):
    g, R = np.array(gammas), np.array(gas_constants)
    cv = R / (g - 1)
    rk = 0.5 + rng.random(len(g))
    v = 0.3 * rng.standard_normal(nd)
    p = 1.0 + 0.3 * rng.random()
    gam = np.sum(rk * cv * g) / np.sum(rk * cv)
    rho = rk.sum()
    return np.concatenate([rho * v, [p / (gam - 1) + 0.5 * rho * v @ v], rk])


@pytest.mark.parametrize("nd", [1, 2])
@pytest.mark.parametrize("gas", MULTI_GASES, ids=["equal", "single", "distinct"])
def test_multicomponent_fluxes_consistency_symmetry(oracle, nd, gas):
    O = oracle
    cfg = O.make_cfg(nd, O.EQ_EULER_MULTI, gammas=gas[0], gas_constants=gas[1])
    rng = np.random.default_rng(11)
    for _ in range(20):
        ul, ur = multi_state(nd, *gas, rng), multi_state(nd, *gas, rng)
        for o in range(1, nd + 1):
            f = O.flux(cfg, ul, o)
            for k in (O.FLUX_RANOCHA, O.FLUX_LLF, O.FLUX_LLF_NAIVE, O.FLUX_CENTRAL):
                np.testing.assert_allclose(O.two_point_flux(cfg, k, ul, ul, o), f, rtol=1e-13, atol=1e-14)
            np.testing.assert_allclose(O.two_point_flux(cfg, O.FLUX_RANOCHA, ul, ur, o),
                                       O.two_point_flux(cfg, O.FLUX_RANOCHA, ur, ul, o), rtol=1e-13, atol=1e-14)
            # partial mass fluxes add up to the momentum-carrying mass flux
            fs = O.two_point_flux(cfg, O.FLUX_RANOCHA, ul, ur, o)
            vavg = [0.5 * (ul[d] / ul[nd + 1:].sum() + ur[d] / ur[nd + 1:].sum()) for d in range(nd)]
            pl = (O.flux(cfg, ul, o)[o - 1] - ul[o - 1] ** 2 / ul[nd + 1:].sum())
            pr = (O.flux(cfg, ur, o)[o - 1] - ur[o - 1] ** 2 / ur[nd + 1:].sum())
            for d in range(nd):
                expect = fs[nd + 1:].sum() * vavg[d] + (0.5 * (pl + pr) if d == o - 1 else 0.0)
                assert fs[d] == pytest.approx(expect, rel=1e-12, abs=1e-14)


@pytest.mark.parametrize("nd", [1, 2])
def test_multicomponent_ranocha_entropy_conservation(oracle, nd):
    """Tadmor's condition [[w]].f* = [[psi]], psi = w.f - U v_o, with the
    mixture entropy -sum_k rho_k (c_v,k log T - R_k log rho_k) and w by complex
    step.  It holds to round-off for one component and for components with
    equal gamma and R (the reference tests' gases); with distinct gases the
    adapted flux is not entropy conservative."""
    O = oracle
    rng = np.random.default_rng(12)
    for gas, holds in zip(MULTI_GASES, (True, True, False)):
        cfg = O.make_cfg(nd, O.EQ_EULER_MULTI, gammas=gas[0], gas_constants=gas[1])
        worst = 0.0
        for it in range(30):
            ul, ur = multi_state(nd, *gas, rng), multi_state(nd, *gas, rng)
            if it % 3 == 0:
                ur = ul * (1 + 1e-4 * rng.standard_normal(len(ul)))  # series branch of ln_mean
            wl, wr = multi_entropy_vars(ul, nd, *gas), multi_entropy_vars(ur, nd, *gas)
            for o in range(1, nd + 1):
                def psi(u, w):
                    return w @ O.flux(cfg, u, o) - multi_entropy(u, nd, *gas) * u[o - 1] / u[nd + 1:].sum()
                fs = O.two_point_flux(cfg, O.FLUX_RANOCHA, ul, ur, o)
                err = abs((wr - wl) @ fs - (psi(ur, wr) - psi(ul, wl)))
                worst = max(worst, err / (1 + np.abs(wr - wl) @ np.abs(fs)))
        if holds:
            assert worst < 1e-13, (gas, worst)
        else:
            assert worst > 1e-6, (gas, worst)


@pytest.mark.parametrize("nd", [1, 2])
def test_multicomponent_single_gas_reduces_to_euler(oracle, nd):
    """One component: the multicomponent fluxes, wave speeds and flux_ranocha
    equal compressible Euler's on the reordered state (rho, rho v, rho e)."""
    O = oracle
    gamma = 1.4
    cm = O.make_cfg(nd, O.EQ_EULER_MULTI, gammas=(gamma,), gas_constants=(0.4,))
    ce = O.make_cfg(nd, O.EQ_EULER, gamma=gamma)
    perm = [nd + 1] + list(range(nd + 1))  # multi -> euler order
    rng = np.random.default_rng(13)
    for _ in range(20):
        ul, ur = multi_state(nd, (gamma,), (0.4,), rng), multi_state(nd, (gamma,), (0.4,), rng)
        el, er = ul[perm], ur[perm]
        np.testing.assert_allclose(O.max_abs_speeds(cm, ul), O.max_abs_speeds(ce, el), rtol=1e-13)
        for o in range(1, nd + 1):
            np.testing.assert_allclose(O.flux(cm, ul, o)[perm], O.flux(ce, el, o), rtol=1e-13, atol=1e-14)
            for k in (O.FLUX_RANOCHA, O.FLUX_LLF, O.FLUX_LLF_NAIVE):
                np.testing.assert_allclose(O.two_point_flux(cm, k, ul, ur, o)[perm],
                                           O.two_point_flux(ce, k, el, er, o), rtol=1e-12, atol=1e-14)


def test_mhd_hlle_structure_and_euler_limit(oracle):
    """flux_hlle (FluxHLL with Einfeldt speeds) of GLM-MHD: consistent, and
    with B = psi = 0 it is the textbook HLLE flux of the Euler system (Roe
    averages with sqrt(rho) weights of v and H, c_roe^2 = (gamma-1)(H_roe -
    |v_roe|^2/2); lambda_min = min(v_l - c_l, v_roe - c_roe), lambda_max =
    max(v_r + c_r, v_roe + c_roe)), computed here independently."""
    O = oracle
    g = 5.0 / 3.0
    rng = np.random.default_rng(40)
    for nd in (2, 3):
        cfg = O.make_cfg(nd, O.EQ_GLM_MHD, gamma=g, c_h=1.0)
        for _ in range(20):
            ul, ur = rng_state_mhd(O, cfg, rng), rng_state_mhd(O, cfg, rng)
            for o in range(1, nd + 1):
                np.testing.assert_allclose(O.two_point_flux(cfg, O.FLUX_HLLE, ul, ul, o), O.flux(cfg, ul, o),
                                           rtol=1e-13, atol=1e-14)
            for u in (ul, ur):
                u[4] -= 0.5 * (u[5:8] @ u[5:8] + u[8] ** 2)
                u[5:9] = 0.0
            for o in range(1, nd + 1):
                def prim(u):
                    rho = u[0]
                    v = u[1:4] / rho
                    p = (g - 1) * (u[4] - 0.5 * rho * v @ v)
                    return rho, v, p, (u[4] + p) / rho
                rl, vl, pl, Hl = prim(ul)
                rr, vr, pr, Hr = prim(ur)
                wl, wr = np.sqrt(rl) / (np.sqrt(rl) + np.sqrt(rr)), np.sqrt(rr) / (np.sqrt(rl) + np.sqrt(rr))
                v_roe, H_roe = wl * vl + wr * vr, wl * Hl + wr * Hr
                c_roe = np.sqrt((g - 1) * (H_roe - 0.5 * v_roe @ v_roe))
                cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
                lmin = min(vl[o - 1] - cl, v_roe[o - 1] - c_roe)
                lmax = max(vr[o - 1] + cr, v_roe[o - 1] + c_roe)
                fl, fr = O.flux(cfg, ul, o), O.flux(cfg, ur, o)
                if lmin >= 0:
                    ref = fl
                elif lmax <= 0:
                    ref = fr
                else:
                    ref = (lmax * fl - lmin * fr + lmin * lmax * (ur - ul)) / (lmax - lmin)
                np.testing.assert_allclose(O.two_point_flux(cfg, O.FLUX_HLLE, ul, ur, o), ref, rtol=1e-12,
                                           atol=1e-13)
