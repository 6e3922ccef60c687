"""GPU parity on the reference's own test configurations, as written.

For each configuration of test/tree_dgsem_{1,2,3}d/ (tests/reference_configs.py)
this checks what the reference test checks (u0 and one rhs! at t = 0 against
Trixi's CPU path, e.g. test/tree_dgsem_3d/euler_ec.jl:52-121), with the FP64
oracle in Trixi's place: u0 to 1e-13 relative (bit for bit when the initial
condition is a host function), du to 1e-12 relative max-norm in FP64, and du
of the mixed-precision kernels within the error stated per case.
"""
import dataclasses

import numpy as np
import pytest

from cases import oracle_cfg, oracle_ic, oracle_mesh, oracle_state, product_semi, rel_max
from reference_configs import BENCHMARK_CASES, REFERENCE_CASES

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 2e-5
# Mixed-precision bars above MIXED_TOL, each with its reason.  At a
# manufactured state with its source term du is the small difference of a flux
# divergence and a source much larger than du, so FP32 rounding of the fluxes
# is amplified relative to |du| by that ratio.
MIXED_TOL_CASE = {
    "1d/euler_source_terms": 2e-3,
    "1d/euler_source_terms_nonperiodic": 2e-3,
    "1d/eulerquasi_source_terms": 2e-3,
    "1d/burgers_basic": 2e-3,
    "2d/euler_source_terms": 2e-3,
    "2d/euler_source_terms_nonperiodic": 2e-3,
    "3d/euler_source_terms": 2e-3,
    "3d/euler_mortar": 2e-3,
    "3d/euler_convergence": 2e-3,
    "1d/hypdiff_nonperiodic": 2e-3,
    "2d/hypdiff_nonperiodic": 2e-3,
    "3d/hypdiff_nonperiodic": 2e-3,
    "1d/hypdiff_harmonic_nonperiodic": 2e-3,
    # (the Alfven wave's steady density and pressure: du is a cancellation of
    # O(1) flux terms in every variable)
    "1d/mhd_alfven_wave": 2e-3,
    "2d/mhd_alfven_wave": 2e-3,
    "2d/mhd_alfven_wave_mortar": 2e-3,
    "3d/mhd_alfven_wave": 2e-3,
    "3d/mhd_alfven_wave_mortar": 2e-3,
    "bench/mhd3d_alfven_level4": 2e-3,
    # advection of the convergence-test state: linear, no cancellation beyond
    # the flux difference itself
}

# Initial states that are exact steady states of the discretisation (the
# harmonic state at t = 0: phi = 5, q = 0): du vanishes up to round-off, so
# its error is measured relative to the state's magnitude instead.
STEADY = ("1d/hypdiff_harmonic_nonperiodic",)
# 3D advection with velocity (0.2, -0.7, 0.5) of the convergence-test state
# sin(pi (x + y + z)): the velocity components sum to zero, the exact du
# vanishes, and du is only the discretisation error (|du| = 0.008 against
# direction terms of size 2).  Its error is measured relative to the du of
# the same state advected with |a| = (0.2, 0.7, 0.5), where the direction
# terms add up instead of cancelling.
CANCELLING = ("3d/advection_basic", "3d/advection_mortar")

ALL = REFERENCE_CASES + BENCHMARK_CASES


def _err(O, case, du, du_ref, u_ref):
    if case.name in STEADY:
        return float(np.abs(du - du_ref).max() / np.abs(u_ref).max())
    if case.name in CANCELLING:
        adding = dataclasses.replace(case, advection=tuple(abs(a) for a in case.advection))
        scale = np.abs(O.rhs(oracle_cfg(O, adding), oracle_mesh(O, adding), u_ref, t=0.0)).max()
        return float(np.abs(du - du_ref).max() / scale)
    return rel_max(du, du_ref)


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _run(th, O, torch, case, precision):
    cfg = oracle_cfg(O, case)
    mesh = oracle_mesh(O, case)
    u_ref = oracle_state(O, case, mesh)
    du_ref = O.rhs(cfg, mesh, u_ref, t=0.0)
    semi = product_semi(th, case, precision)
    try:
        u = th.compute_coefficients(0.0, semi)
        u0 = th.wrap_array_native(u, semi).cpu().numpy()
        du_dev = torch.empty_like(u)
        th.rhs_hip_(du_dev, u, semi, 0.0)
        du = th.wrap_array_native(du_dev, semi).cpu().numpy()
        torch.cuda.synchronize()
    finally:
        semi.close()
    return u0, u_ref, du, du_ref


@pytest.mark.parametrize("case", ALL, ids=lambda c: c.name)
def test_reference_configuration_fp64(th, oracle, torch, case):
    u0, u_ref, du, du_ref = _run(th, oracle, torch, case, "fp64")
    if case.ic_fn is not None:
        np.testing.assert_array_equal(u0, u_ref)
    else:
        np.testing.assert_allclose(u0, u_ref, rtol=1e-13, atol=1e-14 * max(1.0, np.abs(u_ref).max()))
    err = _err(oracle, case, du, du_ref, u_ref)
    print(f"{case.name}: fp64 rel max-norm {err:.3e}", flush=True)
    assert np.all(np.isfinite(du))
    assert err <= FP64_TOL, err


@pytest.mark.parametrize("case", ALL, ids=lambda c: c.name)
def test_reference_configuration_mixed(th, oracle, torch, case):
    _, u_ref, du, du_ref = _run(th, oracle, torch, case, "mixed")
    err = _err(oracle, case, du, du_ref, u_ref)
    l2 = float(np.linalg.norm(du - du_ref) / np.linalg.norm(du_ref))
    tol = MIXED_TOL_CASE.get(case.name, MIXED_TOL)
    print(f"{case.name}: mixed rel max-norm {err:.3e}, rel L2 {l2:.3e} (tol {tol:g})", flush=True)
    assert np.all(np.isfinite(du))
    assert err <= tol, err
