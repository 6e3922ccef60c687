"""The node-parallel form of the rhs! kernel (small periodic meshes) against
the line form: du, and k and u_out of a fused Runge-Kutta stage, in FP64 and
mixed precision.

Every node receives the same flux terms in the same order in both forms;
they differ only where the compiler fuses a multiply and an add in one code
shape and not the other (measured: the energy flux everywhere, the momentum
at the face nodes), so they agree to a few units in the last place of the
largest flux terms (relative max-norm of du 1e-13 in FP64, 2e-5 in mixed
precision, whose error against the FP64 oracle is itself up to 2e-5).  Each form is also checked against the oracle by the parity
tests (the node-parallel form on the small meshes, the line form on the
full-size configurations)."""
import os

import numpy as np
import pytest

from cases import Case, product_semi

pytestmark = pytest.mark.gpu

NP_CASES = [
    Case("euler2d_fd_ranocha_llf", 2, "euler", 4, "fd", "ranocha"),
    Case("euler2d_fd_shima_llfnaive", 2, "euler", 3, "fd", "shima_etal", "llf_naive"),
    Case("euler2d_hg_ranocha_blast_seeded", 2, "euler", 4, "hg", "ranocha", ic="weak_blast_wave_seeded",
         alpha_min=0.0),
    Case("euler3d_fd_ranocha_llf", 3, "euler", 2, "fd", "ranocha", ic="smooth_periodic"),
    Case("euler3d_hg_ranocha_ranocha", 3, "euler", 2, "hg", "ranocha", "ranocha", "ranocha",
         ic="weak_blast_wave_seeded"),
    Case("mhd3d_fd_hg_llf_alfven", 3, "mhd", 2, "fd", "hindenlang_gassner", cmin=-1.0, cmax=1.0,
         ic="convergence_test", gamma=5 / 3),
    Case("mhd2d_hg_ec_blast_seeded", 2, "mhd", 3, "hg", "hindenlang_gassner", "hindenlang_gassner",
         "hindenlang_gassner", ic="weak_blast_wave_seeded", gamma=1.4),
    Case("eulermulti2d_distinct_random", 2, "eulermulti", 3, "fd", "ranocha", "llf",
         gammas=(1.4, 1.648), gas_constants=(0.287, 1.578)),
    Case("euler2d_fd_ranocha_llf_p1", 2, "euler", 3, "fd", "ranocha", polydeg=1),
    Case("euler3d_fd_ranocha_llf_p2", 3, "euler", 2, "fd", "ranocha", polydeg=2),
    Case("mhd2d_hg_ec_blast_p4", 2, "mhd", 3, "hg", "hindenlang_gassner", "hindenlang_gassner",
         "hindenlang_gassner", polydeg=4, ic="weak_blast_wave", gamma=1.4),
]


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _semi(th, case, precision, node_parallel):
    """A semidiscretization whose context picks the node-parallel form for
    every mesh (threshold above any lane count) or never (threshold 0)."""
    old = os.environ.get("TRIXI_HIP_NP_LANES")
    os.environ["TRIXI_HIP_NP_LANES"] = str(1 << 40) if node_parallel else "0"
    try:
        return product_semi(th, case, precision)
    finally:
        if old is None:
            del os.environ["TRIXI_HIP_NP_LANES"]
        else:
            os.environ["TRIXI_HIP_NP_LANES"] = old


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", NP_CASES, ids=lambda c: c.name)
def test_node_parallel_equals_line_form(th, torch, case, precision):
    out = {}
    for npf in (True, False):
        semi = _semi(th, case, precision, npf)
        try:
            u = th.compute_coefficients(0.0, semi)
            du = torch.empty_like(u)
            th.rhs_hip_(du, u, semi, 0.0)
            # one fused RK stage (stage 1: k and u_out both depend on du)
            k = torch.linspace(-1.0, 1.0, u.numel(), dtype=torch.float64, device="cuda")
            uo = torch.empty_like(u)
            th.rk_stage(u, k, uo, 1, 0.0, 1e-3, semi)
            torch.cuda.synchronize()
            out[npf] = (du.cpu().numpy(), k.cpu().numpy(), uo.cpu().numpy())
        finally:
            semi.close()
    tol = 1e-13 if precision == "fp64" else 2e-5
    for name, a, b in zip(("du", "k", "u_out"), out[True], out[False]):
        assert np.all(np.isfinite(a)), name
        err = float(np.abs(a - b).max() / np.abs(b).max())
        print(f"{case.name} {precision} {name}: node-parallel vs line rel max-norm {err:.3e}", flush=True)
        assert err <= tol, (name, err)


# the cases built with an FP32 state (Trixi's RealT = uEltype = Float32)
FP32_NP_CASES = [NP_CASES[0], NP_CASES[2], NP_CASES[3]]


@pytest.mark.parametrize("case", FP32_NP_CASES, ids=lambda c: c.name)
def test_node_parallel_equals_line_form_fp32_state(th, torch, case):
    """The same with an FP32 state: du, and u and k after one CK2N54 step
    of fused stages (the line form's update requests u and k ahead of its
    last barrier) against the node-parallel form, within the FP32-state
    parity tolerance (1e-4, each form's bound against the FP64 oracle).  k is
    measured against the scale of its terms, dt max|du|:
    the step's k is a cancellation of the stages' terms, so its own maximum
    is smaller and would magnify the forms' FP32 rounding differences."""
    out = {}
    dt = 1e-3
    for npf in (True, False):
        semi = _semi(th, case, "fp32", npf)
        try:
            u = th.compute_coefficients(0.0, semi)
            assert u.dtype == torch.float32
            du = torch.empty_like(u)
            th.rhs_hip_(du, u, semi, 0.0)
            k = torch.zeros_like(u)
            th.rk_step(u, k, 0.0, dt, semi, use_graph=False)
            torch.cuda.synchronize()
            out[npf] = (du.double().cpu().numpy(), u.double().cpu().numpy(), k.double().cpu().numpy())
        finally:
            semi.close()
    scale = {"du": np.abs(out[False][0]).max(), "u": np.abs(out[False][1]).max(),
             "k": dt * np.abs(out[False][0]).max()}
    for name, a, b in zip(("du", "u", "k"), out[True], out[False]):
        assert np.all(np.isfinite(a)), name
        err = float(np.abs(a - b).max() / scale[name])
        print(f"{case.name} fp32 state {name}: node-parallel vs line rel max-norm {err:.3e}", flush=True)
        assert err <= 1e-4, (name, err)
