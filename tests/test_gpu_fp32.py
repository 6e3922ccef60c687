"""GPU parity of the FP32-state precision (Trixi's RealT = uEltype = Float32,
the precision of the reference's benchmarks, benchmark/euler_ec_3d.jl:6 and
benchmark/euler_shock_2d.jl:6) against the FP64 oracle.

The oracle evaluates the same FP32 state (rounded, then widened exactly), so
the comparison measures the FP32 arithmetic alone.  Stated bars: rhs! du
relative max-norm 1e-4 (FP32 arithmetic of O(1) flux terms whose differences
give du, with FP32 logarithms in the logarithmic means), and the increment
u_{n+1} - u_n of one CarpenterKennedy2N54 step 1e-4 relative max-norm."""
import numpy as np
import pytest

from cases import Case, oracle_cfg, oracle_mesh, oracle_state, product_semi, rel_max

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-4

FP32_CASES = [
    # the reference's benchmark configurations (test/tree_dgsem_3d/euler_ec.jl
    # and test/tree_dgsem_2d/euler_shock.jl shapes at their benchmark level 3)
    Case("euler3d_ec_blast", 3, "euler", 3, "fd", "ranocha", "ranocha", ic="weak_blast_wave"),
    Case("euler2d_shock_shima_blast", 2, "euler", 3, "hg", "shima_etal", "llf", "llf", ic="weak_blast_wave"),
    # the measured 3D configuration's shape, the 2D flux-differencing
    # configuration at its size, and the 2D shock-capturing one
    Case("euler3d_fd_ranocha_llf_smooth", 3, "euler", 3, "fd", "ranocha", ic="smooth_periodic"),
    Case("euler2d_fd_ranocha_llf_random", 2, "euler", 5, "fd", "ranocha"),
    Case("euler2d_ec_random", 2, "euler", 4, "fd", "ranocha", "ranocha"),
    Case("euler2d_hg_ranocha_blast_seeded", 2, "euler", 5, "hg", "ranocha", ic="weak_blast_wave_seeded"),
]


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _fp32_state(O, case):
    mesh = oracle_mesh(O, case)
    u = oracle_state(O, case, mesh).astype(np.float32).astype(np.float64)
    return mesh, u


@pytest.mark.parametrize("case", FP32_CASES, ids=lambda c: c.name)
def test_fp32_state_rhs_matches_oracle(th, oracle, torch, case):
    O = oracle
    cfg = oracle_cfg(O, case)
    mesh, u = _fp32_state(O, case)
    du_ref = O.rhs(cfg, mesh, u)
    semi = product_semi(th, case, "fp32")
    try:
        ud = th.pack(torch.from_numpy(u).cuda(), semi)
        assert ud.dtype == torch.float32
        dud = torch.empty_like(ud)
        th.rhs_hip_(dud, ud, semi, 0.0)
        du = th.wrap_array_native(dud, semi).double().cpu().numpy()
    finally:
        semi.close()
    err = rel_max(du, du_ref)
    l2 = float(np.linalg.norm(du - du_ref) / np.linalg.norm(du_ref))
    print(f"{case.name}: fp32 state rel max-norm {err:.3e}, rel L2 {l2:.3e}", flush=True)
    assert np.all(np.isfinite(du))
    assert err <= FP32_TOL, err


@pytest.mark.parametrize("form", ["default", "line"])
@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("case", [FP32_CASES[0], FP32_CASES[5]], ids=lambda c: c.name)
def test_fp32_state_rk_step_matches_oracle(th, oracle, torch, case, use_graph, form):
    """Two CK2N54 steps in FP32 state against the oracle's FP64 steps from the
    same FP32 initial state (graph replays checked after each step); `line`
    forces the line form of the fused stage (the form of large meshes) on
    these small ones, which otherwise take the node-parallel form."""
    import os
    O = oracle
    cfg = oracle_cfg(O, case)
    mesh, u = _fp32_state(O, case)
    old = os.environ.get("TRIXI_HIP_NP_LANES")
    if form == "line":
        os.environ["TRIXI_HIP_NP_LANES"] = "0"
    try:
        semi = product_semi(th, case, "fp32")
    finally:
        if old is None:
            os.environ.pop("TRIXI_HIP_NP_LANES", None)
        else:
            os.environ["TRIXI_HIP_NP_LANES"] = old
    try:
        ud = th.pack(torch.from_numpy(u).cuda(), semi)
        kd = torch.zeros_like(ud)
        dt = 0.5 * O.max_dt(cfg, mesh, u)
        k = np.zeros_like(u)
        uo = u.copy()
        stream = torch.cuda.Stream()
        stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(stream):
            for step in range(2):
                th.rk_step(ud, kd, step * dt, dt, semi, use_graph=use_graph, stream=stream)
                torch.cuda.current_stream().synchronize()
                O.ck2n54_step(cfg, mesh, uo, k, step * dt, dt)
                got = th.wrap_array_native(ud, semi).double().cpu().numpy()
                err = rel_max(got - u, uo - u)
                print(f"{case.name} graph={use_graph} step {step}: fp32 rel err of increment {err:.3e}", flush=True)
                assert err <= FP32_TOL, err
        torch.cuda.synchronize()
    finally:
        semi.close()


def test_fp32_state_api(th, torch):
    """RealT=Float32 selects the FP32 state; its vectors are float32, the FP64
    entry points reject them, the callbacks widen them, and configurations
    without FP32 kernels are rejected at construction."""
    c = Case("euler3d_fd", 3, "euler", 2, "fd", "ranocha", ic="smooth_periodic")
    semi = product_semi(th, c, "fp32")
    try:
        u = th.compute_coefficients(0.0, semi)
        assert u.dtype == torch.float32
        u64 = th.compute_coefficients(0.0, product_semi(th, c, "fp64"))
        assert torch.equal(u, u64.float())
        with pytest.raises(th.TrixiHipError):
            th.rhs_hip_(torch.empty_like(u64), u64, semi, 0.0)  # a float64 vector
        lib = th._lib.lib()
        import ctypes
        du = torch.empty_like(u)
        rc = lib.trixi_hip_rhs(semi.ctx, ctypes.c_void_p(du.data_ptr()), ctypes.c_void_p(u.data_ptr()), 0.0,
                               ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        assert rc == th._lib.EINVAL
        assert th.max_dt(u, semi) == pytest.approx(th.max_dt(u.double(), product_semi(th, c, "fp64")), rel=1e-6)
        with pytest.raises(th.TrixiHipError):
            th.rhs_vjp(torch.empty_like(u), u, torch.ones_like(u), semi)
    finally:
        semi.close()
    with pytest.raises(th.TrixiHipError):
        product_semi(th, Case("mhd3d", 3, "mhd", 1, "fd", "hindenlang_gassner", cmin=-1.0, cmax=1.0), "fp32")
