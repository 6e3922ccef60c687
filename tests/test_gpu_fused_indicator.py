"""Fused shock-capturing RK stages: on meshes that take the node-parallel
rhs! kernel, each fused Runge-Kutta stage also computes the raw
Hennemann-Gassner blending factors of its output (the next stage's input), so
a CarpenterKennedy2N54 step launches the indicator kernel once instead of
five times.  The fused factors come from the same values through the same
steps as the indicator kernel's, so a step equals the unfused one
(TRIXI_HIP_FUSE_INDICATOR=0: an indicator launch before every stage) bit for
bit, eagerly and from the hipGraph, in every precision built for shock
capturing."""
import os

import pytest

from cases import Case, product_semi

pytestmark = pytest.mark.gpu

CASES = [
    # (case, precision)
    (Case("euler2d_hg_ranocha_blast_seeded_l4", 2, "euler", 4, "hg", "ranocha", ic="weak_blast_wave_seeded"),
     "fp64"),
    (Case("euler2d_hg_ranocha_blast_seeded_l4", 2, "euler", 4, "hg", "ranocha", ic="weak_blast_wave_seeded"),
     "mixed"),
    # the 2D shock-capturing benchmark configuration (level 6, 4096 elements)
    (Case("euler2d_hg_ranocha_blast_seeded_l6", 2, "euler", 6, "hg", "ranocha", ic="weak_blast_wave_seeded"),
     "mixed"),
    (Case("euler2d_hg_shima_blast_l4", 2, "euler", 4, "hg", "shima_etal", ic="weak_blast_wave_seeded"), "fp32"),
    (Case("euler3d_hg_ranocha_ranocha_l2", 3, "euler", 2, "hg", "ranocha", "ranocha", "ranocha",
          ic="weak_blast_wave_seeded"), "fp64"),
    (Case("euler2d_hg_pressure_l4", 2, "euler", 4, "hg", "ranocha", ic="weak_blast_wave_seeded",
          indicator="pressure"), "fp64"),
]


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _semi(th, case, precision, fuse):
    old = os.environ.get("TRIXI_HIP_FUSE_INDICATOR")
    os.environ["TRIXI_HIP_FUSE_INDICATOR"] = "1" if fuse else "0"
    try:
        return product_semi(th, case, precision)
    finally:
        if old is None:
            del os.environ["TRIXI_HIP_FUSE_INDICATOR"]
        else:
            os.environ["TRIXI_HIP_FUSE_INDICATOR"] = old


@pytest.mark.parametrize("use_graph", [False, True], ids=["eager", "graph"])
@pytest.mark.parametrize("case,precision", CASES, ids=lambda x: x.name if hasattr(x, "name") else x)
def test_fused_indicator_steps_equal_unfused(th, torch, case, precision, use_graph):
    out = {}
    stream = torch.cuda.Stream()
    for fuse in (True, False):
        semi = _semi(th, case, precision, fuse)
        u0 = th.compute_coefficients(0.0, semi)
        dt = 0.5 * th.max_dt(u0, semi)
        u, k = u0.clone(), torch.zeros_like(u0)
        with torch.cuda.stream(stream):
            for i in range(3):
                th.rk_step(u, k, i * dt, dt, semi, use_graph=use_graph, stream=stream)
        torch.cuda.synchronize()
        out[fuse] = (u.clone(), k.clone(), u0.clone())
        semi.close()
    (uf, kf, u0), (uu, ku, _) = out[True], out[False]
    print(f"{case.name} {precision} graph={use_graph}: max |u_fused - u_unfused| "
          f"{float((uf - uu).abs().max()):.3e}, changed {float((uf - u0).abs().max()):.3e}", flush=True)
    assert torch.isfinite(uf).all()
    assert torch.equal(uf, uu)
    assert torch.equal(kf, ku)


# shock-capturing adjoints are built with LLF surface and subcell fluxes
ADJ_CASES = [CASES[0][0], CASES[2][0],
             Case("euler3d_hg_ranocha_llf_l2", 3, "euler", 2, "hg", "ranocha", ic="weak_blast_wave_seeded")]


@pytest.mark.parametrize("case", ADJ_CASES, ids=lambda c: c.name)
def test_fused_indicator_adjoint_step_equals_unfused(th, torch, case):
    """The reverse RK step recomputes its stage inputs with the fused stages
    and keeps their blending factors for the stage VJPs (no indicator launch
    per stage); the cotangent equals the unfused one bit for bit."""
    out = {}
    for fuse in (True, False):
        semi = _semi(th, case, "fp64", fuse)
        th.enable_adjoint(semi)
        u0 = th.compute_coefficients(0.0, semi)
        dt = 0.5 * th.max_dt(u0, semi)
        g = torch.Generator(device=u0.device).manual_seed(20261202)
        ubar = torch.rand(u0.shape, generator=g, device=u0.device, dtype=u0.dtype) - 0.5
        for i in range(2):
            th.rk_step_vjp(u0, ubar, 0.1 + i * dt, dt, semi)
        torch.cuda.synchronize()
        out[fuse] = ubar.clone()
        semi.close()
    print(f"{case.name}: |ubar| {float(out[True].abs().max()):.3e}, "
          f"max diff {float((out[True] - out[False]).abs().max()):.3e}", flush=True)
    assert torch.isfinite(out[True]).all()
    assert torch.equal(out[True], out[False])
