"""Adjoint of the semidiscretization: VJP of rhs!, reverse RK steps, and
multi-step gradients with RK-step checkpointing.

TrixiCUDA.jl has no adjoint; these functions back the rrule that the Julia
binding would attach to rhs_hip! (a ChainRulesCore.rrule) and, in Python, a
torch.autograd.Function, so rhs_hip_ composes with reverse-mode AD.  All
device work is in libtrixihip.so (trixi_hip_rhs_vjp, trixi_hip_rk_step_vjp).
"""
import ctypes

from . import _lib as L
from .semidiscretization import _ptr, _stream


def enable_adjoint(semi):
    """Allocate the adjoint work buffers of a semidiscretization (once)."""
    L.check(L.lib().trixi_hip_enable_adjoint(semi.ctx))


def rhs_vjp(ubar, u, vbar, semi, t=0.0, alpha=1.0, beta=0.0):
    """ubar = beta * ubar + (d rhs/du)(u)^T (alpha * vbar), on the current stream."""
    for x, n in ((ubar, "ubar"), (u, "u"), (vbar, "vbar")):
        semi.check_vector(x, n)
    L.check(L.lib().trixi_hip_rhs_vjp(semi.ctx, _ptr(ubar), _ptr(u), _ptr(vbar), float(t), float(alpha),
                                      float(beta), _stream()))
    return ubar


def rk_step_vjp(u_n, ubar, t, dt, semi):
    """Reverse one CarpenterKennedy2N54 step from its initial state u_n: ubar
    (cotangent of u_{n+1}) is overwritten by the cotangent of u_n."""
    semi.check_vector(u_n, "u_n")
    semi.check_vector(ubar, "ubar")
    L.check(L.lib().trixi_hip_rk_step_vjp(semi.ctx, _ptr(u_n), _ptr(ubar), float(t), float(dt), _stream()))
    return ubar


class Trajectory:
    """Forward CK2N54 trajectory that keeps one checkpoint per step."""

    def __init__(self, semi, u0, times, checkpoints, u_final):
        self.semi, self.u0, self.times, self.checkpoints, self.u_final = semi, u0, times, checkpoints, u_final


def forward_with_checkpoints(semi, u0, dt, nsteps, t0=0.0, use_graph=False):
    """nsteps CK2N54 steps of size dt (a number or a per-step list), storing
    each step's initial state (RK-step checkpointing; stage inputs are
    recomputed in the reverse sweep)."""
    import torch
    from .time_integration import rk_step

    dts = list(dt) if hasattr(dt, "__len__") else [float(dt)] * nsteps
    u = u0.clone()
    k = torch.zeros_like(u)
    cps, times = [], []
    t = t0
    for n in range(nsteps):
        cps.append(u.clone())
        times.append(t)
        rk_step(u, k, t, dts[n], semi, use_graph=use_graph, stream=torch.cuda.current_stream())
        t += dts[n]
    traj = Trajectory(semi, u0, times, cps, u)
    traj.dts = dts
    return traj


def backward(traj, ubar_final):
    """Cotangent of u0 given the cotangent of the final state."""
    enable_adjoint(traj.semi)
    ubar = ubar_final.clone()
    for n in range(len(traj.checkpoints) - 1, -1, -1):
        rk_step_vjp(traj.checkpoints[n], ubar, traj.times[n], traj.dts[n], traj.semi)
    return ubar


def l2_norm_and_gradient(u, semi):
    """Discrete L2 norm sqrt(sum_e J_e^-d sum_nodes w u^2 / |Omega|) over all
    variables (the quadrature of Trixi's analysis) and its gradient."""
    import torch
    w = quadrature_weights(semi, u.device)
    s = (w * u * u).sum() / semi.mesh.total_volume
    norm = torch.sqrt(s)
    return norm, w * u / (semi.mesh.total_volume * norm)


def quadrature_weights(semi, device):
    """J_e * w_i w_j w_k per value, native layout."""
    import numpy as np
    import torch
    b = semi.solver.basis
    nd, N = semi.mesh.ndims, b.nnodes
    wn = np.ones(N ** nd)
    for n in range(N ** nd):
        r = n
        for d in range(nd):
            wn[n] *= b.weights[r % N]
            r //= N
    vol = (1.0 / semi.mesh.inverse_jacobian) ** nd
    w = np.einsum("n,e->ne", wn, vol)[None].repeat(semi.nvariables, axis=0)
    return torch.from_numpy(np.ascontiguousarray(w).reshape(-1)).to(device)


def _make_rhs_function():
    import torch

    class RhsFunction(torch.autograd.Function):
        """du = rhs(u) with a VJP backward: the Python counterpart of an rrule."""

        @staticmethod
        def forward(ctx, u, semi, t):
            from .semidiscretization import rhs_hip_
            du = torch.empty_like(u)
            rhs_hip_(du, u.detach().contiguous(), semi, t)
            ctx.semi, ctx.t = semi, t
            ctx.save_for_backward(u.detach())
            return du

        @staticmethod
        def backward(ctx, g):
            (u,) = ctx.saved_tensors
            ubar = torch.empty_like(u)
            rhs_vjp(ubar, u, g.contiguous(), ctx.semi, ctx.t)
            return ubar, None, None

    return RhsFunction


_RHS_FN = None


def rhs(u, semi, t=0.0):
    """Differentiable rhs!: rhs(u) as a torch op whose backward is the VJP kernel."""
    global _RHS_FN
    if _RHS_FN is None:
        _RHS_FN = _make_rhs_function()
    return _RHS_FN.apply(u, semi, float(t))
