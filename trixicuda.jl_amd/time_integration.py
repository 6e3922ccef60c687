"""CarpenterKennedy2N54 time integration on the device.

The reference hands the ODE problem to OrdinaryDiffEq
(`solve(ode, CarpenterKennedy2N54(williamson_condition = false); dt, callback)`,
e.g. examples/euler_ec_3d.jl:53-55).  Here each step is one library call:
five fused rhs+register-update launches, replayed from a hipGraph.
"""
from . import _lib as L
from .callbacks import AnalysisCallback, StepsizeCallback
from .save_solution import SaveSolutionCallback
from .semidiscretization import _ptr


class CarpenterKennedy2N54:
    def __init__(self, williamson_condition=False):
        self.williamson_condition = williamson_condition


class Solution:
    def __init__(self, u, t, nsteps, semi):
        self.u, self.t, self.nsteps, self.semi = u, t, nsteps, semi

    def __repr__(self):
        return f"Solution(t={self.t}, nsteps={self.nsteps})"


def rk_step(u, k, t, dt, semi, use_graph=True, stream=None):
    """Advance u by one CK2N54 step in place; k is the 2N register."""
    import ctypes
    import torch
    semi.check_vector(u, "u")
    semi.check_vector(k, "k")
    s = stream if stream is not None else torch.cuda.current_stream()
    if use_graph and s.cuda_stream == 0:
        raise L.TrixiHipError(L.EINVAL, "graph replay needs a non-default stream")
    fn = L.lib().trixi_hip_rk_step_f32 if semi.fp32_state else L.lib().trixi_hip_rk_step
    L.check(fn(semi.ctx, _ptr(u), _ptr(k), float(t), float(dt), 1 if use_graph else 0, ctypes.c_void_p(s.cuda_stream)))


def rk_stage(u, k, u_out, stage, t, dt, semi, stream=None):
    """Stage `stage` (0..4) of the CK2N54 step starting at t: k <- A_s k +
    dt rhs(u, t + c_s dt), u_out <- u + B_s k (u_out must not alias u)."""
    import ctypes
    import torch
    for x, n in ((u, "u"), (k, "k"), (u_out, "u_out")):
        semi.check_vector(x, n)
    s = stream if stream is not None else torch.cuda.current_stream()
    L.check(L.lib().trixi_hip_rk_stage(semi.ctx, _ptr(u), _ptr(k), _ptr(u_out), int(stage), float(t), float(dt),
                                       ctypes.c_void_p(s.cuda_stream)))


def synchronize(stream=None):
    """Wait for the stream's work; asynchronous device errors raise here."""
    import ctypes
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    L.check(L.lib().trixi_hip_sync(ctypes.c_void_p(s.cuda_stream)))


def select_step(t, h, t1, save_times=()):
    """Step from t with the proposed size h: shortened to end at t1, or at the
    earliest save time strictly inside (t, t + h].  Returns (h, t_next), where
    t_next is exactly that end time for a shortened step (no rounding of
    t + h) and t + h otherwise."""
    end = t1 if t + h >= t1 else None
    for ts in save_times:
        if ts is not None and t < ts < t + h and (end is None or ts < end):
            end = ts
    if end is None:
        return h, t + h
    return end - t, end


def solve(ode, alg=None, dt=None, callback=(), use_graph=True, maxiters=10**7):
    """Integrate ode over its tspan; returns a Solution with the final state."""
    import torch
    if alg is not None and not isinstance(alg, CarpenterKennedy2N54):
        raise NotImplementedError("only CarpenterKennedy2N54 is built")
    semi = ode.p
    cbs = list(callback) if isinstance(callback, (list, tuple)) else [callback]
    stepsize = next((c for c in cbs if isinstance(c, StepsizeCallback)), None)
    analysis = [c for c in cbs if isinstance(c, AnalysisCallback)]
    saves = [c for c in cbs if isinstance(c, SaveSolutionCallback)]
    if dt is None and stepsize is None:
        raise ValueError("give dt or a StepsizeCallback")
    t0, t1 = ode.tspan
    u = ode.u0.clone()
    k = torch.zeros_like(u)
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    t, nsteps, h = t0, 0, 0.0
    with torch.cuda.stream(stream):
        for c in saves:
            c.initialize(u, t, stepsize.dt(u, semi) if stepsize is not None else dt, semi)
        while t < t1 and nsteps < maxiters:
            h = stepsize.dt(u, semi) if stepsize is not None else dt
            h, t_next = select_step(t, h, t1, [c.next_time for c in saves])
            rk_step(u, k, t, h, semi, use_graph=use_graph, stream=stream)
            t = t_next
            nsteps += 1
            for a in analysis:
                a(u, t, nsteps)
            for c in saves:
                c(u, t, h, nsteps, semi)
        for a in analysis:
            a(u, t, nsteps, final=True)
        for c in saves:
            c(u, t, h, nsteps, semi, final=True)
    torch.cuda.current_stream().wait_stream(stream)
    stream.synchronize()
    return Solution(u, t, nsteps, semi)
