"""GPU-resident callbacks: max_dt, error norms, integrals, state checks.

Twins of the reference's max_dt (src/callbacks_step/stepsize_dg_{1,2,3}d.jl),
calc_error_norms / integrate (src/callbacks_step/analysis_dg_{1,2,3}d.jl)
and the StepsizeCallback / AnalysisCallback usage of its examples.  The
reductions run on the device; only scalars come back to the host.
"""
import ctypes

import numpy as np

from . import _lib as L
from .semidiscretization import _ptr, _stream


def max_dt(u_ode, semi):
    u_ode = semi.as_f64(u_ode, "u_ode")
    out = ctypes.c_double()
    L.check(L.lib().trixi_hip_max_dt(semi.ctx, _ptr(u_ode), ctypes.byref(out), _stream()))
    return out.value


def calc_error_norms(u_ode, t, semi, initial_condition=None):
    """(l2, linf) per variable against the initial condition at time t."""
    u_ode = semi.as_f64(u_ode, "u_ode")
    ic = initial_condition or semi.initial_condition
    if not hasattr(ic, "ic_id"):
        raise NotImplementedError("error norms are computed on the device against a built initial condition")
    l2 = np.zeros(semi.nvariables)
    linf = np.zeros(semi.nvariables)
    L.check(L.lib().trixi_hip_error_norms(semi.ctx, _ptr(u_ode), ic.ic_id, float(t), ic.seed,
                                          l2.ctypes.data_as(ctypes.c_void_p),
                                          linf.ctypes.data_as(ctypes.c_void_p), _stream()))
    return l2, linf


def integrate(u_ode, semi):
    u_ode = semi.as_f64(u_ode, "u_ode")
    out = np.zeros(semi.nvariables)
    L.check(L.lib().trixi_hip_integrate(semi.ctx, _ptr(u_ode), out.ctypes.data_as(ctypes.c_void_p), _stream()))
    return out


def blending_factors(u_ode, semi):
    """Hennemann-Gassner blending factors (after smoothing) as a device tensor."""
    import torch
    u_ode = semi.as_f64(u_ode, "u_ode")
    alpha = torch.empty(semi.nelements, dtype=torch.float64, device=u_ode.device)
    L.check(L.lib().trixi_hip_indicator(semi.ctx, _ptr(u_ode), _ptr(alpha), _stream()))
    return alpha


def check_state(u_ode, semi):
    """(nbad, flags): per-element flags bit0 non-finite, bit1 rho<=0, bit2 p<=0."""
    import torch
    u_ode = semi.as_f64(u_ode, "u_ode")
    flags = torch.zeros(semi.nelements, dtype=torch.int32, device=u_ode.device)
    nbad = ctypes.c_int64()
    rc = L.lib().trixi_hip_check_state(semi.ctx, _ptr(u_ode), _ptr(flags), ctypes.byref(nbad), _stream())
    if rc not in (L.OK, L.ESTATE):
        L.check(rc)
    return nbad.value, flags


class StepsizeCallback:
    def __init__(self, cfl=1.0):
        self.cfl = float(cfl)

    def dt(self, u, semi):
        return self.cfl * max_dt(u, semi)


class AnalysisCallback:
    """Records (t, l2, linf) every `interval` steps and at the final time."""

    def __init__(self, semi, interval=100):
        self.semi, self.interval = semi, int(interval)
        self.history = []

    def __call__(self, u, t, step, final=False):
        if final or (self.interval > 0 and step % self.interval == 0):
            l2, linf = calc_error_norms(u, t, self.semi)
            self.history.append((t, l2, linf))
