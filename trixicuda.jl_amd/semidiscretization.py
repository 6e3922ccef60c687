"""SemidiscretizationHyperbolicHIP, semidiscretizeHIP and rhs_hip_.

HIP-named twins of SemidiscretizationHyperbolicGPU
(src/semidiscretization/semidiscretization_hyperbolic.jl:60-87),
semidiscretizeGPU and rhs_gpu! (src/solvers/solvers.jl:18-56).  State
vectors are torch tensors on the GPU in the library's native layout
(element innermost); torch is only the device-memory and stream plumbing,
every computation is a call into libtrixihip.so.
"""
import ctypes

import numpy as np

from . import _lib as L


def _torch():
    import torch
    return torch


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class SemidiscretizationHyperbolicHIP:
    def __init__(self, mesh, equations, initial_condition, solver, source_terms=None,
                 boundary_conditions="periodic"):
        if mesh.ndims != equations.ndims:
            raise ValueError("mesh and equations dimensions differ")
        if source_terms is not None:
            raise NotImplementedError("source terms are not built on the HIP path")
        if boundary_conditions not in ("periodic", None):
            raise NotImplementedError("only boundary_condition_periodic is supported")
        self.mesh, self.equations, self.initial_condition, self.solver = mesh, equations, initial_condition, solver
        self.nvariables = equations.nvariables
        self.nnodes = solver.basis.nnodes
        self.nelements = mesh.nelements
        self.nnodes_per_element = self.nnodes ** mesh.ndims
        self.ndofs = self.nelements * self.nnodes_per_element
        self.nvalues = self.ndofs * self.nvariables
        self.config = solver.config(equations, ic_length=mesh.length)
        self._node_coordinates = np.ascontiguousarray(mesh.node_coordinates(solver.basis.nodes))
        self._nbr = np.ascontiguousarray(mesh.neighbor_ids, dtype=np.int32)
        self._ori = np.ascontiguousarray(mesh.orientations, dtype=np.int32)
        self._invj = np.ascontiguousarray(mesh.inverse_jacobian, dtype=np.float64)
        md = L.MeshDesc()
        md.nelements, md.ninterfaces = mesh.nelements, mesh.ninterfaces
        md.neighbor_ids = self._nbr.ctypes.data
        md.orientations = self._ori.ctypes.data
        md.inverse_jacobian = self._invj.ctypes.data
        md.node_coordinates = self._node_coordinates.ctypes.data
        md.total_volume = mesh.total_volume
        handle = ctypes.c_void_p()
        L.check(L.lib().trixi_hip_create(ctypes.byref(self.config), ctypes.byref(md), ctypes.byref(handle)))
        self._ctx = handle

    # -- lifetime ------------------------------------------------------------
    def close(self):
        if getattr(self, "_ctx", None):
            L.lib().trixi_hip_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ctx(self):
        if not self._ctx:
            raise L.TrixiHipError(L.EINVAL, "semidiscretization has been closed")
        return self._ctx

    # -- shapes ----------------------------------------------------------------
    @property
    def native_shape(self):
        return (self.nvariables, self.nnodes_per_element, self.nelements)

    @property
    def trixi_shape(self):
        """Trixi's u[v, i, j, k, e] as a C-ordered array (axes reversed)."""
        return (self.nelements,) + (self.nnodes,) * self.mesh.ndims + (self.nvariables,)

    def allocate_coefficients(self):
        torch = _torch()
        return torch.empty(self.nvalues, dtype=torch.float64, device="cuda")

    def check_vector(self, x, name):
        torch = _torch()
        if not (isinstance(x, torch.Tensor) and x.is_cuda and x.dtype == torch.float64 and x.is_contiguous()
                and x.numel() == self.nvalues):
            raise L.TrixiHipError(L.EINVAL, f"{name} must be a contiguous float64 CUDA tensor with "
                                            f"{self.nvalues} values")

    def __repr__(self):
        return (f"SemidiscretizationHyperbolicHIP({self.mesh}, {type(self.equations).__name__}, "
                f"{self.initial_condition}, {self.solver})")


def compute_coefficients(t, semi, initial_condition=None):
    """u(t) of the initial condition at every node (native layout)."""
    ic = initial_condition or semi.initial_condition
    u = semi.allocate_coefficients()
    L.check(L.lib().trixi_hip_compute_coefficients(semi.ctx, _ptr(u), ic.ic_id, float(t), ic.seed, _stream()))
    return u


def rhs_hip_(du_ode, u_ode, semi, t):
    """du_ode = rhs(u_ode, t), asynchronous on the current stream (rhs_gpu!)."""
    semi.check_vector(du_ode, "du_ode")
    semi.check_vector(u_ode, "u_ode")
    L.check(L.lib().trixi_hip_rhs(semi.ctx, _ptr(du_ode), _ptr(u_ode), float(t), _stream()))


class ODEProblem:
    """Minimal in-place ODE problem (SciMLBase.ODEProblem{true}) for rhs_hip_."""

    def __init__(self, f, u0, tspan, p):
        self.f, self.u0, self.tspan, self.p = f, u0, (float(tspan[0]), float(tspan[1])), p

    def __repr__(self):
        return f"ODEProblem(f={self.f.__name__}, tspan={self.tspan}, nvalues={self.u0.numel()})"


def semidiscretizeHIP(semi, tspan):
    u0 = compute_coefficients(tspan[0], semi)
    return ODEProblem(rhs_hip_, u0, tspan, semi)


def wrap_array(u_ode, semi):
    """Native view (nvariables, N^d, nelements) of a state vector (no copy)."""
    return u_ode.view(semi.native_shape)


def pack(u_trixi, semi):
    """Trixi-layout device tensor -> new native-layout state vector."""
    torch = _torch()
    src = u_trixi.contiguous().view(-1)
    if src.numel() != semi.nvalues or src.dtype != torch.float64 or not src.is_cuda:
        raise L.TrixiHipError(L.EINVAL, "u_trixi has the wrong size, dtype or device")
    out = semi.allocate_coefficients()
    L.check(L.lib().trixi_hip_pack(semi.ctx, _ptr(out), _ptr(src), _stream()))
    return out


def wrap_array_native(u_ode, semi):
    """Trixi-layout copy u[e, (k, j,) i, v] of a native state vector."""
    torch = _torch()
    semi.check_vector(u_ode, "u_ode")
    out = torch.empty(semi.trixi_shape, dtype=torch.float64, device=u_ode.device)
    L.check(L.lib().trixi_hip_unpack(semi.ctx, _ptr(out), _ptr(u_ode), _stream()))
    return out
