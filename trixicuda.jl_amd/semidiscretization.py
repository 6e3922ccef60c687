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
from .equations import (BoundaryConditionDirichlet, InitialCondition, SourceTerms, boundary_condition_outflow,
                        boundary_condition_slip_wall,
                        boundary_condition_periodic, boundary_condition_poisson_nonperiodic,
                        initial_condition_constant, initial_condition_poisson_nonperiodic)


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
                 boundary_conditions=boundary_condition_periodic):
        if mesh.ndims != equations.ndims:
            raise ValueError("mesh and equations dimensions differ")
        self.mesh, self.equations, self.initial_condition, self.solver = mesh, equations, initial_condition, solver
        self.source_terms, self.boundary_conditions = source_terms, boundary_conditions
        self.nvariables = equations.nvariables
        self.nnodes = solver.basis.nnodes
        self.nelements = mesh.nelements
        self.nnodes_per_element = self.nnodes ** mesh.ndims
        self.ndofs = self.nelements * self.nnodes_per_element
        self.nvalues = self.ndofs * self.nvariables
        self.config = solver.config(equations, ic_length=mesh.length)
        self._set_boundaries_and_sources(mesh, boundary_conditions, source_terms)
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
        md.nmortars = mesh.nmortars
        if mesh.nmortars:
            self._mids = np.ascontiguousarray(mesh.mortar_neighbor_ids, dtype=np.int32)
            self._mori = np.ascontiguousarray(mesh.mortar_orientations, dtype=np.int32)
            self._mls = np.ascontiguousarray(mesh.mortar_large_sides, dtype=np.int32)
            md.mortar_neighbor_ids = self._mids.ctypes.data
            md.mortar_orientations = self._mori.ctypes.data
            md.mortar_large_sides = self._mls.ctypes.data
        handle = ctypes.c_void_p()
        L.check(L.lib().trixi_hip_create(ctypes.byref(self.config), ctypes.byref(md), ctypes.byref(handle)))
        self._ctx = handle
        # the library's state size must be the host's: every device vector is
        # sized here and addressed there
        nv, nn, nval = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        L.check(L.lib().trixi_hip_sizes(handle, ctypes.byref(nv), ctypes.byref(nn), ctypes.byref(nval)))
        if (nv.value, nn.value, nval.value) != (self.nvariables, self.nnodes, self.nvalues):
            self.close()
            raise L.TrixiHipError(L.EINVAL, f"state size mismatch: library ({nv.value}, {nn.value}, {nval.value}), "
                                            f"host ({self.nvariables}, {self.nnodes}, {self.nvalues})")

    def _set_boundaries_and_sources(self, mesh, bcs, source_terms):
        """Trixi's boundary_conditions (one condition, or one per side as a
        dict/namedtuple x_neg, x_pos, ...: all sides must then be the same
        Dirichlet condition) and source_terms, into the C-ABI configuration."""
        c = self.config
        if source_terms is None:
            c.source_terms = L.SOURCE_NONE
        elif isinstance(source_terms, SourceTerms):
            c.source_terms = source_terms.source_id
        else:
            raise NotImplementedError(f"source terms {source_terms!r} are not built")
        periodic_mesh = all(mesh.periodicity)
        if bcs is None or bcs is boundary_condition_periodic or bcs == "periodic":
            if not periodic_mesh:
                raise ValueError("a non-periodic mesh needs boundary_conditions")
            c.boundary_condition = L.BC_PERIODIC
            return
        names = ("x_neg", "x_pos", "y_neg", "y_pos", "z_neg", "z_pos")[:2 * mesh.ndims]
        if isinstance(bcs, dict) or hasattr(bcs, "_asdict"):
            per_side = dict(bcs) if isinstance(bcs, dict) else bcs._asdict()
            if set(per_side) - set(names):
                raise ValueError(f"unknown boundary names {sorted(set(per_side) - set(names))}")
            sides = [per_side.get(n, boundary_condition_periodic) for n in names]
        else:
            sides = [bcs] * len(names)
        if any(b is boundary_condition_slip_wall for b in sides):
            if not all(b is boundary_condition_slip_wall or b is boundary_condition_periodic for b in sides):
                raise NotImplementedError("slip walls cannot be combined with other boundary conditions")
            c.boundary_condition = L.BC_SLIP_WALL
            return
        outflow = 0
        for f_dir, b in enumerate(sides):
            if b is boundary_condition_outflow:
                outflow |= 1 << f_dir
        if any(isinstance(b, BoundaryConditionDirichlet) and
               b.boundary_value_function is initial_condition_poisson_nonperiodic for b in sides):
            raise NotImplementedError("use boundary_condition_poisson_nonperiodic (the steady state at any time)")
        sides = [BoundaryConditionDirichlet(initial_condition_poisson_nonperiodic, _steady=True)
                 if b is boundary_condition_poisson_nonperiodic else b for b in sides]
        dirichlet = [b for b in sides if b is not boundary_condition_periodic and b is not boundary_condition_outflow]
        if not all(isinstance(b, BoundaryConditionDirichlet) for b in dirichlet):
            raise NotImplementedError("only BoundaryConditionDirichlet, boundary_condition_outflow and periodic "
                                      "boundaries are built")
        if dirichlet:
            f = dirichlet[0].boundary_value_function
            if any(b.boundary_value_function != f for b in dirichlet):
                raise NotImplementedError("all Dirichlet boundary faces must carry the same state")
        else:
            f = initial_condition_constant
        c.boundary_condition = L.BC_DIRICHLET
        c.boundary_ic = f.ic_id
        c.boundary_seed = f.seed
        c.boundary_outflow_faces = outflow

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

    @property
    def fp32_state(self):
        """True for the FP32-state precision (float32 state vectors)."""
        return self.solver.precision == "fp32"

    @property
    def state_dtype(self):
        torch = _torch()
        return torch.float32 if self.fp32_state else torch.float64

    def allocate_coefficients(self, dtype=None):
        torch = _torch()
        return torch.empty(self.nvalues, dtype=dtype or self.state_dtype, device="cuda")

    def check_vector(self, x, name, dtype=None):
        torch = _torch()
        dtype = dtype or self.state_dtype
        if not (isinstance(x, torch.Tensor) and x.is_cuda and x.dtype == dtype and x.is_contiguous()
                and x.numel() == self.nvalues):
            raise L.TrixiHipError(L.EINVAL, f"{name} must be a contiguous {str(dtype).split('.')[-1]} CUDA tensor "
                                            f"with {self.nvalues} values")

    def as_f64(self, x, name):
        """A state vector as float64 (the callbacks, pack and unpack work on
        FP64 vectors; an FP32 state is widened, a copy)."""
        self.check_vector(x, name)
        return x.double() if self.fp32_state else x

    def __repr__(self):
        return (f"SemidiscretizationHyperbolicHIP({self.mesh}, {type(self.equations).__name__}, "
                f"{self.initial_condition}, {self.solver})")


def compute_coefficients(t, semi, initial_condition=None):
    """u(t) of the initial condition at every node (native layout).

    A built initial condition is evaluated on the device.  A Python function
    ``initial_condition(x, t, equations)`` returning the conservative variables
    (Trixi's signature; ``prim2cons`` converts primitive states) is evaluated
    on the host at every node, as Trixi's compute_coefficients! does (in 1D
    the element end nodes move one ulp inwards, so that a discontinuity on an
    element face is sampled from the element's own side), and packed into the
    device state."""
    ic = initial_condition or semi.initial_condition
    if not isinstance(ic, InitialCondition):
        if not callable(ic):
            raise TypeError(f"initial condition {ic!r} is neither built nor callable")
        torch = _torch()
        host = host_initial_state(ic, t, semi)
        return pack(torch.from_numpy(host).to("cuda"), semi)
    torch = _torch()
    u = semi.allocate_coefficients(torch.float64)
    L.check(L.lib().trixi_hip_compute_coefficients(semi.ctx, _ptr(u), ic.ic_id, float(t), ic.seed, _stream()))
    # an FP32 state is the FP64 one rounded (Trixi evaluates the initial
    # condition in RealT; rounding the FP64 value is within its last bit)
    return u.float() if semi.fp32_state else u


def host_initial_state(initial_condition, t, semi):
    """Trixi-layout numpy array [E, (N,)*d, nvars] of initial_condition(x, t,
    equations) at the node coordinates (Trixi's 1D end-node shift included)."""
    x = semi._node_coordinates
    nd, N, nv = semi.mesh.ndims, semi.nnodes, semi.nvariables
    if nd == 1:
        x = x.copy()
        x[:, 0, 0] = np.nextafter(x[:, 0, 0], np.inf)
        x[:, N - 1, 0] = np.nextafter(x[:, N - 1, 0], -np.inf)
    pts = x.reshape(-1, nd)
    out = np.empty((pts.shape[0], nv))
    for i, xi in enumerate(pts):
        val = np.asarray(initial_condition(tuple(xi), float(t), semi.equations), dtype=np.float64)
        if val.shape != (nv,):
            raise ValueError(f"initial condition returned {val.shape}, expected ({nv},)")
        out[i] = val
    return out.reshape(x.shape[:-1] + (nv,))


def rhs_hip_(du_ode, u_ode, semi, t):
    """du_ode = rhs(u_ode, t), asynchronous on the current stream (rhs_gpu!)."""
    semi.check_vector(du_ode, "du_ode")
    semi.check_vector(u_ode, "u_ode")
    fn = L.lib().trixi_hip_rhs_f32 if semi.fp32_state else L.lib().trixi_hip_rhs
    L.check(fn(semi.ctx, _ptr(du_ode), _ptr(u_ode), float(t), _stream()))


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
    if semi.fp32_state and src.dtype in (torch.float32, torch.float64):
        src = src.double()
    if src.numel() != semi.nvalues or src.dtype != torch.float64 or not src.is_cuda:
        raise L.TrixiHipError(L.EINVAL, "u_trixi has the wrong size, dtype or device")
    out = semi.allocate_coefficients(torch.float64)
    L.check(L.lib().trixi_hip_pack(semi.ctx, _ptr(out), _ptr(src), _stream()))
    return out.float() if semi.fp32_state else out


def wrap_array_native(u_ode, semi):
    """Trixi-layout copy u[e, (k, j,) i, v] of a native state vector."""
    torch = _torch()
    u64 = semi.as_f64(u_ode, "u_ode")
    out = torch.empty(semi.trixi_shape, dtype=torch.float64, device=u_ode.device)
    L.check(L.lib().trixi_hip_unpack(semi.ctx, _ptr(out), _ptr(u64), _stream()))
    return out.float() if semi.fp32_state else out
