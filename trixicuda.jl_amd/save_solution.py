"""HDF5 solution files: Trixi's SaveSolutionCallback for this semidiscretization.

The reference keeps Trixi's own SaveSolutionCallback and reaches the state
through `wrap_array_native` (src/solvers/dg.jl:24-32); its examples save with
`solution_variables = cons2prim` (examples/euler_ec_1d.jl:37-39,
examples/eulermulti_ec_2d.jl:40-43, examples/hypdiff_nonperiodic_2d.jl:50-53).
Here the state is unpacked on the device (trixi_hip_unpack), copied to the
host, converted with `cons2prim` / `cons2cons` and written by libtrixih5.so
(include/trixi_h5.h) in the layout of Trixi.jl's `save_solution_file`
(src/callbacks_step/save_solution_dg.jl): file `solution_%09d.h5`, one
float64 dataset per variable with node index i fastest, then j, k, then the
element.  The variable names and conversions are Trixi.jl's `varnames`,
`cons2prim` and `cons2cons` of each equation type.
"""
import ctypes
import os

import numpy as np

from . import _lib as L

H5_LIBRARY_PATH = os.environ.get("TRIXI_H5_LIBRARY") or os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                                      "libtrixih5.so")
_h5 = None


class _Info(ctypes.Structure):
    _fields_ = [("ndims", ctypes.c_int32), ("polydeg", ctypes.c_int32), ("n_vars", ctypes.c_int32),
                ("n_element_vars", ctypes.c_int32), ("n_elements", ctypes.c_int64), ("timestep", ctypes.c_int64),
                ("time", ctypes.c_double), ("dt", ctypes.c_double), ("equations", ctypes.c_char_p),
                ("mesh_type", ctypes.c_char_p), ("mesh_file", ctypes.c_char_p)]


class _MeshInfo(ctypes.Structure):
    _fields_ = [("ndims", ctypes.c_int32), ("periodicity", ctypes.c_int32 * 3), ("n_cells", ctypes.c_int64),
                ("capacity", ctypes.c_int64), ("n_leaf_cells", ctypes.c_int64), ("minimum_level", ctypes.c_int64),
                ("maximum_level", ctypes.c_int64), ("center_level_0", ctypes.c_double * 3),
                ("length_level_0", ctypes.c_double)]


def h5lib():
    """libtrixih5.so; raises when it was not built (no HDF5 headers at build time)."""
    global _h5
    if _h5 is None:
        if not os.path.exists(H5_LIBRARY_PATH):
            raise L.TrixiHipError(L.EINVAL, f"{H5_LIBRARY_PATH} is not built (it needs the HDF5 C library)")
        lib = ctypes.CDLL(H5_LIBRARY_PATH)
        lib.trixi_h5_save_solution.restype = ctypes.c_int
        lib.trixi_h5_save_solution.argtypes = [ctypes.c_char_p, ctypes.POINTER(_Info), ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        lib.trixi_h5_save_mesh.restype = ctypes.c_int
        lib.trixi_h5_save_mesh.argtypes = [ctypes.c_char_p, ctypes.POINTER(_MeshInfo)] + [ctypes.c_void_p] * 5
        lib.trixi_h5_last_error.restype = ctypes.c_char_p
        _h5 = lib
    return _h5


def save_mesh_file(mesh, output_directory="out"):
    """Trixi's save_mesh_file for a TreeMesh: `mesh.h5` in output_directory
    with the cell tree (TreeMesh.tree_arrays); returns the path."""
    t = mesh.tree_arrays()
    nd = mesh.ndims
    info = _MeshInfo()
    info.ndims = nd
    for d in range(nd):
        info.periodicity[d] = 1 if t["periodicity"][d] else 0
        info.center_level_0[d] = t["center_level_0"][d]
    info.n_cells, info.capacity, info.n_leaf_cells = t["n_cells"], t["capacity"], t["n_leaf_cells"]
    info.minimum_level, info.maximum_level = t["minimum_level"], t["maximum_level"]
    info.length_level_0 = t["length_level_0"]
    arrays = [np.ascontiguousarray(t[k], dtype=np.int64) for k in ("parent_ids", "child_ids", "neighbor_ids",
                                                                   "levels")]
    coords = np.ascontiguousarray(t["coordinates"], dtype=np.float64)
    os.makedirs(output_directory, exist_ok=True)
    path = os.path.join(output_directory, "mesh.h5")
    lib = h5lib()
    rc = lib.trixi_h5_save_mesh(os.fsencode(path), ctypes.byref(info), *[a.ctypes.data for a in arrays],
                                coords.ctypes.data)
    if rc != 0:
        raise L.TrixiHipError(L.EINVAL, lib.trixi_h5_last_error().decode())
    return path


def cons2cons(u, equations):
    """Identity (Trixi's cons2cons); u[..., v]."""
    return np.asarray(u, dtype=np.float64)


def cons2prim(u, equations):
    """Trixi's cons2prim of each equation type, vectorized over u[..., v]."""
    u = np.asarray(u, dtype=np.float64)
    eq, nd = equations.equation_id, equations.ndims
    if eq in (L.EQ_LINADV, L.EQ_BURGERS, L.EQ_HYPDIFF):
        return u.copy()
    out = np.empty_like(u)
    if eq == L.EQ_EULER:
        rho = u[..., 0]
        v = u[..., 1:1 + nd] / rho[..., None]
        out[..., 0] = rho
        out[..., 1:1 + nd] = v
        out[..., 1 + nd] = (equations.gamma - 1) * (u[..., 1 + nd] - 0.5 * rho * np.sum(v * v, axis=-1))
        return out
    if eq == L.EQ_GLM_MHD:  # (rho, rho v1..3, rho e, B1..3[, psi])
        rho = u[..., 0]
        v = u[..., 1:4] / rho[..., None]
        B = u[..., 5:8]
        p = u[..., 4] - 0.5 * rho * np.sum(v * v, axis=-1) - 0.5 * np.sum(B * B, axis=-1)
        if u.shape[-1] == 9:
            p = p - 0.5 * u[..., 8] * u[..., 8]
        out[..., 0] = rho
        out[..., 1:4] = v
        out[..., 4] = (equations.gamma - 1) * p
        out[..., 5:] = u[..., 5:]
        return out
    if eq == L.EQ_EULER_MULTI:  # (rho v.., rho e, rho_1..rho_n) -> (v.., p, rho_1..rho_n)
        rk = u[..., nd + 1:]
        g = np.asarray(equations.gammas)
        cv = np.asarray(equations.gas_constants) / (g - 1)
        gamma = np.sum(rk * cv * g, axis=-1) / np.sum(rk * cv, axis=-1)
        rho = np.sum(rk, axis=-1)
        v = u[..., :nd] / rho[..., None]
        out[..., :nd] = v
        out[..., nd] = (gamma - 1) * (u[..., nd] - 0.5 * rho * np.sum(v * v, axis=-1))
        out[..., nd + 1:] = rk
        return out
    if eq == L.EQ_EULER_QUASI1D:  # (a rho, a rho v, a e, a) -> (rho, v, p, a)
        a = u[..., 3]
        rho = u[..., 0] / a
        v = u[..., 1] / u[..., 0]
        out[..., 0] = rho
        out[..., 1] = v
        out[..., 2] = (equations.gamma - 1) * (u[..., 2] / a - 0.5 * rho * v * v)
        out[..., 3] = a
        return out
    if eq == L.EQ_SHALLOW_WATER:  # (h, h v.., b) -> (H, v.., b)
        h, b = u[..., 0], u[..., -1]
        out[..., 0] = h + b
        out[..., 1:1 + nd] = u[..., 1:1 + nd] / h[..., None]
        out[..., -1] = b
        return out
    raise NotImplementedError(f"cons2prim is not provided for {type(equations).__name__}")


def varnames(solution_variables, equations):
    """Trixi's varnames(cons2cons | cons2prim, equations)."""
    eq, nd = equations.equation_id, equations.ndims
    xyz = [str(d + 1) for d in range(nd)]
    prim = solution_variables is cons2prim
    if solution_variables not in (cons2cons, cons2prim):
        raise ValueError("solution_variables must be cons2prim or cons2cons")
    if eq in (L.EQ_LINADV, L.EQ_BURGERS):
        return ["scalar"]
    if eq == L.EQ_HYPDIFF:
        return ["phi"] + ["q" + d for d in xyz]
    if eq == L.EQ_EULER:
        return ["rho"] + [("v" if prim else "rho_v") + d for d in xyz] + ["p" if prim else "rho_e"]
    if eq == L.EQ_GLM_MHD:
        names = ["rho"] + [("v" if prim else "rho_v") + d for d in "123"] + ["p" if prim else "rho_e"]
        return names + ["B1", "B2", "B3"] + (["psi"] if nd > 1 else [])
    if eq == L.EQ_EULER_MULTI:
        rhos = ["rho" + str(k + 1) for k in range(equations.ncomponents)]
        return [("v" if prim else "rho_v") + d for d in xyz] + ["p" if prim else "rho_e"] + rhos
    if eq == L.EQ_EULER_QUASI1D:
        return ["rho", "v1", "p", "a"] if prim else ["a_rho", "a_rho_v1", "a_e", "a"]
    if eq == L.EQ_SHALLOW_WATER:
        vel = ["v"] if nd == 1 else ["v" + d for d in xyz]
        return (["H"] + vel if prim else ["h"] + ["h_" + v for v in vel]) + ["b"]
    raise NotImplementedError(f"varnames is not provided for {type(equations).__name__}")


def write_solution_file(filename, data, equations, polydeg, t=0.0, dt=0.0, timestep=0,
                        solution_variables=cons2prim, element_variables=None, mesh_type="TreeMesh",
                        mesh_file=""):
    """Write a host array data[e, (k, j,) i, v] (Trixi's u[v, i, j, k, e]
    column-major, already converted to `solution_variables`)."""
    data = np.ascontiguousarray(data, dtype=np.float64)
    nd, nv = equations.ndims, data.shape[-1]
    if data.ndim != nd + 2 or any(n != polydeg + 1 for n in data.shape[1:-1]):
        raise ValueError(f"data must have shape (nelements,{' nnodes,' * nd} nvariables)")
    names = varnames(solution_variables, equations)
    if len(names) != nv:
        raise ValueError(f"{type(equations).__name__} has {len(names)} variables, data has {nv}")
    slices = np.ascontiguousarray(np.moveaxis(data, -1, 0))  # [v][e][(k, j,) i]: i fastest, element slowest
    ev = list((element_variables or {}).items())
    ev_data = np.ascontiguousarray(np.array([np.asarray(a, dtype=np.float64).reshape(-1) for _, a in ev]
                                            or np.zeros((0, data.shape[0]))))
    if ev and ev_data.shape != (len(ev), data.shape[0]):
        raise ValueError("each element variable needs one value per element")
    vn = (ctypes.c_char_p * nv)(*[n.encode() for n in names])
    en = (ctypes.c_char_p * max(len(ev), 1))(*[str(k).encode() for k, _ in ev])
    info = _Info(nd, int(polydeg), nv, len(ev), data.shape[0], int(timestep), float(t), float(dt),
                 type(equations).__name__.encode(), mesh_type.encode(), mesh_file.encode())
    lib = h5lib()
    rc = lib.trixi_h5_save_solution(os.fsencode(filename), ctypes.byref(info), slices.ctypes.data,
                                    ctypes.cast(vn, ctypes.c_void_p), ev_data.ctypes.data if ev else None,
                                    ctypes.cast(en, ctypes.c_void_p))
    if rc != 0:
        raise L.TrixiHipError(L.EINVAL, lib.trixi_h5_last_error().decode())
    return filename


def save_solution_file(u_ode, semi, t, dt, timestep, output_directory="out", solution_variables=cons2prim,
                       mesh_file=""):
    """Trixi's save_solution_file for a device state vector; returns the path.
    `mesh_file` is the name of the mesh file the solution refers to (Trixi
    writes the mesh file's name, without its directory).
    Shock capturing adds the blending factors as `indicator_shock_capturing`,
    the element variable Trixi saves for VolumeIntegralShockCapturingHG."""
    from .semidiscretization import wrap_array_native
    u = wrap_array_native(u_ode, semi).cpu().numpy()
    data = solution_variables(u, semi.equations)
    element_variables = {}
    if semi.solver.volume_integral.kind == L.VI_HG:
        from .callbacks import blending_factors
        element_variables["indicator_shock_capturing"] = blending_factors(u_ode, semi).cpu().numpy()
    os.makedirs(output_directory, exist_ok=True)
    path = os.path.join(output_directory, "solution_%09d.h5" % int(timestep))
    return write_solution_file(path, data, semi.equations, semi.solver.polydeg, t, dt, timestep,
                               solution_variables, element_variables, mesh_file=mesh_file)


class SaveSolutionCallback:
    """Saves every `interval` steps or, with `dt`, at the times t0 + n dt
    (solve() shortens the step that would pass such a time so that it lands on
    it, as Trixi does with stop times), and at the initial and final times when
    asked (Trixi's SaveSolutionCallback keywords)."""

    def __init__(self, interval=0, dt=None, save_initial_solution=True, save_final_solution=True,
                 output_directory="out", solution_variables=cons2prim):
        if solution_variables not in (cons2cons, cons2prim):
            raise ValueError("solution_variables must be cons2prim or cons2cons")
        if dt is not None and (int(interval) > 0 or not float(dt) > 0):
            raise ValueError("give either interval or a positive dt")
        self.save_dt, self.next_time = (float(dt) if dt is not None else None), None
        self.interval, self.output_directory = int(interval), output_directory
        self.save_initial_solution, self.save_final_solution = save_initial_solution, save_final_solution
        self.solution_variables = solution_variables
        self.files = []
        self.mesh_file = None

    def _save(self, u, t, dt, step, semi):
        # the mesh file once, before the first solution file (Trixi writes it
        # when the callback first saves and the mesh has unsaved changes)
        if self.mesh_file is None:
            self.mesh_file = save_mesh_file(semi.mesh, self.output_directory)
        self.files.append(save_solution_file(u, semi, t, dt, step, self.output_directory, self.solution_variables,
                                             mesh_file=os.path.basename(self.mesh_file)))

    def initialize(self, u, t, dt, semi):
        self.t0, self.nsaved = t, 0
        if self.save_dt is not None:
            self.next_time = t + self.save_dt
        if self.save_initial_solution:
            self._save(u, t, dt, 0, semi)

    def __call__(self, u, t, dt, step, semi, final=False):
        if final:
            if self.save_final_solution and not (self.files and self.files[-1].endswith("%09d.h5" % step)):
                self._save(u, t, dt, step, semi)
        elif self.save_dt is not None:
            if self.next_time is None:
                raise RuntimeError("SaveSolutionCallback(dt=...) needs initialize(u, t0, dt, semi) before the first step")
            if t >= self.next_time:
                self._save(u, t, dt, step, semi)
                self.nsaved += 1
                self.next_time = self.t0 + (self.nsaved + 1) * self.save_dt
        elif self.interval > 0 and step % self.interval == 0:
            self._save(u, t, dt, step, semi)
