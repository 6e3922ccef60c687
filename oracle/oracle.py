"""ctypes wrapper of the CPU oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the CPU-baseline leg of ``bench.py`` as the checker.  The product path
(the HIP library and its host package) never imports this module.

Arrays use the Trixi.jl layout ``u[v, i, j, k, e]`` (Julia column-major), i.e.
a C-contiguous numpy array of shape ``(E, N, N, N, nv)`` for 3D (reversed
axes), ``(E, N, N, nv)`` for 2D and ``(E, N, nv)`` for 1D.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

EQ_LINADV, EQ_EULER, EQ_GLM_MHD, EQ_BURGERS, EQ_HYPDIFF, EQ_EULER_MULTI, EQ_EULER_QUASI1D, EQ_SHALLOW_WATER = range(8)
VI_WEAK, VI_FD, VI_HG = 0, 1, 2
(FLUX_CENTRAL, FLUX_LLF, FLUX_LLF_NAIVE, FLUX_RANOCHA, FLUX_SHIMA_ETAL, FLUX_HINDENLANG_GASSNER, FLUX_EC, FLUX_HLL,
 FLUX_CHAN_ETAL, FLUX_HLLE, FLUX_WINTERMEYER_ETAL, FLUX_FJORDHOLM_ETAL, FLUX_HYDROSTATIC_LLF) = range(13)
(IC_CONSTANT, IC_CONVERGENCE_TEST, IC_WEAK_BLAST_WAVE, IC_SEEDED_RANDOM, IC_SMOOTH_PERIODIC, IC_WEAK_BLAST_WAVE_SEEDED,
 IC_SHOCK, IC_RAREFACTION, IC_EOC_GRAVITY, IC_POISSON_NONPERIODIC, IC_HARMONIC_NONPERIODIC) = range(11)
BC_PERIODIC, BC_DIRICHLET, BC_SLIP_WALL = 0, 1, 2
SOURCE_NONE, SOURCE_CONVERGENCE_TEST, SOURCE_EOC_TEST_EULER, SOURCE_POISSON_NONPERIODIC, SOURCE_HARMONIC = range(5)
IND_DENSITY_PRESSURE, IND_PRESSURE, IND_DENSITY = 0, 1, 2


class OracleCfg(ctypes.Structure):
    _fields_ = [
        ("ndims", ctypes.c_int32),
        ("equation", ctypes.c_int32),
        ("polydeg", ctypes.c_int32),
        ("volume_integral", ctypes.c_int32),
        ("volume_flux", ctypes.c_int32),
        ("surface_flux", ctypes.c_int32),
        ("fv_flux", ctypes.c_int32),
        ("alpha_smooth", ctypes.c_int32),
        ("gamma", ctypes.c_double),
        ("c_h", ctypes.c_double),
        ("advection_velocity", ctypes.c_double * 3),
        ("alpha_max", ctypes.c_double),
        ("alpha_min", ctypes.c_double),
        ("ic_length", ctypes.c_double),
        ("boundary_condition", ctypes.c_int32),
        ("boundary_ic", ctypes.c_int32),
        ("boundary_seed", ctypes.c_uint64),
        ("source_terms", ctypes.c_int32),
        ("indicator_variable", ctypes.c_int32),
        ("node_coordinates", ctypes.c_void_p),
        ("time", ctypes.c_double),
        ("nmortars", ctypes.c_int64),
        ("mortar_neighbor_ids", ctypes.c_void_p),
        ("mortar_orientations", ctypes.c_void_p),
        ("mortar_large_sides", ctypes.c_void_p),
        ("boundary_outflow_faces", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("nu", ctypes.c_double),
        ("Lr", ctypes.c_double),
        ("ncomponents", ctypes.c_int32),
        ("reserved2", ctypes.c_int32),
        ("gammas", ctypes.c_double * 4),
        ("gas_constants", ctypes.c_double * 4),
        ("gravity", ctypes.c_double),
        ("H0", ctypes.c_double),
    ]


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        i64, i32, dbl, u64 = ctypes.c_int64, ctypes.c_int32, ctypes.c_double, ctypes.c_uint64
        cfgp = ctypes.POINTER(OracleCfg)
        sig = {
            "oracle_nvars": (ctypes.c_int, [cfgp]),
            "oracle_basis": (ctypes.c_int, [ctypes.c_int] + [P] * 8),
            "oracle_interpolation_matrix": (ctypes.c_int, [ctypes.c_int, P, ctypes.c_int, P, P]),
            "oracle_ln_mean": (dbl, [dbl, dbl]),
            "oracle_inv_ln_mean": (dbl, [dbl, dbl]),
            "oracle_flux": (None, [cfgp, P, ctypes.c_int, P]),
            "oracle_two_point_flux": (None, [cfgp, ctypes.c_int, P, P, ctypes.c_int, P]),
            "oracle_noncons_flux": (None, [cfgp, P, P, ctypes.c_int, P]),
            "oracle_max_abs_speeds": (None, [cfgp, P, P]),
            "oracle_entropy": (dbl, [cfgp, P]),
            "oracle_cons2entropy": (None, [cfgp, P, P]),
            "oracle_entropy_potential": (None, [cfgp, P, ctypes.c_int, P]),
            "oracle_mesh_nelements": (i64, [ctypes.c_int, ctypes.c_int]),
            "oracle_mesh_uniform": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, dbl, dbl, ctypes.c_int, P, P, P, P, P]),
            "oracle_mesh_ninterfaces": (i64, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "oracle_mesh": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, dbl, dbl, ctypes.c_int, ctypes.c_int,
                                           P, P, P, P, P]),
            "oracle_source_convergence_test": (None, [cfgp, P, dbl, P]),
            "oracle_initial_condition": (None, [cfgp, ctypes.c_int, P, dbl, u64, u64, P]),
            "oracle_compute_coefficients": (None, [cfgp, ctypes.c_int, i64, P, dbl, u64, P]),
            "oracle_rhs": (None, [cfgp, i64, i64, P, P, P, P, P, P]),
            "oracle_indicator_hg": (None, [cfgp, i64, i64, P, P, P]),
            "oracle_max_dt": (dbl, [cfgp, i64, P, P]),
            "oracle_error_norms": (None, [cfgp, ctypes.c_int, i64, P, P, P, dbl, dbl, u64, P, P]),
            "oracle_integrate": (None, [cfgp, i64, P, P, P]),
            "oracle_ck2n54_coefficients": (None, [P, P, P]),
            "oracle_ck2n54_step": (None, [cfgp, i64, i64, P, P, P, P, P, P, dbl, dbl]),
            "oracle_rhs_jvp": (None, [cfgp, i64, i64, P, P, P, P, P, P, P]),
            "oracle_set_threads": (ctypes.c_int, [ctypes.c_int]),
            "oracle_ck2n54_step_jvp": (None, [cfgp, i64, i64, P, P, P, P, P, dbl, P, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def make_cfg(ndims, equation, polydeg=3, volume_integral=VI_WEAK, volume_flux=FLUX_CENTRAL,
             surface_flux=FLUX_LLF, fv_flux=FLUX_LLF, gamma=1.4, c_h=1.0,
             advection_velocity=(1.0, 1.0, 1.0), alpha_max=0.5, alpha_min=0.001,
             alpha_smooth=True, ic_length=4.0, boundary_condition=BC_PERIODIC, boundary_ic=IC_CONSTANT,
             boundary_seed=0, source_terms=SOURCE_NONE, indicator_variable=0, boundary_outflow_faces=0,
             nu=1.0, Lr=1.0 / (2.0 * np.pi), gammas=(), gas_constants=(), gravity=9.81):
    c = OracleCfg()
    c.gravity = gravity
    c.nu, c.Lr = nu, Lr
    c.ncomponents = len(gammas)
    for k, (g, r) in enumerate(zip(gammas, gas_constants)):
        c.gammas[k], c.gas_constants[k] = g, r
    c.boundary_outflow_faces = boundary_outflow_faces
    c.indicator_variable = indicator_variable
    c.boundary_condition, c.boundary_ic, c.boundary_seed = boundary_condition, boundary_ic, boundary_seed
    c.source_terms = source_terms
    c.ndims, c.equation, c.polydeg = ndims, equation, polydeg
    c.volume_integral, c.volume_flux, c.surface_flux, c.fv_flux = (
        volume_integral, volume_flux, surface_flux, fv_flux)
    c.alpha_smooth = 1 if alpha_smooth else 0
    c.gamma, c.c_h = gamma, c_h
    for d in range(3):
        c.advection_velocity[d] = advection_velocity[d] if d < len(advection_velocity) else 0.0
    c.alpha_max, c.alpha_min, c.ic_length = alpha_max, alpha_min, ic_length
    return c


def nvars(cfg):
    return lib().oracle_nvars(ctypes.byref(cfg))


def basis(nnodes):
    N = nnodes
    out = {k: np.zeros(N) for k in ("nodes", "weights", "lhat_minus", "lhat_plus")}
    for k in ("D", "Dsplit", "Dhat", "Vinv"):
        out[k] = np.zeros((N, N))
    rc = lib().oracle_basis(N, _p(out["nodes"]), _p(out["weights"]), _p(out["D"]), _p(out["Dsplit"]),
                            _p(out["Dhat"]), _p(out["lhat_minus"]), _p(out["lhat_plus"]), _p(out["Vinv"]))
    assert rc == 0
    return out


def interpolation_matrix(nodes_in, nodes_out):
    nin, nout = len(nodes_in), len(nodes_out)
    V = np.zeros((nout, nin))
    lib().oracle_interpolation_matrix(nin, _p(np.ascontiguousarray(nodes_in, dtype=np.float64)), nout,
                                      _p(np.ascontiguousarray(nodes_out, dtype=np.float64)), _p(V))
    return V


def ln_mean(x, y):
    return lib().oracle_ln_mean(float(x), float(y))


def inv_ln_mean(x, y):
    return lib().oracle_inv_ln_mean(float(x), float(y))


def flux(cfg, u, orientation):
    u = np.ascontiguousarray(u, dtype=np.float64)
    f = np.zeros(nvars(cfg))
    lib().oracle_flux(ctypes.byref(cfg), _p(u), orientation, _p(f))
    return f


def two_point_flux(cfg, kind, ull, urr, orientation):
    ull = np.ascontiguousarray(ull, dtype=np.float64)
    urr = np.ascontiguousarray(urr, dtype=np.float64)
    f = np.zeros(nvars(cfg))
    lib().oracle_two_point_flux(ctypes.byref(cfg), kind, _p(ull), _p(urr), orientation, _p(f))
    return f


def noncons_flux(cfg, ull, urr, orientation):
    ull = np.ascontiguousarray(ull, dtype=np.float64)
    urr = np.ascontiguousarray(urr, dtype=np.float64)
    f = np.zeros(nvars(cfg))
    lib().oracle_noncons_flux(ctypes.byref(cfg), _p(ull), _p(urr), orientation, _p(f))
    return f


def cons2entropy(cfg, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    w = np.zeros(nvars(cfg))
    lib().oracle_cons2entropy(ctypes.byref(cfg), _p(u), _p(w))
    return w


def entropy(cfg, u):
    return lib().oracle_entropy(ctypes.byref(cfg), _p(np.ascontiguousarray(u, dtype=np.float64)))


def entropy_potential(cfg, u, orientation):
    out = np.zeros(1)
    lib().oracle_entropy_potential(ctypes.byref(cfg), _p(np.ascontiguousarray(u, dtype=np.float64)),
                                   orientation, _p(out))
    return out[0]


def max_abs_speeds(cfg, u):
    lam = np.zeros(3)
    lib().oracle_max_abs_speeds(ctypes.byref(cfg), _p(np.ascontiguousarray(u, dtype=np.float64)), _p(lam))
    return lam[: cfg.ndims]


class Mesh:
    """Uniform TreeMesh as built by the oracle (periodic unless periodic=False)."""

    def __init__(self, ndims, level, cmin, cmax, polydeg, periodic=True):
        E = lib().oracle_mesh_nelements(ndims, level)
        per = tuple(map(bool, periodic)) if isinstance(periodic, (tuple, list)) else (bool(periodic),) * ndims
        mask = sum(1 << d for d in range(ndims) if per[d])
        I = lib().oracle_mesh_ninterfaces(ndims, level, mask)
        self.periodic = all(per)
        self.periodicity = per
        N = polydeg + 1
        self.ndims, self.level, self.polydeg, self.nelements, self.ninterfaces = ndims, level, polydeg, E, I
        self.coordinates_min, self.coordinates_max = cmin, cmax
        self.neighbor_ids = np.zeros((I, 2), dtype=np.int32)
        self.orientations = np.zeros(I, dtype=np.int32)
        self.inverse_jacobian = np.zeros(E)
        self.node_coordinates = np.zeros((E,) + (N,) * ndims + (ndims,))
        self.cell_index = np.zeros((E, ndims), dtype=np.int32)
        rc = lib().oracle_mesh(ndims, level, cmin, cmax, polydeg, mask, _p(self.neighbor_ids),
                               _p(self.orientations), _p(self.inverse_jacobian), _p(self.node_coordinates),
                               _p(self.cell_index))
        assert rc == 0
        self.total_volume = (cmax - cmin) ** ndims
        self.nmortars = 0
        self.mortar_neighbor_ids = np.zeros((0, (1 << (ndims - 1)) + 1), dtype=np.int32)
        self.mortar_orientations = np.zeros(0, dtype=np.int32)
        self.mortar_large_sides = np.zeros(0, dtype=np.int32)


class RefinedMesh:
    """TreeMesh with box refinement patches, built geometrically.

    Leaves are (level, integer coordinates) cells.  Starting from the uniform
    level-``level`` grid, each box patch refines the leaves whose centre lies
    strictly inside it; then any leaf with a face neighbour leaf two or more
    levels finer is refined until none is left (the unique minimal 2:1 face
    balance, which Trixi's rebalance! reaches one refined cell at a time,
    abstract_tree.jl).  The element order is the depth-first order of the
    tree with children in Morton order (Trixi inserts a cell's children
    directly behind it), i.e. leaves sorted by the Morton code of their
    lower-left corner at the finest level.  Interfaces join equal-level face
    neighbours (left = element, + direction), an L2 mortar joins a leaf to the
    2^(d-1) leaves of its refined same-level neighbour facing it (small
    elements in position order, first tangential direction fastest, then the
    large element; large_side 1 when the large element is on the left), and
    faces outside a non-periodic domain are boundary faces.
    """

    def __init__(self, ndims, level, cmin, cmax, polydeg, patches, periodic=True):
        if ndims == 1:
            raise ValueError("1D trees have no mortars")
        nd = ndims
        per = periodic if isinstance(periodic, tuple) else (bool(periodic),) * nd
        leaves = {(level, k) for k in np.ndindex(*((1 << level),) * nd)}

        def inside(lv, k, lo, hi):
            length = cmax - cmin
            for d in range(nd):
                c = cmin + (k[d] + 0.5) * length / (1 << lv)
                if not (lo[d] < c < hi[d]):
                    return False
            return True

        def split(lv, k):
            leaves.discard((lv, k))
            for ch in range(1 << nd):
                leaves.add((lv + 1, tuple(2 * k[d] + ((ch >> d) & 1) for d in range(nd))))

        def wrap(lv, k, d, s):
            kk = list(k)
            kk[d] += s
            n1 = 1 << lv
            if not 0 <= kk[d] < n1:
                if not per[d]:
                    return None
                kk[d] %= n1
            return tuple(kk)

        def covering_leaf(lv, k):
            # the leaf containing cell (lv, k), if it is at level <= lv
            for l in range(lv, -1, -1):
                kk = tuple(x >> (lv - l) for x in k)
                if (l, kk) in leaves:
                    return (l, kk)
            return None

        for lo, hi in patches:
            for lv, k in [c for c in leaves if inside(c[0], c[1], lo, hi)]:
                split(lv, k)
            changed = True
            while changed:
                changed = False
                for lv, k in sorted(leaves, reverse=True):
                    if (lv, k) not in leaves:
                        continue
                    for d in range(nd):
                        for s in (-1, 1):
                            nk = wrap(lv, k, d, s)
                            if nk is None:
                                continue
                            cov = covering_leaf(lv, nk)
                            if cov is not None and cov[0] < lv - 1:
                                split(*cov)
                                changed = True
        lmax = max(lv for lv, _ in leaves)

        def morton(c):
            lv, k = c
            code = 0
            for b in range(lmax - 1, -1, -1):
                for d in range(nd - 1, -1, -1):
                    code = (code << 1) | (((k[d] << (lmax - lv)) >> b) & 1)
            return code

        elems = sorted(leaves, key=morton)
        E = len(elems)
        eid = {c: i for i, c in enumerate(elems)}
        N = polydeg + 1
        nodes = basis(N)["nodes"]
        self.ndims, self.polydeg, self.level, self.nelements = nd, polydeg, lmax, E
        self.coordinates_min, self.coordinates_max, self.periodic = cmin, cmax, per
        self.levels = np.array([lv for lv, _ in elems], dtype=np.int32)
        self.cell_index = np.array([k for _, k in elems], dtype=np.int32).reshape(E, nd)
        self.inverse_jacobian = np.array([1.0 / ((cmax - cmin) / (1 << lv) / 2.0) for lv, _ in elems])
        self.node_coordinates = np.zeros((E,) + (N,) * nd + (nd,))
        for e, (lv, k) in enumerate(elems):
            jac = (cmax - cmin) / (1 << lv) / 2.0
            for d in range(nd):
                cc, ln = 0.5 * (cmin + cmax), cmax - cmin
                for b in range(lv - 1, -1, -1):
                    ln *= 0.5
                    cc += (0.5 if (k[d] >> b) & 1 else -0.5) * ln
                shp = [1] * nd
                shp[nd - 1 - d] = N
                self.node_coordinates[e, ..., d] = cc + jac * nodes.reshape(shp)
        ifc, mort, bnd = [], [], []
        for e, (lv, k) in enumerate(elems):
            for d in range(nd):
                for s in (-1, 1):
                    nk = wrap(lv, k, d, s)
                    if nk is None:
                        bnd.append((e, 2 * d + (s > 0)))
                        continue
                    if (lv, nk) in leaves:
                        if s > 0:
                            ifc.append((e, eid[(lv, nk)], d + 1))
                        continue
                    if covering_leaf(lv, nk) is not None:
                        continue  # a coarser leaf: this element is a small side of its mortar
                    tdims = [q for q in range(nd) if q != d]
                    small = []
                    for pos in range(1 << (nd - 1)):
                        kk = [2 * x for x in nk]
                        if s < 0:
                            kk[d] += 1
                        for t, q in enumerate(tdims):
                            kk[q] += (pos >> t) & 1
                        small.append(eid[(lv + 1, tuple(kk))])
                    mort.append((small + [e], d + 1, 1 if s > 0 else 2))
        self.neighbor_ids = np.array([(a, b) for a, b, _ in ifc], dtype=np.int32).reshape(-1, 2)
        self.orientations = np.array([o for _, _, o in ifc], dtype=np.int32)
        self.ninterfaces = len(ifc)
        self.boundary_faces = np.array(sorted(bnd), dtype=np.int32).reshape(-1, 2)
        self.nmortars = len(mort)
        self.mortar_neighbor_ids = np.array([m[0] for m in mort], dtype=np.int32).reshape(-1, (1 << (nd - 1)) + 1)
        self.mortar_orientations = np.array([m[1] for m in mort], dtype=np.int32)
        self.mortar_large_sides = np.array([m[2] for m in mort], dtype=np.int32)
        self.total_volume = (cmax - cmin) ** nd


def state_shape(cfg, nelements):
    N = cfg.polydeg + 1
    return (nelements,) + (N,) * cfg.ndims + (nvars(cfg),)


def compute_coefficients(cfg, mesh, ic, t=0.0, seed=20261202):
    u = np.zeros(state_shape(cfg, mesh.nelements))
    lib().oracle_compute_coefficients(ctypes.byref(cfg), ic, mesh.nelements, _p(mesh.node_coordinates),
                                      float(t), seed, _p(u))
    return u


def initial_condition(cfg, ic, x, t=0.0, seed=20261202, node_index=0):
    u = np.zeros(nvars(cfg))
    x = np.ascontiguousarray(np.atleast_1d(x), dtype=np.float64)
    lib().oracle_initial_condition(ctypes.byref(cfg), ic, _p(x), float(t), seed, node_index, _p(u))
    return u


def _bind(cfg, mesh, t=0.0):
    """A copy of cfg that carries the mesh's node coordinates (boundary states,
    sources) and the time."""
    c = OracleCfg.from_buffer_copy(cfg)
    c.node_coordinates = mesh.node_coordinates.ctypes.data
    c.time = float(t)
    c.nmortars = mesh.nmortars
    if mesh.nmortars:
        c.mortar_neighbor_ids = mesh.mortar_neighbor_ids.ctypes.data
        c.mortar_orientations = mesh.mortar_orientations.ctypes.data
        c.mortar_large_sides = mesh.mortar_large_sides.ctypes.data
    return c


def mortar_matrices(polydeg):
    """(forward_lower, forward_upper, reverse_lower, reverse_upper), N x N each."""
    N = polydeg + 1
    m = [np.zeros((N, N)) for _ in range(4)]
    assert lib().oracle_mortar_matrices(N, *[_p(a) for a in m]) == 0
    return tuple(m)


def source_convergence_test(cfg, x, t):
    s = np.zeros(nvars(cfg))
    lib().oracle_source_convergence_test(ctypes.byref(cfg), _p(np.ascontiguousarray(np.atleast_1d(x), dtype=np.float64)),
                                         float(t), _p(s))
    return s


def rhs(cfg, mesh, u, return_alpha=False, t=0.0):
    cfg = _bind(cfg, mesh, t)
    u = np.ascontiguousarray(u, dtype=np.float64)
    du = np.zeros_like(u)
    alpha = np.zeros(mesh.nelements) if return_alpha else None
    lib().oracle_rhs(ctypes.byref(cfg), mesh.nelements, mesh.ninterfaces, _p(mesh.neighbor_ids),
                     _p(mesh.orientations), _p(mesh.inverse_jacobian), _p(u), _p(du), _p(alpha))
    return (du, alpha) if return_alpha else du


def rhs_jvp(cfg, mesh, u, udot, t=0.0):
    cfg = _bind(cfg, mesh, t)
    u = np.ascontiguousarray(u, dtype=np.float64)
    udot = np.ascontiguousarray(udot, dtype=np.float64)
    du = np.zeros_like(u)
    dudot = np.zeros_like(u)
    lib().oracle_rhs_jvp(ctypes.byref(cfg), mesh.nelements, mesh.ninterfaces, _p(mesh.neighbor_ids),
                         _p(mesh.orientations), _p(mesh.inverse_jacobian), _p(u), _p(udot), _p(du), _p(dudot))
    return du, dudot


def indicator_hg(cfg, mesh, u):
    cfg = _bind(cfg, mesh)
    u = np.ascontiguousarray(u, dtype=np.float64)
    alpha = np.zeros(mesh.nelements)
    lib().oracle_indicator_hg(ctypes.byref(cfg), mesh.nelements, mesh.ninterfaces, _p(mesh.neighbor_ids),
                              _p(u), _p(alpha))
    return alpha


def max_dt(cfg, mesh, u):
    return lib().oracle_max_dt(ctypes.byref(cfg), mesh.nelements, _p(mesh.inverse_jacobian),
                               _p(np.ascontiguousarray(u, dtype=np.float64)))


def error_norms(cfg, mesh, u, t, ic, seed=20261202):
    nv = nvars(cfg)
    l2, linf = np.zeros(nv), np.zeros(nv)
    lib().oracle_error_norms(ctypes.byref(cfg), ic, mesh.nelements, _p(mesh.inverse_jacobian),
                             _p(mesh.node_coordinates), _p(np.ascontiguousarray(u, dtype=np.float64)),
                             float(t), float(mesh.total_volume), seed, _p(l2), _p(linf))
    return l2, linf


def integrate(cfg, mesh, u):
    out = np.zeros(nvars(cfg))
    lib().oracle_integrate(ctypes.byref(cfg), mesh.nelements, _p(mesh.inverse_jacobian),
                           _p(np.ascontiguousarray(u, dtype=np.float64)), _p(out))
    return out


def ck2n54_coefficients():
    A, B, c = np.zeros(5), np.zeros(5), np.zeros(5)
    lib().oracle_ck2n54_coefficients(_p(A), _p(B), _p(c))
    return A, B, c


def ck2n54_step(cfg, mesh, u, k, t, dt):
    """Advance ``u`` (in place) by one CK2N54 step; ``k`` is the 2N register."""
    cfg = _bind(cfg, mesh, t)
    du = np.zeros_like(u)
    lib().oracle_ck2n54_step(ctypes.byref(cfg), mesh.nelements, mesh.ninterfaces, _p(mesh.neighbor_ids),
                             _p(mesh.orientations), _p(mesh.inverse_jacobian), _p(u), _p(k), _p(du),
                             float(t), float(dt))


def ck2n54_step_jvp(cfg, mesh, u, udot, dt, t=0.0):
    """(u_next, d u_next) of one CK2N54 step from (t, u) along the tangent udot."""
    cfg = _bind(cfg, mesh, t)
    u = np.ascontiguousarray(u, dtype=np.float64)
    udot = np.ascontiguousarray(udot, dtype=np.float64)
    un, und = np.zeros_like(u), np.zeros_like(u)
    lib().oracle_ck2n54_step_jvp(ctypes.byref(cfg), mesh.nelements, mesh.ninterfaces, _p(mesh.neighbor_ids),
                                 _p(mesh.orientations), _p(mesh.inverse_jacobian), _p(u), _p(udot), float(dt),
                                 _p(un), _p(und))
    return un, und


def jacobian(cfg, mesh, u, t=0.0):
    """Dense Jacobian of rhs! (tiny meshes only), column by column from the dual JVP."""
    n = u.size
    J = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        J[:, j] = rhs_jvp(cfg, mesh, u, e.reshape(u.shape), t)[1].ravel()
    return J


def set_threads(n):
    return lib().oracle_set_threads(int(n))
