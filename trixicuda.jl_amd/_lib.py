"""ctypes binding of libtrixihip.so (the C ABI declared in include/trixi_hip.h).

The library is built in-tree by ``build.py``.  If it is missing, importing
works but every call raises: there is no CPU fallback for any device path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.environ.get("TRIXI_HIP_LIBRARY") or os.path.join(HERE, "libtrixihip.so")

OK, EINVAL, EUNSUPPORTED, EDEVICE, ENOMEM, ESTATE = 0, -1, -2, -3, -4, -5

EQ_LINADV, EQ_EULER, EQ_GLM_MHD, EQ_BURGERS, EQ_HYPDIFF, EQ_EULER_MULTI, EQ_EULER_QUASI1D, EQ_SHALLOW_WATER = range(8)
VI_WEAK, VI_FD, VI_HG = 0, 1, 2
(FLUX_CENTRAL, FLUX_LLF, FLUX_LLF_NAIVE, FLUX_RANOCHA, FLUX_SHIMA_ETAL, FLUX_HINDENLANG_GASSNER, FLUX_EC, FLUX_HLL,
 FLUX_CHAN_ETAL, FLUX_HLLE, FLUX_WINTERMEYER_ETAL, FLUX_FJORDHOLM_ETAL, FLUX_HYDROSTATIC_LLF) = range(13)
PREC_FP64, PREC_MIXED, PREC_MIXED_F32LOG, PREC_FP32 = 0, 1, 2, 3
BC_PERIODIC, BC_DIRICHLET, BC_SLIP_WALL = 0, 1, 2
SOURCE_NONE, SOURCE_CONVERGENCE_TEST, SOURCE_EOC_TEST_EULER, SOURCE_POISSON_NONPERIODIC, SOURCE_HARMONIC = range(5)
(IC_CONSTANT, IC_CONVERGENCE_TEST, IC_WEAK_BLAST_WAVE, IC_SEEDED_RANDOM, IC_SMOOTH_PERIODIC,
 IC_WEAK_BLAST_WAVE_SEEDED, IC_SHOCK, IC_RAREFACTION, IC_EOC_TEST_COUPLED_EULER_GRAVITY, IC_POISSON_NONPERIODIC,
 IC_HARMONIC_NONPERIODIC) = range(11)


class TrixiHipError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"trixi_hip error {code}: {message}")
        self.code = code


class Config(ctypes.Structure):
    _fields_ = [
        ("ndims", ctypes.c_int32),
        ("equation", ctypes.c_int32),
        ("polydeg", ctypes.c_int32),
        ("volume_integral", ctypes.c_int32),
        ("volume_flux", ctypes.c_int32),
        ("surface_flux", ctypes.c_int32),
        ("fv_flux", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("alpha_smooth", ctypes.c_int32),
        ("indicator_variable", ctypes.c_int32),
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
        ("boundary_outflow_faces", ctypes.c_int32),
        ("nu", ctypes.c_double),
        ("Lr", ctypes.c_double),
        ("ncomponents", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("gammas", ctypes.c_double * 4),
        ("gas_constants", ctypes.c_double * 4),
        ("gravity", ctypes.c_double),
        ("H0", ctypes.c_double),
    ]


class MeshDesc(ctypes.Structure):
    _fields_ = [
        ("nelements", ctypes.c_int64),
        ("ninterfaces", ctypes.c_int64),
        ("neighbor_ids", ctypes.c_void_p),
        ("orientations", ctypes.c_void_p),
        ("inverse_jacobian", ctypes.c_void_p),
        ("node_coordinates", ctypes.c_void_p),
        ("total_volume", ctypes.c_double),
        ("nmortars", ctypes.c_int64),
        ("mortar_neighbor_ids", ctypes.c_void_p),
        ("mortar_orientations", ctypes.c_void_p),
        ("mortar_large_sides", ctypes.c_void_p),
    ]


# name -> (restype, argtypes); the exported surface of include/trixi_hip.h
_P, _I32, _I64, _U64, _D = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double
SIGNATURES = {
    "trixi_hip_version": (ctypes.c_char_p, []),
    "trixi_hip_last_error": (ctypes.c_char_p, []),
    "trixi_hip_basis": (ctypes.c_int, [_I32, _P, _P, _P, _P, _P, _P, _P]),
    "trixi_hip_create": (ctypes.c_int, [ctypes.POINTER(Config), ctypes.POINTER(MeshDesc), ctypes.POINTER(_P)]),
    "trixi_hip_destroy": (ctypes.c_int, [_P]),
    "trixi_hip_sizes": (ctypes.c_int, [_P, _P, _P, _P]),
    "trixi_hip_rhs": (ctypes.c_int, [_P, _P, _P, _D, _P]),
    "trixi_hip_rhs_f32": (ctypes.c_int, [_P, _P, _P, _D, _P]),
    "trixi_hip_pack": (ctypes.c_int, [_P, _P, _P, _P]),
    "trixi_hip_unpack": (ctypes.c_int, [_P, _P, _P, _P]),
    "trixi_hip_compute_coefficients": (ctypes.c_int, [_P, _P, _I32, _D, _U64, _P]),
    "trixi_hip_rk_step": (ctypes.c_int, [_P, _P, _P, _D, _D, _I32, _P]),
    "trixi_hip_rk_step_f32": (ctypes.c_int, [_P, _P, _P, _D, _D, _I32, _P]),
    "trixi_hip_rk_stage": (ctypes.c_int, [_P, _P, _P, _P, _I32, _D, _D, _P]),
    "trixi_hip_sync": (ctypes.c_int, [_P]),
    "trixi_hip_max_dt": (ctypes.c_int, [_P, _P, _P, _P]),
    "trixi_hip_max_dt_noalloc": (ctypes.c_int, [_P, _P, _P, _P]),
    "trixi_hip_error_norms": (ctypes.c_int, [_P, _P, _I32, _D, _U64, _P, _P, _P]),
    "trixi_hip_integrate": (ctypes.c_int, [_P, _P, _P, _P]),
    "trixi_hip_indicator": (ctypes.c_int, [_P, _P, _P, _P]),
    "trixi_hip_check_state": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "trixi_hip_rhs_vjp": (ctypes.c_int, [_P, _P, _P, _P, _D, _D, _D, _P]),
    "trixi_hip_enable_adjoint": (ctypes.c_int, [_P]),
    "trixi_hip_rk_step_vjp": (ctypes.c_int, [_P, _P, _P, _D, _D, _P]),
}

_lib = None


def lib():
    """Load libtrixihip.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY_PATH):
            raise TrixiHipError(EINVAL, f"{LIBRARY_PATH} is missing; build it with "
                                        f"`python trixicuda.jl_amd/build.py` (there is no fallback)")
        L = ctypes.CDLL(LIBRARY_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != OK:
        msg = lib().trixi_hip_last_error()
        raise TrixiHipError(rc, msg.decode() if msg else "")
    return rc


def version():
    return lib().trixi_hip_version().decode()
