"""DGSEM solver description: basis, volume integrals, indicator, DGSEMHIP.

HIP-named twins of the reference constructors LobattoLegendreBasisGPU
(src/solvers/basis_lobatto_legendre.jl:25-100) and DGSEMGPU
(src/solvers/dgsem_gpu.jl:6-53).  The matrices are computed by the library's
host code (trixi_hip_basis) so host and device agree bit for bit.
"""
import ctypes

import numpy as np

from . import _lib as L
from . import equations as E
from .equations import density, density_pressure, first, flux_central, flux_ranocha, pressure, split_flux


class LobattoLegendreBasisHIP:
    def __init__(self, polydeg, RealT=np.float64):
        if polydeg < 1 or polydeg > 7:
            raise ValueError("polydeg must be in 1..7")
        N = polydeg + 1
        self.polydeg, self.nnodes, self.RealT = polydeg, N, RealT
        self.nodes = np.zeros(N)
        self.weights = np.zeros(N)
        self.derivative_matrix = np.zeros((N, N))
        self.derivative_split = np.zeros((N, N))
        self.derivative_dhat = np.zeros((N, N))
        self.boundary_interpolation = np.zeros((N, 2))
        self.inverse_vandermonde_legendre = np.zeros((N, N))
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        L.check(L.lib().trixi_hip_basis(N, p(self.nodes), p(self.weights), p(self.derivative_matrix),
                                        p(self.derivative_split), p(self.derivative_dhat),
                                        p(self.boundary_interpolation), p(self.inverse_vandermonde_legendre)))
        self.inverse_weights = 1.0 / self.weights
        self.derivative_split_transpose = self.derivative_split.T.copy()

    def __repr__(self):
        return f"LobattoLegendreBasisHIP(polydeg={self.polydeg})"


class VolumeIntegralWeakForm:
    kind = L.VI_WEAK


class VolumeIntegralFluxDifferencing:
    kind = L.VI_FD

    def __init__(self, volume_flux=flux_ranocha):
        self.volume_flux = volume_flux


class IndicatorHennemannGassner:
    def __init__(self, equations, basis, alpha_max=0.5, alpha_min=0.001, alpha_smooth=True,
                 variable=density_pressure):
        if variable not in (density_pressure, pressure, density, first, E.waterheight_pressure):
            raise NotImplementedError("the indicator variable must be density_pressure, pressure, density, first "
                                      "or waterheight_pressure")
        self.alpha_max, self.alpha_min, self.alpha_smooth = float(alpha_max), float(alpha_min), bool(alpha_smooth)
        self.variable = variable


class VolumeIntegralShockCapturingHG:
    kind = L.VI_HG

    def __init__(self, indicator, volume_flux_dg=flux_ranocha, volume_flux_fv=None):
        self.indicator = indicator
        self.volume_flux_dg = volume_flux_dg
        self.volume_flux_fv = volume_flux_fv


# precision name -> enum trixi_hip_precision
PRECISIONS = {"fp64": L.PREC_FP64, "mixed": L.PREC_MIXED, "mixed_f32log": L.PREC_MIXED_F32LOG,
              "fp32": L.PREC_FP32}


class DGSEMHIP:
    """DGSEMHIP(; polydeg, surface_flux=flux_central, volume_integral=VolumeIntegralWeakForm(),
    RealT=np.float64, precision=None).

    precision "fp64" computes everything in FP64; "mixed" keeps the state,
    the logarithms of the logarithmic means, the RK update and the shock
    indicator in FP64 and does the volume/surface arithmetic in FP32;
    "mixed_f32log" is "mixed" with the ln-mean logarithms in FP32 too (built
    for the benchmark's periodic flux_ranocha / flux_hindenlang_gassner
    configurations); "fp32" is the reference's RealT = Float32 mode: float32
    state vectors and FP32 arithmetic throughout (built for the reference's
    benchmark configurations, periodic 2D/3D Euler flux differencing and 2D
    shock capturing).  RealT=np.float32 selects "fp32"."""

    def __init__(self, polydeg, surface_flux=flux_central, volume_integral=None, RealT=np.float64,
                 precision=None):
        self.basis = LobattoLegendreBasisHIP(polydeg, RealT)
        self.surface_flux = surface_flux
        self.volume_integral = volume_integral if volume_integral is not None else VolumeIntegralWeakForm()
        if precision is None:
            precision = "fp32" if RealT == np.float32 else "fp64"
        if precision not in PRECISIONS:
            raise ValueError("precision must be one of " + ", ".join(repr(p) for p in PRECISIONS))
        self.precision = precision

    @property
    def polydeg(self):
        return self.basis.polydeg

    def config(self, equations, ic_length=4.0):
        """Fill the C-ABI configuration for these equations."""
        c = L.Config()
        c.ndims = equations.ndims
        c.equation = equations.equation_id
        c.polydeg = self.polydeg
        vi = self.volume_integral
        c.volume_integral = vi.kind
        sflux, s_nc = split_flux(self.surface_flux)
        c.surface_flux = sflux.flux_id
        quasi = equations.equation_id == L.EQ_EULER_QUASI1D
        swe = equations.equation_id == L.EQ_SHALLOW_WATER
        nc_needed = (equations.equation_id == L.EQ_GLM_MHD and equations.ndims > 1) or quasi or swe
        # the nonconservative term each position takes (shallow water: Wintermeyer
        # in the volume, Fjordholm at the surface, the reference tests' pairs)
        hr = getattr(sflux, "flux_id", None) == L.FLUX_HYDROSTATIC_LLF
        nc_kind = ((E.flux_nonconservative_audusse_etal if hr else E.flux_nonconservative_fjordholm_etal) if swe else
                   E.flux_nonconservative_chan_etal if quasi else E.flux_nonconservative_powell)
        nc_vol = E.flux_nonconservative_wintermeyer_etal if swe else nc_kind
        if nc_needed != (s_nc is not None) or (s_nc is not None and s_nc is not nc_kind):
            raise NotImplementedError("GLM-MHD in 2D/3D needs (flux, flux_nonconservative_powell) tuples, quasi-1D "
                                      "Euler (flux_chan_etal, flux_nonconservative_chan_etal), other equations "
                                      "(and 1D MHD) plain fluxes")
        c.volume_flux = L.FLUX_CENTRAL
        c.fv_flux = L.FLUX_CENTRAL
        c.alpha_smooth = 0
        c.alpha_max, c.alpha_min = 0.5, 0.001
        if vi.kind == L.VI_FD:
            vflux, v_nc = split_flux(vi.volume_flux)
            if nc_needed != (v_nc is not None) or (v_nc is not None and v_nc is not nc_vol):
                raise NotImplementedError("the volume flux must pair the flux with the equation's nonconservative "
                                          "term (flux_nonconservative_powell, flux_nonconservative_chan_etal)")
            c.volume_flux = vflux.flux_id
        elif vi.kind == L.VI_HG:
            dg_flux, dg_nc = split_flux(vi.volume_flux_dg)
            fv_flux, fv_nc = split_flux(vi.volume_flux_fv if vi.volume_flux_fv is not None else self.surface_flux)
            if nc_needed != (dg_nc is not None) or nc_needed != (fv_nc is not None):
                raise NotImplementedError("equations with nonconservative terms need (flux, nonconservative flux) "
                                          "pairs as DG and FV volume fluxes")
            if nc_needed and (dg_nc is not nc_vol or
                              (swe and fv_nc is not (E.flux_nonconservative_audusse_etal
                                                     if fv_flux.flux_id == L.FLUX_HYDROSTATIC_LLF
                                                     else E.flux_nonconservative_fjordholm_etal))):
                raise NotImplementedError("the DG / FV volume fluxes do not pair with the equation's "
                                          "nonconservative terms")
            c.volume_flux = dg_flux.flux_id
            c.fv_flux = fv_flux.flux_id
            c.alpha_smooth = 1 if vi.indicator.alpha_smooth else 0
            # first (the solution of a scalar equation) and density both take u[0]
            c.indicator_variable = {density_pressure: 0, pressure: 1, density: 2, first: 2,
                                    E.waterheight_pressure: 0}[vi.indicator.variable]
            c.alpha_max, c.alpha_min = vi.indicator.alpha_max, vi.indicator.alpha_min
        c.precision = PRECISIONS[self.precision]
        c.gamma = getattr(equations, "gamma", 1.4)
        c.c_h = getattr(equations, "c_h", 0.0)
        vel = getattr(equations, "advection_velocity", ())
        for d in range(3):
            c.advection_velocity[d] = vel[d] if d < len(vel) else 0.0
        c.ic_length = ic_length
        c.nu = getattr(equations, "nu", 0.0)
        c.Lr = getattr(equations, "Lr", 0.0)
        gammas = getattr(equations, "gammas", ())
        if len(gammas) > 4:
            raise NotImplementedError("at most four gas components")
        c.ncomponents = len(gammas)
        c.gravity = getattr(equations, "gravity", 0.0)
        c.H0 = getattr(equations, "H0", 0.0)
        for k, (g, r) in enumerate(zip(gammas, getattr(equations, "gas_constants", ()))):
            c.gammas[k], c.gas_constants[k] = g, r
        return c

    def __repr__(self):
        return (f"DGSEMHIP(polydeg={self.polydeg}, surface_flux={self.surface_flux}, "
                f"volume_integral={type(self.volume_integral).__name__}, precision={self.precision})")
