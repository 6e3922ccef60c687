"""Algorithmic work per DOF-update of the fused rhs! kernel, counted from the
flux source in csrc/physics.hpp and the kernel in csrc/rhs_kernel.hpp.

Convention: +, -, *, /, sqrt, log each count 1; a fused multiply-add counts
2; comparisons, max/min and selects are free.  Work is counted once per
algorithmic quantity (each symmetric two-point flux once, each interface flux
once, primitive variables and logarithms once per node), not per executed
instruction: FP64 divisions and logarithms expand to many instructions,
which the rocprof VALU counters show separately.

Bytes per DOF-update: read u once and write du once, 2 * nvars * 8 bytes
(the state and du are FP64 in both precision modes).
"""

# ---- 3D compressible Euler, flux differencing (flux_ranocha) + LLF ----------
# node data (per node): 3 velocity divisions, kinetic energy 3 mul + 2 add,
# pressure 3, log(rho), log(p)                                              = 13
NODE_EULER3D = 3 + 5 + 3 + 2
# ln_mean(rho_l, rho_r): num 5, den 3, closed branch (y-x)/(ly-lx) 3         = 11
LN_MEAN = 11
# p_l p_r inv_ln_mean(rho_l p_r, rho_r p_l): 2 products, 2 log sums, 11, 2   = 17
INV_LN_MEAN_P = 2 + 2 + 11 + 2
# averages: v_avg 3x2, |v|^2 avg 3 mul + 2 add + 1, p_avg 2                  = 14
AVERAGES_3D = 6 + 6 + 2
# flux: f1 1, momentum 3 mul + 1 add, energy 8                               = 13
FLUX_TERMS_3D = 1 + 4 + 8
RANOCHA_3D = LN_MEAN + INV_LN_MEAN_P + AVERAGES_3D + FLUX_TERMS_3D          # = 55
# two-sided accumulation D_split[i,j] f and D_split[j,i] f: 2 x 5 FMA
PAIR_ACCUM_3D = 2 * 5 * 2
# symmetric pairs per DOF: ndims (N-1)/2 = 4.5 at N = 4
PAIRS_PER_DOF_3D = 3 * 3 / 2
# LLF (max_abs_speed): two physical fluxes 2x6, two sound speeds 2x3,
# two |v|+c 2, blend 5x5 + 1                                                = 46
LLF_EULER3D = 12 + 6 + 2 + 26
# interface fluxes per DOF: ndims N^(d-1) per element / N^d = 3/4
IFACE_PER_DOF_3D = 3 / 4
# surface integral: 6 faces x 16 nodes x 5 FMA / 64 nodes                   = 15
SURFACE_INTEGRAL_3D = 6 * 16 * 5 * 2 / 64
JACOBIAN_3D = 5

# ---- 2D compressible Euler, flux differencing (flux_ranocha) + LLF ----------
# node: 2 velocity divisions, kinetic energy 2 mul + 1 add, pressure 3, 2 logs = 10
NODE_EULER2D = 2 + 3 + 3 + 2
# averages: v_avg 2x2, |v|^2 avg 2 mul + 1 add + 1, p_avg 2                  = 10
AVERAGES_2D = 4 + 4 + 2
# flux: f1 1, momentum 2 mul + 1 add, energy 8                               = 12
FLUX_TERMS_2D = 1 + 3 + 8
RANOCHA_2D = LN_MEAN + INV_LN_MEAN_P + AVERAGES_2D + FLUX_TERMS_2D          # = 50
PAIR_ACCUM_2D = 2 * 4 * 2
PAIRS_PER_DOF_2D = 2 * 3 / 2
# LLF: two physical fluxes 2x5, two sound speeds 2x3, two |v|+c 2, blend 4x5 + 1 = 39
LLF_EULER2D = 10 + 6 + 2 + 21
IFACE_PER_DOF_2D = 2 / 4
# surface integral: 4 faces x 4 nodes x 4 FMA / 16 nodes                    = 8
SURFACE_INTEGRAL_2D = 4 * 4 * 4 * 2 / 16
JACOBIAN_2D = 4

# ---- 2D Euler HG shock capturing (flux_ranocha DG, LLF FV and surface) ------
# the DG part as above with the (1 - alpha) factors of the two D_split weights
# per pair (2 mul); the indicator per element: density_pressure 5 per node,
# two 4x4 modal transforms (2 x 16 x 4 FMA), energies 16 x 2, logistic ~6,
# over 16 nodes.  The FV subcell fluxes (per DOF 1.5 LLF fluxes of 39 plus
# 2 x 4 FMA and 2 weights) run only on blended elements and are not counted:
# the figure is a lower bound on a state with few blended elements.
HG_BLEND_PER_PAIR_2D = 2
INDICATOR_PER_DOF_2D = (16 * 5 + 2 * 16 * 4 * 2 + 16 * 2 + 6) / 16
FV_SUBCELL_PER_DOF_2D = 1.5 * (LLF_EULER2D + 2 * 4 * 2 + 2)

# ---- 3D ideal GLM-MHD, (flux_hindenlang_gassner, powell) + (LLF, powell) ----
# node: 1 division, 3 velocities, v.u 5, |B|^2 5, psi^2 2, pressure 2, 2 logs = 20
NODE_MHD3D = 1 + 3 + 5 + 5 + 2 + 2 + 2
# flux_hindenlang_gassner: ln-means 11 + 17, averages v 6, p 2, psi 2, |v|^2 6,
# |B|^2 6, f1 1, momentum 20, induction 17, psi 3, energy sum 32 + 5       = 128
HINDENLANG_GASSNER_3D = LN_MEAN + INV_LN_MEAN_P + 6 + 2 + 2 + 6 + 6 + 1 + 20 + 17 + 3 + 32 + 5
PAIR_ACCUM_MHD = 2 * 9 * 2
# Powell + GLM nonconservative term per ordered (i, j) pair of a line (all
# N^2, diagonal included): v.B 5, psi v_o 1, nine components 12, accumulate 9 FMA
POWELL_PAIR = 5 + 1 + 12 + 9 * 2
POWELL_PAIRS_PER_DOF_3D = 3 * 4
POWELL_HALF_PER_DOF_3D = 3 * 9
# LLF with Powell terms per interface: two fluxes 2x30, two fast speeds 2x20,
# lambda 4, blend 9x5, conservative energies 2x10, two Powell terms 2x18 + 9 x 2 = 219
LLF_POWELL_MHD3D = 60 + 40 + 4 + 45 + 20 + 36 + 18
SURFACE_INTEGRAL_MHD3D = 6 * 16 * 9 * 2 / 64
JACOBIAN_MHD3D = 9

FLOPS_PER_DOF = {
    "euler3d_fd_ranocha_llf": (NODE_EULER3D + PAIRS_PER_DOF_3D * (RANOCHA_3D + PAIR_ACCUM_3D)
                               + IFACE_PER_DOF_3D * LLF_EULER3D + SURFACE_INTEGRAL_3D + JACOBIAN_3D),
    "euler2d_fd_ranocha_llf": (NODE_EULER2D + PAIRS_PER_DOF_2D * (RANOCHA_2D + PAIR_ACCUM_2D)
                               + IFACE_PER_DOF_2D * LLF_EULER2D + SURFACE_INTEGRAL_2D + JACOBIAN_2D),
    "euler2d_hg_ranocha_llf": (NODE_EULER2D + PAIRS_PER_DOF_2D * (RANOCHA_2D + PAIR_ACCUM_2D + HG_BLEND_PER_PAIR_2D)
                               + INDICATOR_PER_DOF_2D + IFACE_PER_DOF_2D * LLF_EULER2D + SURFACE_INTEGRAL_2D
                               + JACOBIAN_2D),
    "mhd3d_fd_hindenlang_gassner_llf": (NODE_MHD3D + PAIRS_PER_DOF_3D * (HINDENLANG_GASSNER_3D + PAIR_ACCUM_MHD)
                                        + POWELL_PAIRS_PER_DOF_3D * POWELL_PAIR + POWELL_HALF_PER_DOF_3D
                                        + IFACE_PER_DOF_3D * LLF_POWELL_MHD3D + SURFACE_INTEGRAL_MHD3D
                                        + JACOBIAN_MHD3D),
}

# read u once and write du once, FP64 (both precision modes keep u and du in FP64)
BYTES_PER_DOF = {
    "euler3d_fd_ranocha_llf": 2 * 5 * 8,
    "euler2d_fd_ranocha_llf": 2 * 4 * 8,
    "euler2d_hg_ranocha_llf": 2 * 4 * 8,
    "mhd3d_fd_hindenlang_gassner_llf": 2 * 9 * 8,
}

# peak rates of one AMD Instinct MI355X (public datasheet)
PEAK = {"hbm_GBps": 8000.0, "fp64_TFLOPs": 78.6, "fp32_TFLOPs": 157.3}

if __name__ == "__main__":
    for k, v in FLOPS_PER_DOF.items():
        print(f"{k}: {v:.2f} flop/DOF, {BYTES_PER_DOF[k]} B/DOF")
