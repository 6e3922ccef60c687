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

FLOPS_PER_DOF = {
    "euler3d_fd_ranocha_llf": (NODE_EULER3D + PAIRS_PER_DOF_3D * (RANOCHA_3D + PAIR_ACCUM_3D)
                               + IFACE_PER_DOF_3D * LLF_EULER3D + SURFACE_INTEGRAL_3D + JACOBIAN_3D),
}

BYTES_PER_DOF = {
    "euler3d_fd_ranocha_llf": 2 * 5 * 8,
}

# peak rates of one AMD Instinct MI355X (public datasheet)
PEAK = {"hbm_GBps": 8000.0, "fp64_TFLOPs": 78.6, "fp32_TFLOPs": 157.3}

if __name__ == "__main__":
    for k, v in FLOPS_PER_DOF.items():
        print(f"{k}: {v:.2f} flop/DOF, {BYTES_PER_DOF[k]} B/DOF")
