// Kernel instantiations of the FP32-state precision (Trixi's RealT = uEltype =
// Float32): the reference's benchmark configurations, 3D Euler flux_ranocha
// volume and surface flux (benchmark/euler_ec_3d.jl) and 2D Euler HG shock
// capturing with flux_shima_etal + LLF (benchmark/euler_shock_2d.jl), and the
// flux_ranocha + LLF variants of the measured 2D/3D configurations.
#include "registry.hpp"

namespace trixi_hip {

void register_fp32() {
    RegistrarF32State<3, 4, Euler<3>, VI_FD, FLUX_RANOCHA, FLUX_RANOCHA, FLUX_CENTRAL>::add(EQ_EULER);
    RegistrarF32State<3, 4, Euler<3>, VI_FD, FLUX_RANOCHA, FLUX_LLF, FLUX_CENTRAL>::add(EQ_EULER);
    RegistrarF32State<2, 4, Euler<2>, VI_FD, FLUX_RANOCHA, FLUX_RANOCHA, FLUX_CENTRAL>::add(EQ_EULER);
    RegistrarF32State<2, 4, Euler<2>, VI_FD, FLUX_RANOCHA, FLUX_LLF, FLUX_CENTRAL>::add(EQ_EULER);
    RegistrarF32State<2, 4, Euler<2>, VI_HG, FLUX_SHIMA_ETAL, FLUX_LLF, FLUX_LLF>::add(EQ_EULER);
    RegistrarF32State<2, 4, Euler<2>, VI_HG, FLUX_RANOCHA, FLUX_LLF, FLUX_LLF>::add(EQ_EULER);
}

}  // namespace trixi_hip
