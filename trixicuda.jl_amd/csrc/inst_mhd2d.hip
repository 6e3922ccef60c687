// Kernel instantiations: ideal GLM-MHD in 2D (IdealGlmMhdEquations2D), flux
// differencing with the Powell / Galilean nonconservative terms, polydeg 3,
// as in test/tree_dgsem_2d/mhd_alfven_wave.jl (volume flux_central, surface
// LLF) and test/tree_dgsem_2d/mhd_ec.jl (flux_hindenlang_gassner throughout).
#include "registry.hpp"

namespace trixi_hip {

void register_mhd2d() {
    constexpr int C = FLUX_CENTRAL;
    using EQ = GlmMhd;
    Registrar<2, 4, EQ, VI_FD, C, FLUX_LLF, C>::add(EQ_GLM_MHD);
    Registrar<2, 4, EQ, VI_FD, FLUX_HINDENLANG_GASSNER, FLUX_LLF, C>::add(EQ_GLM_MHD);
    Registrar<2, 4, EQ, VI_FD, FLUX_HINDENLANG_GASSNER, FLUX_LLF_NAIVE, C>::add(EQ_GLM_MHD);
    Registrar<2, 4, EQ, VI_FD, FLUX_HINDENLANG_GASSNER, FLUX_HINDENLANG_GASSNER, C>::add(EQ_GLM_MHD);
}

}  // namespace trixi_hip
