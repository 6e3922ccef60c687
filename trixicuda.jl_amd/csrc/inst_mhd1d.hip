// Kernel instantiations: ideal MHD in 1D (IdealGlmMhdEquations1D, eight
// variables), flux differencing with flux_hindenlang_gassner, as in
// test/tree_dgsem_1d/mhd_alfven_wave.jl (LLF surface, polydeg 4) and
// test/tree_dgsem_1d/mhd_ec.jl (EC surface, polydeg 3), plus the weak form.
#include "registry.hpp"

namespace trixi_hip {

void register_mhd1d() {
    constexpr int C = FLUX_CENTRAL;
    using EQ = GlmMhd1D;
    Registrar<1, 4, EQ, VI_FD, FLUX_HINDENLANG_GASSNER, FLUX_LLF, C>::add(EQ_GLM_MHD);
    Registrar<1, 4, EQ, VI_FD, FLUX_HINDENLANG_GASSNER, FLUX_HINDENLANG_GASSNER, C>::add(EQ_GLM_MHD);
    Registrar<1, 4, EQ, VI_WEAK, C, FLUX_LLF, C>::add(EQ_GLM_MHD);
    Registrar<1, 5, EQ, VI_FD, FLUX_HINDENLANG_GASSNER, FLUX_LLF, C>::add(EQ_GLM_MHD);
}

}  // namespace trixi_hip
