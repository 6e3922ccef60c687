// Kernel instantiations: ideal GLM-MHD 3D, flux differencing with the
// nonconservative Powell/GLM terms.
#include "registry.hpp"

namespace trixi_hip {

void register_mhd3d() {
    using EQ = GlmMhd;
    constexpr int C = FLUX_CENTRAL;
    Registrar<3, 4, EQ, VI_FD, FLUX_HINDENLANG_GASSNER, FLUX_LLF, C>::add(EQ_GLM_MHD);
    RegistrarF32Log<3, 4, EQ, VI_FD, FLUX_HINDENLANG_GASSNER, FLUX_LLF, C>::add(EQ_GLM_MHD);
    Registrar<3, 4, EQ, VI_FD, FLUX_HINDENLANG_GASSNER, FLUX_LLF_NAIVE, C>::add(EQ_GLM_MHD);
    Registrar<3, 4, EQ, VI_FD, FLUX_HINDENLANG_GASSNER, FLUX_HINDENLANG_GASSNER, C>::add(EQ_GLM_MHD);
}

}  // namespace trixi_hip
