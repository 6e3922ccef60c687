// Adjoint kernel instantiations: ideal GLM-MHD 2D and 3D, flux differencing
// with flux_hindenlang_gassner (3D; 2D also flux_central) and the
// nonconservative Powell / GLM terms.
#include "registry.hpp"

namespace trixi_hip {

void register_vjp_mhd3d() {
    using EQ = GlmMhdVjp;
    VjpRegistrar<3, 4, EQ, VI_FD, FLUX_HINDENLANG_GASSNER, FLUX_LLF>::add(EQ_GLM_MHD);
    VjpRegistrar<3, 4, EQ, VI_FD, FLUX_HINDENLANG_GASSNER, FLUX_LLF_NAIVE>::add(EQ_GLM_MHD);
    VjpRegistrar<3, 4, EQ, VI_FD, FLUX_HINDENLANG_GASSNER, FLUX_HINDENLANG_GASSNER>::add(EQ_GLM_MHD);
    VjpRegistrar<2, 4, EQ, VI_FD, FLUX_CENTRAL, FLUX_LLF>::add(EQ_GLM_MHD);
    VjpRegistrar<2, 4, EQ, VI_FD, FLUX_HINDENLANG_GASSNER, FLUX_LLF>::add(EQ_GLM_MHD);
    VjpRegistrar<2, 4, EQ, VI_FD, FLUX_HINDENLANG_GASSNER, FLUX_LLF_NAIVE>::add(EQ_GLM_MHD);
    VjpRegistrar<2, 4, EQ, VI_FD, FLUX_HINDENLANG_GASSNER, FLUX_HINDENLANG_GASSNER>::add(EQ_GLM_MHD);
}

}  // namespace trixi_hip
