// Adjoint kernel instantiations: compressible Euler 3D, weak form, flux differencing and
// Hennemann-Gassner shock capturing.
#include "registry.hpp"

namespace trixi_hip {

void register_vjp_euler3d() {
    using EQ = EulerVjp<3>;
    constexpr int C = FLUX_CENTRAL;
    VjpRegistrar<3, 4, EQ, VI_WEAK, C, FLUX_LLF>::add(EQ_EULER);
    VjpRegistrar<3, 4, EQ, VI_WEAK, C, FLUX_LLF_NAIVE>::add(EQ_EULER);
    VjpRegistrar<3, 4, EQ, VI_WEAK, C, FLUX_CENTRAL>::add(EQ_EULER);
    VjpRegistrar<3, 4, EQ, VI_FD, FLUX_RANOCHA, FLUX_LLF>::add(EQ_EULER);
    VjpRegistrar<3, 4, EQ, VI_FD, FLUX_RANOCHA, FLUX_LLF_NAIVE>::add(EQ_EULER);
    VjpRegistrar<3, 4, EQ, VI_FD, FLUX_RANOCHA, FLUX_RANOCHA>::add(EQ_EULER);
    VjpRegistrar<3, 4, EQ, VI_FD, FLUX_SHIMA_ETAL, FLUX_LLF>::add(EQ_EULER);
    VjpRegistrar<3, 4, EQ, VI_FD, FLUX_SHIMA_ETAL, FLUX_LLF_NAIVE>::add(EQ_EULER);
    VjpRegistrar<3, 4, EQ, VI_FD, FLUX_SHIMA_ETAL, FLUX_SHIMA_ETAL>::add(EQ_EULER);
    VjpRegistrar<3, 4, EQ, VI_HG, FLUX_RANOCHA, FLUX_LLF, FLUX_LLF>::add(EQ_EULER);
    VjpRegistrar<3, 4, EQ, VI_HG, FLUX_RANOCHA, FLUX_LLF_NAIVE, FLUX_LLF_NAIVE>::add(EQ_EULER);
    VjpRegistrar<3, 4, EQ, VI_HG, FLUX_SHIMA_ETAL, FLUX_LLF, FLUX_LLF>::add(EQ_EULER);
    VjpRegistrar<3, 4, EQ, VI_HG, FLUX_SHIMA_ETAL, FLUX_LLF_NAIVE, FLUX_LLF_NAIVE>::add(EQ_EULER);
    VjpRegistrar<3, 4, EQ, VI_HG, FLUX_RANOCHA, FLUX_RANOCHA, FLUX_RANOCHA>::add(EQ_EULER);
}

}  // namespace trixi_hip
