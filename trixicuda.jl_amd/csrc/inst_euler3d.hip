// Kernel instantiations: compressible Euler 3D, weak form and flux differencing.
#include "registry.hpp"

namespace trixi_hip {

void register_euler3d() {
    using EQ = Euler<3>;
    constexpr int C = FLUX_CENTRAL;
    Registrar<3, 4, EQ, VI_WEAK, C, FLUX_LLF, C>::add(EQ_EULER);
    Registrar<3, 4, EQ, VI_WEAK, C, FLUX_LLF_NAIVE, C>::add(EQ_EULER);
    Registrar<3, 4, EQ, VI_WEAK, C, FLUX_CENTRAL, C>::add(EQ_EULER);
    Registrar<3, 4, EQ, VI_WEAK, C, FLUX_HLL, C>::add(EQ_EULER);
    Registrar<3, 4, EQ, VI_FD, FLUX_RANOCHA, FLUX_LLF, C>::add(EQ_EULER);
    RegistrarF32Log<3, 4, EQ, VI_FD, FLUX_RANOCHA, FLUX_LLF, C>::add(EQ_EULER);
    Registrar<3, 4, EQ, VI_FD, FLUX_RANOCHA, FLUX_LLF_NAIVE, C>::add(EQ_EULER);
    Registrar<3, 4, EQ, VI_FD, FLUX_RANOCHA, FLUX_RANOCHA, C>::add(EQ_EULER);
    Registrar<3, 4, EQ, VI_FD, FLUX_SHIMA_ETAL, FLUX_LLF, C>::add(EQ_EULER);
    Registrar<3, 4, EQ, VI_FD, FLUX_SHIMA_ETAL, FLUX_LLF_NAIVE, C>::add(EQ_EULER);
    Registrar<3, 4, EQ, VI_FD, FLUX_SHIMA_ETAL, FLUX_SHIMA_ETAL, C>::add(EQ_EULER);
    Registrar<3, 2, EQ, VI_FD, FLUX_RANOCHA, FLUX_LLF, C>::add(EQ_EULER);
    Registrar<3, 3, EQ, VI_FD, FLUX_RANOCHA, FLUX_LLF, C>::add(EQ_EULER);
}

}  // namespace trixi_hip
