// Kernel instantiations: compressible Euler 2D, Hennemann-Gassner shock capturing.
#include "registry.hpp"

namespace trixi_hip {

void register_euler2d_hg() {
    using EQ = Euler<2>;
    Registrar<2, 4, EQ, VI_HG, FLUX_RANOCHA, FLUX_LLF, FLUX_LLF>::add(EQ_EULER);
    RegistrarF32Log<2, 4, EQ, VI_HG, FLUX_RANOCHA, FLUX_LLF, FLUX_LLF>::add(EQ_EULER);
    Registrar<2, 4, EQ, VI_HG, FLUX_SHIMA_ETAL, FLUX_LLF, FLUX_LLF>::add(EQ_EULER);
    Registrar<2, 4, EQ, VI_HG, FLUX_RANOCHA, FLUX_LLF_NAIVE, FLUX_LLF_NAIVE>::add(EQ_EULER);
    Registrar<2, 4, EQ, VI_HG, FLUX_SHIMA_ETAL, FLUX_LLF_NAIVE, FLUX_LLF_NAIVE>::add(EQ_EULER);
}

}  // namespace trixi_hip
