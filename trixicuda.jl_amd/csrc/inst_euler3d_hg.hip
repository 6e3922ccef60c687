// Kernel instantiations: compressible Euler 3D, Hennemann-Gassner shock capturing.
#include "registry.hpp"

namespace trixi_hip {

void register_euler3d_hg() {
    using EQ = Euler<3>;
    Registrar<3, 4, EQ, VI_HG, FLUX_RANOCHA, FLUX_LLF, FLUX_LLF>::add(EQ_EULER);
    Registrar<3, 4, EQ, VI_HG, FLUX_SHIMA_ETAL, FLUX_LLF, FLUX_LLF>::add(EQ_EULER);
    Registrar<3, 4, EQ, VI_HG, FLUX_RANOCHA, FLUX_LLF_NAIVE, FLUX_LLF_NAIVE>::add(EQ_EULER);
    Registrar<3, 4, EQ, VI_HG, FLUX_SHIMA_ETAL, FLUX_LLF_NAIVE, FLUX_LLF_NAIVE>::add(EQ_EULER);
    // test/tree_dgsem_3d/euler_shock.jl: flux_ranocha as DG, FV and surface flux
    Registrar<3, 4, EQ, VI_HG, FLUX_RANOCHA, FLUX_RANOCHA, FLUX_RANOCHA>::add(EQ_EULER);
}

}  // namespace trixi_hip
