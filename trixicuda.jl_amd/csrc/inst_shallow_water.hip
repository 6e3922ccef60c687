// Kernel instantiations: shallow water in 2D, flux differencing with
// (flux_wintermeyer_etal, flux_nonconservative_wintermeyer_etal) in the volume
// and (flux_fjordholm_etal or LLF, flux_nonconservative_fjordholm_etal) at the
// surface, polydeg 3 and 4, as in test/tree_dgsem_2d/shallowwater_ec.jl
// (polydeg 4), shallowwater_source_terms.jl and
// shawllowwater_source_terms_nonperiodic.jl (polydeg 3, sources, Dirichlet).
#include "registry.hpp"

namespace trixi_hip {

void register_shallow_water() {
    constexpr int C = FLUX_CENTRAL;
    constexpr int W = FLUX_WINTERMEYER_ETAL, F = FLUX_FJORDHOLM_ETAL;
    using EQ = ShallowWater2D;
    Registrar<2, 5, EQ, VI_FD, W, F, C>::add(EQ_SHALLOW_WATER);
    Registrar<2, 4, EQ, VI_FD, W, F, C>::add(EQ_SHALLOW_WATER);
    Registrar<2, 4, EQ, VI_FD, W, FLUX_LLF, C>::add(EQ_SHALLOW_WATER);
    RegistrarExt<2, 4, EQ, VI_FD, W, FLUX_LLF, C>::add(EQ_SHALLOW_WATER);
    RegistrarExt<2, 5, EQ, VI_FD, W, F, C>::add(EQ_SHALLOW_WATER);
}

}  // namespace trixi_hip
