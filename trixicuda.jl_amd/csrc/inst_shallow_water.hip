// Kernel instantiations: shallow water.
//  2D: flux differencing with (flux_wintermeyer_etal,
//  flux_nonconservative_wintermeyer_etal) in the volume and
//  (flux_fjordholm_etal or LLF, flux_nonconservative_fjordholm_etal) at the
//  surface, polydeg 3 and 4, as in test/tree_dgsem_2d/shallowwater_ec.jl
//  (polydeg 4), shallowwater_source_terms.jl and
//  shawllowwater_source_terms_nonperiodic.jl (polydeg 3, sources, Dirichlet).
//  1D: test/tree_dgsem_1d/shallowwater_shock.jl, polydeg 4: HG shock
//  capturing with the Wintermeyer pair as DG flux and the hydrostatic
//  reconstruction (LLF + Audusse's term) as FV and surface flux, slip walls;
//  also periodic flux differencing with either surface flux.
#include "registry.hpp"

namespace trixi_hip {

void register_shallow_water() {
    constexpr int C = FLUX_CENTRAL;
    constexpr int W = FLUX_WINTERMEYER_ETAL, F = FLUX_FJORDHOLM_ETAL, HR = FLUX_HYDROSTATIC_LLF;
    using S2 = ShallowWater<2>;
    using S1 = ShallowWater<1>;
    Registrar<2, 5, S2, VI_FD, W, F, C>::add(EQ_SHALLOW_WATER);
    Registrar<2, 4, S2, VI_FD, W, F, C>::add(EQ_SHALLOW_WATER);
    Registrar<2, 4, S2, VI_FD, W, FLUX_LLF, C>::add(EQ_SHALLOW_WATER);
    RegistrarExt<2, 4, S2, VI_FD, W, FLUX_LLF, C>::add(EQ_SHALLOW_WATER);
    RegistrarExt<2, 5, S2, VI_FD, W, F, C>::add(EQ_SHALLOW_WATER);
    Registrar<1, 5, S1, VI_FD, W, HR, C>::add(EQ_SHALLOW_WATER);
    Registrar<1, 5, S1, VI_FD, W, F, C>::add(EQ_SHALLOW_WATER);
    Registrar<1, 5, S1, VI_HG, W, HR, HR>::add(EQ_SHALLOW_WATER);
    RegistrarExt<1, 5, S1, VI_HG, W, HR, HR>::add(EQ_SHALLOW_WATER);
    RegistrarExt<1, 5, S1, VI_FD, W, HR, C>::add(EQ_SHALLOW_WATER);
}

}  // namespace trixi_hip
