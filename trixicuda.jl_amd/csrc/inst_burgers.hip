// Kernel instantiations: the inviscid Burgers equation in 1D, polydeg 3, as in
// test/tree_dgsem_1d/burgers_basic.jl (weak form, LLF, source terms) and
// burgers_shock.jl / burgers_rarefraction.jl (HG shock capturing with the
// entropy-conservative flux_ec, LLF FV and surface fluxes, inflow Dirichlet
// and outflow boundaries); periodic and with boundaries / sources.
#include "registry.hpp"

namespace trixi_hip {

void register_burgers() {
    constexpr int C = FLUX_CENTRAL;
    Registrar<1, 4, Burgers, VI_WEAK, C, FLUX_LLF, C>::add(EQ_BURGERS);
    Registrar<1, 4, Burgers, VI_FD, FLUX_EC, FLUX_LLF, C>::add(EQ_BURGERS);
    Registrar<1, 4, Burgers, VI_FD, FLUX_EC, FLUX_EC, C>::add(EQ_BURGERS);
    Registrar<1, 4, Burgers, VI_HG, FLUX_EC, FLUX_LLF, FLUX_LLF>::add(EQ_BURGERS);
    RegistrarExt<1, 4, Burgers, VI_WEAK, C, FLUX_LLF, C>::add(EQ_BURGERS);
    RegistrarExt<1, 4, Burgers, VI_FD, FLUX_EC, FLUX_LLF, C>::add(EQ_BURGERS);
    RegistrarExt<1, 4, Burgers, VI_HG, FLUX_EC, FLUX_LLF, FLUX_LLF>::add(EQ_BURGERS);
}

}  // namespace trixi_hip
