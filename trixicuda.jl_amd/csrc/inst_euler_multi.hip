// Kernel instantiations: multicomponent compressible Euler in 1D and 2D with
// one or two gas components, polydeg 3, as in test/tree_dgsem_{1,2}d/
// eulermulti_ec.jl (flux differencing, flux_ranocha volume and surface flux)
// and eulermulti_es.jl (flux_ranocha volume flux, LLF surface flux), plus
// the weak form with LLF.
#include "registry.hpp"

namespace trixi_hip {

template <int DIM, int NC>
static void add_multi() {
    constexpr int C = FLUX_CENTRAL;
    using EQ = EulerMulti<DIM, NC>;
    Registrar<DIM, 4, EQ, VI_FD, FLUX_RANOCHA, FLUX_RANOCHA, C>::add(multi_key(NC));
    Registrar<DIM, 4, EQ, VI_FD, FLUX_RANOCHA, FLUX_LLF, C>::add(multi_key(NC));
    Registrar<DIM, 4, EQ, VI_WEAK, C, FLUX_LLF, C>::add(multi_key(NC));
}

void register_euler_multi() {
    add_multi<1, 1>();
    add_multi<1, 2>();
    add_multi<2, 1>();
    add_multi<2, 2>();
}

}  // namespace trixi_hip
