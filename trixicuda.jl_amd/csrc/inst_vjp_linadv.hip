// Adjoint kernel instantiations: linear scalar advection, weak form.
#include "registry.hpp"

namespace trixi_hip {

template <int DIM, int N>
static void add_vjp_linadv() {
    constexpr int C = FLUX_CENTRAL;
    VjpRegistrar<DIM, N, LinAdvVjp<DIM>, VI_WEAK, C, FLUX_LLF>::add(EQ_LINADV);
    VjpRegistrar<DIM, N, LinAdvVjp<DIM>, VI_WEAK, C, FLUX_CENTRAL>::add(EQ_LINADV);
}

void register_vjp_linadv() {
    add_vjp_linadv<1, 4>();
    add_vjp_linadv<2, 4>();
    add_vjp_linadv<3, 4>();
    add_vjp_linadv<1, 2>();
    add_vjp_linadv<1, 3>();
}

}  // namespace trixi_hip
