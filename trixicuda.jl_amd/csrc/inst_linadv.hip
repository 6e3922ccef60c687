// Kernel instantiations: linear scalar advection, weak form, 1D/2D/3D.
#include "registry.hpp"

namespace trixi_hip {

template <int DIM, int N>
static void add_linadv() {
    Registrar<DIM, N, LinAdv<DIM>, VI_WEAK, FLUX_CENTRAL, FLUX_LLF, FLUX_CENTRAL>::add(EQ_LINADV);
    Registrar<DIM, N, LinAdv<DIM>, VI_WEAK, FLUX_CENTRAL, FLUX_CENTRAL, FLUX_CENTRAL>::add(EQ_LINADV);
}

void register_linadv() {
    add_linadv<1, 2>(); add_linadv<1, 3>(); add_linadv<1, 4>(); add_linadv<1, 5>();
    add_linadv<2, 2>(); add_linadv<2, 3>(); add_linadv<2, 4>(); add_linadv<2, 5>();
    add_linadv<3, 2>(); add_linadv<3, 3>(); add_linadv<3, 4>();
}

}  // namespace trixi_hip
