// Adjoint kernel instantiations with boundary faces (non-periodic TreeMesh,
// Dirichlet states): compressible Euler in 1D, 2D and 3D, weak form and flux
// differencing with LLF.  The convergence-test source depends on x and t
// only, so it has no adjoint term.
#include "registry.hpp"

namespace trixi_hip {

void register_vjp_euler_ext() {
    constexpr int C = FLUX_CENTRAL;
    VjpRegistrar<1, 4, EulerVjp<1>, VI_WEAK, C, FLUX_LLF, C, true>::add(EQ_EULER);
    VjpRegistrar<2, 4, EulerVjp<2>, VI_WEAK, C, FLUX_LLF, C, true>::add(EQ_EULER);
    VjpRegistrar<3, 4, EulerVjp<3>, VI_WEAK, C, FLUX_LLF, C, true>::add(EQ_EULER);
    VjpRegistrar<1, 4, EulerVjp<1>, VI_FD, FLUX_RANOCHA, FLUX_LLF, C, true>::add(EQ_EULER);
    VjpRegistrar<2, 4, EulerVjp<2>, VI_FD, FLUX_RANOCHA, FLUX_LLF, C, true>::add(EQ_EULER);
    VjpRegistrar<3, 4, EulerVjp<3>, VI_FD, FLUX_RANOCHA, FLUX_LLF, C, true>::add(EQ_EULER);
}

}  // namespace trixi_hip
