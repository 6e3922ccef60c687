// Kernel instantiations with boundary faces and source terms (non-periodic
// TreeMesh with Dirichlet states, source_terms_convergence_test): compressible
// Euler in 1D, 2D and 3D, weak form and flux differencing with LLF, as in the
// reference tests test/tree_dgsem_{1,2}d/euler_source_terms_nonperiodic.jl and
// test/tree_dgsem_{1,2,3}d/euler_source_terms.jl.
#include "registry.hpp"

namespace trixi_hip {

void register_euler_ext() {
    constexpr int C = FLUX_CENTRAL;
    RegistrarExt<1, 4, Euler<1>, VI_WEAK, C, FLUX_LLF, C>::add(EQ_EULER);
    RegistrarExt<2, 4, Euler<2>, VI_WEAK, C, FLUX_LLF, C>::add(EQ_EULER);
    RegistrarExt<3, 4, Euler<3>, VI_WEAK, C, FLUX_LLF, C>::add(EQ_EULER);
    RegistrarExt<1, 4, Euler<1>, VI_FD, FLUX_RANOCHA, FLUX_LLF, C>::add(EQ_EULER);
    RegistrarExt<2, 4, Euler<2>, VI_FD, FLUX_RANOCHA, FLUX_LLF, C>::add(EQ_EULER);
    RegistrarExt<3, 4, Euler<3>, VI_FD, FLUX_RANOCHA, FLUX_LLF, C>::add(EQ_EULER);
    // test/tree_dgsem_3d/euler_convergence.jl: weak form, flux_hll,
    // source_terms_eoc_test_euler
    RegistrarExt<2, 4, Euler<2>, VI_WEAK, C, FLUX_HLL, C>::add(EQ_EULER);
    RegistrarExt<3, 4, Euler<3>, VI_WEAK, C, FLUX_HLL, C>::add(EQ_EULER);
    // test/tree_dgsem_1d/euler_source_terms.jl: polydeg 4
    RegistrarExt<1, 5, Euler<1>, VI_WEAK, C, FLUX_LLF, C>::add(EQ_EULER);
}

}  // namespace trixi_hip
