// Kernel instantiations: hyperbolic diffusion (HyperbolicDiffusionEquations
// {1,2,3}D), weak form with LLF, polydeg 4 as in
// test/tree_dgsem_{1,2,3}d/hypdiff_nonperiodic.jl (Poisson steady state on
// the x boundaries, periodic elsewhere, source_terms_poisson_nonperiodic) and
// polydeg 3 as in test/tree_dgsem_1d/hypdiff_harmonic_nonperiodic.jl.
#include "registry.hpp"

namespace trixi_hip {

template <int DIM, int N>
static void add_hypdiff() {
    constexpr int C = FLUX_CENTRAL;
    Registrar<DIM, N, HypDiff<DIM>, VI_WEAK, C, FLUX_LLF, C>::add(EQ_HYPDIFF);
    RegistrarExt<DIM, N, HypDiff<DIM>, VI_WEAK, C, FLUX_LLF, C>::add(EQ_HYPDIFF);
}

void register_hypdiff() {
    add_hypdiff<1, 4>();
    add_hypdiff<1, 5>();
    add_hypdiff<2, 4>();
    add_hypdiff<2, 5>();
    add_hypdiff<3, 4>();
    add_hypdiff<3, 5>();
}

}  // namespace trixi_hip
