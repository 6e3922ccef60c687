// Kernel instantiations: quasi-one-dimensional compressible Euler, flux
// differencing with (flux_chan_etal, flux_nonconservative_chan_etal) as volume
// and surface flux, polydeg 4 as in test/tree_dgsem_1d/eulerquasi_ec.jl and
// eulerquasi_source_terms.jl (with source_terms_convergence_test), and polydeg 3.
#include "registry.hpp"

namespace trixi_hip {

void register_euler_quasi1d() {
    constexpr int C = FLUX_CENTRAL;
    using EQ = EulerQuasi1D;
    Registrar<1, 4, EQ, VI_FD, FLUX_CHAN_ETAL, FLUX_CHAN_ETAL, C>::add(EQ_EULER_QUASI1D);
    Registrar<1, 5, EQ, VI_FD, FLUX_CHAN_ETAL, FLUX_CHAN_ETAL, C>::add(EQ_EULER_QUASI1D);
    RegistrarExt<1, 4, EQ, VI_FD, FLUX_CHAN_ETAL, FLUX_CHAN_ETAL, C>::add(EQ_EULER_QUASI1D);
    RegistrarExt<1, 5, EQ, VI_FD, FLUX_CHAN_ETAL, FLUX_CHAN_ETAL, C>::add(EQ_EULER_QUASI1D);
}

}  // namespace trixi_hip
