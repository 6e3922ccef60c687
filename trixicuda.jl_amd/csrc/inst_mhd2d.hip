// Kernel instantiations: ideal GLM-MHD in 2D (IdealGlmMhdEquations2D), flux
// differencing with the Powell / Galilean nonconservative terms, polydeg 3,
// as in test/tree_dgsem_2d/mhd_alfven_wave.jl (volume flux_central, surface
// LLF) and test/tree_dgsem_2d/mhd_ec.jl (flux_hindenlang_gassner throughout);
// MHD shock capturing in 2D/3D at polydeg 3 and 4.
#include "registry.hpp"

namespace trixi_hip {

void register_mhd2d() {
    constexpr int C = FLUX_CENTRAL;
    using EQ = GlmMhd;
    Registrar<2, 4, EQ, VI_FD, C, FLUX_LLF, C>::add(EQ_GLM_MHD);
    Registrar<2, 4, EQ, VI_FD, FLUX_HINDENLANG_GASSNER, FLUX_LLF, C>::add(EQ_GLM_MHD);
    Registrar<2, 4, EQ, VI_FD, FLUX_HINDENLANG_GASSNER, FLUX_LLF_NAIVE, C>::add(EQ_GLM_MHD);
    Registrar<2, 4, EQ, VI_FD, FLUX_HINDENLANG_GASSNER, FLUX_HINDENLANG_GASSNER, C>::add(EQ_GLM_MHD);
    // test/tree_dgsem_{2,3}d/mhd_shock.jl: HG shock capturing with
    // (flux_hindenlang_gassner, flux_nonconservative_powell) as DG, FV and
    // surface flux, indicator density_pressure
    Registrar<2, 4, EQ, VI_HG, FLUX_HINDENLANG_GASSNER, FLUX_HINDENLANG_GASSNER, FLUX_HINDENLANG_GASSNER>::add(EQ_GLM_MHD);
    Registrar<2, 4, EQ, VI_HG, FLUX_HINDENLANG_GASSNER, FLUX_LLF, FLUX_LLF>::add(EQ_GLM_MHD);
    Registrar<3, 4, EQ, VI_HG, FLUX_HINDENLANG_GASSNER, FLUX_HINDENLANG_GASSNER, FLUX_HINDENLANG_GASSNER>::add(EQ_GLM_MHD);
    // the tests' own polydeg 4
    Registrar<2, 5, EQ, VI_HG, FLUX_HINDENLANG_GASSNER, FLUX_HINDENLANG_GASSNER, FLUX_HINDENLANG_GASSNER>::add(EQ_GLM_MHD);
    Registrar<3, 5, EQ, VI_HG, FLUX_HINDENLANG_GASSNER, FLUX_HINDENLANG_GASSNER, FLUX_HINDENLANG_GASSNER>::add(EQ_GLM_MHD);
}

}  // namespace trixi_hip
