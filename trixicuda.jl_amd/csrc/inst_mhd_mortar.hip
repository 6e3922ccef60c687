// Instantiations: ideal GLM-MHD on refined TreeMeshes (L2 mortars with the
// Powell / Galilean nonconservative terms), polydeg 3, as in
// test/tree_dgsem_{2,3}d/mhd_alfven_wave_mortar.jl (volume
// flux_hindenlang_gassner, surface flux_hlle, both with
// flux_nonconservative_powell), with LLF as a second surface flux; and
// flux_hlle on conforming meshes.
#include "registry.hpp"

namespace trixi_hip {

void register_mhd_mortar() {
    constexpr int C = FLUX_CENTRAL;
    constexpr int HG = FLUX_HINDENLANG_GASSNER;
    using EQ = GlmMhd;
    MortarRegistrar<2, 4, EQ, FLUX_HLLE>::add(EQ_GLM_MHD);
    MortarRegistrar<3, 4, EQ, FLUX_HLLE>::add(EQ_GLM_MHD);
    MortarRegistrar<2, 4, EQ, FLUX_LLF>::add(EQ_GLM_MHD);
    MortarRegistrar<3, 4, EQ, FLUX_LLF>::add(EQ_GLM_MHD);
    RegistrarExt<2, 4, EQ, VI_FD, HG, FLUX_HLLE, C>::add(EQ_GLM_MHD);
    RegistrarExt<3, 4, EQ, VI_FD, HG, FLUX_HLLE, C>::add(EQ_GLM_MHD);
    RegistrarExt<2, 4, EQ, VI_FD, HG, FLUX_LLF, C>::add(EQ_GLM_MHD);
    RegistrarExt<3, 4, EQ, VI_FD, HG, FLUX_LLF, C>::add(EQ_GLM_MHD);
    Registrar<2, 4, EQ, VI_FD, HG, FLUX_HLLE, C>::add(EQ_GLM_MHD);
    Registrar<3, 4, EQ, VI_FD, HG, FLUX_HLLE, C>::add(EQ_GLM_MHD);
}

}  // namespace trixi_hip
