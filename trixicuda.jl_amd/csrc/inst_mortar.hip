// Instantiations for refined TreeMeshes with L2 mortars (polydeg 3): the
// mortar passes and the rhs variants that read mortar faces, for the
// reference's mortar tests test/tree_dgsem_{2,3}d/advection_mortar.jl
// (linear advection, weak form, LLF) and test/tree_dgsem_3d/euler_mortar.jl
// (compressible Euler, weak form, LLF, source terms); the compressible Euler
// rhs variants with boundary faces (weak form and flux differencing with
// flux_ranocha, LLF surface flux) also read mortar faces, and so do the
// shock-capturing variants below.
#include "registry.hpp"

namespace trixi_hip {

void register_mortar_kernels() {
    constexpr int C = FLUX_CENTRAL;
    MortarRegistrar<2, 4, LinAdv<2>, FLUX_LLF>::add(EQ_LINADV);
    MortarRegistrar<3, 4, LinAdv<3>, FLUX_LLF>::add(EQ_LINADV);
    MortarRegistrar<2, 4, LinAdv<2>, C>::add(EQ_LINADV);
    MortarRegistrar<3, 4, LinAdv<3>, C>::add(EQ_LINADV);
    MortarRegistrar<2, 4, Euler<2>, FLUX_LLF>::add(EQ_EULER);
    MortarRegistrar<3, 4, Euler<3>, FLUX_LLF>::add(EQ_EULER);
    RegistrarExt<2, 4, LinAdv<2>, VI_WEAK, C, FLUX_LLF, C>::add(EQ_LINADV);
    RegistrarExt<3, 4, LinAdv<3>, VI_WEAK, C, FLUX_LLF, C>::add(EQ_LINADV);
    RegistrarExt<2, 4, LinAdv<2>, VI_WEAK, C, C, C>::add(EQ_LINADV);
    RegistrarExt<3, 4, LinAdv<3>, VI_WEAK, C, C, C>::add(EQ_LINADV);
    // shock capturing on refined or non-periodic meshes
    // (test/tree_dgsem_2d/euler_blob_mortar.jl: DG flux_ranocha, FV and surface LLF)
    RegistrarExt<2, 4, Euler<2>, VI_HG, FLUX_RANOCHA, FLUX_LLF, FLUX_LLF>::add(EQ_EULER);
    RegistrarExt<3, 4, Euler<3>, VI_HG, FLUX_RANOCHA, FLUX_LLF, FLUX_LLF>::add(EQ_EULER);
}

}  // namespace trixi_hip
