// registry.hpp -- table of the compiled rhs kernel instantiations.
#pragma once

#include "mortar_kernel.hpp"
#include "rhs_kernel.hpp"
#include "vjp_kernel.hpp"

namespace trixi_hip {

struct KernelKey {
    int dim, N, eq, vol, vf, sf, fvf, prec, rk;
    int ext = 0;  // 1: boundary faces and source terms (non-periodic meshes)
    bool operator<(const KernelKey& o) const {
        const int a[10] = {dim, N, eq, vol, vf, sf, fvf, prec, rk, ext};
        const int b[10] = {o.dim, o.N, o.eq, o.vol, o.vf, o.sf, o.fvf, o.prec, o.rk, o.ext};
        for (int i = 0; i < 10; ++i)
            if (a[i] != b[i]) return a[i] < b[i];
        return false;
    }
};

// registry equation id of multicomponent Euler with nc components
constexpr int multi_key(int nc) { return 100 + nc; }

using LaunchFn = hipError_t (*)(const RhsArgs&, hipStream_t);
using VjpFn = hipError_t (*)(const VjpArgs&, hipStream_t);
using MortarFn = hipError_t (*)(const MortarArgs&, hipStream_t);

void register_kernel(const KernelKey& key, LaunchFn fn);
LaunchFn find_kernel(const KernelKey& key);
void register_vjp(const KernelKey& key, VjpFn fn);
VjpFn find_vjp(const KernelKey& key);
// mortar passes are keyed by dim, N, eq and sf only (other fields zero)
void register_mortar(const KernelKey& key, MortarFn fn);
MortarFn find_mortar(const KernelKey& key);

template <int DIM, int N, class EQ, int SF>
struct MortarRegistrar {
    static void add(int eq) { register_mortar({DIM, N, eq, 0, 0, SF, 0, 0, 0, 0}, &launch_mortar<DIM, N, EQ, SF>); }
};

template <int DIM, int N, class EQ, int VOL, int VF, int SF, int FVF = FLUX_CENTRAL, bool EXT = false>
struct VjpRegistrar {
    static void add(int eq) {
        constexpr int M = pick_m_vjp<DIM, N, VOL, EQ::NV, EQ::NRF>();
        constexpr int MG = pick_m_vjp<DIM, N, VOL, EQ::NV, EQ::NRF, DIM>();
        register_vjp({DIM, N, eq, VOL, VF, SF, FVF, 0, 0, EXT ? 1 : 0},
                     &launch_vjp<DIM, N, EQ, VOL, VF, SF, FVF, double, M, EXT, MG>);
    }
};

// Workgroup width in elements: fill about 256 lanes and keep the LDS tiles
// (node records with their two logarithms, direction accumulator) within
// 40 KiB (FP64 arithmetic) so that several workgroups share a CU, or 64 KiB
// with FP32 arithmetic and up to five variables, whose wider tiles (3D Euler: 16 elements, whole
// 128-byte rows of an FP64 state) measured faster despite two waves per SIMD
// instead of 2.5: 3D Euler level 5, mixed rhs! 57.2 -> 54.9 us, FP32-state
// rhs! 45.0 -> 40.4 us, RK steps -11 % and -8 % (the FP64 kernel at 8
// elements: 80 -> 101 us).
#ifndef TRIXI_HIP_LDS_BUDGET_KIB
#define TRIXI_HIP_LDS_BUDGET_KIB 40
#endif
#ifndef TRIXI_HIP_LDS_BUDGET_F32_KIB
#define TRIXI_HIP_LDS_BUDGET_F32_KIB 64
#endif
#ifndef TRIXI_HIP_STAGE_NODES
#define TRIXI_HIP_STAGE_NODES 1
#endif
#ifndef TRIXI_HIP_MAX_THREADS
#define TRIXI_HIP_MAX_THREADS 256
#endif
// RS: bytes of the arithmetic type; LS: bytes of the staged logarithms.
template <int DIM, int N, int NV, int NRF, int RS, int LS>
constexpr int pick_m() {
    constexpr int NL = cpow(N, DIM - 1), NN = cpow(N, DIM);
    constexpr int per_elem = NN * ((TRIXI_HIP_STAGE_NODES ? NRF * RS + 2 * LS : 0) + (DIM > 1 ? NV * RS : 0));
    int m = TRIXI_HIP_MAX_THREADS / NL;
    if (m < 1) m = 1;
    // (the nine-variable MHD system keeps the FP64 budget: its wider tiles
    // take 256 VGPRs)
    constexpr int budget = (RS == 4 && NV <= 5 ? TRIXI_HIP_LDS_BUDGET_F32_KIB : TRIXI_HIP_LDS_BUDGET_KIB) * 1024;
    while (m > 1 && m * per_elem > budget) m /= 2;
    return m;
}

template <int DIM, int N, class EQ, int VOL, int VF, int SF, int FVF>
struct Registrar {
    static void add(int eq) {
        constexpr int MD = pick_m<DIM, N, EQ::NV, EQ::NRF, 8, 8>();
        constexpr int MF = pick_m<DIM, N, EQ::NV, EQ::NRF, 4, 8>();
        register_kernel({DIM, N, eq, VOL, VF, SF, FVF, 0, 0},
                        &launch_rhs<DIM, N, EQ, VOL, VF, SF, FVF, double, MD, false>);
        register_kernel({DIM, N, eq, VOL, VF, SF, FVF, 0, 1},
                        &launch_rhs<DIM, N, EQ, VOL, VF, SF, FVF, double, MD, true>);
        register_kernel({DIM, N, eq, VOL, VF, SF, FVF, 1, 0},
                        &launch_rhs<DIM, N, EQ, VOL, VF, SF, FVF, float, MF, false>);
        register_kernel({DIM, N, eq, VOL, VF, SF, FVF, 1, 1},
                        &launch_rhs<DIM, N, EQ, VOL, VF, SF, FVF, float, MF, true>);
    }
};

// The same four variants with boundary faces and source terms.
template <int DIM, int N, class EQ, int VOL, int VF, int SF, int FVF>
struct RegistrarExt {
    static void add(int eq) {
        constexpr int MD = pick_m<DIM, N, EQ::NV, EQ::NRF, 8, 8>();
        constexpr int MF = pick_m<DIM, N, EQ::NV, EQ::NRF, 4, 8>();
        register_kernel({DIM, N, eq, VOL, VF, SF, FVF, 0, 0, 1},
                        &launch_rhs<DIM, N, EQ, VOL, VF, SF, FVF, double, MD, false, true>);
        register_kernel({DIM, N, eq, VOL, VF, SF, FVF, 0, 1, 1},
                        &launch_rhs<DIM, N, EQ, VOL, VF, SF, FVF, double, MD, true, true>);
        register_kernel({DIM, N, eq, VOL, VF, SF, FVF, 1, 0, 1},
                        &launch_rhs<DIM, N, EQ, VOL, VF, SF, FVF, float, MF, false, true>);
        register_kernel({DIM, N, eq, VOL, VF, SF, FVF, 1, 1, 1},
                        &launch_rhs<DIM, N, EQ, VOL, VF, SF, FVF, float, MF, true, true>);
    }
};

// Precision 2 (TRIXI_HIP_MIXED_F32LOG): the mixed mode with the per-node
// logarithms of the ln-means taken and staged in FP32, for the configurations
// whose fluxes take ln-means (periodic meshes).
template <int DIM, int N, class EQ, int VOL, int VF, int SF, int FVF>
struct RegistrarF32Log {
    static void add(int eq) {
        constexpr int MF = pick_m<DIM, N, EQ::NV, EQ::NRF, 4, 4>();
        register_kernel({DIM, N, eq, VOL, VF, SF, FVF, 2, 0},
                        &launch_rhs<DIM, N, EQ, VOL, VF, SF, FVF, float, MF, false, false, float>);
        register_kernel({DIM, N, eq, VOL, VF, SF, FVF, 2, 1},
                        &launch_rhs<DIM, N, EQ, VOL, VF, SF, FVF, float, MF, true, false, float>);
    }
};

// Precision 3 (TRIXI_HIP_FP32): FP32 state vectors and FP32 arithmetic,
// logarithms included (Trixi's RealT = uEltype = Float32), periodic meshes.
template <int DIM, int N, class EQ, int VOL, int VF, int SF, int FVF>
struct RegistrarF32State {
    static void add(int eq) {
        constexpr int MF = pick_m<DIM, N, EQ::NV, EQ::NRF, 4, 4>();
        register_kernel({DIM, N, eq, VOL, VF, SF, FVF, 3, 0},
                        &launch_rhs<DIM, N, EQ, VOL, VF, SF, FVF, float, MF, false, false, float, float>);
        register_kernel({DIM, N, eq, VOL, VF, SF, FVF, 3, 1},
                        &launch_rhs<DIM, N, EQ, VOL, VF, SF, FVF, float, MF, true, false, float, float>);
    }
};

}  // namespace trixi_hip
