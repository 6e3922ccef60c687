// mortar_kernel.hpp -- L2 mortar fluxes of a refined TreeMesh (FP64).
//
// One pass per rhs evaluation, before the fused rhs kernel, replacing the
// reference's prolong2mortars! + calc_mortar_flux! + the mortar part of the
// surface flux copy (dg_2d.jl:549-709, dg_3d.jl:640-900; kernels
// dg_2d_kernel.jl:1167-1390, dg_3d_kernel.jl:1361-1790):
//
//  * the large element's face values are interpolated to every small face
//    (forward_lower / forward_upper in each tangential direction; position
//    p = t1 + 2 t2 with t1, t2 the halves along the first and second
//    tangential direction),
//  * the surface flux is evaluated with the large element on its side of the
//    mortar (large_side 1: large on the left),
//  * the small faces take it as their face flux, and the large face takes the
//    L2 projection (reverse_lower / reverse_upper) of all small-face fluxes.
//
// The fluxes go to mflux[(slot * NL + q) * NV + v], one slot per element
// face on a mortar (slot m*(K+1)+p for small position p, m*(K+1)+K for the
// large face), which the rhs kernel reads at faces encoded nbr = -2 - slot.
// A mortar is handled by K*NL lanes of one workgroup: lanes stage the large
// face in LDS, evaluate one small-face node each, and the first NL lanes of
// the mortar project the K small-face fluxes back to the large face.
#pragma once

#include "rhs_kernel.hpp"

namespace trixi_hip {

struct MortarArgs {
    const double* u;          // SoA state u[(v*NN + node)*E + e]
    double* mflux;            // [slots][NL][NV]
    const int32_t* ids;       // [nmortars][K+1]: small elements by position, large
    const int32_t* ori;       // [nmortars] 1..DIM
    const int32_t* ls;        // [nmortars] 1: large on the left, 2: on the right
    int64_t nmortars;
    int64_t E;
    PhysParams phys;
    double F[2][MAXN * MAXN]; // forward_lower, forward_upper (row-major [small node][large node])
    double R[2][MAXN * MAXN]; // reverse_lower, reverse_upper (row-major [large node][small node])
};

template <int DIM, int N>
__device__ __forceinline__ int face_base(int q, int od) {
    return line_base<DIM, N>(q, od);
}

template <int DIM, int N, class EQ, int SF>
__global__ void __launch_bounds__(256) mortar_kernel(const MortarArgs a) {
    constexpr int NL = cpow(N, DIM - 1);
    constexpr int NN = cpow(N, DIM);
    constexpr int NV = EQ::NV;
    constexpr int K = 1 << (DIM - 1);
    constexpr int LPM = K * NL;             // lanes per mortar
    constexpr int MPB = 256 / LPM;          // mortars per workgroup
    constexpr bool LOGS = flux_needs_logs(SF);
    static_assert(LPM <= 256 && 256 % LPM == 0, "mortar lanes must tile the workgroup");
    using Node = typename EQ::template Node<double>;
    __shared__ double large_face[MPB][NV][NL];
    __shared__ double fstar[MPB][K][NL][NV];

    const int tid = threadIdx.x;
    const int lm = tid / LPM;               // mortar within the workgroup
    const int p = (tid % LPM) / NL;         // small position
    const int q = tid % NL;                 // small-face node
    const int64_t m = (int64_t)blockIdx.x * MPB + lm;
    const bool valid = m < a.nmortars;
    const int64_t E = a.E;
    const uint32_t E32 = (uint32_t)E;
    const int64_t slab = (int64_t)NN * E;
    const PhysR<double> P = PhysR<double>::from(a.phys);

    int od = 0, side = 1;
    int32_t large = 0, small = 0;
    if (valid) {
        od = a.ori[m] - 1;
        side = a.ls[m];
        large = a.ids[m * (K + 1) + K];
        small = a.ids[m * (K + 1) + p];
    }
    const int sod = cpow(N, od);
    const int pos_large = (side == 1) ? N - 1 : 0;
    const int pos_small = (side == 1) ? 0 : N - 1;
    if (valid && p == 0) {
        const int node = face_base<DIM, N>(q, od) + pos_large * sod;
#pragma unroll
        for (int v = 0; v < NV; ++v) large_face[lm][v][q] = a.u[soa(v, node, large, slab, E32)];
    }
    __syncthreads();
    if (valid) {
        const int p1 = p & 1, p2 = (p >> 1) & 1;
        const int q1 = q % N, q2 = q / N;
        double uL[NV], uS[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) uL[v] = 0.0;
        for (int b = 0; b < (DIM == 3 ? N : 1); ++b) {
            const double fb = (DIM == 3) ? a.F[p2][q2 * N + b] : 1.0;
            for (int c = 0; c < N; ++c) {
                const double w = a.F[p1][q1 * N + c] * fb;
#pragma unroll
                for (int v = 0; v < NV; ++v) uL[v] += w * large_face[lm][v][c + N * b];
            }
        }
        const int node = face_base<DIM, N>(q, od) + pos_small * sod;
#pragma unroll
        for (int v = 0; v < NV; ++v) uS[v] = a.u[soa(v, node, small, slab, E32)];
        const Node nl = EQ::template node<double, LOGS>(uL, P);
        const Node ns = EQ::template node<double, LOGS>(uS, P);
        double f[NV], nc_small[NV], nc_large[NV];
        if (side == 1)
            EQ::template two_point<SF>(nl, ns, od, P, f);
        else
            EQ::template two_point<SF>(ns, nl, od, P, f);
        // nonconservative terms in the reference's mortar order
        // (dg_2d_kernel.jl:1318-1337, dg_3d_kernel.jl:1598-1640): the small
        // face takes 1/2 nc(u_large, u_small), the projected large-face flux
        // 1/2 nc(u_small, u_large)
#pragma unroll
        for (int v = 0; v < NV; ++v) nc_small[v] = nc_large[v] = 0.0;
        if constexpr (EQ::NONCONS) {
            EQ::template noncons<SF>(nl, ns, od, P, nc_small);
            EQ::template noncons<SF>(ns, nl, od, P, nc_large);
        }
        double* out = a.mflux + ((m * (K + 1) + p) * NL + q) * NV;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            out[v] = f[v] + 0.5 * nc_small[v];
            fstar[lm][p][q][v] = f[v] + 0.5 * nc_large[v];
        }
    }
    __syncthreads();
    if (valid && p == 0) {
        const int j1 = q % N, j2 = q / N;
        double acc[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] = 0.0;
        for (int pp = 0; pp < K; ++pp) {
            const int p1 = pp & 1, p2 = (pp >> 1) & 1;
            for (int qq = 0; qq < NL; ++qq) {
                const int q1 = qq % N, q2 = qq / N;
                double r = a.R[p1][j1 * N + q1];
                if (DIM == 3) r *= a.R[p2][j2 * N + q2];
#pragma unroll
                for (int v = 0; v < NV; ++v) acc[v] += r * fstar[lm][pp][qq][v];
            }
        }
        double* out = a.mflux + ((m * (K + 1) + K) * NL + q) * NV;
#pragma unroll
        for (int v = 0; v < NV; ++v) out[v] = acc[v];
    }
}

template <int DIM, int N, class EQ, int SF>
hipError_t launch_mortar(const MortarArgs& a, hipStream_t s) {
    constexpr int LPM = (1 << (DIM - 1)) * cpow(N, DIM - 1);
    constexpr int MPB = 256 / LPM;
    const int64_t blocks = (a.nmortars + MPB - 1) / MPB;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL((mortar_kernel<DIM, N, EQ, SF>), dim3((unsigned)blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace trixi_hip
