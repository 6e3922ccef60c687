// indicator.hpp -- the Hennemann-Gassner blending factor per element
// (Trixi's IndicatorHennemannGassner, reference src/solvers/indicators.jl:7-42
// with the external calc_indicator_hennemann_gassner!): the indicator
// variable at the nodes, its modal coefficients through the inverse Legendre
// Vandermonde, the energy fractions of the highest and second-highest modes,
// the logistic and the clipping; used by the indicator kernel and its
// reverse.
#pragma once

#include "physics.hpp"

namespace trixi_hip {

// internal indicator-variable code of shallow water's waterheight_pressure
// (its configuration value is density_pressure, the same position in Trixi's
// API)
constexpr int IND_WATERHEIGHT_PRESSURE = 3;

template <int DIM, int N>
__device__ __forceinline__ int max_mode(int n) {
    int maxc = 0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const int c = (n / cpow(N, d)) % N;
        maxc = c > maxc ? c : maxc;
    }
    return maxc;
}

// modal coefficients of the indicator variable (density * pressure,
// pressure or density) at this lane's node (all lanes of the block take part:
// it synchronises)
// the indicator variable (density * pressure, pressure or density) at node n
// of element ec
template <int DIM, int N, typename S>
__device__ __forceinline__ double indicator_value(const S* __restrict__ u, int64_t ec, int64_t E,
                                                  const PhysR<double>& P, int n, int var, int nv) {
    constexpr int NN = cpow(N, DIM);
    // compressible Euler (DIM + 2 variables) or a scalar equation (nv = 1:
    // the variable is the solution itself)
    double m;
    if (var == IND_WATERHEIGHT_PRESSURE) {
        // shallow water: waterheight_pressure = h g h^2 / 2
        const double h = (double)u[(int64_t)n * E + ec];
        m = h * (0.5 * P.grav * h * h);
    } else if (nv == 1) {
        m = (double)u[(int64_t)n * E + ec];
    } else if (nv == DIM + 2) {
        double uc[DIM + 2];
#pragma unroll
        for (int v = 0; v < DIM + 2; ++v) uc[v] = (double)u[soa(v, n, ec, (int64_t)NN * E, (uint32_t)E)];
        m = Euler<DIM>::indicator_variable(uc, P, var);
    } else {
        // ideal GLM-MHD (nine variables): density_pressure with the magnetic
        // and cleaning energies taken out of the pressure
        double uc[9];
#pragma unroll
        for (int v = 0; v < 9; ++v) uc[v] = (double)u[soa(v, n, ec, (int64_t)NN * E, (uint32_t)E)];
        const double ke = (uc[1] * uc[1] + uc[2] * uc[2] + uc[3] * uc[3]) / uc[0];
        const double me = uc[5] * uc[5] + uc[6] * uc[6] + uc[7] * uc[7] + uc[8] * uc[8];
        m = uc[0] * P.gm1 * (uc[4] - 0.5 * ke - 0.5 * me);
    }
    return m;
}

// modal coefficient at this lane's node of the nodal values m of element
// slot `el` (N^DIM lanes per slot; every lane of the block takes part: it
// synchronises)
template <int DIM, int N>
__device__ __forceinline__ double modal_transform(double m, const double* __restrict__ Vinv, double* buf, int el,
                                                  int n) {
    constexpr int NN = cpow(N, DIM);
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const int s = cpow(N, d);
        const int idx = (n / s) % N, base = n - idx * s;
        buf[threadIdx.x] = m;
        __syncthreads();
        double r = 0.0;
#pragma unroll
        for (int ii = 0; ii < N; ++ii) r += Vinv[idx * N + ii] * buf[el * NN + base + ii * s];
        __syncthreads();
        m = r;
    }
    return m;
}

// The rows of the inverse Vandermonde matrix a lane's modal transform uses
// (row (n / N^d) % N for direction d), loaded before the state so that their
// latency overlaps it instead of following the transform's barriers.
template <int DIM, int N>
struct VinvRows {
    double r[DIM][N];
};
template <int DIM, int N>
__device__ __forceinline__ VinvRows<DIM, N> vinv_rows(const double* __restrict__ Vinv, int n) {
    VinvRows<DIM, N> v;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const int idx = (n / cpow(N, d)) % N;
#pragma unroll
        for (int ii = 0; ii < N; ++ii) v.r[d][ii] = Vinv[idx * N + ii];
    }
    return v;
}
// modal_transform with preloaded rows (the same products in the same order)
template <int DIM, int N>
__device__ __forceinline__ double modal_transform_rows(double m, const VinvRows<DIM, N>& V, double* buf, int el,
                                                       int n) {
    constexpr int NN = cpow(N, DIM);
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const int s = cpow(N, d);
        const int idx = (n / s) % N, base = n - idx * s;
        buf[threadIdx.x] = m;
        __syncthreads();
        double r = 0.0;
#pragma unroll
        for (int ii = 0; ii < N; ++ii) r += V.r[d][ii] * buf[el * NN + base + ii * s];
        __syncthreads();
        m = r;
    }
    return m;
}

// modal coefficients of the indicator variable at this lane's node
template <int DIM, int N, typename S>
__device__ __forceinline__ double modal_indicator(const S* __restrict__ u, int64_t ec, int64_t E,
                                                  const PhysR<double>& P, const double* __restrict__ Vinv,
                                                  double* buf, int el, int n, int var, int nv) {
    return modal_transform<DIM, N>(indicator_value<DIM, N, S>(u, ec, E, P, n, var, nv), Vinv, buf, el, n);
}

template <int DIM, int N>
__device__ __forceinline__ void modal_energies(const double* e2, int el, double& total, double& clip1,
                                               double& clip2) {
    constexpr int NN = cpow(N, DIM);
    total = clip1 = clip2 = 0.0;
    for (int q = 0; q < NN; ++q) {
        const int maxc = max_mode<DIM, N>(q);
        const double v = e2[el * NN + q];
        total += v;
        if (maxc < N - 1) clip1 += v;
        if (maxc < N - 2) clip2 += v;
    }
}

// alpha_raw of one element from its modal energies: the larger of the
// highest-mode and second-highest-mode energy fractions through the
// logistic 1 / (1 + exp(-s / threshold (E - threshold))), then clipped to 0
// below alpha_min, to 1 above 1 - alpha_min, and to alpha_max.
__device__ __forceinline__ double hg_alpha(double total, double clip1, double clip2, double threshold,
                                           double parameter_s, double alpha_min, double alpha_max) {
    const double frac1 = (total != 0.0) ? (total - clip1) / total : 0.0;
    const double frac2 = (clip1 != 0.0) ? (clip1 - clip2) / clip1 : 0.0;
    const double energy = frac2 > frac1 ? frac2 : frac1;
    double a = 1.0 / (1.0 + exp(-parameter_s / threshold * (energy - threshold)));
    if (a < alpha_min) a = 0.0;
    if (a > 1.0 - alpha_min) a = 1.0;
    return a < alpha_max ? a : alpha_max;
}

}  // namespace trixi_hip
