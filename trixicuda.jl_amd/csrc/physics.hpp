// physics.hpp -- device-side equations and numerical fluxes for the gfx950
// DGSEM kernels.  Everything is templated on the arithmetic type R (double in
// the FP64 mode, float in the mixed mode); the logarithmic means are always
// evaluated in double.
//
// The flux algebra follows the Trixi.jl definitions that TrixiCUDA.jl's
// kernels inline (reference: fluxes passed in by the tests, e.g.
// test/tree_dgsem_3d/euler_ec.jl:13-17, test/tree_dgsem_3d/mhd_alfven_wave.jl:13-18;
// call sites dg_3d_kernel.jl:221-234, 1155-1203).  Orientation o is 0-based here.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace trixi_hip {

enum Equation : int { EQ_LINADV = 0, EQ_EULER = 1, EQ_GLM_MHD = 2 };
enum VolumeIntegral : int { VI_WEAK = 0, VI_FD = 1, VI_HG = 2 };
enum Flux : int {
    FLUX_CENTRAL = 0,
    FLUX_LLF = 1,         // FluxLaxFriedrichs(max_abs_speed)
    FLUX_LLF_NAIVE = 2,   // FluxLaxFriedrichs(max_abs_speed_naive)
    FLUX_RANOCHA = 3,
    FLUX_SHIMA_ETAL = 4,
    FLUX_HINDENLANG_GASSNER = 5
};

struct PhysParams {
    double gamma, inv_gm1, gm1, c_h;
    double a[3];
};

template <typename R>
struct PhysR {
    R gamma, inv_gm1, gm1, c_h, a[3];
    __host__ __device__ static PhysR from(const PhysParams& p) {
        PhysR r;
        r.gamma = (R)p.gamma;
        r.inv_gm1 = (R)p.inv_gm1;
        r.gm1 = (R)p.gm1;
        r.c_h = (R)p.c_h;
        for (int d = 0; d < 3; ++d) r.a[d] = (R)p.a[d];
        return r;
    }
};

__device__ __forceinline__ double dsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float dsqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double dabs(double x) { return fabs(x); }
__device__ __forceinline__ float dabs(float x) { return fabsf(x); }
template <typename R> __device__ __forceinline__ R dmax(R a, R b) { return b > a ? b : a; }

// ---------------------------------------------------------------------------
// Logarithmic mean from precomputed logarithms.  ln_mean(x, y) equals
// (y - x) / log(y / x) away from x = y and the 4-term series in
// f2 = ((x - y) / (x + y))^2 below f2 = 1e-4 (Trixi.jl ln_mean).  The branch
// test is done without a division (den > 0 for positive x, y).  In the
// closed branch |log(y/x)| > 0.02, so log(y) - log(x) has a relative error of
// a few 1e-14, far inside the FP64 parity bound.
// ---------------------------------------------------------------------------
struct LnMean {
    double mean;      // ln_mean(x, y)
    double inv_mean;  // inv_ln_mean(x, y)
};

__device__ __forceinline__ double ln_mean_from_logs(double x, double y, double lx, double ly) {
    const double num = x * (x - 2.0 * y) + y * y;
    const double den = x * (x + 2.0 * y) + y * y;
    if (num < 1.0e-4 * den) {
        const double f2 = num / den;
        return (x + y) / (2.0 + f2 * (2.0 / 3.0 + f2 * (2.0 / 5.0 + f2 * (2.0 / 7.0))));
    }
    return (y - x) / (ly - lx);
}

__device__ __forceinline__ double inv_ln_mean_from_logs(double x, double y, double lx, double ly) {
    const double num = x * (x - 2.0 * y) + y * y;
    const double den = x * (x + 2.0 * y) + y * y;
    if (num < 1.0e-4 * den) {
        const double f2 = num / den;
        return (2.0 + f2 * (2.0 / 3.0 + f2 * (2.0 / 5.0 + f2 * (2.0 / 7.0)))) / (x + y);
    }
    return (ly - lx) / (y - x);
}

// ===========================================================================
// Linear scalar advection in DIM dimensions: u_t + a . grad u = 0.
// ===========================================================================
template <int DIM>
struct LinAdv {
    static constexpr int NV = 1;
    static constexpr bool NONCONS = false;
    template <typename R>
    struct Node {
        R u[1];
    };
    template <typename R, bool LOGS = true>
    __device__ static Node<R> node(const R* u, const PhysR<R>&) {
        Node<R> n;
        n.u[0] = u[0];
        return n;
    }
    template <typename R>
    __device__ static void flux(const Node<R>& n, int o, const PhysR<R>& P, R* f) {
        f[0] = P.a[o] * n.u[0];
    }
    template <int KIND, typename R>
    __device__ static void two_point(const Node<R>& l, const Node<R>& r, int o, const PhysR<R>& P,
                                     R* f) {
        R fl = P.a[o] * l.u[0], fr = P.a[o] * r.u[0];
        if (KIND == FLUX_LLF || KIND == FLUX_LLF_NAIVE) {
            R lam = dabs(P.a[o]);
            f[0] = R(0.5) * (fl + fr) + (R(-0.5) * lam * (r.u[0] - l.u[0]));
        } else {
            f[0] = R(0.5) * (fl + fr);
        }
    }
    template <typename R>
    __device__ static void max_abs_speeds(const R* u, const PhysR<R>& P, R* lam) {
        for (int d = 0; d < DIM; ++d) lam[d] = dabs(P.a[d]);
    }
};

// ===========================================================================
// Compressible Euler in DIM dimensions: (rho, rho v_1..DIM, rho e).
// ===========================================================================
template <int DIM>
struct Euler {
    static constexpr int NV = DIM + 2;
    static constexpr bool NONCONS = false;
    template <typename R>
    struct Node {
        R u[NV];
        R rho, v[DIM], p;
        double lrho, lp;  // log(rho), log(p) for the entropy-conservative means
    };

    template <typename R>
    __device__ static void prim(const R* u, const PhysR<R>& P, R& rho, R* v, R& p) {
        rho = u[0];
        R ke = R(0);
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            v[d] = u[1 + d] / rho;
            ke = (d == 0) ? u[1 + d] * v[d] : ke + u[1 + d] * v[d];
        }
        p = P.gm1 * (u[DIM + 1] - R(0.5) * ke);
    }

    template <typename R, bool LOGS = true>
    __device__ static Node<R> node(const R* u, const PhysR<R>& P) {
        Node<R> n;
#pragma unroll
        for (int q = 0; q < NV; ++q) n.u[q] = u[q];
        prim(u, P, n.rho, n.v, n.p);
        if (LOGS) {
            n.lrho = log((double)n.rho);
            n.lp = log((double)n.p);
        } else {
            n.lrho = 0.0;
            n.lp = 0.0;
        }
        return n;
    }

    template <typename R>
    __device__ static void flux(const Node<R>& n, int o, const PhysR<R>& P, R* f) {
        const R rho_vo = n.u[1 + o];
        f[0] = rho_vo;
#pragma unroll
        for (int d = 0; d < DIM; ++d) f[1 + d] = (d == o) ? rho_vo * n.v[d] + n.p : rho_vo * n.v[d];
        f[DIM + 1] = (n.u[DIM + 1] + n.p) * n.v[o];
    }

    template <typename R>
    __device__ static R sound(const Node<R>& n, const PhysR<R>& P) {
        return dsqrt(P.gamma * n.p / n.rho);
    }

    template <int KIND, typename R>
    __device__ static void two_point(const Node<R>& l, const Node<R>& r, int o, const PhysR<R>& P,
                                     R* f) {
        if (KIND == FLUX_RANOCHA) {
            const double rho_mean_d = ln_mean_from_logs((double)l.rho, (double)r.rho, l.lrho, r.lrho);
            // inv_ln_mean(rho_ll p_rr, rho_rr p_ll) with logs log(rho)+log(p)
            const double x = (double)l.rho * (double)r.p, y = (double)r.rho * (double)l.p;
            const double lx = l.lrho + r.lp, ly = r.lrho + l.lp;
            const double inv_rho_p_mean_d = (double)l.p * (double)r.p * inv_ln_mean_from_logs(x, y, lx, ly);
            const R rho_mean = (R)rho_mean_d;
            const R inv_rho_p_mean = (R)inv_rho_p_mean_d;
            R vavg[DIM];
            R vsq = R(0);
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                vavg[d] = R(0.5) * (l.v[d] + r.v[d]);
                vsq = (d == 0) ? l.v[d] * r.v[d] : vsq + l.v[d] * r.v[d];
            }
            const R velocity_square_avg = R(0.5) * vsq;
            const R p_avg = R(0.5) * (l.p + r.p);
            const R f1 = rho_mean * vavg[o];
            f[0] = f1;
#pragma unroll
            for (int d = 0; d < DIM; ++d) f[1 + d] = (d == o) ? f1 * vavg[d] + p_avg : f1 * vavg[d];
            f[DIM + 1] = f1 * (velocity_square_avg + inv_rho_p_mean * P.inv_gm1) +
                         R(0.5) * (l.p * r.v[o] + r.p * l.v[o]);
        } else if (KIND == FLUX_SHIMA_ETAL) {
            const R rho_avg = R(0.5) * (l.rho + r.rho);
            R vavg[DIM];
            R ksum = R(0);
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                vavg[d] = R(0.5) * (l.v[d] + r.v[d]);
                ksum = (d == 0) ? l.v[d] * r.v[d] : ksum + l.v[d] * r.v[d];
            }
            const R kin_avg = R(0.5) * ksum;
            const R p_avg = R(0.5) * (l.p + r.p);
            const R pv_avg = R(0.5) * (l.p * r.v[o] + r.p * l.v[o]);
            const R f1 = rho_avg * vavg[o];
            f[0] = f1;
#pragma unroll
            for (int d = 0; d < DIM; ++d) f[1 + d] = (d == o) ? f1 * vavg[d] + p_avg : f1 * vavg[d];
            f[DIM + 1] = p_avg * vavg[o] * P.inv_gm1 + f1 * kin_avg + pv_avg;
        } else {
            R fl[NV], fr[NV];
            flux(l, o, P, fl);
            flux(r, o, P, fr);
            if (KIND == FLUX_LLF || KIND == FLUX_LLF_NAIVE) {
                const R cl = sound(l, P), cr = sound(r, P);
                R lam;
                if (KIND == FLUX_LLF_NAIVE)
                    lam = dmax(dabs(l.v[o]), dabs(r.v[o])) + dmax(cl, cr);
                else
                    lam = dmax(dabs(l.v[o]) + cl, dabs(r.v[o]) + cr);
#pragma unroll
                for (int q = 0; q < NV; ++q)
                    f[q] = R(0.5) * (fl[q] + fr[q]) + (R(-0.5) * lam * (r.u[q] - l.u[q]));
            } else {
#pragma unroll
                for (int q = 0; q < NV; ++q) f[q] = R(0.5) * (fl[q] + fr[q]);
            }
        }
    }

    template <typename R>
    __device__ static void max_abs_speeds(const R* u, const PhysR<R>& P, R* lam) {
        R rho, v[DIM], p;
        prim(u, P, rho, v, p);
        const R c = dsqrt(P.gamma * p / rho);
#pragma unroll
        for (int d = 0; d < DIM; ++d) lam[d] = dabs(v[d]) + c;
    }

    // density * pressure (Trixi density_pressure), the HG indicator variable
    template <typename R>
    __device__ static R density_pressure(const R* u, const PhysR<R>& P) {
        R ke = R(0);
#pragma unroll
        for (int d = 0; d < DIM; ++d) ke = (d == 0) ? u[1 + d] * u[1 + d] : ke + u[1 + d] * u[1 + d];
        return P.gm1 * (u[0] * u[DIM + 1] - R(0.5) * ke);
    }
};

// ===========================================================================
// Ideal GLM-MHD in 3D: (rho, rho v1..3, rho e, B1..3, psi), Powell and
// Galilean nonconservative terms.
// ===========================================================================
struct GlmMhd3D {
    static constexpr int NV = 9;
    static constexpr bool NONCONS = true;
    template <typename R>
    struct Node {
        R u[9];
        R rho, v[3], p;
        double lrho, lp;
    };
    template <typename R, bool LOGS = true>
    __device__ static Node<R> node(const R* u, const PhysR<R>& P) {
        Node<R> n;
#pragma unroll
        for (int q = 0; q < 9; ++q) n.u[q] = u[q];
        n.rho = u[0];
        n.v[0] = u[1] / n.rho;
        n.v[1] = u[2] / n.rho;
        n.v[2] = u[3] / n.rho;
        n.p = P.gm1 * (u[4] - R(0.5) * (u[1] * n.v[0] + u[2] * n.v[1] + u[3] * n.v[2]) -
                       R(0.5) * (u[5] * u[5] + u[6] * u[6] + u[7] * u[7]) - R(0.5) * u[8] * u[8]);
        if (LOGS) {
            n.lrho = log((double)n.rho);
            n.lp = log((double)n.p);
        } else {
            n.lrho = 0.0;
            n.lp = 0.0;
        }
        return n;
    }

    template <typename R>
    __device__ static void flux(const Node<R>& n, int o, const PhysR<R>& P, R* f) {
        const R* u = n.u;
        const R rho = u[0];
        const R v[3] = {n.v[0], n.v[1], n.v[2]};
        const R B[3] = {u[5], u[6], u[7]};
        const R psi = u[8];
        const R kin_en = R(0.5) * rho * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        const R mag_en = R(0.5) * (B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
        const R p_over_gm1 = u[4] - kin_en - mag_en - R(0.5) * psi * psi;
        const R p = P.gm1 * p_over_gm1;
        const R vB = v[0] * B[0] + v[1] * B[1] + v[2] * B[2];
        f[0] = u[1 + o];
#pragma unroll
        for (int d = 0; d < 3; ++d)
            f[1 + d] = (d == o) ? u[1 + o] * v[d] + p + mag_en - B[o] * B[o] : u[1 + o] * v[d] - B[o] * B[d];
        f[4] = (kin_en + P.gamma * p_over_gm1 + R(2) * mag_en) * v[o] - B[o] * vB + P.c_h * psi * B[o];
#pragma unroll
        for (int d = 0; d < 3; ++d) f[5 + d] = (d == o) ? P.c_h * psi : v[o] * B[d] - v[d] * B[o];
        f[8] = P.c_h * B[o];
    }

    template <typename R>
    __device__ static R fast_speed(const R* u, int o, const PhysR<R>& P) {
        const R rho = u[0];
        const R v1 = u[1] / rho, v2 = u[2] / rho, v3 = u[3] / rho;
        const R B1 = u[5], B2 = u[6], B3 = u[7], psi = u[8];
        const R kin_en = R(0.5) * rho * (v1 * v1 + v2 * v2 + v3 * v3);
        const R mag_en = R(0.5) * (B1 * B1 + B2 * B2 + B3 * B3);
        const R p = P.gm1 * (u[4] - kin_en - mag_en - R(0.5) * psi * psi);
        const R a_square = P.gamma * p / rho;
        const R sqrt_rho = dsqrt(rho);
        const R b1 = B1 / sqrt_rho, b2 = B2 / sqrt_rho, b3 = B3 / sqrt_rho;
        const R b_square = b1 * b1 + b2 * b2 + b3 * b3;
        const R bo = (o == 0) ? b1 : ((o == 1) ? b2 : b3);
        const R s = a_square + b_square;
        return dsqrt(R(0.5) * s + R(0.5) * dsqrt(s * s - R(4) * a_square * bo * bo));
    }

    template <int KIND, typename R>
    __device__ static void two_point(const Node<R>& l, const Node<R>& r, int o, const PhysR<R>& P,
                                     R* f) {
        if (KIND == FLUX_HINDENLANG_GASSNER) {
            const R Bl[3] = {l.u[5], l.u[6], l.u[7]}, Br[3] = {r.u[5], r.u[6], r.u[7]};
            const R psil = l.u[8], psir = r.u[8];
            const double rho_mean_d = ln_mean_from_logs((double)l.rho, (double)r.rho, l.lrho, r.lrho);
            const double x = (double)l.rho * (double)r.p, y = (double)r.rho * (double)l.p;
            const double inv_rho_p_mean_d =
                (double)l.p * (double)r.p * inv_ln_mean_from_logs(x, y, l.lrho + r.lp, r.lrho + l.lp);
            const R rho_mean = (R)rho_mean_d, inv_rho_p_mean = (R)inv_rho_p_mean_d;
            R vavg[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) vavg[d] = R(0.5) * (l.v[d] + r.v[d]);
            const R p_avg = R(0.5) * (l.p + r.p);
            const R psi_avg = R(0.5) * (psil + psir);
            const R velocity_square_avg = R(0.5) * (l.v[0] * r.v[0] + l.v[1] * r.v[1] + l.v[2] * r.v[2]);
            const R magnetic_square_avg = R(0.5) * (Bl[0] * Br[0] + Bl[1] * Br[1] + Bl[2] * Br[2]);
            const R f1 = rho_mean * vavg[o];
            f[0] = f1;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                R t = f1 * vavg[d];
                if (d == o) t = t + p_avg + magnetic_square_avg;
                f[1 + d] = t - R(0.5) * (Bl[o] * Br[d] + Br[o] * Bl[d]);
            }
#pragma unroll
            for (int d = 0; d < 3; ++d)
                f[5 + d] = (d == o) ? P.c_h * psi_avg
                                    : R(0.5) * (l.v[o] * Bl[d] - l.v[d] * Bl[o] + r.v[o] * Br[d] - r.v[d] * Br[o]);
            f[8] = P.c_h * R(0.5) * (Bl[o] + Br[o]);
            R s = l.p * r.v[o] + r.p * l.v[o];
#pragma unroll
            for (int d = 0; d < 3; ++d)
                if (d != o) s = s + (l.v[o] * Bl[d] * Br[d] + r.v[o] * Br[d] * Bl[d]);
#pragma unroll
            for (int d = 0; d < 3; ++d)
                if (d != o) s = s - (l.v[d] * Bl[o] * Br[d] + r.v[d] * Br[o] * Bl[d]);
            s = s + P.c_h * (Bl[o] * psir + Br[o] * psil);
            f[4] = f1 * (velocity_square_avg + inv_rho_p_mean * P.inv_gm1) + R(0.5) * s;
        } else {
            R fl[9], fr[9];
            flux(l, o, P, fl);
            flux(r, o, P, fr);
            if (KIND == FLUX_LLF || KIND == FLUX_LLF_NAIVE) {
                const R cl = fast_speed(l.u, o, P), cr = fast_speed(r.u, o, P);
                R lam;
                if (KIND == FLUX_LLF_NAIVE)
                    lam = dmax(dabs(l.v[o]), dabs(r.v[o])) + dmax(cl, cr);
                else
                    lam = dmax(dabs(l.v[o]) + cl, dabs(r.v[o]) + cr);
#pragma unroll
                for (int q = 0; q < 9; ++q)
                    f[q] = R(0.5) * (fl[q] + fr[q]) + (R(-0.5) * lam * (r.u[q] - l.u[q]));
            } else {
#pragma unroll
                for (int q = 0; q < 9; ++q) f[q] = R(0.5) * (fl[q] + fr[q]);
            }
        }
    }

    // flux_nonconservative_powell(u_ll, u_rr, o): Powell (0, B, v.B, v, 0) and
    // Galilean GLM terms of the left state times the right state's B_o / psi.
    template <typename R>
    __device__ static void noncons(const Node<R>& l, const Node<R>& r, int o, R* f) {
        const R B_ll[3] = {l.u[5], l.u[6], l.u[7]};
        const R psi_ll = l.u[8];
        const R v_dot_B_ll = l.v[0] * B_ll[0] + l.v[1] * B_ll[1] + l.v[2] * B_ll[2];
        const R Bo_rr = r.u[5 + o];
        const R psi_rr = r.u[8];
        const R psi_vo_ll = psi_ll * l.v[o];
        f[0] = R(0);
#pragma unroll
        for (int d = 0; d < 3; ++d) f[1 + d] = B_ll[d] * Bo_rr;
        f[4] = v_dot_B_ll * Bo_rr + psi_vo_ll * psi_rr;
#pragma unroll
        for (int d = 0; d < 3; ++d) f[5 + d] = l.v[d] * Bo_rr;
        f[8] = l.v[o] * psi_rr;
    }

    template <typename R>
    __device__ static void max_abs_speeds(const R* u, const PhysR<R>& P, R* lam) {
#pragma unroll
        for (int d = 0; d < 3; ++d) lam[d] = dabs(u[1 + d] / u[0]) + fast_speed(u, d, P);
    }
};

}  // namespace trixi_hip
