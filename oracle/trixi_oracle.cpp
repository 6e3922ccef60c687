// trixi_oracle.cpp -- CPU oracle (TEST INFRASTRUCTURE ONLY, never shipped).
//
// A plain FP64 restatement of Trixi.jl's CPU semidiscretization for the path
// TrixiCUDA.jl offloads to the GPU.  Every stage mirrors the reference stage
// order of rhs_gpu! (src/solvers/dg_1d.jl:523-548, dg_2d.jl:760-790,
// dg_3d.jl:895-925) and the index conventions of its kernels
// (src/solvers/dg_{1,2,3}d_kernel.jl).  The arithmetic of each stage restates
// the Trixi.jl CPU routine the reference tests compare against (Trixi.jl is an
// external dependency, compat 0.11-0.13.5 in the reference Project.toml:19,
// not present in this tree).  Functions are templated on the scalar type so
// the same code runs in double and in forward-mode dual numbers (used to
// check the adjoint by a dot-product test).
//
// Nothing in the product (the HIP library and its host package) includes,
// links or calls this file.

#include "trixi_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ----------------------------------------------------------------------------
// Scalar types: double, and a forward-mode dual number.
// ----------------------------------------------------------------------------
struct Dual {
    double v, d;
    Dual() : v(0), d(0) {}
    Dual(double x) : v(x), d(0) {}
    Dual(double x, double dx) : v(x), d(dx) {}
};
inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline Dual operator/(Dual a, Dual b) {
    return Dual(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v));
}
inline Dual& operator+=(Dual& a, Dual b) { a = a + b; return a; }
inline Dual& operator-=(Dual& a, Dual b) { a = a - b; return a; }
inline Dual& operator*=(Dual& a, Dual b) { a = a * b; return a; }
inline Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
inline Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
inline Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
inline Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
inline Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
inline Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
inline Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
inline Dual operator/(double a, Dual b) { return Dual(a / b.v, -a * b.d / (b.v * b.v)); }

inline double val(double x) { return x; }
inline double val(const Dual& x) { return x.v; }

inline double tlog(double x) { return log(x); }
inline Dual tlog(Dual x) { return Dual(log(x.v), x.d / x.v); }
inline double tsqrt(double x) { return sqrt(x); }
inline Dual tsqrt(Dual x) {
    double s = sqrt(x.v);
    return Dual(s, x.d / (2.0 * s));
}
inline double texp(double x) { return exp(x); }
inline Dual texp(Dual x) {
    double e = exp(x.v);
    return Dual(e, e * x.d);
}
inline double tabs(double x) { return fabs(x); }
inline Dual tabs(Dual x) { return x.v < 0 ? -x : x; }
template <class T> inline T tmax(T a, T b) { return val(b) > val(a) ? b : a; }
template <class T> inline T tmin(T a, T b) { return val(b) < val(a) ? b : a; }

constexpr int MAXV = 9;

struct Cfg {
    int nd, eq, N, nv, vi, vflux, sflux, fvflux, smooth;
    int nc;                     // multicomponent Euler: components
    double mgamma[4], mR[4], mcv[4];  // gamma_k, R_k, c_v,k = R_k / (gamma_k - 1)
    double grav;                      // shallow water
    double gamma, inv_gm1, c_h, a[3], amax, amin, ic_length, nu, inv_Tr;
    int bc, bc_ic, src, ivar, outflow;
    uint64_t bc_seed;
    const double* xnodes;
    double t;
    int64_t nmortars;
    const int32_t *mids, *mori, *mls;
};

Cfg make_cfg(const oracle_cfg* c) {
    Cfg k;
    k.nd = c->ndims;
    k.eq = c->equation;
    k.N = c->polydeg + 1;
    k.nv = oracle_nvars(c);
    k.vi = c->volume_integral;
    k.vflux = c->volume_flux;
    k.sflux = c->surface_flux;
    k.fvflux = c->fv_flux;
    k.smooth = c->alpha_smooth;
    k.gamma = c->gamma;
    k.inv_gm1 = 1.0 / (c->gamma - 1.0);
    k.c_h = (c->equation == OR_EQ_GLM_MHD && c->ndims == 1) ? 0.0 : c->c_h;  // no cleaning field in 1D
    for (int d = 0; d < 3; ++d) k.a[d] = c->advection_velocity[d];
    k.amax = c->alpha_max;
    k.amin = c->alpha_min;
    k.ic_length = c->ic_length;
    k.bc = c->boundary_condition;
    k.bc_ic = c->boundary_ic;
    k.bc_seed = c->boundary_seed;
    k.src = c->source_terms;
    k.xnodes = c->node_coordinates;
    k.t = c->time;
    k.ivar = c->indicator_variable;
    k.nu = c->nu;
    k.inv_Tr = (c->Lr > 0.0) ? c->nu / (c->Lr * c->Lr) : 0.0;
    k.outflow = c->boundary_outflow_faces;
    k.nmortars = c->nmortars;
    k.mids = c->mortar_neighbor_ids;
    k.mori = c->mortar_orientations;
    k.mls = c->mortar_large_sides;
    k.nc = (c->equation == OR_EQ_EULER_MULTI) ? c->ncomponents : 0;
    k.grav = c->gravity;
    for (int i = 0; i < 4; ++i) {
        k.mgamma[i] = (i < k.nc) ? c->gammas[i] : 0.0;
        k.mR[i] = (i < k.nc) ? c->gas_constants[i] : 0.0;
        k.mcv[i] = (i < k.nc) ? c->gas_constants[i] / (c->gammas[i] - 1.0) : 0.0;
    }
    return k;
}

// ----------------------------------------------------------------------------
// ln_mean / inv_ln_mean (Trixi.jl src/auxiliary/math.jl): series in
// f2 = ((x-y)/(x+y))^2 below 1e-4, else the closed form.
// ----------------------------------------------------------------------------
template <class T> T ln_mean(T x, T y) {
    const double epsilon_f2 = 1.0e-4;
    T f2 = (x * (x - 2.0 * y) + y * y) / (x * (x + 2.0 * y) + y * y);
    if (val(f2) < epsilon_f2) {
        return (x + y) / (2.0 + f2 * (2.0 / 3.0 + f2 * (2.0 / 5.0 + f2 * (2.0 / 7.0))));
    }
    return (y - x) / tlog(y / x);
}
template <class T> T inv_ln_mean(T x, T y) {
    const double epsilon_f2 = 1.0e-4;
    T f2 = (x * (x - 2.0 * y) + y * y) / (x * (x + 2.0 * y) + y * y);
    if (val(f2) < epsilon_f2) {
        return (2.0 + f2 * (2.0 / 3.0 + f2 * (2.0 / 5.0 + f2 * (2.0 / 7.0)))) / (x + y);
    }
    return tlog(y / x) / (y - x);
}

// ----------------------------------------------------------------------------
// Equations.  State layouts (Trixi.jl conventions):
//   LinAdv:  (u)
//   Euler:   (rho, rho_v[1..nd], rho_e)
//   GLM-MHD: (rho, rho_v1..3, rho_e, B1..3, psi)   (2D: orientations 1, 2)
// Orientation o is 1-based as in Trixi.
// ----------------------------------------------------------------------------
template <class T> void euler_prim(const Cfg& c, const T* u, T& rho, T* v, T& p) {
    rho = u[0];
    T ke = 0.0;
    for (int d = 0; d < c.nd; ++d) {
        v[d] = u[1 + d] / rho;
        ke = (d == 0) ? u[1 + d] * v[d] : ke + u[1 + d] * v[d];
    }
    p = (c.gamma - 1.0) * (u[c.nd + 1] - 0.5 * ke);
}

template <class T>
void mhd_prim(const Cfg& c, const T* u, T& rho, T* v, T& p, T* B, T& psi) {
    rho = u[0];
    for (int d = 0; d < 3; ++d) v[d] = u[1 + d] / rho;
    for (int d = 0; d < 3; ++d) B[d] = u[5 + d];
    psi = (c.nv > 8) ? u[8] : T(0.0);  // 1D ideal MHD: no cleaning field
    p = (c.gamma - 1.0) *
        (u[4] - 0.5 * (u[1] * v[0] + u[2] * v[1] + u[3] * v[2]) -
         0.5 * (B[0] * B[0] + B[1] * B[1] + B[2] * B[2]) - 0.5 * psi * psi);
}

// Multicomponent Euler (Trixi CompressibleEulerMulticomponentEquations{1,2}D):
// u = (rho v_1..nd, rho e, rho_1..rho_nc); density = sum of the partial
// densities, totalgamma = sum rho_k c_v,k gamma_k / sum rho_k c_v,k,
// p = (gamma - 1)(rho e - rho |v|^2 / 2).
template <class T> T multi_density(const Cfg& c, const T* u) {
    T rho = u[c.nd + 1];
    for (int k = 1; k < c.nc; ++k) rho = rho + u[c.nd + 1 + k];
    return rho;
}
template <class T> T multi_totalgamma(const Cfg& c, const T* u) {
    T h1 = 0.0, h2 = 0.0;
    for (int k = 0; k < c.nc; ++k) {
        h1 += u[c.nd + 1 + k] * c.mcv[k] * c.mgamma[k];
        h2 += u[c.nd + 1 + k] * c.mcv[k];
    }
    return h1 / h2;
}
template <class T> void multi_prim(const Cfg& c, const T* u, T& rho, T* v, T& p, T& gamma) {
    rho = multi_density(c, u);
    T vsq = 0.0;
    for (int d = 0; d < c.nd; ++d) {
        v[d] = u[d] / rho;
        vsq += v[d] * v[d];
    }
    gamma = multi_totalgamma(c, u);
    p = (gamma - 1.0) * (u[c.nd] - 0.5 * rho * vsq);
}

// Quasi-1D Euler (Trixi CompressibleEulerEquationsQuasi1D): u = (a rho,
// a rho v, a e, a); cons2prim divides by a and applies the Euler cons2prim.
template <class T> void quasi_prim(const Cfg& c, const T* u, T& rho, T& v, T& p, T& a) {
    a = u[3];
    rho = u[0] / a;
    const T rho_v = u[1] / a, rho_e = u[2] / a;
    v = rho_v / rho;
    p = (c.gamma - 1.0) * (rho_e - 0.5 * rho_v * v);
}

template <class T> void phys_flux(const Cfg& c, const T* u, int o, T* f) {
    const int oi = o - 1;
    if (c.eq == OR_EQ_SHALLOW_WATER) {
        // Trixi ShallowWaterEquations{1,2}D flux: (h v_o, h v_d v_o + g h^2/2 [d = o], 0)
        const T h = u[0], p = 0.5 * c.grav * h * h;
        const T vo = u[o] / h;
        f[0] = u[o];
        for (int d = 0; d < c.nd; ++d) f[1 + d] = (d == oi) ? u[1 + d] * vo + p : u[1 + d] * vo;
        f[c.nd + 1] = 0.0;
        return;
    }
    if (c.eq == OR_EQ_EULER_QUASI1D) {
        // Trixi flux: (a rho v, a rho v^2, a v (e + p), 0), e = (a e) / a; the
        // pressure gradient is the nonconservative term a p_x
        T rho, v, p, a;
        quasi_prim(c, u, rho, v, p, a);
        const T e = u[2] / a;
        f[0] = u[1];
        f[1] = u[1] * v;
        f[2] = a * v * (e + p);
        f[3] = 0.0;
        return;
    }
    if (c.eq == OR_EQ_EULER_MULTI) {
        T rho, v[3], p, gamma;
        multi_prim(c, u, rho, v, p, gamma);
        for (int d = 0; d < c.nd; ++d) f[d] = (d == oi) ? u[d] * v[oi] + p : u[d] * v[oi];
        f[c.nd] = (u[c.nd] + p) * v[oi];
        for (int k = 0; k < c.nc; ++k) f[c.nd + 1 + k] = u[c.nd + 1 + k] * v[oi];
        return;
    }
    if (c.eq == OR_EQ_HYPDIFF) {
        // Trixi HyperbolicDiffusionEquations flux: (-nu q_o, -phi / T_r e_o)
        f[0] = -c.nu * u[o];
        for (int d = 0; d < c.nd; ++d) f[1 + d] = (d == oi) ? -u[0] * c.inv_Tr : T(0.0);
    } else if (c.eq == OR_EQ_BURGERS) {
        f[0] = 0.5 * u[0] * u[0];  // Trixi InviscidBurgersEquation1D flux
    } else if (c.eq == OR_EQ_LINADV) {
        f[0] = c.a[oi] * u[0];
    } else if (c.eq == OR_EQ_EULER) {
        T rho, v[3], p;
        euler_prim(c, u, rho, v, p);
        const T rho_vo = u[1 + oi];
        f[0] = rho_vo;
        for (int d = 0; d < c.nd; ++d) f[1 + d] = (d == oi) ? rho_vo * v[d] + p : rho_vo * v[d];
        f[c.nd + 1] = (u[c.nd + 1] + p) * v[oi];
    } else {
        T rho = u[0];
        T v[3] = {u[1] / rho, u[2] / rho, u[3] / rho};
        T B[3] = {u[5], u[6], u[7]};
        T psi = (c.nv > 8) ? u[8] : T(0.0);
        T kin_en = 0.5 * rho * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        T mag_en = 0.5 * (B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
        T p_over_gm1 = u[4] - kin_en - mag_en - 0.5 * psi * psi;
        T p = (c.gamma - 1.0) * p_over_gm1;
        T vB = v[0] * B[0] + v[1] * B[1] + v[2] * B[2];
        f[0] = u[1 + oi];
        for (int d = 0; d < 3; ++d)
            f[1 + d] = (d == oi) ? u[1 + oi] * v[d] + p + mag_en - B[oi] * B[oi]
                                 : u[1 + oi] * v[d] - B[oi] * B[d];
        f[4] = (kin_en + c.gamma * p_over_gm1 + 2.0 * mag_en) * v[oi] - B[oi] * vB +
               c.c_h * psi * B[oi];
        for (int d = 0; d < 3; ++d)
            f[5 + d] = (d == oi) ? c.c_h * psi : v[oi] * B[d] - v[d] * B[oi];
        f[8] = c.c_h * B[oi];
    }
}

template <class T> T mhd_fast_wavespeed(const Cfg& c, const T* u, int o) {
    T rho = u[0];
    T v1 = u[1] / rho, v2 = u[2] / rho, v3 = u[3] / rho;
    T B1 = u[5], B2 = u[6], B3 = u[7], psi = (c.nv > 8) ? u[8] : T(0.0);
    T kin_en = 0.5 * rho * (v1 * v1 + v2 * v2 + v3 * v3);
    T mag_en = 0.5 * (B1 * B1 + B2 * B2 + B3 * B3);
    T p = (c.gamma - 1.0) * (u[4] - kin_en - mag_en - 0.5 * psi * psi);
    T a_square = c.gamma * p / rho;
    T sqrt_rho = tsqrt(rho);
    T b1 = B1 / sqrt_rho, b2 = B2 / sqrt_rho, b3 = B3 / sqrt_rho;
    T b_square = b1 * b1 + b2 * b2 + b3 * b3;
    T bo = (o == 1) ? b1 : (o == 2) ? b2 : b3;
    T s = a_square + b_square;
    return tsqrt(0.5 * s + 0.5 * tsqrt(s * s - 4.0 * a_square * bo * bo));
}

// Normal speed and characteristic speed at one state (for LLF estimates).
template <class T> void normal_speed(const Cfg& c, const T* u, int o, T& vn, T& cs) {
    const int oi = o - 1;
    if (c.eq == OR_EQ_LINADV) {
        vn = c.a[oi];
        cs = 0.0;
    } else if (c.eq == OR_EQ_BURGERS) {
        vn = u[0];  // max_abs_speed(_naive) of Burgers: max(|u_ll|, |u_rr|)
        cs = 0.0;
    } else if (c.eq == OR_EQ_HYPDIFF) {
        vn = 0.0;  // constant speed sqrt(nu / T_r)
        cs = sqrt(c.nu * c.inv_Tr);
    } else if (c.eq == OR_EQ_EULER) {
        T rho, v[3], p;
        euler_prim(c, u, rho, v, p);
        vn = v[oi];
        cs = tsqrt(c.gamma * p / rho);
    } else if (c.eq == OR_EQ_EULER_MULTI) {
        T rho, v[3], p, gamma;
        multi_prim(c, u, rho, v, p, gamma);
        vn = v[oi];
        cs = tsqrt(gamma * p / rho);
    } else if (c.eq == OR_EQ_EULER_QUASI1D) {
        T rho, p, a;
        quasi_prim(c, u, rho, vn, p, a);
        cs = tsqrt(c.gamma * p / rho);
    } else if (c.eq == OR_EQ_SHALLOW_WATER) {
        vn = u[o] / u[0];
        cs = tsqrt(c.grav * u[0]);
    } else {
        vn = u[1 + oi] / u[0];
        cs = mhd_fast_wavespeed(c, u, o);
    }
}

// Trixi calc_fast_wavespeed_roe (ideal GLM-MHD): Roe-averaged normal
// velocity and fast magnetoacoustic speed (Cargo & Gallice 1997), with
// total pressure p + |B|^2/2, enthalpy H = (rho e + p_total) / rho,
// X = |[[B]]|^2 / (2 (sqrt(rho_l) + sqrt(rho_r))^2), b^2 = |B_roe|^2 / sqrt(rho_l rho_r),
// a^2 = (2 - gamma) X + (gamma - 1)(H_roe - |v_roe|^2 / 2 - b^2).
template <class T> void mhd_fast_wavespeed_roe(const Cfg& c, const T* ull, const T* urr, int o, T& vel, T& cf) {
    T rl = ull[0], rr = urr[0];
    T vl[3] = {ull[1] / rl, ull[2] / rl, ull[3] / rl}, vr[3] = {urr[1] / rr, urr[2] / rr, urr[3] / rr};
    T Bl[3] = {ull[5], ull[6], ull[7]}, Br[3] = {urr[5], urr[6], urr[7]};
    T psil = (c.nv > 8) ? ull[8] : T(0.0), psir = (c.nv > 8) ? urr[8] : T(0.0);
    T kel = 0.5 * (ull[1] * vl[0] + ull[2] * vl[1] + ull[3] * vl[2]);
    T ker = 0.5 * (urr[1] * vr[0] + urr[2] * vr[1] + urr[3] * vr[2]);
    T mbl = Bl[0] * Bl[0] + Bl[1] * Bl[1] + Bl[2] * Bl[2], mbr = Br[0] * Br[0] + Br[1] * Br[1] + Br[2] * Br[2];
    T pl = (c.gamma - 1.0) * (ull[4] - kel - 0.5 * mbl - 0.5 * psil * psil);
    T pr = (c.gamma - 1.0) * (urr[4] - ker - 0.5 * mbr - 0.5 * psir * psir);
    T ptl = pl + 0.5 * mbl, ptr = pr + 0.5 * mbr;
    T sl = tsqrt(rl), sr = tsqrt(rr);
    T inv_add = 1.0 / (sl + sr), inv_prod = 1.0 / (sl * sr);
    T wl = sl * inv_add, wr = sr * inv_add;
    T v_roe[3], B_roe[3];
    for (int d = 0; d < 3; ++d) {
        v_roe[d] = vl[d] * wl + vr[d] * wr;
        B_roe[d] = Bl[d] * wl + Br[d] * wr;
    }
    T H_roe = (ull[4] + ptl) / rl * wl + (urr[4] + ptr) / rr * wr;
    T dB0 = Bl[0] - Br[0], dB1 = Bl[1] - Br[1], dB2 = Bl[2] - Br[2];
    T X = 0.5 * (dB0 * dB0 + dB1 * dB1 + dB2 * dB2) * inv_add * inv_add;
    T b2 = (B_roe[0] * B_roe[0] + B_roe[1] * B_roe[1] + B_roe[2] * B_roe[2]) * inv_prod;
    T a2 = (2.0 - c.gamma) * X +
           (c.gamma - 1.0) * (H_roe - 0.5 * (v_roe[0] * v_roe[0] + v_roe[1] * v_roe[1] + v_roe[2] * v_roe[2]) - b2);
    T ca = B_roe[o - 1] * B_roe[o - 1] * inv_prod;
    T astar = tsqrt((a2 + b2) * (a2 + b2) - 4.0 * a2 * ca);
    cf = tsqrt(0.5 * (a2 + b2 + astar));
    vel = v_roe[o - 1];
}

// Trixi max_abs_speed (Davis-type) and max_abs_speed_naive.
template <class T> T lambda_max(const Cfg& c, const T* ull, const T* urr, int o, bool naive) {
    if (c.eq == OR_EQ_LINADV) return T(fabs(c.a[o - 1]));
    T vll, cll, vrr, crr;
    normal_speed(c, ull, o, vll, cll);
    normal_speed(c, urr, o, vrr, crr);
    if (naive) return tmax(tabs(vll), tabs(vrr)) + tmax(cll, crr);
    return tmax(tabs(vll) + cll, tabs(vrr) + crr);
}

template <class T> void flux_ranocha_euler(const Cfg& c, const T* ull, const T* urr, int o, T* f) {
    const int oi = o - 1;
    T rl, vl[3], pl, rr, vr[3], pr;
    euler_prim(c, ull, rl, vl, pl);
    euler_prim(c, urr, rr, vr, pr);
    T rho_mean = ln_mean(rl, rr);
    T inv_rho_p_mean = pl * pr * inv_ln_mean(rl * pr, rr * pl);
    T vavg[3] = {0.0, 0.0, 0.0}, vsq = 0.0;
    for (int d = 0; d < c.nd; ++d) {
        vavg[d] = 0.5 * (vl[d] + vr[d]);
        vsq = (d == 0) ? vl[d] * vr[d] : vsq + vl[d] * vr[d];
    }
    T velocity_square_avg = 0.5 * vsq;
    T p_avg = 0.5 * (pl + pr);
    T f1 = rho_mean * vavg[oi];
    f[0] = f1;
    for (int d = 0; d < c.nd; ++d) f[1 + d] = (d == oi) ? f1 * vavg[d] + p_avg : f1 * vavg[d];
    f[c.nd + 1] = f1 * (velocity_square_avg + inv_rho_p_mean * c.inv_gm1) +
                  0.5 * (pl * vr[oi] + pr * vl[oi]);
}

// Trixi flux_ranocha(u_ll, u_rr, orientation, ::CompressibleEulerMulticomponentEquations{1,2}D):
// ln-means of the partial densities times the mean normal velocity; the
// momentum and energy fluxes of Ranocha's flux with the temperature-based
// pressures p = T sum rho_k R_k, T = (rho e - rho |v|^2 / 2) / sum rho_k c_v,k,
// and the gamma of the averaged state.
template <class T> void flux_ranocha_multi(const Cfg& c, const T* ull, const T* urr, int o, T* f) {
    const int oi = o - 1, nd = c.nd;
    T rl = multi_density(c, ull), rr = multi_density(c, urr);
    T uavg[MAXV];
    for (int v = 0; v < c.nv; ++v) uavg[v] = 0.5 * (ull[v] + urr[v]);
    T gamma = multi_totalgamma(c, uavg);
    T inv_gm1 = 1.0 / (gamma - 1.0);
    T vl[3], vr[3], vavg[3], vsq = 0.0, kel = 0.0, ker = 0.0;
    for (int d = 0; d < nd; ++d) {
        vl[d] = ull[d] / rl;
        vr[d] = urr[d] / rr;
        vavg[d] = 0.5 * (vl[d] + vr[d]);
        vsq += vl[d] * vr[d];
        kel += vl[d] * vl[d];
        ker += vr[d] * vr[d];
    }
    T velocity_square_avg = 0.5 * vsq;
    T enth_l = 0.0, enth_r = 0.0, h_l = 0.0, h_r = 0.0;
    for (int k = 0; k < c.nc; ++k) {
        enth_l += ull[nd + 1 + k] * c.mR[k];
        enth_r += urr[nd + 1 + k] * c.mR[k];
        h_l += ull[nd + 1 + k] * c.mcv[k];
        h_r += urr[nd + 1 + k] * c.mcv[k];
    }
    T Tl = (ull[nd] - 0.5 * rl * kel) / h_l, Tr = (urr[nd] - 0.5 * rr * ker) / h_r;
    T pl = Tl * enth_l, pr = Tr * enth_r;
    T p_avg = 0.5 * (pl + pr);
    T inv_rho_p_mean = pl * pr * inv_ln_mean(rl * pr, rr * pl);
    T f_rho_sum = 0.0;
    for (int k = 0; k < c.nc; ++k) {
        f[nd + 1 + k] = ln_mean(ull[nd + 1 + k], urr[nd + 1 + k]) * vavg[oi];
        f_rho_sum += f[nd + 1 + k];
    }
    for (int d = 0; d < nd; ++d) f[d] = (d == oi) ? f_rho_sum * vavg[d] + p_avg : f_rho_sum * vavg[d];
    f[nd] = f_rho_sum * (velocity_square_avg + inv_rho_p_mean * inv_gm1) + 0.5 * (pl * vr[oi] + pr * vl[oi]);
}

// Trixi flux_chan_etal (CompressibleEulerEquationsQuasi1D; Chan, Shukla, Wu,
// Liu, Nalluri 2023): ln-mean density times the mean of a v, Ranocha's
// energy flux with a inside the pressure work term; the a component is 0.
template <class T> void flux_chan_etal(const Cfg& c, const T* ull, const T* urr, T* f) {
    T rl, vl, pl, al, rr, vr, pr, ar;
    quasi_prim(c, ull, rl, vl, pl, al);
    quasi_prim(c, urr, rr, vr, pr, ar);
    const T rho_mean = ln_mean(rl, rr);
    const T inv_rho_p_mean = pl * pr * inv_ln_mean(rl * pr, rr * pl);
    const T v1_avg = 0.5 * (vl + vr);
    const T a_v1_avg = 0.5 * (al * vl + ar * vr);
    const T velocity_square_avg = 0.5 * (vl * vr);
    f[0] = rho_mean * a_v1_avg;
    f[1] = rho_mean * a_v1_avg * v1_avg;
    f[2] = f[0] * (velocity_square_avg + inv_rho_p_mean * c.inv_gm1) + 0.5 * (pl * ar * vr + pr * al * vl);
    f[3] = 0.0;
}

template <class T> void flux_shima_euler(const Cfg& c, const T* ull, const T* urr, int o, T* f) {
    const int oi = o - 1;
    T rl, vl[3], pl, rr, vr[3], pr;
    euler_prim(c, ull, rl, vl, pl);
    euler_prim(c, urr, rr, vr, pr);
    T rho_avg = 0.5 * (rl + rr);
    T vavg[3] = {0.0, 0.0, 0.0}, ksum = 0.0;
    for (int d = 0; d < c.nd; ++d) {
        vavg[d] = 0.5 * (vl[d] + vr[d]);
        ksum = (d == 0) ? vl[d] * vr[d] : ksum + vl[d] * vr[d];
    }
    T kin_avg = 0.5 * ksum;
    T p_avg = 0.5 * (pl + pr);
    T pv_avg = 0.5 * (pl * vr[oi] + pr * vl[oi]);
    T f1 = rho_avg * vavg[oi];
    f[0] = f1;
    for (int d = 0; d < c.nd; ++d) f[1 + d] = (d == oi) ? f1 * vavg[d] + p_avg : f1 * vavg[d];
    f[c.nd + 1] = p_avg * vavg[oi] * c.inv_gm1 + f1 * kin_avg + pv_avg;
}

template <class T>
void flux_hindenlang_gassner_mhd(const Cfg& c, const T* ull, const T* urr, int o, T* f) {
    const int oi = o - 1;
    T rl, vl[3], pl, Bl[3], psil, rr, vr[3], pr, Br[3], psir;
    mhd_prim(c, ull, rl, vl, pl, Bl, psil);
    mhd_prim(c, urr, rr, vr, pr, Br, psir);
    T rho_mean = ln_mean(rl, rr);
    T inv_rho_p_mean = pl * pr * inv_ln_mean(rl * pr, rr * pl);
    T vavg[3];
    for (int d = 0; d < 3; ++d) vavg[d] = 0.5 * (vl[d] + vr[d]);
    T p_avg = 0.5 * (pl + pr);
    T psi_avg = 0.5 * (psil + psir);
    T velocity_square_avg = 0.5 * (vl[0] * vr[0] + vl[1] * vr[1] + vl[2] * vr[2]);
    T magnetic_square_avg = 0.5 * (Bl[0] * Br[0] + Bl[1] * Br[1] + Bl[2] * Br[2]);
    T f1 = rho_mean * vavg[oi];
    f[0] = f1;
    for (int d = 0; d < 3; ++d) {
        T t = f1 * vavg[d];
        if (d == oi) t = t + p_avg + magnetic_square_avg;
        f[1 + d] = t - 0.5 * (Bl[oi] * Br[d] + Br[oi] * Bl[d]);
    }
    for (int d = 0; d < 3; ++d)
        f[5 + d] = (d == oi) ? c.c_h * psi_avg
                             : 0.5 * (vl[oi] * Bl[d] - vl[d] * Bl[oi] + vr[oi] * Br[d] -
                                      vr[d] * Br[oi]);
    f[8] = c.c_h * 0.5 * (Bl[oi] + Br[oi]);
    T s = pl * vr[oi] + pr * vl[oi];
    for (int d = 0; d < 3; ++d)
        if (d != oi) s = s + (vl[oi] * Bl[d] * Br[d] + vr[oi] * Br[d] * Bl[d]);
    for (int d = 0; d < 3; ++d)
        if (d != oi) s = s - (vl[d] * Bl[oi] * Br[d] + vr[d] * Br[oi] * Bl[d]);
    s = s + c.c_h * (Bl[oi] * psir + Br[oi] * psil);
    f[4] = f1 * (velocity_square_avg + inv_rho_p_mean * c.inv_gm1) + 0.5 * s;
}

// flux_nonconservative_powell (Trixi ideal_glm_mhd_3d.jl, orientation form):
// Powell term (0, B, v.B, v, 0) and Galilean GLM term, multiplied by the
// right state's normal B (resp. psi); callers apply the factor 0.5.
// Nonconservative two-point terms.  `pos` is -1 for the volume term, else the
// surface (or shock-capturing subcell) flux the term is paired with; the
// terms differ by position only for shallow water.
template <class T> void noncons_flux(const Cfg& c, const T* ull, const T* urr, int o, T* f, int pos) {
    if (c.eq == OR_EQ_SHALLOW_WATER) {
        // volume: flux_nonconservative_wintermeyer_etal g h_ll b_rr; with the
        // hydrostatic reconstruction: flux_nonconservative_audusse_etal
        // g h_ll b_ll - g (h_ll* + h_ll)(h_ll* - h_ll), h* = max(0, h + b - max(b_ll, b_rr))
        // (half of it plus the reconstructed flux g h*^2 / 2 gives the left
        // cell g h_ll^2 / 2 + g h_ll b_ll / 2 at a lake at rest);
        // otherwise flux_nonconservative_fjordholm_etal g h_ll b_ll + g h_avg (b_rr - b_ll)
        const int nd = c.nd;
        const T hl = ull[0], hr = urr[0], bl = ull[nd + 1], br = urr[nd + 1];
        T t;
        if (pos < 0) {
            t = c.grav * hl * br;
        } else if (pos == OR_FLUX_HYDROSTATIC_LLF) {
            const T hs = tmax(T(0.0), hl + bl - tmax(bl, br));
            t = c.grav * hl * bl - c.grav * (hs + hl) * (hs - hl);
        } else {
            t = c.grav * hl * bl + c.grav * (0.5 * (hl + hr)) * (br - bl);
        }
        for (int v = 0; v < c.nv; ++v) f[v] = 0.0;
        f[o] = t;
        return;
    }
    if (c.eq == OR_EQ_EULER_QUASI1D) {
        // flux_nonconservative_chan_etal: (0, a_ll p_rr, 0, 0), the pressure
        // gradient a p_x in the same left-state-times-right-state form
        T rl, vl, pl, al, rr, vr, pr, ar;
        quasi_prim(c, ull, rl, vl, pl, al);
        quasi_prim(c, urr, rr, vr, pr, ar);
        f[0] = 0.0;
        f[1] = al * pr;
        f[2] = 0.0;
        f[3] = 0.0;
        return;
    }
    const int oi = o - 1;
    T rho_ll = ull[0];
    T v_ll[3] = {ull[1] / rho_ll, ull[2] / rho_ll, ull[3] / rho_ll};
    T B_ll[3] = {ull[5], ull[6], ull[7]};
    T psi_ll = ull[8];
    T v_dot_B_ll = v_ll[0] * B_ll[0] + v_ll[1] * B_ll[1] + v_ll[2] * B_ll[2];
    T Bo_rr = urr[5 + oi];
    T psi_rr = urr[8];
    T psi_vo_ll = psi_ll * v_ll[oi];
    f[0] = 0.0;
    for (int d = 0; d < 3; ++d) f[1 + d] = B_ll[d] * Bo_rr;
    f[4] = v_dot_B_ll * Bo_rr + psi_vo_ll * psi_rr;
    for (int d = 0; d < 3; ++d) f[5 + d] = v_ll[d] * Bo_rr;
    f[8] = v_ll[oi] * psi_rr;
}

template <class T> void two_point(const Cfg& c, int kind, const T* ull, const T* urr, int o, T* f) {
    switch (kind) {
        case OR_FLUX_RANOCHA:
            if (c.eq == OR_EQ_EULER_MULTI)
                flux_ranocha_multi(c, ull, urr, o, f);
            else
                flux_ranocha_euler(c, ull, urr, o, f);
            return;
        case OR_FLUX_SHIMA_ETAL:
            flux_shima_euler(c, ull, urr, o, f);
            return;
        case OR_FLUX_CHAN_ETAL:
            flux_chan_etal(c, ull, urr, f);
            return;
        case OR_FLUX_WINTERMEYER_ETAL:
        case OR_FLUX_FJORDHOLM_ETAL: {
            // Trixi flux_wintermeyer_etal: mean h v_o, p = g h_ll h_rr / 2;
            // flux_fjordholm_etal: h_avg v_o,avg, p = g (h_ll^2 + h_rr^2) / 4
            const int nd = c.nd;
            const T hl = ull[0], hr = urr[0];
            T vavg[3];
            for (int d = 0; d < nd; ++d) vavg[d] = 0.5 * (ull[1 + d] / hl + urr[1 + d] / hr);
            T f1, p_avg;
            if (kind == OR_FLUX_WINTERMEYER_ETAL) {
                f1 = 0.5 * (ull[o] + urr[o]);
                p_avg = 0.5 * c.grav * hl * hr;
            } else {
                f1 = 0.5 * (hl + hr) * vavg[o - 1];
                p_avg = 0.25 * c.grav * (hl * hl + hr * hr);
            }
            f[0] = f1;
            for (int d = 0; d < nd; ++d) f[1 + d] = f1 * vavg[d] + ((d == o - 1) ? p_avg : T(0.0));
            f[nd + 1] = 0.0;
            return;
        }
        case OR_FLUX_HYDROSTATIC_LLF: {
            // Trixi FluxHydrostaticReconstruction(flux_lax_friedrichs,
            // hydrostatic_reconstruction_audusse_etal): LLF of the states with
            // h* = max(0, h + b - max(b_ll, b_rr)) and h* v (the velocities kept)
            const int nd = c.nd;
            const T bmax = tmax(ull[nd + 1], urr[nd + 1]);
            T ls[MAXV], rs[MAXV];
            const T hls = tmax(T(0.0), ull[0] + ull[nd + 1] - bmax), hrs = tmax(T(0.0), urr[0] + urr[nd + 1] - bmax);
            ls[0] = hls;
            rs[0] = hrs;
            for (int d = 0; d < nd; ++d) {
                ls[1 + d] = hls * (ull[1 + d] / ull[0]);
                rs[1 + d] = hrs * (urr[1 + d] / urr[0]);
            }
            ls[nd + 1] = ull[nd + 1];
            rs[nd + 1] = urr[nd + 1];
            two_point(c, OR_FLUX_LLF, ls, rs, o, f);
            return;
        }
        case OR_FLUX_HINDENLANG_GASSNER:
            flux_hindenlang_gassner_mhd(c, ull, urr, o, f);
            return;
        case OR_FLUX_HLL: {
            // Trixi FluxHLL with min_max_speed_davis (the flux_hll default since
            // Trixi 0.7): lambda_min = min(v_ll - c_ll, v_rr - c_rr),
            // lambda_max = max(v_ll + c_ll, v_rr + c_rr)
            T fl[MAXV], fr[MAXV], vl, cl, vr, cr;
            phys_flux(c, ull, o, fl);
            phys_flux(c, urr, o, fr);
            normal_speed(c, ull, o, vl, cl);
            normal_speed(c, urr, o, vr, cr);
            const T lmin = tmin(vl - cl, vr - cr), lmax = tmax(vl + cl, vr + cr);
            if (val(lmin) >= 0.0 && val(lmax) >= 0.0) {
                for (int v = 0; v < c.nv; ++v) f[v] = fl[v];
            } else if (val(lmax) <= 0.0 && val(lmin) <= 0.0) {
                for (int v = 0; v < c.nv; ++v) f[v] = fr[v];
            } else {
                for (int v = 0; v < c.nv; ++v)
                    f[v] = (lmax * fl[v] - lmin * fr[v] + lmin * lmax * (urr[v] - ull[v])) / (lmax - lmin);
            }
            return;
        }
        case OR_FLUX_HLLE: {
            // Trixi flux_hlle = FluxHLL(min_max_speed_einfeldt), ideal GLM-MHD:
            // lambda_min = min(v_ll - c_f,ll, v_roe - c_f,roe),
            // lambda_max = max(v_rr + c_f,rr, v_roe + c_f,roe)
            T fl[MAXV], fr[MAXV], vl, cl, vr, cr, vroe, croe;
            phys_flux(c, ull, o, fl);
            phys_flux(c, urr, o, fr);
            normal_speed(c, ull, o, vl, cl);
            normal_speed(c, urr, o, vr, cr);
            mhd_fast_wavespeed_roe(c, ull, urr, o, vroe, croe);
            const T lmin = tmin(vl - cl, vroe - croe), lmax = tmax(vr + cr, vroe + croe);
            if (val(lmin) >= 0.0 && val(lmax) >= 0.0) {
                for (int v = 0; v < c.nv; ++v) f[v] = fl[v];
            } else if (val(lmax) <= 0.0 && val(lmin) <= 0.0) {
                for (int v = 0; v < c.nv; ++v) f[v] = fr[v];
            } else {
                for (int v = 0; v < c.nv; ++v)
                    f[v] = (lmax * fl[v] - lmin * fr[v] + lmin * lmax * (urr[v] - ull[v])) / (lmax - lmin);
            }
            return;
        }
        case OR_FLUX_EC:  // Trixi flux_ec(u_ll, u_rr, orientation, ::InviscidBurgersEquation1D)
            f[0] = (ull[0] * ull[0] + ull[0] * urr[0] + urr[0] * urr[0]) / 6.0;
            return;
        case OR_FLUX_LLF:
        case OR_FLUX_LLF_NAIVE: {
            T fl[MAXV], fr[MAXV];
            phys_flux(c, ull, o, fl);
            phys_flux(c, urr, o, fr);
            T lam = lambda_max(c, ull, urr, o, kind == OR_FLUX_LLF_NAIVE);
            for (int v = 0; v < c.nv; ++v)
                f[v] = 0.5 * (fl[v] + fr[v]) + (-0.5 * lam * (urr[v] - ull[v]));
            // Trixi's shallow-water LLF dissipation leaves the bottom topography out
            if (c.eq == OR_EQ_SHALLOW_WATER) f[c.nd + 1] = 0.0;
            return;
        }
        case OR_FLUX_CENTRAL:
        default: {
            T fl[MAXV], fr[MAXV];
            phys_flux(c, ull, o, fl);
            phys_flux(c, urr, o, fr);
            for (int v = 0; v < c.nv; ++v) f[v] = 0.5 * (fl[v] + fr[v]);
            return;
        }
    }
}

// (1D ideal MHD has no nonconservative terms: B1 is constant)
inline bool has_noncons(const Cfg& c) {
    return (c.eq == OR_EQ_GLM_MHD && c.nd > 1) || c.eq == OR_EQ_EULER_QUASI1D || c.eq == OR_EQ_SHALLOW_WATER;
}

// ----------------------------------------------------------------------------
// Basis (Trixi.jl basis_lobatto_legendre.jl; used by the reference at
// src/solvers/basis_lobatto_legendre.jl:60-100).
// ----------------------------------------------------------------------------
void calc_q_and_l(int N, double x, double& q, double& qder, double& L) {
    double L_Nm2 = 1.0, L_Nm1 = x, Lder_Nm2 = 0.0, Lder_Nm1 = 1.0;
    L = x;
    for (int i = 2; i <= N; ++i) {
        L = ((2 * i - 1) * x * L_Nm1 - (i - 1) * L_Nm2) / i;
        double Lder = Lder_Nm2 + (2 * i - 1) * L_Nm1;
        L_Nm2 = L_Nm1;
        L_Nm1 = L;
        Lder_Nm2 = Lder_Nm1;
        Lder_Nm1 = Lder;
    }
    q = (2.0 * N + 1) / (N + 1) * (x * L - L_Nm2);
    qder = (2.0 * N + 1) * L;
}

void gauss_lobatto_nodes_weights(int n_nodes, double* nodes, double* weights) {
    const double tolerance = 4 * 2.220446049250313e-16;
    const int n_iterations = 20;
    const int N = n_nodes - 1;
    nodes[0] = -1.0;
    nodes[N] = 1.0;
    weights[0] = 2.0 / (N * (N + 1));
    weights[N] = weights[0];
    if (N > 1) {
        const double cont1 = M_PI / N;
        const double cont2 = 3.0 / (8.0 * N * M_PI);
        for (int i = 1; i <= (N + 1) / 2 - 1; ++i) {
            nodes[i] = -cos(cont1 * (i + 0.25) - cont2 / (i + 0.25));
            for (int k = 0; k < n_iterations; ++k) {
                double q, qder, L;
                calc_q_and_l(N, nodes[i], q, qder, L);
                double dx = -q / qder;
                nodes[i] += dx;
                if (fabs(dx) < tolerance * fabs(nodes[i])) break;
            }
            double q, qder, L;
            calc_q_and_l(N, nodes[i], q, qder, L);
            weights[i] = weights[0] / (L * L);
            nodes[N - i] = -nodes[i];
            weights[N - i] = weights[i];
        }
    }
    if (n_nodes % 2 == 1) {
        double q, qder, L;
        calc_q_and_l(N, 0.0, q, qder, L);
        nodes[N / 2] = 0.0;
        weights[N / 2] = weights[0] / (L * L);
    }
}

void barycentric_weights(int n, const double* x, double* w) {
    for (int j = 0; j < n; ++j) w[j] = 1.0;
    for (int j = 1; j < n; ++j)
        for (int k = 0; k < j; ++k) {
            w[k] *= x[k] - x[j];
            w[j] *= x[j] - x[k];
        }
    for (int j = 0; j < n; ++j) w[j] = 1.0 / w[j];
}

void lagrange_polys(double x, int n, const double* nodes, const double* wbary, double* out) {
    for (int j = 0; j < n; ++j)
        if (fabs(x - nodes[j]) <= 2.220446049250313e-16 * fabs(x) || x == nodes[j]) {
            for (int k = 0; k < n; ++k) out[k] = (k == j) ? 1.0 : 0.0;
            return;
        }
    double total = 0.0;
    for (int j = 0; j < n; ++j) {
        out[j] = wbary[j] / (x - nodes[j]);
        total += out[j];
    }
    for (int j = 0; j < n; ++j) out[j] /= total;
}

// Legendre-Gauss nodes and weights (Newton iteration on P_n from the
// Chebyshev-like first guess; w_i = 2 / ((1 - x_i^2) P_n'(x_i)^2)).
void gauss_nodes_weights(int n, double* x, double* w) {
    for (int i = 0; i < n; ++i) {
        double xi = -cos(M_PI * (i + 0.75) / (n + 0.5));
        double dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = xi;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2 * k - 1) * xi * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            const double pn = (n == 1) ? xi : p1, pm = (n == 1) ? 1.0 : p0;
            dp = n * (xi * pn - pm) / (xi * xi - 1.0);
            const double dx = pn / dp;
            xi -= dx;
            if (fabs(dx) < 1e-16) break;
        }
        double p0 = 1.0, p1 = xi;
        for (int k = 2; k <= n; ++k) {
            const double p2 = ((2 * k - 1) * xi * p1 - (k - 1) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        dp = n * (xi * ((n == 1) ? xi : p1) - ((n == 1) ? 1.0 : p0)) / (xi * xi - 1.0);
        x[i] = xi;
        w[i] = 2.0 / ((1.0 - xi * xi) * dp * dp);
    }
}

// Mortar operators of Trixi's MortarL2 (external to the reference, used
// through MortarL2GPU, basis_lobatto_legendre.jl:155-173):
//   forward_{lower,upper}[j][i] = l_i((x_j -+ 1) / 2)  (Lobatto interpolation),
//   reverse_{lower,upper} = G2L * P * L2G with the Gauss-basis projection
//   P[i][j] = 1/2 * lg_i((g_j -+ 1)/2) * wg_j / wg_i.
void mortar_matrices(int N, double* fl, double* fu, double* rl, double* ru) {
    double x[16], w[16], wb[16], g[16], wg[16], gb[16], poly[16];
    gauss_lobatto_nodes_weights(N, x, w);
    barycentric_weights(N, x, wb);
    gauss_nodes_weights(N, g, wg);
    barycentric_weights(N, g, gb);
    for (int j = 0; j < N; ++j) {
        lagrange_polys(0.5 * (x[j] - 1.0), N, x, wb, poly);
        for (int i = 0; i < N; ++i) fl[j * N + i] = poly[i];
        lagrange_polys(0.5 * (x[j] + 1.0), N, x, wb, poly);
        for (int i = 0; i < N; ++i) fu[j * N + i] = poly[i];
    }
    double G2L[256], L2G[256];
    for (int j = 0; j < N; ++j) {
        lagrange_polys(x[j], N, g, gb, poly);  // Gauss -> Lobatto
        for (int i = 0; i < N; ++i) G2L[j * N + i] = poly[i];
        lagrange_polys(g[j], N, x, wb, poly);  // Lobatto -> Gauss
        for (int i = 0; i < N; ++i) L2G[j * N + i] = poly[i];
    }
    for (int half = 0; half < 2; ++half) {
        double P[256], T[256];
        for (int j = 0; j < N; ++j) {
            lagrange_polys(0.5 * (g[j] + (half ? 1.0 : -1.0)), N, g, gb, poly);
            for (int i = 0; i < N; ++i) P[i * N + j] = 0.5 * poly[i] * wg[j] / wg[i];
        }
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) {
                double acc = 0.0;
                for (int k = 0; k < N; ++k) acc += P[i * N + k] * L2G[k * N + j];
                T[i * N + j] = acc;
            }
        double* R = half ? ru : rl;
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) {
                double acc = 0.0;
                for (int k = 0; k < N; ++k) acc += G2L[i * N + k] * T[k * N + j];
                R[i * N + j] = acc;
            }
    }
}

double legendre_normalized(int N, double x) {
    double poly;
    if (N == 0) {
        poly = 1.0;
    } else if (N == 1) {
        poly = x;
    } else {
        double p2 = 1.0, p1 = x;
        poly = 0.0;
        for (int i = 2; i <= N; ++i) {
            poly = ((2 * i - 1) * x * p1 - (i - 1) * p2) / i;
            p2 = p1;
            p1 = poly;
        }
    }
    return poly * sqrt(N + 0.5);
}

bool invert_matrix(int n, const double* A, double* Ainv) {
    std::vector<double> M(A, A + n * n), I(n * n, 0.0);
    for (int i = 0; i < n; ++i) I[i * n + i] = 1.0;
    for (int col = 0; col < n; ++col) {
        int piv = col;
        for (int r = col + 1; r < n; ++r)
            if (fabs(M[r * n + col]) > fabs(M[piv * n + col])) piv = r;
        if (M[piv * n + col] == 0.0) return false;
        if (piv != col)
            for (int k = 0; k < n; ++k) {
                std::swap(M[piv * n + k], M[col * n + k]);
                std::swap(I[piv * n + k], I[col * n + k]);
            }
        double d = M[col * n + col];
        for (int k = 0; k < n; ++k) {
            M[col * n + k] /= d;
            I[col * n + k] /= d;
        }
        for (int r = 0; r < n; ++r)
            if (r != col) {
                double f = M[r * n + col];
                if (f != 0.0)
                    for (int k = 0; k < n; ++k) {
                        M[r * n + k] -= f * M[col * n + k];
                        I[r * n + k] -= f * I[col * n + k];
                    }
            }
    }
    memcpy(Ainv, I.data(), sizeof(double) * n * n);
    return true;
}

struct Basis {
    int N;
    double nodes[16], weights[16], invw[16];
    double D[256], Ds[256], Dhat[256], lm[16], lp[16], Vinv[256];
};

Basis make_basis(int N) {
    Basis b;
    b.N = N;
    oracle_basis(N, b.nodes, b.weights, b.D, b.Ds, b.Dhat, b.lm, b.lp, b.Vinv);
    for (int i = 0; i < N; ++i) b.invw[i] = 1.0 / b.weights[i];
    return b;
}

// ----------------------------------------------------------------------------
// Node indexing helpers: node n = i + N*j + N^2*k, stride of direction d = N^d.
// ----------------------------------------------------------------------------
inline int ipow(int b, int e) {
    int r = 1;
    for (int i = 0; i < e; ++i) r *= b;
    return r;
}

// Face node f (0..N^(nd-1)-1) of direction od (0-based) at coordinate idx.
inline int face_node(int N, int nd, int od, int f, int idx) {
    int coords[3] = {0, 0, 0};
    int rem = f;
    for (int d = 0; d < nd; ++d) {
        if (d == od) continue;
        coords[d] = rem % N;
        rem /= N;
    }
    coords[od] = idx;
    return coords[0] + N * (coords[1] + N * coords[2]);
}

// ----------------------------------------------------------------------------
// Volume integrals (Trixi dg_{1,2,3}d.jl weak_form_kernel!,
// flux_differencing_kernel!, fv_kernel!; reference GPU counterparts at
// dg_3d_kernel.jl:65-120, 188-255, 340-423, 589-728).
// ----------------------------------------------------------------------------
template <class T>
void weak_form_element(const Cfg& c, const Basis& b, const T* ue, T* due) {
    const int N = c.N, nd = c.nd, nv = c.nv, nn = ipow(N, nd);
    T f[MAXV];
    for (int n = 0; n < nn; ++n) {
        for (int d = 0; d < nd; ++d) {
            const int s = ipow(N, d), idx = (n / s) % N;
            phys_flux(c, ue + nv * n, d + 1, f);
            for (int ii = 0; ii < N; ++ii) {
                const int m = n + (ii - idx) * s;
                const double fac = b.Dhat[ii * N + idx];
                for (int v = 0; v < nv; ++v) due[nv * m + v] = due[nv * m + v] + fac * f[v];
            }
        }
    }
}

template <class T>
void flux_differencing_element(const Cfg& c, const Basis& b, int kind, const T* ue, T* due,
                               T alpha) {
    const int N = c.N, nd = c.nd, nv = c.nv, nn = ipow(N, nd);
    T f[MAXV];
    for (int n = 0; n < nn; ++n) {
        for (int d = 0; d < nd; ++d) {
            const int s = ipow(N, d), idx = (n / s) % N;
            for (int ii = idx + 1; ii < N; ++ii) {
                const int m = n + (ii - idx) * s;
                two_point(c, kind, ue + nv * n, ue + nv * m, d + 1, f);
                T fa = alpha * b.Ds[idx * N + ii];
                T fb = alpha * b.Ds[ii * N + idx];
                for (int v = 0; v < nv; ++v) {
                    due[nv * n + v] = due[nv * n + v] + fa * f[v];
                    due[nv * m + v] = due[nv * m + v] + fb * f[v];
                }
            }
        }
    }
    if (has_noncons(c)) {
        for (int n = 0; n < nn; ++n) {
            T integral[MAXV];
            for (int v = 0; v < nv; ++v) integral[v] = 0.0;
            for (int d = 0; d < nd; ++d) {
                const int s = ipow(N, d), idx = (n / s) % N;
                for (int ii = 0; ii < N; ++ii) {
                    const int m = n + (ii - idx) * s;
                    noncons_flux(c, ue + nv * n, ue + nv * m, d + 1, f, -1);
                    for (int v = 0; v < nv; ++v)
                        integral[v] = integral[v] + b.Ds[idx * N + ii] * f[v];
                }
            }
            T fac = alpha * 0.5;
            for (int v = 0; v < nv; ++v) due[nv * n + v] = due[nv * n + v] + fac * integral[v];
        }
    }
}

template <class T>
void fv_element(const Cfg& c, const Basis& b, const T* ue, T* due, T alpha) {
    // Trixi calcflux_fv! / fv_kernel! (reference dg_2d_kernel.jl:646-654,
    // 744-748): at the subcell interface between nodes idx - 1 and idx,
    // fstar_L = f + nc(u_{idx-1}, u_idx)/2 (used by the left node) and
    // fstar_R = f + nc(u_idx, u_{idx-1})/2 (used by the right node); without
    // nonconservative terms both are f.  Zero at the element ends.
    const int N = c.N, nd = c.nd, nv = c.nv, nn = ipow(N, nd);
    std::vector<T> fl((size_t)nd * nv * (N + 1) * nn, T(0.0)), fr(fl.size(), T(0.0));
    auto at = [&](int d, int v, int pos, int n_other) -> size_t {
        return (((size_t)d * nv + v) * (N + 1) + pos) * nn + n_other;
    };
    T f[MAXV], ncl[MAXV], ncr[MAXV];
    for (int d = 0; d < nd; ++d) {
        const int s = ipow(N, d);
        for (int n = 0; n < nn; ++n) {
            const int idx = (n / s) % N;
            if (idx == 0) continue;
            two_point(c, c.fvflux, ue + nv * (n - s), ue + nv * n, d + 1, f);
            const int base = n - idx * s;
            if (has_noncons(c)) {
                noncons_flux(c, ue + nv * (n - s), ue + nv * n, d + 1, ncl, c.fvflux);
                noncons_flux(c, ue + nv * n, ue + nv * (n - s), d + 1, ncr, c.fvflux);
            }
            for (int v = 0; v < nv; ++v) {
                fl[at(d, v, idx, base)] = has_noncons(c) ? f[v] + 0.5 * ncl[v] : f[v];
                fr[at(d, v, idx, base)] = has_noncons(c) ? f[v] + 0.5 * ncr[v] : f[v];
            }
        }
    }
    for (int n = 0; n < nn; ++n) {
        for (int v = 0; v < nv; ++v) {
            T acc = 0.0;
            for (int d = 0; d < nd; ++d) {
                const int s = ipow(N, d), idx = (n / s) % N, base = n - idx * s;
                T term = b.invw[idx] * (fl[at(d, v, idx + 1, base)] - fr[at(d, v, idx, base)]);
                acc = (d == 0) ? term : acc + term;
            }
            due[nv * n + v] = due[nv * n + v] + alpha * acc;
        }
    }
}

// ----------------------------------------------------------------------------
// Hennemann-Gassner indicator (reference callable src/solvers/indicators.jl:7-42
// with Trixi calc_indicator_hennemann_gassner! and apply_smoothing!).
// ----------------------------------------------------------------------------
// The indicator variable of compressible Euler: Trixi's density_pressure
// rho p = (gamma - 1)(rho rho_e - |rho v|^2 / 2), pressure
// p = (gamma - 1)(rho_e - |rho v|^2 / (2 rho)) or density rho.
template <class T> T indicator_variable(const Cfg& c, const T* u) {
    if (c.eq == OR_EQ_SHALLOW_WATER) {
        // Trixi waterheight_pressure: h * (g h^2 / 2)
        return u[0] * (0.5 * c.grav * u[0] * u[0]);
    }
    if (c.eq == OR_EQ_GLM_MHD) {
        // Trixi density_pressure(u, ::IdealGlmMhdEquations): rho p
        T rho, v[3], p, B[3], psi;
        mhd_prim(c, u, rho, v, p, B, psi);
        return rho * p;
    }
    if (c.ivar == 2) return u[0];
    T ke = 0.0;
    for (int d = 0; d < c.nd; ++d) ke = (d == 0) ? u[1 + d] * u[1 + d] : ke + u[1 + d] * u[1 + d];
    if (c.ivar == 1) return (c.gamma - 1.0) * (u[c.nd + 1] - 0.5 * ke / u[0]);
    return (c.gamma - 1.0) * (u[0] * u[c.nd + 1] - 0.5 * ke);
}

template <class T>
T indicator_element(const Cfg& c, const Basis& b, const T* ue, double threshold,
                    double parameter_s) {
    const int N = c.N, nd = c.nd, nv = c.nv, nn = ipow(N, nd);
    std::vector<T> ind(nn), tmp(nn);
    for (int n = 0; n < nn; ++n) ind[n] = indicator_variable(c, ue + nv * n);
    // multiply_scalar_dimensionwise!: apply Vinv along x, then y, then z.
    for (int d = 0; d < nd; ++d) {
        const int s = ipow(N, d);
        for (int n = 0; n < nn; ++n) {
            const int idx = (n / s) % N, base = n - idx * s;
            T r = 0.0;
            for (int ii = 0; ii < N; ++ii) r = r + b.Vinv[idx * N + ii] * ind[base + ii * s];
            tmp[n] = r;
        }
        ind.swap(tmp);
    }
    T total = 0.0, clip1 = 0.0, clip2 = 0.0;
    for (int n = 0; n < nn; ++n) {
        int maxc = 0;
        for (int d = 0; d < nd; ++d) maxc = std::max(maxc, (n / ipow(N, d)) % N);
        T e2 = ind[n] * ind[n];
        total = total + e2;
        if (maxc < N - 1) clip1 = clip1 + e2;
        if (maxc < N - 2) clip2 = clip2 + e2;
    }
    T frac1 = (val(total) != 0.0) ? (total - clip1) / total : T(0.0);
    T frac2 = (val(clip1) != 0.0) ? (clip1 - clip2) / clip1 : T(0.0);
    T energy = tmax(frac1, frac2);
    T a = 1.0 / (1.0 + texp(-parameter_s / threshold * (energy - threshold)));
    if (val(a) < c.amin) a = 0.0;
    if (val(a) > 1.0 - c.amin) a = 1.0;
    return tmin(T(c.amax), a);
}

template <class T>
void indicator_hg(const Cfg& c, const Basis& b, int64_t E, int64_t I, const int32_t* nbr,
                  const T* u, T* alpha) {
    const int nn = ipow(c.N, c.nd), nv = c.nv;
    const double threshold = 0.5 * pow(10.0, -1.8 * pow((double)c.N, 0.25));
    const double parameter_s = log((1.0 - 0.0001) / 0.0001);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < E; ++e)
        alpha[e] = indicator_element(c, b, u + (size_t)e * nn * nv, threshold, parameter_s);
    if (c.smooth) {
        std::vector<T> tmp(alpha, alpha + E);
        for (int64_t i = 0; i < I; ++i) {
            const int32_t l = nbr[2 * i], r = nbr[2 * i + 1];
            alpha[l] = tmax(tmax(tmp[l], 0.5 * tmp[r]), alpha[l]);
            alpha[r] = tmax(tmax(tmp[r], 0.5 * tmp[l]), alpha[r]);
        }
        // L2 mortars: each small element with the large one (Trixi's
        // apply_smoothing! loops over the mortars after the interfaces)
        const int K = 1 << (c.nd - 1);
        for (int64_t m = 0; m < c.nmortars; ++m) {
            const int32_t* ids = c.mids + m * (K + 1);
            const int32_t large = ids[K];
            for (int p = 0; p < K; ++p) {
                const int32_t s = ids[p];
                alpha[s] = tmax(tmax(tmp[s], 0.5 * tmp[large]), alpha[s]);
                alpha[large] = tmax(tmax(tmp[large], 0.5 * tmp[s]), alpha[large]);
            }
        }
    }
}

// ----------------------------------------------------------------------------
// Full rhs! (reset, volume, prolong2interfaces, interface flux, surface
// integral, Jacobian); periodic meshes only, no mortars, no sources -- the
// reference asserts the same (dg_3d.jl:463-466, 542-544).
// ----------------------------------------------------------------------------
void initial_condition(const Cfg& c, int ic, const double* x, double t, uint64_t seed, uint64_t node_index,
                       double* u);
void hypdiff_state(const Cfg& c, int ic, const double* x, double t, bool steady, uint64_t seed, uint64_t idx,
                   double* u);

// source_terms_convergence_test (compressible Euler): with
// g(s) = 2 + 0.1 sin(pi s), s = x_1 + ... + x_d - t, the convergence-test state
// rho = rho v_i = g, rho e = g^2 (v_i = 1, p = (gamma - 1)(g^2 - d g / 2)) solves
// u_t + div f(u) = S with S the terms below (dg/ds = 0.1 pi cos(pi s)):
//   S_rho      = d g' - g'                                (continuity)
//   S_{rho v_i} = d g' - g' + p'                           (momentum, i = 1..d)
//   S_{rho e}  = d (2 g g' + p') - 2 g g'                 (energy)
// with p' = (gamma - 1)(2 g - d/2) g'.
void source_convergence_test(const Cfg& c, const double* x, double t, double* sv) {
    if (c.src == OR_SOURCE_EOC_TEST_EULER) {
        // source_terms_eoc_test_euler (Trixi, d = 2, 3) in its own form:
        // rho_x = 0.1 pi cos(pi (sum x - t)), C_grav = -4 G / (d pi) (G = 1),
        // du_rho = (d-1) rho_x, du_{rho v} = rho_x ((d-1) - C_grav rho),
        // du_E = rho_x (d (d-1)/2 - ((d-1)/(gamma-1) + d) C_grav rho)
        const int d = c.nd;
        double s = -t;
        for (int q = 0; q < d; ++q) s += x[q];
        const double rhox = 0.1 * M_PI * cos(M_PI * s), rho = 2.0 + 0.1 * sin(M_PI * s);
        const double C_grav = -4.0 / (d * M_PI);
        sv[0] = (d - 1) * rhox;
        for (int q = 0; q < d; ++q) sv[1 + q] = rhox * ((d - 1) - C_grav * rho);
        sv[d + 1] = rhox * (0.5 * d * (d - 1) - ((d - 1) * c.inv_gm1 + d) * C_grav * rho);
        return;
    }
    if (c.eq == OR_EQ_SHALLOW_WATER) {
        // Trixi's ShallowWaterEquations2D source_terms_convergence_test for the
        // state H = 7 + cos(w x) sin(w y) cos(w_t t), v = (0.5, 1.5),
        // b = 2 + (sin(w_b x) + sin(w_b y)) / 2, w = 2 pi sqrt 2, w_t = 2 pi,
        // w_b = sqrt 2 pi: du1 = H_t + v1 (H_x - b_x) + v2 (H_y - b_y),
        // du2 = v1 du1 + g (H - b) H_x, du3 = v2 du1 + g (H - b) H_y
        const double w = 2.0 * M_PI * sqrt(2.0), wt = 2.0 * M_PI, wb = sqrt(2.0) * M_PI, v1 = 0.5, v2 = 1.5;
        const double H = 7.0 + cos(w * x[0]) * sin(w * x[1]) * cos(wt * t);
        const double H_x = -w * sin(w * x[0]) * sin(w * x[1]) * cos(wt * t);
        const double H_y = w * cos(w * x[0]) * cos(w * x[1]) * cos(wt * t);
        const double H_t = -wt * cos(w * x[0]) * sin(w * x[1]) * sin(wt * t);
        const double b = 2.0 + 0.5 * sin(wb * x[0]) + 0.5 * sin(wb * x[1]);
        const double b_x = 0.5 * wb * cos(wb * x[0]), b_y = 0.5 * wb * cos(wb * x[1]);
        sv[0] = H_t + v1 * (H_x - b_x) + v2 * (H_y - b_y);
        sv[1] = v1 * sv[0] + c.grav * (H - b) * H_x;
        sv[2] = v2 * sv[0] + c.grav * (H - b) * H_y;
        sv[3] = 0.0;
        return;
    }
    if (c.eq == OR_EQ_EULER_QUASI1D) {
        // Trixi's quasi-1D source_terms_convergence_test: the residual of the
        // manufactured state rho = ini, v = 1, e = ini^2 / rho,
        // p = (gamma - 1)(e - rho v^2 / 2), a = 1.5 - 0.5 cos(pi x):
        //   du1 = (a rho)_t + (a rho v)_x, du2 = (a rho v)_t + (a rho v v)_x + a p_x,
        //   du3 = (a e)_t + (a v (e + p))_x,
        // by the product rule with ini_t = -w A cos, ini_x = w A cos (w = pi, A = 0.1)
        const double A = 0.1, w = M_PI, arg = w * (x[0] - t);
        const double ini = 2.0 + A * sin(arg), ini_x = w * A * cos(arg), ini_t = -ini_x;
        const double rho = ini, v = 1.0, e = ini * ini / rho;
        const double rho_x = ini_x, rho_t = ini_t;
        const double e_x = 2.0 * ini * ini_x / rho - ini * ini * rho_x / (rho * rho);
        const double e_t = 2.0 * ini * ini_t / rho - ini * ini * rho_t / (rho * rho);
        const double p = (c.gamma - 1.0) * (e - 0.5 * rho * v * v);
        const double p_x = (c.gamma - 1.0) * (e_x - 0.5 * rho_x * v * v);
        const double a = 1.5 - 0.5 * cos(x[0] * M_PI), a_x = 0.5 * M_PI * sin(x[0] * M_PI);
        sv[0] = a * rho_t + (a_x * rho * v + a * rho_x * v);
        sv[1] = a * rho_t * v + (a_x * rho * v * v + a * rho_x * v * v) + a * p_x;
        sv[2] = a * e_t + (a_x * v * (e + p) + a * v * (e_x + p_x));
        sv[3] = 0.0;
        return;
    }
    if (c.eq == OR_EQ_BURGERS) {
        // u = 2 + sin(2 pi (x - t)): u_t + u u_x = 2 pi cos(.) (u - 1)
        const double w = 2.0 * M_PI, arg = w * (x[0] - t);
        sv[0] = w * cos(arg) * (2.0 + sin(arg) - 1.0);
        return;
    }
    const int d = c.nd;
    double s = -t;
    for (int q = 0; q < d; ++q) s += x[q];
    const double g = 2.0 + 0.1 * sin(M_PI * s), gp = 0.1 * M_PI * cos(M_PI * s);
    const double pp = (c.gamma - 1.0) * (2.0 * g - 0.5 * d) * gp;
    sv[0] = d * gp - gp;
    for (int q = 0; q < d; ++q) sv[1 + q] = d * gp - gp + pp;
    sv[d + 1] = d * (2.0 * g * gp + pp) - 2.0 * g * gp;
}

template <class T>
void rhs_impl(const Cfg& c, int64_t E, int64_t I, const int32_t* nbr, const int32_t* ori,
              const double* invJ, const T* u, T* du, T* alpha_out) {
    const Basis b = make_basis(c.N);
    const int N = c.N, nd = c.nd, nv = c.nv, nn = ipow(N, nd), nf = ipow(N, nd - 1);
    const size_t esz = (size_t)nn * nv;

    // reset_du!
    for (size_t q = 0; q < (size_t)E * esz; ++q) du[q] = 0.0;

    // calc_volume_integral!
    if (c.vi == OR_VI_WEAK) {
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < E; ++e) weak_form_element(c, b, u + e * esz, du + e * esz);
    } else if (c.vi == OR_VI_FD) {
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < E; ++e)
            flux_differencing_element(c, b, c.vflux, u + e * esz, du + e * esz, T(1.0));
    } else {
        std::vector<T> alpha(E);
        indicator_hg(c, b, E, I, nbr, u, alpha.data());
        const double eps = 2.220446049250313e-16;
        const double atol = std::max(100 * eps, pow(eps, 0.75));
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < E; ++e) {
            T ae = alpha[e];
            if (fabs(val(ae)) <= atol) {
                flux_differencing_element(c, b, c.vflux, u + e * esz, du + e * esz, T(1.0));
            } else {
                flux_differencing_element(c, b, c.vflux, u + e * esz, du + e * esz, 1.0 - ae);
                fv_element(c, b, u + e * esz, du + e * esz, ae);
            }
        }
        if (alpha_out)
            for (int64_t e = 0; e < E; ++e) alpha_out[e] = alpha[e];
    }

    // prolong2interfaces!: iu[s][v][f][iface]
    std::vector<T> iu((size_t)2 * nv * nf * I);
    auto IU = [&](int s, int v, int f, int64_t i) -> T& {
        return iu[(((size_t)i * nf + f) * nv + v) * 2 + s];
    };
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < I; ++i) {
        const int od = ori[i] - 1;
        const int32_t L = nbr[2 * i], R = nbr[2 * i + 1];
        for (int f = 0; f < nf; ++f) {
            const int nL = face_node(N, nd, od, f, N - 1), nR = face_node(N, nd, od, f, 0);
            for (int v = 0; v < nv; ++v) {
                IU(0, v, f, i) = u[L * esz + nv * nL + v];
                IU(1, v, f, i) = u[R * esz + nv * nR + v];
            }
        }
    }

    // calc_interface_flux!: sfv[v][f][dir][e]
    std::vector<T> sfv((size_t)nv * nf * 2 * nd * E);
    auto SF = [&](int v, int f, int dir, int64_t e) -> T& {
        return sfv[(((size_t)e * 2 * nd + dir) * nf + f) * nv + v];
    };
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < I; ++i) {
        const int o = ori[i];
        const int32_t L = nbr[2 * i], R = nbr[2 * i + 1];
        const int ldir = 2 * o - 1, rdir = 2 * o - 2;  // 0-based (Julia 2o, 2o-1)
        T ull[MAXV], urr[MAXV], flux[MAXV], ncl[MAXV], ncr[MAXV];
        for (int f = 0; f < nf; ++f) {
            for (int v = 0; v < nv; ++v) {
                ull[v] = IU(0, v, f, i);
                urr[v] = IU(1, v, f, i);
            }
            two_point(c, c.sflux, ull, urr, o, flux);
            if (has_noncons(c)) {
                noncons_flux(c, ull, urr, o, ncl, c.sflux);
                noncons_flux(c, urr, ull, o, ncr, c.sflux);
                for (int v = 0; v < nv; ++v) {
                    SF(v, f, ldir, L) = flux[v] + 0.5 * ncl[v];
                    SF(v, f, rdir, R) = flux[v] + 0.5 * ncr[v];
                }
            } else {
                for (int v = 0; v < nv; ++v) {
                    SF(v, f, ldir, L) = flux[v];
                    SF(v, f, rdir, R) = flux[v];
                }
            }
        }
    }

    // prolong2mortars! / calc_mortar_flux! / mortar_fluxes_to_elements! (L2
    // mortars; reference dg_2d_kernel.jl:1167-1390, dg_3d_kernel.jl:1361-1790):
    // the large element's face values are interpolated to each small face
    // (forward_lower / forward_upper per tangential direction, position bits
    // first tangential direction fastest), the surface flux is evaluated with
    // the large side on its side of the mortar, the small elements take it as
    // their face flux, and the large element takes the L2 projection
    // (reverse_lower / reverse_upper) of all small-face fluxes.
    std::vector<char> mortar_face((size_t)2 * nd * E, 0);
    if (c.nmortars > 0) {
        const int K = 1 << (nd - 1);
        double Fm[2][256], Rm[2][256];
        mortar_matrices(N, Fm[0], Fm[1], Rm[0], Rm[1]);
        for (int64_t m = 0; m < c.nmortars; ++m) {
            const int o = c.mori[m], ls = c.mls[m];
            const int32_t* ids = c.mids + m * (K + 1);
            const int32_t large = ids[K];
            const int dir_large = 2 * (o - 1) + (ls == 1 ? 1 : 0), dir_small = 2 * (o - 1) + (ls == 1 ? 0 : 1);
            const int pos_large = (ls == 1) ? N - 1 : 0, pos_small = (ls == 1) ? 0 : N - 1;
            mortar_face[(size_t)dir_large * E + large] = 1;
            std::vector<T> fstar((size_t)K * nf * nv);
            for (int p = 0; p < K; ++p) {
                const int32_t small = ids[p];
                mortar_face[(size_t)dir_small * E + small] = 1;
                const int p1 = p & 1, p2 = (p >> 1) & 1;
                for (int q = 0; q < nf; ++q) {
                    const int q1 = q % N, q2 = q / N;
                    T uL[MAXV], uS[MAXV], flux[MAXV];
                    for (int v = 0; v < nv; ++v) {
                        T acc = 0.0;
                        for (int a = 0; a < N; ++a) {
                            if (nd == 2) {
                                acc = acc + Fm[p1][q1 * N + a] * u[large * esz + nv * face_node(N, nd, o - 1, a, pos_large) + v];
                            } else {
                                for (int bb = 0; bb < N; ++bb)
                                    acc = acc + Fm[p1][q1 * N + a] * Fm[p2][q2 * N + bb] *
                                                    u[large * esz + nv * face_node(N, nd, o - 1, a + N * bb, pos_large) + v];
                            }
                        }
                        uL[v] = acc;
                        uS[v] = u[small * esz + nv * face_node(N, nd, o - 1, q, pos_small) + v];
                    }
                    if (ls == 1)
                        two_point(c, c.sflux, uL, uS, o, flux);
                    else
                        two_point(c, c.sflux, uS, uL, o, flux);
                    // nonconservative terms as the reference's mortar_flux_kernel!
                    // (dg_2d_kernel.jl:1318-1337, dg_3d_kernel.jl:1598-1640): the small
                    // faces take 1/2 nc(u_large, u_small), the projected large-face
                    // fluxes 1/2 nc(u_small, u_large)
                    T nc_small[MAXV], nc_large[MAXV];
                    for (int v = 0; v < nv; ++v) nc_small[v] = nc_large[v] = 0.0;
                    if (has_noncons(c)) {
                        noncons_flux(c, uL, uS, o, nc_small, c.sflux);
                        noncons_flux(c, uS, uL, o, nc_large, c.sflux);
                    }
                    for (int v = 0; v < nv; ++v) {
                        SF(v, q, dir_small, small) = flux[v] + 0.5 * nc_small[v];
                        fstar[((size_t)p * nf + q) * nv + v] = flux[v] + 0.5 * nc_large[v];
                    }
                }
            }
            for (int j = 0; j < nf; ++j) {
                const int j1 = j % N, j2 = j / N;
                for (int v = 0; v < nv; ++v) {
                    T acc = 0.0;
                    for (int p = 0; p < K; ++p) {
                        const int p1 = p & 1, p2 = (p >> 1) & 1;
                        for (int q = 0; q < nf; ++q) {
                            const int q1 = q % N, q2 = q / N;
                            const double r = (nd == 2) ? Rm[p1][j1 * N + q1] : Rm[p1][j1 * N + q1] * Rm[p2][j2 * N + q2];
                            acc = acc + r * fstar[((size_t)p * nf + q) * nv + v];
                        }
                    }
                    SF(v, j, dir_large, large) = acc;
                }
            }
        }
    }

    // calc_boundary_flux! (Trixi BoundaryConditionDirichlet, as in the reference
    // test test/tree_dgsem_2d/euler_source_terms_nonperiodic.jl:12-17 and the
    // GPU kernel boundary_flux_kernel!): every element face without an
    // interface is a boundary face; the state outside is boundary_ic(x, t) at
    // the face node; the flux is surface_flux(u_inner, u_b) when the boundary
    // lies on the element's + side (Trixi direction even), else
    // surface_flux(u_b, u_inner).
    if (c.bc == OR_BC_DIRICHLET || c.bc == OR_BC_SLIP_WALL) {
        std::vector<char> has_interface((size_t)2 * nd * E, 0);
        for (int64_t i = 0; i < I; ++i) {
            const int o = ori[i];
            has_interface[(size_t)(2 * o - 1) * E + nbr[2 * i]] = 1;
            has_interface[(size_t)(2 * o - 2) * E + nbr[2 * i + 1]] = 1;
        }
        for (size_t q = 0; q < has_interface.size(); ++q)
            if (mortar_face[q]) has_interface[q] = 1;
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < E; ++e)
            for (int dir = 0; dir < 2 * nd; ++dir) {
                if (has_interface[(size_t)dir * E + e]) continue;
                const int od = dir / 2;
                const bool plus = (dir % 2) == 1;
                T uin[MAXV], ub[MAXV], flux[MAXV];
                for (int f = 0; f < nf; ++f) {
                    const int node = face_node(N, nd, od, f, plus ? N - 1 : 0);
                    double x[3] = {0, 0, 0}, ubd[MAXV];
                    for (int d = 0; d < nd; ++d) x[d] = c.xnodes[((size_t)e * nn + node) * nd + d];
                    if (c.bc == OR_BC_DIRICHLET) {  // (slip walls: the mirrored inner state below)
                        if (c.eq == OR_EQ_HYPDIFF)
                            hypdiff_state(c, c.bc_ic, x, c.t, c.bc_ic == OR_IC_POISSON_NONPERIODIC, c.bc_seed,
                                          (uint64_t)(e * nn + node), ubd);
                        else
                            initial_condition(c, c.bc_ic, x, c.t, c.bc_seed, (uint64_t)(e * nn + node), ubd);
                    }
                    for (int v = 0; v < nv; ++v) {
                        uin[v] = u[e * esz + nv * node + v];
                        ub[v] = (c.bc == OR_BC_SLIP_WALL) ? uin[v] : T(ubd[v]);
                    }
                    // boundary_condition_slip_wall (shallow water): (h, -h v_n, b) outside
                    if (c.bc == OR_BC_SLIP_WALL) ub[1 + od] = -uin[1 + od];
                    if ((c.outflow >> dir) & 1)
                        phys_flux(c, uin, od + 1, flux);  // outflow: flux(u_inner)
                    else if (plus)
                        two_point(c, c.sflux, uin, ub, od + 1, flux);
                    else
                        two_point(c, c.sflux, ub, uin, od + 1, flux);
                    // nonconservative terms (reference boundary_flux_kernel!,
                    // dg_2d_kernel.jl:1144-1159): half of Trixi's Dirichlet
                    // nonconservative flux, inner state first on either side
                    if (has_noncons(c) && !((c.outflow >> dir) & 1)) {
                        T nc[MAXV];
                        noncons_flux(c, uin, ub, od + 1, nc, c.sflux);
                        for (int v = 0; v < nv; ++v) flux[v] = flux[v] + 0.5 * nc[v];
                    }
                    for (int v = 0; v < nv; ++v) SF(v, f, dir, e) = flux[v];
                }
            }
    }

    // calc_surface_integral!
    const double factor_1 = b.lm[0], factor_2 = b.lp[N - 1];
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < E; ++e) {
        T* due = du + e * esz;
        for (int f = 0; f < nf; ++f)
            for (int v = 0; v < nv; ++v)
                for (int od = 0; od < nd; ++od) {
                    const int nm = face_node(N, nd, od, f, 0), np = face_node(N, nd, od, f, N - 1);
                    due[nv * nm + v] = due[nv * nm + v] - SF(v, f, 2 * od, e) * factor_1;
                    due[nv * np + v] = due[nv * np + v] + SF(v, f, 2 * od + 1, e) * factor_2;
                }
    }

    // apply_jacobian!
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < E; ++e) {
        const double factor = -invJ[e];
        for (size_t q = 0; q < esz; ++q) du[e * esz + q] = du[e * esz + q] * factor;
    }

    // calc_sources! (after the Jacobian, as in Trixi and the reference
    // cuda_sources!, dg_2d.jl:746-757)
    if (c.src != OR_SOURCE_NONE) {
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < E; ++e)
            for (int n = 0; n < nn; ++n) {
                double x[3] = {0, 0, 0}, sv[MAXV];
                for (int d = 0; d < nd; ++d) x[d] = c.xnodes[((size_t)e * nn + n) * nd + d];
                if (c.eq == OR_EQ_HYPDIFF) {
                    // source_terms_poisson_nonperiodic / source_terms_harmonic:
                    // f = -nu lap phi of the steady state (sum of the second
                    // derivatives along each direction), and -q / T_r
                    double f = 0.0;
                    if (c.src == OR_SOURCE_POISSON_NONPERIODIC) {
                        if (nd == 1) {
                            f = c.nu * M_PI * M_PI * sin(M_PI * x[0]);
                        } else {
                            double k2 = M_PI * M_PI;
                            for (int k = 1; k < nd; ++k) k2 += 4.0 * M_PI * M_PI;
                            double prod = 2.0 * cos(M_PI * x[0]);
                            for (int k = 1; k < nd; ++k) prod *= sin(2.0 * M_PI * x[k]);
                            f = c.nu * k2 * prod;
                        }
                    }
                    sv[0] = f;
                    for (int k = 0; k < nd; ++k) sv[1 + k] = -c.inv_Tr * val(u[e * esz + nv * n + 1 + k]);
                } else {
                    source_convergence_test(c, x, c.t, sv);
                }
                for (int v = 0; v < nv; ++v) du[e * esz + nv * n + v] = du[e * esz + nv * n + v] + sv[v];
            }
    }
}

// ----------------------------------------------------------------------------
// Initial conditions.
// ----------------------------------------------------------------------------
uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
// Uniform in (0,1): counter-based, keyed by (seed, index, stream).
double u01(uint64_t seed, uint64_t index, uint64_t stream) {
    uint64_t r = splitmix64(splitmix64(seed) ^ (index * 16 + stream));
    return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
double normal01(uint64_t seed, uint64_t index, uint64_t stream) {
    double a = u01(seed, index, stream), bb = u01(seed, index, stream + 1);
    return sqrt(-2.0 * log(a)) * cos(2.0 * M_PI * bb);
}

void euler_prim2cons(const Cfg& c, double rho, const double* v, double p, double* u) {
    u[0] = rho;
    double ke = 0.0;
    for (int d = 0; d < c.nd; ++d) {
        u[1 + d] = rho * v[d];
        ke = (d == 0) ? u[1 + d] * v[d] : ke + u[1 + d] * v[d];
    }
    u[c.nd + 1] = p * c.inv_gm1 + 0.5 * ke;
}

void mhd_prim2cons(const Cfg& c, const double* prim, double* u) {
    const double rho = prim[0], p = prim[4], psi = prim[8];
    const double* v = prim + 1;
    const double* B = prim + 5;
    u[0] = rho;
    for (int d = 0; d < 3; ++d) u[1 + d] = rho * v[d];
    u[4] = p * c.inv_gm1 + 0.5 * (u[1] * v[0] + u[2] * v[1] + u[3] * v[2]) +
           0.5 * (B[0] * B[0] + B[1] * B[1] + B[2] * B[2]) + 0.5 * psi * psi;
    for (int d = 0; d < 3; ++d) u[5 + d] = B[d];
    if (c.nv > 8) u[8] = psi;
}

// Seeded smooth periodic field: sum over 3 modes of amp*trig(2 pi k.x/L + phase),
// with k_d in [-3,3], amp in (-1/3,1/3), phase in (0, 2pi); parameters from the
// counter generator at reserved indices so the field is layout independent.
double smooth_field(uint64_t seed, int field, int nd, const double* x, double L, bool use_cos) {
    double s = 0.0;
    for (int m = 0; m < 3; ++m) {
        const uint64_t base = 0xFFFFFFFF00000000ULL + (uint64_t)(field * 3 + m) * 16;
        double arg = 0.0;
        for (int d = 0; d < nd; ++d) {
            int k = (int)floor(u01(seed, base, d) * 7.0) - 3;
            arg += k * x[d];
        }
        double amp = (2.0 * u01(seed, base, 3) - 1.0) / 3.0;
        double ph = 2.0 * M_PI * u01(seed, base, 4);
        double th = 2.0 * M_PI * arg / L + ph;
        s += amp * (use_cos ? cos(th) : sin(th));
    }
    return s;
}

// Hyperbolic diffusion states (Trixi initial_condition_poisson_nonperiodic:
// the steady solution phi of -lap phi = f, q = grad phi; the initial guess
// phi = q = 1 at t = 0; boundary_condition_poisson_nonperiodic always the
// steady solution) and the 1D harmonic state phi = A + B x, q = B (A = 3,
// B = e; phi = 5, q = 0 at t = 0).
void hypdiff_state(const Cfg& c, int ic, const double* x, double t, bool steady, uint64_t seed, uint64_t idx,
                   double* u) {
    const int d = c.nd;
    if (ic == OR_IC_SEEDED_RANDOM) {
        for (int k = 0; k <= d; ++k) u[k] = 1.0 + 0.5 * (2.0 * u01(seed, idx, k) - 1.0);
    } else if (ic == OR_IC_POISSON_NONPERIODIC) {
        if (t == 0.0 && !steady) {
            for (int k = 0; k <= d; ++k) u[k] = 1.0;
        } else if (d == 1) {
            u[0] = sin(M_PI * x[0]);
            u[1] = M_PI * cos(M_PI * x[0]);
        } else {
            // phi = 2 cos(pi x) prod_{k>0} sin(2 pi x_k) + 2 and its gradient
            double fac[3], dfac[3];
            fac[0] = cos(M_PI * x[0]);
            dfac[0] = -M_PI * sin(M_PI * x[0]);
            for (int k = 1; k < d; ++k) {
                fac[k] = sin(2.0 * M_PI * x[k]);
                dfac[k] = 2.0 * M_PI * cos(2.0 * M_PI * x[k]);
            }
            double prod = 2.0;
            for (int k = 0; k < d; ++k) prod *= fac[k];
            u[0] = prod + 2.0;
            for (int j = 0; j < d; ++j) {
                double g = 2.0 * dfac[j];
                for (int k = 0; k < d; ++k)
                    if (k != j) g *= fac[k];
                u[1 + j] = g;
            }
        }
    } else if (ic == OR_IC_HARMONIC_NONPERIODIC) {
        const double A = 3.0, B = exp(1.0);
        u[0] = (t == 0.0) ? 5.0 : A + B * x[0];
        u[1] = (t == 0.0) ? 0.0 : B;
        for (int k = 2; k <= d; ++k) u[k] = 0.0;
    } else {
        u[0] = 1.0;
        for (int k = 1; k <= d; ++k) u[k] = 0.0;
    }
}

// Weak blast wave primitives (Trixi initial_condition_weak_blast_wave of
// CompressibleEulerEquations{1,2,3}D, Hennemann & Gassner 2020 Sec. 6.3);
// the seeded variant adds U(-0.01, 0.01) per node.
void weak_blast_prim(const Cfg& c, int ic, const double* x, uint64_t seed, uint64_t idx, double& rho,
                     double* v, double& p) {
    double r2 = 0.0;
    for (int d = 0; d < c.nd; ++d) r2 += x[d] * x[d];
    const double r = sqrt(r2);
    const bool inside = !(r > 0.5);
    if (c.nd == 1) {
        const double phi = x[0] > 0 ? 0.0 : M_PI;
        v[0] = inside ? 0.1882 * cos(phi) : 0.0;
    } else if (c.nd == 2) {
        const double phi = atan2(x[1], x[0]);
        v[0] = inside ? 0.1882 * cos(phi) : 0.0;
        v[1] = inside ? 0.1882 * sin(phi) : 0.0;
    } else {
        const double phi = atan2(x[1], x[0]);
        const double theta = (r == 0.0) ? 0.0 : acos(x[2] / r);
        v[0] = inside ? 0.1882 * cos(phi) * sin(theta) : 0.0;
        v[1] = inside ? 0.1882 * sin(phi) * sin(theta) : 0.0;
        v[2] = inside ? 0.1882 * cos(theta) : 0.0;
    }
    rho = inside ? 1.1691 : 1.0;
    p = inside ? 1.245 : 1.0;
    if (ic == OR_IC_WEAK_BLAST_WAVE_SEEDED) {
        // breaks the exact symmetries (v = 0, equal neighbour states)
        rho += 0.01 * (2.0 * u01(seed, idx, 0) - 1.0);
        p += 0.01 * (2.0 * u01(seed, idx, 1) - 1.0);
        for (int d = 0; d < c.nd; ++d) v[d] += 0.01 * (2.0 * u01(seed, idx, 2 + d) - 1.0);
    }
}

void initial_condition(const Cfg& c, int ic, const double* x, double t, uint64_t seed,
                       uint64_t idx, double* u) {
    if (c.eq == OR_EQ_HYPDIFF) {
        hypdiff_state(c, ic, x, t, false, seed, idx, u);
        return;
    }
    if (c.eq == OR_EQ_BURGERS) {
        // Trixi initial_condition_convergence_test(x, t, ::InviscidBurgersEquation1D):
        // c + A sin(omega (x - t)), c = 2, A = 1, L = 1, omega = 2 pi / L; and
        // the shock / rarefaction steps of the reference's Burgers tests
        if (ic == OR_IC_CONVERGENCE_TEST) {
            const double cc = 2.0, A = 1.0, L = 1.0, omega = 2.0 * M_PI / L;
            u[0] = cc + A * sin(omega * (x[0] - t));
        } else if (ic == OR_IC_SHOCK) {
            u[0] = x[0] < 0.5 ? 1.5 : 0.5;
        } else if (ic == OR_IC_RAREFACTION) {
            u[0] = x[0] < 0.5 ? 0.5 : 1.5;
        } else if (ic == OR_IC_SEEDED_RANDOM) {
            u[0] = 1.0 + 0.5 * (2.0 * u01(seed, idx, 0) - 1.0);
        } else {
            u[0] = 2.0;
        }
        return;
    }
    if (c.eq == OR_EQ_LINADV) {
        if (ic == OR_IC_CONVERGENCE_TEST) {
            double xs = 0.0;
            for (int d = 0; d < c.nd; ++d) xs += x[d] - c.a[d] * t;
            const double cc = 1.0, A = 0.5, Ld = 2.0, f = 1.0 / Ld, omega = 2.0 * M_PI * f;
            u[0] = cc + A * sin(omega * xs);
        } else if (ic == OR_IC_SEEDED_RANDOM) {
            u[0] = 1.0 + 0.1 * (2.0 * u01(seed, idx, 0) - 1.0);
        } else {
            u[0] = 2.0;
        }
        return;
    }
    if (c.eq == OR_EQ_SHALLOW_WATER) {
        // Trixi's ShallowWaterEquations2D initial conditions via prim2cons
        // (H, v1, v2, b) -> (H - b, (H - b) v1, (H - b) v2, b): the convergence
        // test (periodic on [0, sqrt 2]^2) and the weak blast wave (centre
        // (0.7, 0.7), H = 4 / 3.25, b = 0); seeded: b = 0.1 U, h = 1 + 0.1 U,
        // v = 0.1 N(0, 1); constant: h = 1, v = 0, b = 0
        double H = 1.0, v1 = 0.0, v2 = 0.0, b = 0.0;
        if (c.nd == 1) {
            // 1D: seeded h = 1 + 0.1 U, v = 0.1 N(0, 1), b = 0.1 U; else h = 1 at rest
            if (ic == OR_IC_SEEDED_RANDOM) {
                b = 0.1 * u01(seed, idx, 6);
                H = 1.0 + 0.1 * (2.0 * u01(seed, idx, 0) - 1.0) + b;
                v1 = 0.1 * normal01(seed, idx, 2);
            }
            u[0] = H - b;
            u[1] = (H - b) * v1;
            u[2] = b;
            return;
        }
        if (ic == OR_IC_CONVERGENCE_TEST) {
            const double cc = 7.0, omega_x = 2.0 * M_PI * sqrt(2.0), omega_t = 2.0 * M_PI;
            H = cc + cos(omega_x * x[0]) * sin(omega_x * x[1]) * cos(omega_t * t);
            v1 = 0.5;
            v2 = 1.5;
            b = 2.0 + 0.5 * sin(sqrt(2.0) * M_PI * x[0]) + 0.5 * sin(sqrt(2.0) * M_PI * x[1]);
        } else if (ic == OR_IC_WEAK_BLAST_WAVE) {
            const double xn = x[0] - 0.7, yn = x[1] - 0.7;
            const double r = sqrt(xn * xn + yn * yn), phi = atan2(yn, xn);
            H = r > 0.5 ? 3.25 : 4.0;
            v1 = r > 0.5 ? 0.0 : 0.1882 * cos(phi);
            v2 = r > 0.5 ? 0.0 : 0.1882 * sin(phi);
        } else if (ic == OR_IC_SEEDED_RANDOM) {
            b = 0.1 * u01(seed, idx, 6);
            H = 1.0 + 0.1 * (2.0 * u01(seed, idx, 0) - 1.0) + b;
            v1 = 0.1 * normal01(seed, idx, 2);
            v2 = 0.1 * normal01(seed, idx, 4);
        }
        const double h = H - b;
        u[0] = h;
        u[1] = h * v1;
        u[2] = h * v2;
        u[3] = b;
        return;
    }
    if (c.eq == OR_EQ_EULER_QUASI1D) {
        // Trixi's quasi-1D initial_condition_convergence_test (c = 2, A = 0.1,
        // L = 2, omega = 2 pi / L); seeded: rho = 1 + 0.1 U, p = 1 + 0.05 U,
        // v = 0.1 N(0, 1), a = 1 + 0.2 U; prim2cons (a rho, a rho v, a (p/(gamma-1)
        // + rho v^2 / 2), a)
        double rho = 1.0, v = 0.0, p = 1.0, a = 1.0;
        if (ic == OR_IC_CONVERGENCE_TEST) {
            const double cc = 2.0, A = 0.1, L = 2.0, f = 1.0 / L, omega = 2.0 * M_PI * f;
            const double ini = cc + A * sin(omega * (x[0] - t));
            rho = ini;
            v = 1.0;
            const double e = ini * ini / rho;
            p = (c.gamma - 1.0) * (e - 0.5 * rho * v * v);
            a = 1.5 - 0.5 * cos(x[0] * M_PI);
        } else if (ic == OR_IC_SEEDED_RANDOM) {
            rho = 1.0 + 0.1 * (2.0 * u01(seed, idx, 0) - 1.0);
            p = 1.0 + 0.05 * (2.0 * u01(seed, idx, 1) - 1.0);
            v = 0.1 * normal01(seed, idx, 2);
            a = 1.0 + 0.2 * (2.0 * u01(seed, idx, 4) - 1.0);
        }
        u[0] = a * rho;
        u[1] = a * rho * v;
        u[2] = a * (p * c.inv_gm1 + 0.5 * rho * v * v);
        u[3] = a;
        return;
    }
    if (c.eq == OR_EQ_EULER_MULTI) {
        // Trixi's multicomponent initial_condition_weak_blast_wave: the Euler
        // blast with rho_k = 2^(k-1) (1 - 2) rho / (1 - 2^nc) (prim2cons with
        // the totalgamma of the partial densities); the seeded state
        // rho_k = 2^(k-1) / (2^nc - 1) (1 + 0.1 U), v = 0.1 N(0, 1), p = 1 + 0.05 U.
        double v[3] = {0, 0, 0}, rho = 1.0, p = 1.0, prim[MAXV];
        if (ic == OR_IC_WEAK_BLAST_WAVE || ic == OR_IC_WEAK_BLAST_WAVE_SEEDED)
            weak_blast_prim(c, ic, x, seed, idx, rho, v, p);
        for (int k = 0; k < c.nc; ++k)
            prim[c.nd + 1 + k] = pow(2.0, k) * (1.0 - 2.0) * rho / (1.0 - pow(2.0, c.nc));
        if (ic == OR_IC_SEEDED_RANDOM) {
            for (int k = 0; k < c.nc; ++k) prim[c.nd + 1 + k] *= 1.0 + 0.1 * (2.0 * u01(seed, idx, 8 + k) - 1.0);
            p = 1.0 + 0.05 * (2.0 * u01(seed, idx, 1) - 1.0);
            for (int d = 0; d < c.nd; ++d) v[d] = 0.1 * normal01(seed, idx, 2 + 2 * d);
        }
        const double rho_t = multi_density(c, prim), gamma = multi_totalgamma(c, prim);
        double ke = 0.0;
        for (int d = 0; d < c.nd; ++d) {
            u[d] = rho_t * v[d];
            ke += u[d] * v[d];
        }
        u[c.nd] = p / (gamma - 1.0) + 0.5 * ke;
        for (int k = 0; k < c.nc; ++k) u[c.nd + 1 + k] = prim[c.nd + 1 + k];
        return;
    }
    if (c.eq == OR_EQ_EULER) {
        double v[3] = {0, 0, 0}, rho = 1.0, p = 1.0;
        if (ic == OR_IC_EOC_GRAVITY) {
            // initial_condition_eoc_test_coupled_euler_gravity: c = 2, A = 0.1,
            // G = 1; rho = c + A sin(pi (sum x - t)), v = 1, p = rho^2 G 2/(d pi)
            double s = -t;
            for (int d = 0; d < c.nd; ++d) s += x[d];
            const double ini = 2.0 + 0.1 * sin(M_PI * s);
            rho = ini;
            for (int d = 0; d < c.nd; ++d) v[d] = 1.0;
            p = ini * ini * 1.0 * 2.0 / (c.nd * M_PI);
            euler_prim2cons(c, rho, v, p, u);
            return;
        }
        if (ic == OR_IC_WEAK_BLAST_WAVE || ic == OR_IC_WEAK_BLAST_WAVE_SEEDED) {
            weak_blast_prim(c, ic, x, seed, idx, rho, v, p);
        } else if (ic == OR_IC_SEEDED_RANDOM) {
            rho = 1.0 + 0.1 * (2.0 * u01(seed, idx, 0) - 1.0);
            p = 1.0 + 0.05 * (2.0 * u01(seed, idx, 1) - 1.0);
            for (int d = 0; d < c.nd; ++d) v[d] = 0.1 * normal01(seed, idx, 2 + 2 * d);
        } else if (ic == OR_IC_SMOOTH_PERIODIC) {
            const double L = c.ic_length > 0 ? c.ic_length : 4.0;
            rho = 1.0 + 0.1 * smooth_field(seed, 0, c.nd, x, L, false);
            for (int d = 0; d < c.nd; ++d) v[d] = 0.1 * smooth_field(seed, 1 + d, c.nd, x, L, false);
            p = 1.0 + 0.05 * smooth_field(seed, 4, c.nd, x, L, true);
        } else if (ic == OR_IC_CONVERGENCE_TEST) {
            // Trixi initial_condition_convergence_test for Euler (density wave
            // form used with source terms): rho = 2 + 0.1 sin(pi (sum x - t)).
            double xs = 0.0;
            for (int d = 0; d < c.nd; ++d) xs += x[d];
            const double ini = 2.0 + 0.1 * sin(M_PI * (xs - t));
            double uu[5];
            uu[0] = ini;
            for (int d = 0; d < c.nd; ++d) uu[1 + d] = ini;
            uu[c.nd + 1] = ini * ini;
            for (int q = 0; q < c.nd + 2; ++q) u[q] = uu[q];
            return;
        }
        euler_prim2cons(c, rho, v, p, u);
        return;
    }
    // GLM-MHD
    double prim[9] = {1.0, 0, 0, 0, 1.0, 0, 0, 0, 0};
    if (ic == OR_IC_CONVERGENCE_TEST && c.nd == 1) {
        // Alfven wave of Trixi's IdealGlmMhdEquations1D convergence test:
        // rho = 1, v1 = 0, p = 0.1, B1 = 1, v_perp = B_perp = 0.1 (sin, cos)(2 pi (x + t))
        const double arg = 2.0 * M_PI * (x[0] + t);
        const double pr[9] = {1.0, 0.0, 0.1 * sin(arg), 0.1 * cos(arg), 0.1, 1.0, 0.1 * sin(arg), 0.1 * cos(arg), 0.0};
        memcpy(prim, pr, sizeof(pr));
    } else if (ic == OR_IC_CONVERGENCE_TEST && c.nd == 2) {
        // Alfven wave of Trixi's IdealGlmMhdEquations2D convergence test
        // (Derigs et al. 2016): rho = 1, p = 0.1, along x_perp = (x + y)/sqrt 2
        const double alpha = 0.25 * M_PI;
        const double x_perp = x[0] * cos(alpha) + x[1] * sin(alpha);
        const double B_perp = 0.1 * sin(2.0 * M_PI * (x_perp + t));
        const double v1 = -B_perp * sin(alpha), v2 = B_perp * cos(alpha);
        const double v3 = 0.1 * cos(2.0 * M_PI * (x_perp + t));
        const double pr[9] = {1.0, v1, v2, v3, 0.1, cos(alpha) + v1, sin(alpha) + v2, v3, 0.0};
        memcpy(prim, pr, sizeof(pr));
    } else if (ic == OR_IC_CONVERGENCE_TEST) {
        // Alfven wave (Trixi initial_condition_convergence_test, ideal GLM-MHD).
        const double p = 1.0, omega = 2.0 * M_PI, r = 2.0, e = 0.2;
        const double nx = 1.0 / sqrt(r * r + 1.0), ny = r / sqrt(r * r + 1.0), sqr = 1.0;
        const double Va = omega / (ny * sqr);
        const double phi_alv = omega / ny * (nx * (x[0] - 0.5 * r) + ny * (x[1] - 0.5 * r)) - Va * t;
        const double rho = 1.0;
        const double v1 = -e * ny * cos(phi_alv) / rho;
        const double v2 = e * nx * cos(phi_alv) / rho;
        const double v3 = e * sin(phi_alv) / rho;
        const double B1 = nx - rho * v1 * sqr, B2 = ny - rho * v2 * sqr, B3 = -rho * v3 * sqr;
        const double pr[9] = {rho, v1, v2, v3, p, B1, B2, B3, 0.0};
        memcpy(prim, pr, sizeof(pr));
    } else if (ic == OR_IC_WEAK_BLAST_WAVE || ic == OR_IC_WEAK_BLAST_WAVE_SEEDED) {
        // Trixi's MHD weak blast wave: the Euler blast (Hennemann & Gassner
        // 2020) with B = (1, 1, 1) and psi = 0 (via the Euler state of the
        // same dimension)
        Cfg ce = c;
        ce.eq = OR_EQ_EULER;
        ce.nv = c.nd + 2;
        double ue[5];
        initial_condition(ce, ic, x, t, seed, idx, ue);
        double ke = 0.0;
        prim[0] = ue[0];
        for (int d = 0; d < c.nd; ++d) {
            prim[1 + d] = ue[1 + d] / ue[0];
            ke += ue[1 + d] * prim[1 + d];
        }
        prim[4] = (c.gamma - 1.0) * (ue[c.nd + 1] - 0.5 * ke);
        prim[5] = prim[6] = prim[7] = 1.0;
        prim[8] = 0.0;
    } else if (ic == OR_IC_SEEDED_RANDOM) {
        prim[0] = 1.0 + 0.1 * (2.0 * u01(seed, idx, 0) - 1.0);
        prim[4] = 1.0 + 0.05 * (2.0 * u01(seed, idx, 1) - 1.0);
        for (int d = 0; d < 3; ++d) prim[1 + d] = 0.1 * normal01(seed, idx, 2 + 2 * d);
        for (int d = 0; d < 3; ++d) prim[5 + d] = 0.5 + 0.1 * (2.0 * u01(seed, idx, 8 + d) - 1.0);
        prim[8] = 0.01 * (2.0 * u01(seed, idx, 11) - 1.0);
    }
    mhd_prim2cons(c, prim, u);
}

// Entropy and entropy variables (Trixi entropy / cons2entropy).
template <class T> T entropy_math(const Cfg& c, const T* u) {
    if (c.eq == OR_EQ_LINADV || c.eq == OR_EQ_BURGERS) return 0.5 * u[0] * u[0];
    T rho, v[3], p;
    if (c.eq == OR_EQ_EULER) {
        euler_prim(c, u, rho, v, p);
    } else {
        T B[3], psi;
        mhd_prim(c, u, rho, v, p, B, psi);
    }
    T s = tlog(p) - c.gamma * tlog(rho);
    return -rho * s * c.inv_gm1;
}

}  // namespace

// ============================================================================
// extern "C" surface
// ============================================================================
extern "C" {

int oracle_nvars(const oracle_cfg* cfg) {
    if (cfg->equation == OR_EQ_LINADV || cfg->equation == OR_EQ_BURGERS) return 1;
    if (cfg->equation == OR_EQ_HYPDIFF) return cfg->ndims + 1;
    if (cfg->equation == OR_EQ_EULER) return cfg->ndims + 2;
    if (cfg->equation == OR_EQ_EULER_MULTI) return cfg->ndims + 1 + cfg->ncomponents;
    if (cfg->equation == OR_EQ_EULER_QUASI1D) return 4;
    if (cfg->equation == OR_EQ_SHALLOW_WATER) return cfg->ndims + 2;
    return cfg->ndims == 1 ? 8 : 9;  // ideal MHD: (rho, rho v, rho e, B) in 1D, + psi in 2D/3D
}

int oracle_basis(int N, double* nodes, double* weights, double* D, double* Dsplit, double* Dhat,
                 double* lhat_minus, double* lhat_plus, double* Vinv) {
    if (N < 2 || N > 16) return -1;
    double x[16], w[16], wb[16];
    gauss_lobatto_nodes_weights(N, x, w);
    barycentric_weights(N, x, wb);
    if (nodes) memcpy(nodes, x, sizeof(double) * N);
    if (weights) memcpy(weights, w, sizeof(double) * N);
    double Dm[256];
    // polynomial_derivative_matrix (barycentric form, negative-sum diagonal)
    for (int i = 0; i < N; ++i) {
        Dm[i * N + i] = 0.0;
        for (int j = 0; j < N; ++j)
            if (j != i) {
                Dm[i * N + j] = wb[j] / wb[i] / (x[i] - x[j]);
                Dm[i * N + i] -= Dm[i * N + j];
            }
    }
    if (D) memcpy(D, Dm, sizeof(double) * N * N);
    if (Dsplit) {
        for (int q = 0; q < N * N; ++q) Dsplit[q] = 2.0 * Dm[q];
        Dsplit[0] += 1.0 / w[0];
        Dsplit[N * N - 1] -= 1.0 / w[N - 1];
    }
    if (Dhat)
        for (int n = 0; n < N; ++n)
            for (int j = 0; j < N; ++j) Dhat[j * N + n] = Dm[n * N + j] * (-w[n] / w[j]);
    if (lhat_minus) {
        lagrange_polys(-1.0, N, x, wb, lhat_minus);
        for (int j = 0; j < N; ++j) lhat_minus[j] /= w[j];
    }
    if (lhat_plus) {
        lagrange_polys(1.0, N, x, wb, lhat_plus);
        for (int j = 0; j < N; ++j) lhat_plus[j] /= w[j];
    }
    if (Vinv) {
        double V[256];
        for (int i = 0; i < N; ++i)
            for (int m = 0; m < N; ++m) V[i * N + m] = legendre_normalized(m, x[i]);
        if (!invert_matrix(N, V, Vinv)) return -2;
    }
    return 0;
}

int oracle_interpolation_matrix(int nin, const double* xin, int nout, const double* xout,
                                double* Vdm) {
    double wb[32];
    if (nin > 32) return -1;
    barycentric_weights(nin, xin, wb);
    for (int k = 0; k < nout; ++k) {
        bool match = false;
        for (int j = 0; j < nin; ++j) Vdm[k * nin + j] = 0.0;
        for (int j = 0; j < nin; ++j)
            if (fabs(xout[k] - xin[j]) <= 1.4901161193847656e-08 * std::max(fabs(xout[k]), fabs(xin[j]))) {
                match = true;
                Vdm[k * nin + j] = 1.0;
            }
        if (!match) {
            double s = 0.0;
            for (int j = 0; j < nin; ++j) {
                double t = wb[j] / (xout[k] - xin[j]);
                Vdm[k * nin + j] = t;
                s += t;
            }
            for (int j = 0; j < nin; ++j) Vdm[k * nin + j] /= s;
        }
    }
    return 0;
}

double oracle_ln_mean(double x, double y) { return ln_mean(x, y); }
double oracle_inv_ln_mean(double x, double y) { return inv_ln_mean(x, y); }

void oracle_flux(const oracle_cfg* cfg, const double* u, int o, double* f) {
    phys_flux(make_cfg(cfg), u, o, f);
}
void oracle_two_point_flux(const oracle_cfg* cfg, int kind, const double* ull, const double* urr,
                           int o, double* f) {
    two_point(make_cfg(cfg), kind, ull, urr, o, f);
}
void oracle_noncons_flux(const oracle_cfg* cfg, const double* ull, const double* urr, int o,
                         double* f) {
    const Cfg c = make_cfg(cfg);
    noncons_flux(c, ull, urr, o, f, c.sflux);
}

void oracle_max_abs_speeds(const oracle_cfg* cfg, const double* u, double* lam) {
    Cfg c = make_cfg(cfg);
    for (int d = 0; d < c.nd; ++d) {
        if (c.eq == OR_EQ_LINADV) {
            lam[d] = fabs(c.a[d]);
        } else if (c.eq == OR_EQ_BURGERS) {
            lam[d] = fabs(u[0]);
        } else if (c.eq == OR_EQ_HYPDIFF) {
            lam[d] = sqrt(c.nu * c.inv_Tr);
        } else if (c.eq == OR_EQ_EULER) {
            double rho, v[3], p;
            euler_prim(c, u, rho, v, p);
            lam[d] = fabs(v[d]) + sqrt(c.gamma * p / rho);
        } else if (c.eq == OR_EQ_EULER_MULTI) {
            double rho, v[3], p, gamma;
            multi_prim(c, u, rho, v, p, gamma);
            lam[d] = fabs(v[d]) + sqrt(gamma * p / rho);
        } else if (c.eq == OR_EQ_EULER_QUASI1D) {
            double rho, v, p, a;
            quasi_prim(c, u, rho, v, p, a);
            lam[d] = fabs(v) + sqrt(c.gamma * p / rho);
        } else if (c.eq == OR_EQ_SHALLOW_WATER) {
            lam[d] = fabs(u[1 + d] / u[0]) + sqrt(c.grav * u[0]);  // any dimension
        } else {
            lam[d] = fabs(u[1 + d] / u[0]) + mhd_fast_wavespeed(c, u, d + 1);
        }
    }
}

double oracle_entropy(const oracle_cfg* cfg, const double* u) {
    return entropy_math(make_cfg(cfg), u);
}

void oracle_cons2entropy(const oracle_cfg* cfg, const double* u, double* w) {
    Cfg c = make_cfg(cfg);
    if (c.eq == OR_EQ_LINADV || c.eq == OR_EQ_BURGERS) {
        w[0] = u[0];
        return;
    }
    double rho, v[3] = {0, 0, 0}, p, B[3] = {0, 0, 0}, psi = 0;
    if (c.eq == OR_EQ_EULER)
        euler_prim(c, u, rho, v, p);
    else
        mhd_prim(c, u, rho, v, p, B, psi);
    const int nvel = (c.eq == OR_EQ_EULER) ? c.nd : 3;
    double v_square = 0.0;
    for (int d = 0; d < nvel; ++d) v_square += v[d] * v[d];
    double s = log(p) - c.gamma * log(rho);
    double rho_p = rho / p;
    w[0] = (c.gamma - s) * c.inv_gm1 - 0.5 * rho_p * v_square;
    for (int d = 0; d < nvel; ++d) w[1 + d] = rho_p * v[d];
    w[nvel + 1] = -rho_p;
    if (c.eq == OR_EQ_GLM_MHD) {
        for (int d = 0; d < 3; ++d) w[5 + d] = rho_p * B[d];
        if (c.nv > 8) w[8] = rho_p * psi;
    }
}

// Entropy flux potential psi_o with [[w]].f* = [[psi_o]] for EC fluxes:
// Euler: rho v_o; LinAdv (U = u^2/2): a_o u^2 / 2; Burgers (U = u^2/2,
// F = u^3/3): w f - F = u^3 / 6.
void oracle_entropy_potential(const oracle_cfg* cfg, const double* u, int o, double* psi) {
    Cfg c = make_cfg(cfg);
    if (c.eq == OR_EQ_BURGERS) {
        *psi = u[0] * u[0] * u[0] / 6.0;
    } else if (c.eq == OR_EQ_LINADV) {
        *psi = 0.5 * c.a[o - 1] * u[0] * u[0];
    } else {
        *psi = u[o];
    }
}

int64_t oracle_mesh_nelements(int ndims, int level) { return (int64_t)1 << (ndims * level); }

// Uniform periodic TreeMesh: cells are created level by level, each refined
// cell appending its 2^d children (child c has offset bit d set iff
// ((c-1) >> d) & 1), so the leaf order is the Morton order with the coarsest
// level most significant.  Elements are the leaves in cell order; interfaces
// are created per element for the +x, +y, +z neighbours in that order
// (Trixi TreeMesh + init_elements!/init_interfaces!).
int64_t oracle_mesh_ninterfaces(int nd, int level, int periodic) {
    const int64_t E = oracle_mesh_nelements(nd, level);
    // per direction: E interfaces when periodic (bit d), else the faces on
    // the + boundary have none
    int64_t I = 0;
    for (int d = 0; d < nd; ++d) I += ((periodic >> d) & 1) ? E : E - (E >> level);
    return I;
}

int oracle_mesh_uniform(int nd, int level, double cmin, double cmax, int polydeg,
                        int32_t* neighbor_ids, int32_t* orientations, double* invJ,
                        double* node_coordinates, int32_t* cell_index) {
    return oracle_mesh(nd, level, cmin, cmax, polydeg, (1 << nd) - 1, neighbor_ids, orientations, invJ,
                       node_coordinates, cell_index);
}

int oracle_mesh(int nd, int level, double cmin, double cmax, int polydeg, int periodic,
                int32_t* neighbor_ids, int32_t* orientations, double* invJ,
                double* node_coordinates, int32_t* cell_index) {
    if (nd < 1 || nd > 3 || level < 0 || level > 10) return -1;
    const int64_t E = oracle_mesh_nelements(nd, level);
    const int64_t n1 = (int64_t)1 << level;
    // Build the leaf integer coordinates by literal refinement.
    std::vector<std::vector<int64_t>> cells(1, std::vector<int64_t>(nd, 0));
    for (int l = 0; l < level; ++l) {
        std::vector<std::vector<int64_t>> next;
        next.reserve(cells.size() << nd);
        for (auto& cc : cells)
            for (int ch = 0; ch < (1 << nd); ++ch) {
                std::vector<int64_t> k(nd);
                for (int d = 0; d < nd; ++d) k[d] = 2 * cc[d] + ((ch >> d) & 1);
                next.push_back(k);
            }
        cells.swap(next);
    }
    // Lookup from integer coords to element id.
    std::vector<int32_t> lut((size_t)E);
    for (int64_t e = 0; e < E; ++e) {
        int64_t key = 0;
        for (int d = nd - 1; d >= 0; --d) key = key * n1 + cells[e][d];
        lut[key] = (int32_t)e;
    }
    const double length = cmax - cmin;
    const double dx = length / (double)n1;
    const double jac = dx / 2.0;
    double nodes[16], w[16];
    const int N = polydeg + 1;
    gauss_lobatto_nodes_weights(N, nodes, w);
    const int nn = ipow(N, nd);
    int64_t count = 0;
    for (int64_t e = 0; e < E; ++e) {
        if (cell_index)
            for (int d = 0; d < nd; ++d) cell_index[e * nd + d] = (int32_t)cells[e][d];
        if (invJ) invJ[e] = 1.0 / jac;
        if (node_coordinates) {
            double center[3];
            for (int d = 0; d < nd; ++d) {
                // Cell centre by halving from the root centre, as the tree does.
                double cc = 0.5 * (cmin + cmax), len = length;
                for (int l = level - 1; l >= 0; --l) {
                    const int bit = (int)((cells[e][d] >> l) & 1);
                    len *= 0.5;
                    cc += (bit ? 0.5 : -0.5) * len;
                }
                center[d] = cc;
            }
            for (int n = 0; n < nn; ++n)
                for (int d = 0; d < nd; ++d) {
                    const int id = (n / ipow(N, d)) % N;
                    node_coordinates[((size_t)e * nn + n) * nd + d] = center[d] + jac * nodes[id];
                }
        }
        for (int d = 0; d < nd; ++d) {
            if (!((periodic >> d) & 1) && cells[e][d] + 1 == n1) continue;  // boundary face
            std::vector<int64_t> k = cells[e];
            k[d] = (k[d] + 1) % n1;
            int64_t key = 0;
            for (int q = nd - 1; q >= 0; --q) key = key * n1 + k[q];
            if (neighbor_ids) {
                neighbor_ids[2 * count] = (int32_t)e;
                neighbor_ids[2 * count + 1] = lut[key];
            }
            if (orientations) orientations[count] = d + 1;
            ++count;
        }
    }
    return 0;
}

int oracle_mortar_matrices(int N, double* fl, double* fu, double* rl, double* ru) {
    if (N < 2 || N > 16) return -1;
    mortar_matrices(N, fl, fu, rl, ru);
    return 0;
}

void oracle_source_convergence_test(const oracle_cfg* cfg, const double* x, double t, double* s) {
    source_convergence_test(make_cfg(cfg), x, t, s);
}

void oracle_initial_condition(const oracle_cfg* cfg, int ic, const double* x, double t,
                              uint64_t seed, uint64_t node_index, double* u) {
    Cfg c = make_cfg(cfg);
    initial_condition(c, ic, x, t, seed, node_index, u);
}

// compute_coefficients (reference dg.jl:46-160, including the 1D nudge of the
// element end nodes with nextfloat/prevfloat at dg.jl:54-58).
void oracle_compute_coefficients(const oracle_cfg* cfg, int ic, int64_t E, const double* xc,
                                 double t, uint64_t seed, double* u) {
    Cfg c = make_cfg(cfg);
    const int nn = ipow(c.N, c.nd);
    for (int64_t e = 0; e < E; ++e)
        for (int n = 0; n < nn; ++n) {
            double x[3];
            for (int d = 0; d < c.nd; ++d) x[d] = xc[((size_t)e * nn + n) * c.nd + d];
            if (c.nd == 1) {
                if (n == 0) x[0] = nextafter(x[0], INFINITY);
                if (n == c.N - 1) x[0] = nextafter(x[0], -INFINITY);
            }
            initial_condition(c, ic, x, t, seed, (uint64_t)(e * nn + n),
                              u + ((size_t)e * nn + n) * c.nv);
        }
}

void oracle_rhs(const oracle_cfg* cfg, int64_t E, int64_t I, const int32_t* nbr,
                const int32_t* ori, const double* invJ, const double* u, double* du,
                double* alpha_out) {
    rhs_impl<double>(make_cfg(cfg), E, I, nbr, ori, invJ, u, du, alpha_out);
}

void oracle_indicator_hg(const oracle_cfg* cfg, int64_t E, int64_t I, const int32_t* nbr,
                         const double* u, double* alpha) {
    Cfg c = make_cfg(cfg);
    Basis b = make_basis(c.N);
    indicator_hg<double>(c, b, E, I, nbr, u, alpha);
}

void oracle_rhs_jvp(const oracle_cfg* cfg, int64_t E, int64_t I, const int32_t* nbr,
                    const int32_t* ori, const double* invJ, const double* u, const double* udot,
                    double* du, double* dudot) {
    Cfg c = make_cfg(cfg);
    const size_t n = (size_t)E * ipow(c.N, c.nd) * c.nv;
    std::vector<Dual> ud(n), dud(n);
    for (size_t q = 0; q < n; ++q) ud[q] = Dual(u[q], udot[q]);
    rhs_impl<Dual>(c, E, I, nbr, ori, invJ, ud.data(), dud.data(), (Dual*)nullptr);
    for (size_t q = 0; q < n; ++q) {
        if (du) du[q] = dud[q].v;
        dudot[q] = dud[q].d;
    }
}

// max_dt (reference stepsize_dg_{1,2,3}d.jl): 2 / (N * max_e invJ_e * sum_d lambda_d).
double oracle_max_dt(const oracle_cfg* cfg, int64_t E, const double* invJ, const double* u) {
    Cfg c = make_cfg(cfg);
    const int nn = ipow(c.N, c.nd);
    double max_scaled_speed = nextafter(0.0, 1.0);
    for (int64_t e = 0; e < E; ++e) {
        double lam_max[3] = {0, 0, 0};
        for (int n = 0; n < nn; ++n) {
            double lam[3];
            oracle_max_abs_speeds(cfg, u + ((size_t)e * nn + n) * c.nv, lam);
            for (int d = 0; d < c.nd; ++d) lam_max[d] = std::max(lam_max[d], lam[d]);
        }
        double s = 0.0;
        for (int d = 0; d < c.nd; ++d) s += lam_max[d];
        max_scaled_speed = std::max(max_scaled_speed, invJ[e] * s);
    }
    return 2.0 / (c.N * max_scaled_speed);
}

// calc_error_norms (reference analysis_dg_{1,2,3}d.jl, e.g. 3D :45-89):
// interpolate u and x to the 2*polydeg analysis nodes, then
// L2 = sqrt(sum w J diff^2 / |Omega|), Linf = max |diff|.
void oracle_error_norms(const oracle_cfg* cfg, int ic, int64_t E, const double* invJ,
                        const double* xc, const double* u, double t, double total_volume,
                        uint64_t seed, double* l2, double* linf) {
    Cfg c = make_cfg(cfg);
    const int N = c.N, nd = c.nd, nv = c.nv, nn = ipow(N, nd);
    const int NA = 2 * (N - 1) + 1, nna = ipow(NA, nd);
    double xa[32], wa[32], xb[16], wbb[16], Vdm[32 * 16];
    gauss_lobatto_nodes_weights(NA, xa, wa);
    gauss_lobatto_nodes_weights(N, xb, wbb);
    oracle_interpolation_matrix(N, xb, NA, xa, Vdm);
    for (int v = 0; v < nv; ++v) {
        l2[v] = 0.0;
        linf[v] = 0.0;
    }
    std::vector<double> uin(nn * nv), xin(nn * nd), ua, xa_loc;
    auto interp = [&](const std::vector<double>& in, int ncomp, std::vector<double>& out) {
        // x, then y, then z (multiply_dimensionwise!)
        std::vector<double> cur = in;
        int nx[3] = {N, N, N};
        for (int d = nd; d < 3; ++d) nx[d] = 1;
        for (int d = 0; d < nd; ++d) {
            int ny[3] = {nx[0], nx[1], nx[2]};
            ny[d] = NA;
            std::vector<double> nxt((size_t)ny[0] * ny[1] * ny[2] * ncomp, 0.0);
            for (int k = 0; k < ny[2]; ++k)
                for (int j = 0; j < ny[1]; ++j)
                    for (int i = 0; i < ny[0]; ++i) {
                        int o3[3] = {i, j, k};
                        for (int q = 0; q < ncomp; ++q) {
                            double r = 0.0;
                            for (int ii = 0; ii < nx[d]; ++ii) {
                                int s3[3] = {i, j, k};
                                s3[d] = ii;
                                r += Vdm[o3[d] * N + ii] *
                                     cur[((size_t)(s3[0] + nx[0] * (s3[1] + nx[1] * s3[2]))) * ncomp + q];
                            }
                            nxt[((size_t)(i + ny[0] * (j + ny[1] * k))) * ncomp + q] = r;
                        }
                    }
            cur.swap(nxt);
            nx[d] = NA;
        }
        out.swap(cur);
    };
    for (int64_t e = 0; e < E; ++e) {
        for (int q = 0; q < nn * nv; ++q) uin[q] = u[(size_t)e * nn * nv + q];
        for (int q = 0; q < nn * nd; ++q) xin[q] = xc[(size_t)e * nn * nd + q];
        interp(uin, nv, ua);
        interp(xin, nd, xa_loc);
        const double vj = pow(1.0 / invJ[e], nd);
        for (int n = 0; n < nna; ++n) {
            double uex[MAXV];
            initial_condition(c, ic, &xa_loc[(size_t)n * nd], t, seed, 0, uex);
            double w = 1.0;
            for (int d = 0; d < nd; ++d) w *= wa[(n / ipow(NA, d)) % NA];
            for (int v = 0; v < nv; ++v) {
                double diff = uex[v] - ua[(size_t)n * nv + v];
                l2[v] += diff * diff * (w * vj);
                linf[v] = std::max(linf[v], fabs(diff));
            }
        }
    }
    for (int v = 0; v < nv; ++v) l2[v] = sqrt(l2[v] / total_volume);
}

void oracle_integrate(const oracle_cfg* cfg, int64_t E, const double* invJ, const double* u,
                      double* integral) {
    Cfg c = make_cfg(cfg);
    const int N = c.N, nd = c.nd, nv = c.nv, nn = ipow(N, nd);
    double x[16], w[16];
    gauss_lobatto_nodes_weights(N, x, w);
    for (int v = 0; v < nv; ++v) integral[v] = 0.0;
    for (int64_t e = 0; e < E; ++e) {
        const double vj = pow(1.0 / invJ[e], nd);
        for (int n = 0; n < nn; ++n) {
            double wt = 1.0;
            for (int d = 0; d < nd; ++d) wt *= w[(n / ipow(N, d)) % N];
            for (int v = 0; v < nv; ++v) integral[v] += vj * wt * u[((size_t)e * nn + n) * nv + v];
        }
    }
}

// Carpenter & Kennedy (1994) five-stage fourth-order 2N-storage scheme
// (OrdinaryDiffEq CarpenterKennedy2N54, used by the reference examples, e.g.
// examples/euler_ec_3d.jl:53-55).
void oracle_ck2n54_coefficients(double* A, double* B, double* c) {
    const double a[5] = {0.0, -567301805773.0 / 1357537059087.0,
                         -2404267990393.0 / 2016746695238.0,
                         -3550918686646.0 / 2091501179385.0,
                         -1275806237668.0 / 842570457699.0};
    const double b[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                         1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                         2277821191437.0 / 14882151754819.0};
    const double cc[5] = {0.0, 1432997174477.0 / 9575080441755.0,
                          2526269341429.0 / 6820363962896.0, 2006345519317.0 / 3224310063776.0,
                          2802321613138.0 / 2924317926251.0};
    for (int s = 0; s < 5; ++s) {
        A[s] = a[s];
        B[s] = b[s];
        c[s] = cc[s];
    }
}

void oracle_ck2n54_step(const oracle_cfg* cfg, int64_t E, int64_t I, const int32_t* nbr,
                        const int32_t* ori, const double* invJ, double* u, double* k, double* du,
                        double t, double dt) {
    double A[5], B[5], C[5];
    oracle_ck2n54_coefficients(A, B, C);
    Cfg c = make_cfg(cfg);
    const size_t n = (size_t)E * ipow(c.N, c.nd) * c.nv;
    for (int s = 0; s < 5; ++s) {
        c.t = t + C[s] * dt;
        rhs_impl<double>(c, E, I, nbr, ori, invJ, u, du, (double*)nullptr);
        for (size_t q = 0; q < n; ++q) {
            k[q] = (s == 0) ? dt * du[q] : A[s] * k[q] + dt * du[q];
            u[q] = u[q] + B[s] * k[q];
        }
    }
}

// Forward-mode derivative of one CK2N54 step: (u, k) -> (u_next) with
// tangent udot; k starts at zero as in the low-storage scheme.
void oracle_ck2n54_step_jvp(const oracle_cfg* cfg, int64_t E, int64_t I, const int32_t* nbr,
                            const int32_t* ori, const double* invJ, const double* u, const double* udot,
                            double dt, double* u_next, double* u_next_dot) {
    double A[5], B[5], C[5];
    oracle_ck2n54_coefficients(A, B, C);
    Cfg c = make_cfg(cfg);
    const size_t n = (size_t)E * ipow(c.N, c.nd) * c.nv;
    std::vector<Dual> uu(n), k(n), du(n);
    for (size_t q = 0; q < n; ++q) uu[q] = Dual(u[q], udot[q]);
    const double t0 = c.t;  // start of the step
    for (int s = 0; s < 5; ++s) {
        c.t = t0 + C[s] * dt;
        rhs_impl<Dual>(c, E, I, nbr, ori, invJ, uu.data(), du.data(), (Dual*)nullptr);
        for (size_t q = 0; q < n; ++q) {
            k[q] = (s == 0) ? dt * du[q] : A[s] * k[q] + dt * du[q];
            uu[q] = uu[q] + B[s] * k[q];
        }
    }
    for (size_t q = 0; q < n; ++q) {
        u_next[q] = uu[q].v;
        u_next_dot[q] = uu[q].d;
    }
}

int oracle_set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
    return omp_get_max_threads();
#else
    (void)nthreads;
    return 1;
#endif
}

}  // extern "C"
