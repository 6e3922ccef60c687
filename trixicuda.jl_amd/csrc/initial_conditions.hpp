// initial_conditions.hpp -- initial conditions on the device: the Trixi
// initial_condition_* functions used by the reference tests and examples
// (also the Dirichlet boundary states), plus the seeded benchmark states.
#pragma once

#include "../../include/trixi_hip.h"
#include "physics.hpp"

namespace trixi_hip {

// ---------------------------------------------------------------------------
// counter-based generator for the seeded initial states (splitmix64)
// ---------------------------------------------------------------------------
__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__host__ __device__ inline double u01(uint64_t seed, uint64_t index, uint64_t stream) {
    const uint64_t r = splitmix64(splitmix64(seed) ^ (index * 16 + stream));
    return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
__host__ __device__ inline double normal01(uint64_t seed, uint64_t index, uint64_t stream) {
    const double a = u01(seed, index, stream), b = u01(seed, index, stream + 1);
    return sqrt(-2.0 * log(a)) * cos(2.0 * M_PI * b);
}

struct IcParams {
    int dim, eq, nv, ic;
    double gamma, inv_gm1, c_h, a[3], length;
    uint64_t seed;
    int nc;                          // multicomponent Euler: components
    double cv[MAXC], cvg[MAXC];      // c_v,k and c_v,k gamma_k
    double grav;                     // shallow water: gravitational acceleration
};

// Shallow water 2D states (h, h v1, h v2, b) from (H = h + b, v1, v2, b):
//   convergence test (Trixi initial_condition_convergence_test, periodic on
//   [0, sqrt 2]^2): H = 7 + cos(w x) sin(w y) cos(2 pi t), w = 2 pi sqrt 2,
//   v = (0.5, 1.5), b = 2 + 0.5 sin(sqrt 2 pi x) + 0.5 sin(sqrt 2 pi y);
//   weak blast wave: centre (0.7, 0.7), H = 4 inside r <= 0.5 with the outward
//   velocity 0.1882, else H = 3.25 at rest, b = 0;
//   seeded: h = 1 + 0.1 U, v = 0.1 N(0, 1), b = 0.1 U; constant: h = 1, v = 0, b = 0.
__host__ __device__ inline void shallow_water_state(const IcParams& q, const double* x, double t, uint64_t idx,
                                                    double* u) {
    double H = 1.0, v1 = 0.0, v2 = 0.0, b = 0.0;
    if (q.ic == TRIXI_HIP_IC_CONVERGENCE_TEST) {
        const double wx = 2.0 * M_PI * sqrt(2.0), wt = 2.0 * M_PI;
        H = 7.0 + cos(wx * x[0]) * sin(wx * x[1]) * cos(wt * t);
        v1 = 0.5;
        v2 = 1.5;
        b = 2.0 + 0.5 * sin(sqrt(2.0) * M_PI * x[0]) + 0.5 * sin(sqrt(2.0) * M_PI * x[1]);
    } else if (q.ic == TRIXI_HIP_IC_WEAK_BLAST_WAVE) {
        const double xn = x[0] - 0.7, yn = x[1] - 0.7;
        const double r = sqrt(xn * xn + yn * yn), phi = atan2(yn, xn);
        const bool inside = !(r > 0.5);
        H = inside ? 4.0 : 3.25;
        v1 = inside ? 0.1882 * cos(phi) : 0.0;
        v2 = inside ? 0.1882 * sin(phi) : 0.0;
    } else if (q.ic == TRIXI_HIP_IC_SEEDED_RANDOM) {
        b = 0.1 * u01(q.seed, idx, 6);
        H = 1.0 + 0.1 * (2.0 * u01(q.seed, idx, 0) - 1.0) + b;
        v1 = 0.1 * normal01(q.seed, idx, 2);
        v2 = 0.1 * normal01(q.seed, idx, 4);
    }
    const double h = H - b;
    u[0] = h;
    u[1] = h * v1;
    u[2] = h * v2;
    u[3] = b;
}

// source_terms_convergence_test of ShallowWaterEquations2D: the residual of
// the convergence-test state, with H_t = -2 pi cos(w x) sin(w y) sin(2 pi t),
// H_x, H_y its gradient and b_x, b_y the bottom slope (v is constant):
//   S_h = H_t + v1 (H_x - b_x) + v2 (H_y - b_y),  S_{h v_i} = v_i S_h + g (H - b) H_{x_i}.
__host__ __device__ inline void source_shallow_water_convergence_test(double g, const double* x, double t,
                                                                      double* s) {
    const double wx = 2.0 * M_PI * sqrt(2.0), wt = 2.0 * M_PI, wb = sqrt(2.0) * M_PI;
    const double v1 = 0.5, v2 = 1.5;
    const double sX = sin(wx * x[0]), cX = cos(wx * x[0]), sY = sin(wx * x[1]), cY = cos(wx * x[1]);
    const double sT = sin(wt * t), cT = cos(wt * t);
    const double H = 7.0 + cX * sY * cT, H_x = -wx * sX * sY * cT, H_y = wx * cX * cY * cT;
    const double H_t = -wt * cX * sY * sT;
    const double b = 2.0 + 0.5 * sin(wb * x[0]) + 0.5 * sin(wb * x[1]);
    const double b_x = 0.5 * wb * cos(wb * x[0]), b_y = 0.5 * wb * cos(wb * x[1]);
    s[0] = H_t + v1 * (H_x - b_x) + v2 * (H_y - b_y);
    s[1] = v1 * s[0] + g * (H - b) * H_x;
    s[2] = v2 * s[0] + g * (H - b) * H_y;
    s[3] = 0.0;
}

__device__ inline double smooth_field(uint64_t seed, int field, int nd, const double* x, double L,
                               bool use_cos) {
    double s = 0.0;
    for (int m = 0; m < 3; ++m) {
        const uint64_t base = 0xFFFFFFFF00000000ULL + (uint64_t)(field * 3 + m) * 16;
        double arg = 0.0;
        for (int d = 0; d < nd; ++d) {
            const int k = (int)floor(u01(seed, base, d) * 7.0) - 3;
            arg += k * x[d];
        }
        const double amp = (2.0 * u01(seed, base, 3) - 1.0) / 3.0;
        const double ph = 2.0 * M_PI * u01(seed, base, 4);
        const double th = 2.0 * M_PI * arg / L + ph;
        s += amp * (use_cos ? cos(th) : sin(th));
    }
    return s;
}

// Hyperbolic diffusion states (phi, q_1..d): initial_condition_poisson_nonperiodic
// is the steady solution of -lap phi = f with q = grad phi,
//   1D: phi = sin(pi x);  2D: 2 cos(pi x) sin(2 pi y) + 2;
//   3D: 2 cos(pi x) sin(2 pi y) sin(2 pi z) + 2,
// and at t = 0 the initial guess phi = q = 1; as a boundary state (steady =
// true) it is always the steady solution.  The harmonic state of the 1D test
// is phi = 5, q = 0 at t = 0, else phi = 3 + e x, q = e.
__host__ __device__ inline void hypdiff_state(const IcParams& q, const double* x, double t, bool steady, double* u) {
    const int d = q.dim;
    if (q.ic == TRIXI_HIP_IC_POISSON_NONPERIODIC) {
        if (t == 0.0 && !steady) {
            for (int k = 0; k <= d; ++k) u[k] = 1.0;
            return;
        }
        if (d == 1) {
            u[0] = sin(M_PI * x[0]);
            u[1] = M_PI * cos(M_PI * x[0]);
        } else {
            const double sx = sin(M_PI * x[0]), cx = cos(M_PI * x[0]);
            const double sy = sin(2.0 * M_PI * x[1]), cy = cos(2.0 * M_PI * x[1]);
            const double sz = d == 3 ? sin(2.0 * M_PI * x[2]) : 1.0, cz = d == 3 ? cos(2.0 * M_PI * x[2]) : 0.0;
            u[0] = 2.0 * cx * sy * sz + 2.0;
            u[1] = -2.0 * M_PI * sx * sy * sz;
            u[2] = 4.0 * M_PI * cx * cy * sz;
            if (d == 3) u[3] = 4.0 * M_PI * cx * sy * cz;
        }
        return;
    }
    if (q.ic == TRIXI_HIP_IC_HARMONIC_NONPERIODIC) {
        const double A = 3.0, B = exp(1.0);
        u[0] = (t == 0.0) ? 5.0 : A + B * x[0];
        for (int k = 1; k <= d; ++k) u[k] = (t == 0.0 || k > 1) ? 0.0 : B;
        return;
    }
    u[0] = 1.0;
    for (int k = 1; k <= d; ++k) u[k] = 0.0;
}

// Weak blast wave (Hennemann & Gassner 2020, Sec. 6.3): inside r <= 0.5
// rho = 1.1691, p = 1.245 and an outward velocity of 0.1882; outside at rest
// with rho = p = 1; the seeded variant adds U(-0.01, 0.01) per node.
__device__ inline void weak_blast_prim(const IcParams& q, const double* x, uint64_t idx, double& rho, double* v,
                                       double& p) {
    double r2 = 0.0;
    for (int d = 0; d < q.dim; ++d) r2 += x[d] * x[d];
    const double r = sqrt(r2);
    const bool inside = !(r > 0.5);
    if (q.dim == 1) {
        const double phi = x[0] > 0 ? 0.0 : M_PI;
        v[0] = inside ? 0.1882 * cos(phi) : 0.0;
    } else if (q.dim == 2) {
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
    if (q.ic == TRIXI_HIP_IC_WEAK_BLAST_WAVE_SEEDED) {
        // breaks the exact symmetries (v = 0, equal neighbour states)
        rho += 0.01 * (2.0 * u01(q.seed, idx, 0) - 1.0);
        p += 0.01 * (2.0 * u01(q.seed, idx, 1) - 1.0);
        for (int d = 0; d < q.dim; ++d) v[d] += 0.01 * (2.0 * u01(q.seed, idx, 2 + d) - 1.0);
    }
}

// Initial conditions (Trixi initial_condition_* functions used by the
// reference tests and examples, plus the seeded benchmark states).
__device__ inline void initial_condition(const IcParams& q, const double* x, double t, uint64_t idx,
                                  double* u) {
    if (q.eq == EQ_HYPDIFF) {
        if (q.ic == TRIXI_HIP_IC_SEEDED_RANDOM) {
            for (int k = 0; k <= q.dim; ++k) u[k] = 1.0 + 0.5 * (2.0 * u01(q.seed, idx, k) - 1.0);
            return;
        }
        hypdiff_state(q, x, t, false, u);
        return;
    }
    if (q.eq == EQ_BURGERS) {
        if (q.ic == TRIXI_HIP_IC_CONVERGENCE_TEST) {
            // c + A sin(omega (x - t)), c = 2, A = 1, omega = 2 pi
            u[0] = 2.0 + sin(2.0 * M_PI * (x[0] - t));
        } else if (q.ic == TRIXI_HIP_IC_SHOCK) {
            u[0] = x[0] < 0.5 ? 1.5 : 0.5;
        } else if (q.ic == TRIXI_HIP_IC_RAREFACTION) {
            u[0] = x[0] < 0.5 ? 0.5 : 1.5;
        } else if (q.ic == TRIXI_HIP_IC_SEEDED_RANDOM) {
            u[0] = 1.0 + 0.5 * (2.0 * u01(q.seed, idx, 0) - 1.0);
        } else {
            u[0] = 2.0;
        }
        return;
    }
    if (q.eq == EQ_SHALLOW_WATER) {
        shallow_water_state(q, x, t, idx, u);
        return;
    }
    if (q.eq == EQ_EULER_QUASI1D) {
        // initial_condition_convergence_test of CompressibleEulerEquationsQuasi1D:
        // ini = 2 + 0.1 sin(pi (x - t)), rho = ini, v = 1, e = ini^2 / rho,
        // p = (gamma - 1)(e - rho v^2 / 2), a = 1.5 - 0.5 cos(pi x)
        double rho = 1.0, v = 0.0, p = 1.0, a = 1.0;
        if (q.ic == TRIXI_HIP_IC_CONVERGENCE_TEST) {
            const double ini = 2.0 + 0.1 * sin(M_PI * (x[0] - t));
            rho = ini;
            v = 1.0;
            const double e = ini * ini / rho;
            p = (q.gamma - 1.0) * (e - 0.5 * rho * v * v);
            a = 1.5 - 0.5 * cos(x[0] * M_PI);
        } else if (q.ic == TRIXI_HIP_IC_SEEDED_RANDOM) {
            rho = 1.0 + 0.1 * (2.0 * u01(q.seed, idx, 0) - 1.0);
            p = 1.0 + 0.05 * (2.0 * u01(q.seed, idx, 1) - 1.0);
            v = 0.1 * normal01(q.seed, idx, 2);
            a = 1.0 + 0.2 * (2.0 * u01(q.seed, idx, 4) - 1.0);
        }
        u[0] = a * rho;
        u[1] = a * rho * v;
        u[2] = a * (p * q.inv_gm1 + 0.5 * rho * v * v);
        u[3] = a;
        return;
    }
    if (q.eq == EQ_EULER_MULTI) {
        // (rho v, rho e, rho_1..rho_nc).  Weak blast wave (Trixi's
        // multicomponent initial_condition_weak_blast_wave): the Euler blast
        // with rho_k = 2^(k-1) (1 - 2) rho / (1 - 2^nc); the seeded state:
        // rho_k = 2^(k-1) / (2^nc - 1) (1 + 0.1 U), v = 0.1 N(0, 1), p = 1 + 0.05 U.
        double v[3] = {0, 0, 0}, rho = 1.0, p = 1.0, rk[MAXC];
        const double den = 1.0 - (double)(1 << q.nc);
        if (q.ic == TRIXI_HIP_IC_WEAK_BLAST_WAVE || q.ic == TRIXI_HIP_IC_WEAK_BLAST_WAVE_SEEDED)
            weak_blast_prim(q, x, idx, rho, v, p);
        for (int k = 0; k < q.nc; ++k) rk[k] = (double)(1 << k) * (1.0 - 2.0) * rho / den;
        if (q.ic == TRIXI_HIP_IC_SEEDED_RANDOM) {
            for (int k = 0; k < q.nc; ++k) rk[k] *= 1.0 + 0.1 * (2.0 * u01(q.seed, idx, 8 + k) - 1.0);
            p = 1.0 + 0.05 * (2.0 * u01(q.seed, idx, 1) - 1.0);
            for (int d = 0; d < q.dim; ++d) v[d] = 0.1 * normal01(q.seed, idx, 2 + 2 * d);
        }
        double rho_t = 0.0, h1 = 0.0, h2 = 0.0;
        for (int k = 0; k < q.nc; ++k) {
            rho_t += rk[k];
            h1 += rk[k] * q.cvg[k];
            h2 += rk[k] * q.cv[k];
        }
        const double gamma = h1 / h2;
        double ke = 0.0;
        for (int d = 0; d < q.dim; ++d) {
            u[d] = rho_t * v[d];
            ke += u[d] * v[d];
        }
        u[q.dim] = p / (gamma - 1.0) + 0.5 * ke;
        for (int k = 0; k < q.nc; ++k) u[q.dim + 1 + k] = rk[k];
        return;
    }
    if (q.eq == EQ_LINADV) {
        if (q.ic == TRIXI_HIP_IC_CONVERGENCE_TEST) {
            double xs = 0.0;
            for (int d = 0; d < q.dim; ++d) xs += x[d] - q.a[d] * t;
            const double omega = 2.0 * M_PI * (1.0 / 2.0);
            u[0] = 1.0 + 0.5 * sin(omega * xs);
        } else if (q.ic == TRIXI_HIP_IC_SEEDED_RANDOM) {
            u[0] = 1.0 + 0.1 * (2.0 * u01(q.seed, idx, 0) - 1.0);
        } else {
            u[0] = 2.0;
        }
        return;
    }
    if (q.eq == EQ_EULER) {
        double v[3] = {0, 0, 0}, rho = 1.0, p = 1.0;
        if (q.ic == TRIXI_HIP_IC_WEAK_BLAST_WAVE || q.ic == TRIXI_HIP_IC_WEAK_BLAST_WAVE_SEEDED) {
            weak_blast_prim(q, x, idx, rho, v, p);
        } else if (q.ic == TRIXI_HIP_IC_EOC_TEST_COUPLED_EULER_GRAVITY) {
            // c = 2, A = 0.1, G = 1: rho = c + A sin(pi (sum x - t)), v = 1,
            // p = rho^2 G 2 / (d pi)
            double xs = -t;
            for (int d = 0; d < q.dim; ++d) xs += x[d];
            rho = 2.0 + 0.1 * sin(M_PI * xs);
            for (int d = 0; d < q.dim; ++d) v[d] = 1.0;
            p = rho * rho * 2.0 / (q.dim * M_PI);
        } else if (q.ic == TRIXI_HIP_IC_SEEDED_RANDOM) {
            rho = 1.0 + 0.1 * (2.0 * u01(q.seed, idx, 0) - 1.0);
            p = 1.0 + 0.05 * (2.0 * u01(q.seed, idx, 1) - 1.0);
            for (int d = 0; d < q.dim; ++d) v[d] = 0.1 * normal01(q.seed, idx, 2 + 2 * d);
        } else if (q.ic == TRIXI_HIP_IC_SMOOTH_PERIODIC) {
            const double L = q.length > 0 ? q.length : 4.0;
            rho = 1.0 + 0.1 * smooth_field(q.seed, 0, q.dim, x, L, false);
            for (int d = 0; d < q.dim; ++d) v[d] = 0.1 * smooth_field(q.seed, 1 + d, q.dim, x, L, false);
            p = 1.0 + 0.05 * smooth_field(q.seed, 4, q.dim, x, L, true);
        } else if (q.ic == TRIXI_HIP_IC_CONVERGENCE_TEST) {
            double xs = 0.0;
            for (int d = 0; d < q.dim; ++d) xs += x[d];
            const double ini = 2.0 + 0.1 * sin(M_PI * (xs - t));
            u[0] = ini;
            for (int d = 0; d < q.dim; ++d) u[1 + d] = ini;
            u[q.dim + 1] = ini * ini;
            return;
        }
        u[0] = rho;
        double ke = 0.0;
        for (int d = 0; d < q.dim; ++d) {
            u[1 + d] = rho * v[d];
            ke = (d == 0) ? u[1 + d] * v[d] : ke + u[1 + d] * v[d];
        }
        u[q.dim + 1] = p * q.inv_gm1 + 0.5 * ke;
        return;
    }
    double prim[9] = {1.0, 0, 0, 0, 1.0, 0, 0, 0, 0};
    if (q.ic == TRIXI_HIP_IC_CONVERGENCE_TEST && q.dim == 1) {
        // circularly polarised Alfven wave of IdealGlmMhdEquations1D on [0, 1]:
        // rho = 1, p = 0.1, B1 = 1, (v2, v3) = (B2, B3) = 0.1 (sin, cos)(2 pi (x + t))
        const double arg = 2.0 * M_PI * (x[0] + t);
        prim[1] = 0.0;
        prim[2] = 0.1 * sin(arg);
        prim[3] = 0.1 * cos(arg);
        prim[4] = 0.1;
        prim[5] = 1.0;
        prim[6] = prim[2];
        prim[7] = prim[3];
    } else if (q.ic == TRIXI_HIP_IC_CONVERGENCE_TEST && q.dim == 2) {
        // Alfven wave of IdealGlmMhdEquations2D (Derigs et al. 2016): along
        // x_perp = x cos(pi/4) + y sin(pi/4), B_perp = 0.1 sin(2 pi (x_perp + t))
        const double ca = cos(0.25 * M_PI), sa = sin(0.25 * M_PI);
        const double xp = x[0] * ca + x[1] * sa;
        const double Bp = 0.1 * sin(2.0 * M_PI * (xp + t));
        prim[0] = 1.0;
        prim[1] = -Bp * sa;
        prim[2] = Bp * ca;
        prim[3] = 0.1 * cos(2.0 * M_PI * (xp + t));
        prim[4] = 0.1;
        prim[5] = ca + prim[1];
        prim[6] = sa + prim[2];
        prim[7] = prim[3];
        prim[8] = 0.0;
    } else if (q.ic == TRIXI_HIP_IC_CONVERGENCE_TEST) {
        const double omega = 2.0 * M_PI, r = 2.0, e = 0.2;
        const double nx = 1.0 / sqrt(r * r + 1.0), ny = r / sqrt(r * r + 1.0), sqr = 1.0;
        const double Va = omega / (ny * sqr);
        const double phi_alv = omega / ny * (nx * (x[0] - 0.5 * r) + ny * (x[1] - 0.5 * r)) - Va * t;
        const double rho = 1.0;
        const double v1 = -e * ny * cos(phi_alv) / rho;
        const double v2 = e * nx * cos(phi_alv) / rho;
        const double v3 = e * sin(phi_alv) / rho;
        prim[0] = rho;
        prim[1] = v1;
        prim[2] = v2;
        prim[3] = v3;
        prim[4] = 1.0;
        prim[5] = nx - rho * v1 * sqr;
        prim[6] = ny - rho * v2 * sqr;
        prim[7] = -rho * v3 * sqr;
        prim[8] = 0.0;
    } else if (q.ic == TRIXI_HIP_IC_WEAK_BLAST_WAVE || q.ic == TRIXI_HIP_IC_WEAK_BLAST_WAVE_SEEDED) {
        // the MHD weak blast: the Euler blast with B = (1, 1, 1), psi = 0
        double rho, v[3] = {0.0, 0.0, 0.0}, p;
        weak_blast_prim(q, x, idx, rho, v, p);
        prim[0] = rho;
        for (int d = 0; d < 3; ++d) prim[1 + d] = v[d];
        prim[4] = p;
        for (int d = 0; d < 3; ++d) prim[5 + d] = 1.0;
        prim[8] = 0.0;
    } else if (q.ic == TRIXI_HIP_IC_SEEDED_RANDOM) {
        prim[0] = 1.0 + 0.1 * (2.0 * u01(q.seed, idx, 0) - 1.0);
        prim[4] = 1.0 + 0.05 * (2.0 * u01(q.seed, idx, 1) - 1.0);
        for (int d = 0; d < 3; ++d) prim[1 + d] = 0.1 * normal01(q.seed, idx, 2 + 2 * d);
        for (int d = 0; d < 3; ++d) prim[5 + d] = 0.5 + 0.1 * (2.0 * u01(q.seed, idx, 8 + d) - 1.0);
        prim[8] = 0.01 * (2.0 * u01(q.seed, idx, 11) - 1.0);
    }
    const double rho = prim[0], p = prim[4], psi = prim[8];
    u[0] = rho;
    for (int d = 0; d < 3; ++d) u[1 + d] = rho * prim[1 + d];
    u[4] = p * q.inv_gm1 + 0.5 * (u[1] * prim[1] + u[2] * prim[2] + u[3] * prim[3]) +
           0.5 * (prim[5] * prim[5] + prim[6] * prim[6] + prim[7] * prim[7]) + 0.5 * psi * psi;
    for (int d = 0; d < 3; ++d) u[5 + d] = prim[5 + d];
    if (q.nv > 8) u[8] = psi;  // 1D ideal MHD has no cleaning field
}


// Dirichlet boundary states built into the non-periodic kernels: the
// convergence-test state of compressible Euler (initial_condition_convergence_test)
// and the constant state; for the Burgers equation its convergence-test wave
// and the shock / rarefaction steps.  The library rejects other boundary states.
__host__ __device__ inline void boundary_value(const IcParams& q, const double* x, double t, double* u) {
    if (q.eq == EQ_SHALLOW_WATER) {
        shallow_water_state(q, x, t, 0, u);
        return;
    }
    if (q.eq == EQ_HYPDIFF) {
        hypdiff_state(q, x, t, q.ic == TRIXI_HIP_IC_POISSON_NONPERIODIC, u);
        return;
    }
    if (q.eq == EQ_BURGERS) {
        if (q.ic == TRIXI_HIP_IC_CONVERGENCE_TEST)
            u[0] = 2.0 + sin(2.0 * M_PI * (x[0] - t));
        else if (q.ic == TRIXI_HIP_IC_SHOCK)
            u[0] = x[0] < 0.5 ? 1.5 : 0.5;
        else if (q.ic == TRIXI_HIP_IC_RAREFACTION)
            u[0] = x[0] < 0.5 ? 0.5 : 1.5;
        else
            u[0] = 2.0;
        return;
    }
    if (q.ic == TRIXI_HIP_IC_CONVERGENCE_TEST) {
        double xs = 0.0;
        for (int d = 0; d < q.dim; ++d) xs += x[d];
        const double ini = 2.0 + 0.1 * sin(M_PI * (xs - t));
        u[0] = ini;
        for (int d = 0; d < q.dim; ++d) u[1 + d] = ini;
        u[q.dim + 1] = ini * ini;
    } else {
        u[0] = 1.0;
        for (int d = 0; d < q.dim; ++d) u[1 + d] = 0.0;
        u[q.dim + 1] = q.inv_gm1;
    }
}

// source_terms_convergence_test (compressible Euler, d dimensions): the
// source that makes initial_condition_convergence_test,
//   rho = rho v_k = g,  rho e = g^2,  g = 2 + 0.1 sin(pi (sum_k x_k - t)),
// an exact solution.  With g' = 0.1 pi cos(pi (sum_k x_k - t)), v_k = 1 and
// p = (gamma - 1)(g^2 - d g / 2), S = d_t u + div f(u) gives
//   S_rho = (d - 1) g',  S_{rho v_k} = (d - 1) g' + p',
//   S_{rho e} = 2 (d - 1) g g' + d p',  p' = (gamma - 1)(2 g - d / 2) g'.
__host__ __device__ inline void source_convergence_test(int dim, double gm1, const double* x, double t, double* s) {
    double xs = 0.0;
    for (int d = 0; d < dim; ++d) xs += x[d];
    const double arg = M_PI * (xs - t);
    const double g = 2.0 + 0.1 * sin(arg);
    const double gp = 0.1 * M_PI * cos(arg);
    const double pp = gm1 * (2.0 * g - 0.5 * dim) * gp;
    s[0] = (dim - 1) * gp;
    for (int d = 0; d < dim; ++d) s[1 + d] = (dim - 1) * gp + pp;
    s[dim + 1] = 2.0 * (dim - 1) * g * gp + dim * pp;
}

// source_terms_convergence_test of CompressibleEulerEquationsQuasi1D: with
// ini = 2 + 0.1 sin(pi (x - t)), g' = 0.1 pi cos(pi (x - t)) (so ini_t = -g',
// ini_x = g'), rho = e = ini, v = 1, p = (gamma - 1) ini / 2,
// a = 1.5 - 0.5 cos(pi x), a_x = pi/2 sin(pi x):
//   S1 = (a rho)_t + (a rho v)_x              = a_x ini,
//   S2 = (a rho v)_t + (a rho v^2)_x + a p_x   = a_x ini + a p_x,
//   S3 = (a e)_t + (a v (e + p))_x             = a_x (ini + p) + a p_x,
// p_x = (gamma - 1) g' / 2, S4 = 0.
__host__ __device__ inline void source_quasi1d_convergence_test(double gamma, const double* x, double t, double* s) {
    const double arg = M_PI * (x[0] - t);
    const double ini = 2.0 + 0.1 * sin(arg);
    const double gp = 0.1 * M_PI * cos(arg);
    const double a = 1.5 - 0.5 * cos(x[0] * M_PI);
    const double a_x = 0.5 * M_PI * sin(x[0] * M_PI);
    const double p = 0.5 * (gamma - 1.0) * ini, p_x = 0.5 * (gamma - 1.0) * gp;
    s[0] = a_x * ini;
    s[1] = a_x * ini + a * p_x;
    s[2] = a_x * (ini + p) + a * p_x;
    s[3] = 0.0;
}

// source_terms_eoc_test_euler (compressible Euler, d dimensions) for the state
// of initial_condition_eoc_test_coupled_euler_gravity: rho = g =
// 2 + 0.1 sin(pi (sum_k x_k - t)), v_k = 1, p = K g^2 with K = 2 / (d pi).
// With g' = 0.1 pi cos(pi (sum_k x_k - t)) and p' = 2 K g g', S = d_t u + div f(u):
//   S_rho = (d - 1) g',  S_{rho v_k} = (d - 1) g' + p',
//   S_{rho e} = (d - 1) e' + d p',  e' = p' / (gamma - 1) + d g' / 2.
__host__ __device__ inline void source_eoc_test_euler(int dim, double gm1, const double* x, double t, double* s) {
    double xs = -t;
    for (int d = 0; d < dim; ++d) xs += x[d];
    const double g = 2.0 + 0.1 * sin(M_PI * xs);
    const double gp = 0.1 * M_PI * cos(M_PI * xs);
    const double pp = 2.0 * (2.0 / (dim * M_PI)) * g * gp;
    s[0] = (dim - 1) * gp;
    for (int d = 0; d < dim; ++d) s[1 + d] = (dim - 1) * gp + pp;
    const double ep = pp / gm1 + 0.5 * dim * gp;
    s[dim + 1] = (dim - 1) * ep + dim * pp;
}

// Hyperbolic diffusion sources: source_terms_poisson_nonperiodic
// (f = -nu lap phi of the steady state: nu pi^2 sin(pi x) in 1D,
// 10 nu pi^2 cos(pi x) sin(2 pi y) in 2D, 18 nu pi^2 cos(pi x) sin(2 pi y) sin(2 pi z)
// in 3D) and source_terms_harmonic (f = 0); both relax q: -q_k / T_r.
__host__ __device__ inline void source_hypdiff(int dim, int src, double nu, double inv_Tr, const double* x,
                                               const double* qv, double* s) {
    double f = 0.0;
    if (src == TRIXI_HIP_SOURCE_POISSON_NONPERIODIC) {
        if (dim == 1)
            f = nu * M_PI * M_PI * sin(M_PI * x[0]);
        else if (dim == 2)
            f = 10.0 * nu * M_PI * M_PI * cos(M_PI * x[0]) * sin(2.0 * M_PI * x[1]);
        else
            f = 18.0 * nu * M_PI * M_PI * cos(M_PI * x[0]) * sin(2.0 * M_PI * x[1]) * sin(2.0 * M_PI * x[2]);
    }
    s[0] = f;
    for (int k = 0; k < dim; ++k) s[1 + k] = -inv_Tr * qv[k];
}

// source_terms_convergence_test of the inviscid Burgers equation: with
// u = c + A sin(omega (x - t)) (c = 2, A = 1, omega = 2 pi),
// S = u_t + u u_x = omega A cos(omega (x - t)) (c - 1 + A sin(omega (x - t))).
__host__ __device__ inline double source_burgers_convergence_test(double x, double t) {
    const double arg = 2.0 * M_PI * (x - t);
    return 2.0 * M_PI * cos(arg) * (1.0 + sin(arg));
}

}  // namespace trixi_hip
