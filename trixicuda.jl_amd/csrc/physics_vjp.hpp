// physics_vjp.hpp -- hand-written reverse-mode (vector-Jacobian) versions of
// the point and two-point fluxes, for the adjoint of rhs!.
//
// Cotangents are accumulated per node in primitive space (rho, v, p): every
// volume and surface flux here is a function of the node primitives (and of
// log(rho), log(p), whose derivatives are folded in through the exact
// derivatives of the logarithmic means), so a node's contributions from all
// pairs, faces and directions add up in primitive space and are turned into
// conservative cotangents once per node at the end (to_cons_bar).
// Conservative-space cotangents that do arise (the LLF dissipation term and
// the energy inside the physical flux) are mapped to primitive space with
// (d cons / d prim)^T (add_cons_bar), which to_cons_bar inverts exactly.
#pragma once

#include "physics.hpp"

namespace trixi_hip {

// ln_mean(x, y) and its partial derivatives (both branches of the Trixi
// definition; ix = 1/x, iy = 1/y).
//   closed: L = (y - x)/D, D = log y - log x:  L_x = (L/x - 1)/D, L_y = (1 - L/y)/D
//   series: L = s Q(f2), s = x + y, z = (y - x)/s, f2 = z^2:
//           L_x = Q - 4 y z Q'(f2)/s,  L_y = Q + 4 x z Q'(f2)/s
template <typename R>
__device__ __forceinline__ void ln_mean_grad(R x, R y, double lx, double ly, R ix, R iy, R& L, R& Lx, R& Ly) {
    const R dxy = y - x, s = x + y;
    const bool series = dabs(dxy) < R(0.01) * s;  // the forward mean's branch test
    if (series) {
        const R r = rrcp(s);
        const R z = dxy * r, f2 = z * z;
        const R q = R(0.5) + f2 * (R(-1.0 / 6.0) + f2 * (R(-2.0 / 45.0) + f2 * R(-22.0 / 945.0)));
        const R qp = R(-1.0 / 6.0) + f2 * (R(-4.0 / 45.0) + f2 * R(-66.0 / 945.0));
        L = s * q;
        Lx = q - R(4) * y * z * qp * r;
        Ly = q + R(4) * x * z * qp * r;
    } else {
        const R r = rrcp((R)(ly - lx));
        L = dxy * r;
        Lx = (L * ix - R(1)) * r;
        Ly = (R(1) - L * iy) * r;
    }
}

// inv_ln_mean(x, y) = 1/ln_mean and its partial derivatives.
//   closed: I = D/(y - x):  I_x = (I - 1/x)/(y - x), I_y = (1/y - I)/(y - x)
//   series: I = P(f2)/s:    I_x = -(4 y z P' /s + P)/s^2, I_y = (4 x z P'/s - P)/s^2
template <typename R>
__device__ __forceinline__ void inv_ln_mean_grad(R x, R y, double lx, double ly, R ix, R iy, R& I, R& Ix,
                                                 R& Iy) {
    const R dxy = y - x, s = x + y;
    const bool series = dabs(dxy) < R(0.01) * s;  // the forward mean's branch test
    if (series) {
        const R r = rrcp(s);
        const R z = dxy * r, f2 = z * z;
        const R pp = R(2) + f2 * (R(2.0 / 3.0) + f2 * (R(2.0 / 5.0) + f2 * R(2.0 / 7.0)));
        const R ppd = R(2.0 / 3.0) + f2 * (R(4.0 / 5.0) + f2 * R(6.0 / 7.0));
        const R r2 = r * r;
        I = pp * r;
        Ix = -r2 * (R(4) * y * z * ppd * r + pp);
        Iy = r2 * (R(4) * x * z * ppd * r - pp);
    } else {
        const R r = rrcp(dxy);
        I = (R)(ly - lx) * r;
        Ix = (I - ix) * r;
        Iy = (iy - I) * r;
    }
}

// ===========================================================================
// Linear scalar advection.
// ===========================================================================
template <int DIM>
struct LinAdvVjp {
    using F = LinAdv<DIM>;
    static constexpr int NV = 1;
    static constexpr bool NONCONS = false;
    static constexpr int NRF = 1;
    template <typename R>
    struct Node {
        R u;
        double lrho, lp;
    };
    template <typename R>
    struct Bar {
        R u;
    };
    template <typename R>
    __device__ static Node<R> node(const R* u, const PhysR<R>&) {
        Node<R> n;
        n.u = u[0];
        n.lrho = n.lp = 0.0;
        return n;
    }
    template <typename R>
    __device__ static void to_rec(const Node<R>& n, R* f) { f[0] = n.u; }
    template <typename R>
    __device__ static Node<R> from_rec(const R* f, double, double) {
        Node<R> n;
        n.u = f[0];
        n.lrho = n.lp = 0.0;
        return n;
    }
    template <typename R>
    __device__ static void zero(Bar<R>& b) { b.u = R(0); }
    template <typename R>
    __device__ static void bar_to_array(const Bar<R>& b, R* a) { a[0] = b.u; }
    template <typename R>
    __device__ static void array_to_bar(const R* a, Bar<R>& b) { b.u = a[0]; }
    template <typename R>
    __device__ static void flux_vjp(const Node<R>&, int o, const R* g, Bar<R>& b, const PhysR<R>& P) {
        b.u += P.a[o] * g[0];
    }
    template <int KIND, typename R>
    __device__ static void two_point_vjp(const Node<R>&, const Node<R>&, int o, const R* g, Bar<R>& bl,
                                         Bar<R>& br, const PhysR<R>& P) {
        const R h = R(0.5) * P.a[o] * g[0];
        if (KIND == FLUX_LLF || KIND == FLUX_LLF_NAIVE) {
            const R d = R(0.5) * dabs(P.a[o]) * g[0];
            bl.u += h + d;
            br.u += h - d;
        } else {
            bl.u += h;
            br.u += h;
        }
    }
    template <typename R>
    __device__ static void to_cons_bar(const Node<R>&, const Bar<R>& b, R* ub, const PhysR<R>&) { ub[0] = b.u; }
};

// ===========================================================================
// Compressible Euler in DIM dimensions.
// ===========================================================================
template <int DIM>
struct EulerVjp {
    static constexpr int NV = DIM + 2;
    static constexpr bool NONCONS = false;
    static constexpr int NRF = DIM + 5;  // rho, v[DIM], p, c, 1/rho, 1/p
    template <typename R>
    struct Node {
        R rho, v[DIM], p, c, irho, ip;
        double lrho, lp;
    };
    template <typename R>
    struct Bar {
        R rho, v[DIM], p;
    };

    template <typename R>
    __device__ static Node<R> node(const R* u, const PhysR<R>& P) {
        Node<R> n;
        n.rho = u[0];
        n.irho = rrcp(n.rho);
        R ke = R(0);
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            n.v[d] = u[1 + d] * n.irho;
            ke = (d == 0) ? u[1 + d] * n.v[d] : ke + u[1 + d] * n.v[d];
        }
        n.p = P.gm1 * (u[DIM + 1] - R(0.5) * ke);
        n.ip = rrcp(n.p);
        n.c = rsqrt_(P.gamma * n.p * n.irho);
        n.lrho = node_log((double)n.rho);
        n.lp = node_log((double)n.p);
        return n;
    }
    template <typename R>
    __device__ static void to_rec(const Node<R>& n, R* f) {
        f[0] = n.rho;
#pragma unroll
        for (int d = 0; d < DIM; ++d) f[1 + d] = n.v[d];
        f[DIM + 1] = n.p;
        f[DIM + 2] = n.c;
        f[DIM + 3] = n.irho;
        f[DIM + 4] = n.ip;
    }
    template <typename R>
    __device__ static Node<R> from_rec(const R* f, double lrho, double lp) {
        Node<R> n;
        n.rho = f[0];
#pragma unroll
        for (int d = 0; d < DIM; ++d) n.v[d] = f[1 + d];
        n.p = f[DIM + 1];
        n.c = f[DIM + 2];
        n.irho = f[DIM + 3];
        n.ip = f[DIM + 4];
        n.lrho = lrho;
        n.lp = lp;
        return n;
    }
    template <typename R>
    __device__ static void zero(Bar<R>& b) {
        b.rho = R(0);
#pragma unroll
        for (int d = 0; d < DIM; ++d) b.v[d] = R(0);
        b.p = R(0);
    }
    template <typename R>
    __device__ static void bar_to_array(const Bar<R>& b, R* a) {
        a[0] = b.rho;
#pragma unroll
        for (int d = 0; d < DIM; ++d) a[1 + d] = b.v[d];
        a[DIM + 1] = b.p;
    }
    template <typename R>
    __device__ static void array_to_bar(const R* a, Bar<R>& b) {
        b.rho = a[0];
#pragma unroll
        for (int d = 0; d < DIM; ++d) b.v[d] = a[1 + d];
        b.p = a[DIM + 1];
    }
    template <typename R>
    __device__ static R vsq(const Node<R>& n) {
        R s = R(0);
#pragma unroll
        for (int d = 0; d < DIM; ++d) s += n.v[d] * n.v[d];
        return s;
    }
    template <typename R>
    __device__ static R cons(const Node<R>& n, int q, const PhysR<R>& P) {
        if (q == 0) return n.rho;
        if (q <= DIM) return n.rho * n.v[q - 1];
        return n.p * P.inv_gm1 + R(0.5) * n.rho * vsq(n);
    }
    // (d cons / d prim)^T c: conservative cotangent c into primitive space.
    template <typename R>
    __device__ static void add_cons_bar(const Node<R>& n, const R* c, Bar<R>& b, const PhysR<R>& P) {
        const R cE = c[DIM + 1];
        R t = c[0] + cE * R(0.5) * vsq(n);
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            t += c[1 + d] * n.v[d];
            b.v[d] += (c[1 + d] + cE * n.v[d]) * n.rho;
        }
        b.rho += t;
        b.p += cE * P.inv_gm1;
    }
    // (d prim / d cons)^T b: primitive cotangent into conservative space.
    template <typename R>
    __device__ static void to_cons_bar(const Node<R>& n, const Bar<R>& b, R* ub, const PhysR<R>& P) {
        R t = b.rho + b.p * P.gm1 * R(0.5) * vsq(n);
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            t -= b.v[d] * n.v[d] * n.irho;
            ub[1 + d] = b.v[d] * n.irho - b.p * P.gm1 * n.v[d];
        }
        ub[0] = t;
        ub[DIM + 1] = b.p * P.gm1;
    }

    // physical flux f_o(u): (rho v_o, rho v_o v + p e_o, (E + p) v_o)
    template <typename R>
    __device__ static void flux_vjp(const Node<R>& n, int o, const R* g, Bar<R>& b, const PhysR<R>& P) {
        const R vo = n.v[o];
        const R E = cons(n, DIM + 1, P);
        R gv = R(0);
#pragma unroll
        for (int d = 0; d < DIM; ++d) gv += g[1 + d] * n.v[d];
        const R gE = g[DIM + 1];
        b.rho += g[0] * vo + gv * vo;
        R vob = g[0] * n.rho + gv * n.rho + gE * (E + n.p);
#pragma unroll
        for (int d = 0; d < DIM; ++d) b.v[d] += g[1 + d] * n.rho * vo;
        b.v[o] += vob;
        b.p += g[1 + o] + gE * vo;
        // E = p/(gamma-1) + rho |v|^2 / 2
        const R Eb = gE * vo;
        b.rho += Eb * R(0.5) * vsq(n);
#pragma unroll
        for (int d = 0; d < DIM; ++d) b.v[d] += Eb * n.rho * n.v[d];
        b.p += Eb * P.inv_gm1;
    }

    // forward two-point flux value (the forward kernel's formulas), used for
    // the cotangent of the shock-capturing blending factor
    template <int KIND, typename R>
    __device__ static void two_point_value(const Node<R>& l, const Node<R>& r, int o, const PhysR<R>& P, R* f) {
        using F = Euler<DIM>;
        typename F::template Node<R> a, b;
        a.rho = l.rho;
        b.rho = r.rho;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            a.v[d] = l.v[d];
            b.v[d] = r.v[d];
        }
        a.p = l.p;
        b.p = r.p;
        a.c = l.c;
        b.c = r.c;
        a.lrho = l.lrho;
        b.lrho = r.lrho;
        a.lp = l.lp;
        b.lp = r.lp;
        a.inv_gm1 = b.inv_gm1 = P.inv_gm1;
        F::template two_point<KIND>(a, b, o, P, f);
    }

    // sound speed c = sqrt(gamma p / rho): dc/dp = c/(2p), dc/drho = -c/(2 rho)
    template <typename R>
    __device__ static void sound_vjp(const Node<R>& n, R cb, Bar<R>& b) {
        const R h = R(0.5) * cb * n.c;
        b.p += h * n.ip;
        b.rho -= h * n.irho;
    }

    template <int KIND, typename R>
    __device__ static void two_point_vjp(const Node<R>& l, const Node<R>& r, int o, const R* g, Bar<R>& bl,
                                         Bar<R>& br, const PhysR<R>& P) {
        if (KIND == FLUX_RANOCHA) {
            R rho_mean, dLx, dLy;
            ln_mean_grad<R>(l.rho, r.rho, l.lrho, r.lrho, l.irho, r.irho, rho_mean, dLx, dLy);
            const R X = l.rho * r.p, Y = r.rho * l.p;
            R I, dIx, dIy;
            inv_ln_mean_grad<R>(X, Y, l.lrho + r.lp, r.lrho + l.lp, l.irho * r.ip, r.irho * l.ip, I, dIx, dIy);
            const R B = l.p * r.p * I;
            R va[DIM], vs = R(0);
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                va[d] = R(0.5) * (l.v[d] + r.v[d]);
                vs += l.v[d] * r.v[d];
            }
            vs = R(0.5) * vs;
            const R f1 = rho_mean * va[o];
            const R gE = g[DIM + 1];
            R f1b = g[0] + gE * (vs + B * P.inv_gm1);
#pragma unroll
            for (int d = 0; d < DIM; ++d) f1b += g[1 + d] * va[d];
            const R rmb = f1b * va[o];
            const R vsb = gE * f1;
            const R Bb = gE * f1 * P.inv_gm1;
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                R vab = g[1 + d] * f1;
                if (d == o) vab += f1b * rho_mean;
                bl.v[d] += R(0.5) * (vab + vsb * r.v[d]);
                br.v[d] += R(0.5) * (vab + vsb * l.v[d]);
            }
            const R pab = g[1 + o];
            bl.p += R(0.5) * (pab + gE * r.v[o]);
            br.p += R(0.5) * (pab + gE * l.v[o]);
            bl.v[o] += R(0.5) * gE * r.p;
            br.v[o] += R(0.5) * gE * l.p;
            const R Ib = Bb * l.p * r.p;
            bl.p += Bb * r.p * I;
            br.p += Bb * l.p * I;
            const R Xb = Ib * dIx, Yb = Ib * dIy;
            bl.rho += Xb * r.p + rmb * dLx;
            br.p += Xb * l.rho;
            br.rho += Yb * l.p + rmb * dLy;
            bl.p += Yb * r.rho;
        } else if (KIND == FLUX_SHIMA_ETAL) {
            const R rho_avg = R(0.5) * (l.rho + r.rho);
            R va[DIM], kin = R(0);
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                va[d] = R(0.5) * (l.v[d] + r.v[d]);
                kin += l.v[d] * r.v[d];
            }
            kin = R(0.5) * kin;
            const R p_avg = R(0.5) * (l.p + r.p);
            const R f1 = rho_avg * va[o];
            const R gE = g[DIM + 1];
            R f1b = g[0] + gE * kin;
#pragma unroll
            for (int d = 0; d < DIM; ++d) f1b += g[1 + d] * va[d];
            const R rab = f1b * va[o];
            const R pab = g[1 + o] + gE * va[o] * P.inv_gm1;
            const R kinb = gE * f1;
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                R vab = g[1 + d] * f1;
                if (d == o) vab += f1b * rho_avg + gE * p_avg * P.inv_gm1;
                bl.v[d] += R(0.5) * (vab + kinb * r.v[d]);
                br.v[d] += R(0.5) * (vab + kinb * l.v[d]);
            }
            bl.rho += R(0.5) * rab;
            br.rho += R(0.5) * rab;
            // pv_avg = (p_l v_r,o + p_r v_l,o)/2
            bl.p += R(0.5) * (pab + gE * r.v[o]);
            br.p += R(0.5) * (pab + gE * l.v[o]);
            br.v[o] += R(0.5) * gE * l.p;
            bl.v[o] += R(0.5) * gE * r.p;
        } else {
            R h[NV];
#pragma unroll
            for (int q = 0; q < NV; ++q) h[q] = R(0.5) * g[q];
            flux_vjp(l, o, h, bl, P);
            flux_vjp(r, o, h, br, P);
            if (KIND == FLUX_LLF || KIND == FLUX_LLF_NAIVE) {
                const R al = dabs(l.v[o]), ar = dabs(r.v[o]);
                R lam;
                bool right_v, right_c;
                if (KIND == FLUX_LLF_NAIVE) {
                    right_v = ar > al;
                    right_c = r.c > l.c;
                    lam = (right_v ? ar : al) + (right_c ? r.c : l.c);
                } else {
                    right_v = (ar + r.c) > (al + l.c);
                    right_c = right_v;
                    lam = right_v ? ar + r.c : al + l.c;
                }
                // F += -lam/2 (u_r - u_l)
                R cl[NV], cr[NV], dot = R(0);
#pragma unroll
                for (int q = 0; q < NV; ++q) {
                    const R hl = R(0.5) * lam * g[q];
                    cl[q] = hl;
                    cr[q] = -hl;
                    dot += g[q] * (cons(r, q, P) - cons(l, q, P));
                }
                add_cons_bar(l, cl, bl, P);
                add_cons_bar(r, cr, br, P);
                const R lb = R(-0.5) * dot;
                if (right_v)
                    br.v[o] += lb * (r.v[o] >= R(0) ? R(1) : R(-1));
                else
                    bl.v[o] += lb * (l.v[o] >= R(0) ? R(1) : R(-1));
                if (right_c)
                    sound_vjp(r, lb, br);
                else
                    sound_vjp(l, lb, bl);
            }
        }
    }
};

// ===========================================================================
// Ideal GLM-MHD (2D and 3D) with the nonconservative Powell / GLM terms.
// Primitive variables (rho, v[3], p, B[3], psi).
// ===========================================================================
struct GlmMhdVjp {
    static constexpr int NV = 9;
    static constexpr bool NONCONS = true;
    static constexpr int NRF = 11;  // rho, v[3], p, B[3], psi, 1/rho, 1/p
    template <typename R>
    struct Node {
        R rho, v[3], p, B[3], psi, irho, ip;
        double lrho, lp;
    };
    template <typename R>
    struct Bar {
        R rho, v[3], p, B[3], psi;
    };

    template <typename R>
    __device__ static Node<R> node(const R* u, const PhysR<R>& P) {
        Node<R> n;
        n.rho = u[0];
        n.irho = rrcp(n.rho);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            n.v[d] = u[1 + d] * n.irho;
            n.B[d] = u[5 + d];
        }
        n.psi = u[8];
        n.p = P.gm1 * (u[4] - R(0.5) * (u[1] * n.v[0] + u[2] * n.v[1] + u[3] * n.v[2]) -
                       R(0.5) * (u[5] * u[5] + u[6] * u[6] + u[7] * u[7]) - R(0.5) * u[8] * u[8]);
        n.ip = rrcp(n.p);
        n.lrho = node_log((double)n.rho);
        n.lp = node_log((double)n.p);
        return n;
    }
    template <typename R>
    __device__ static void to_rec(const Node<R>& n, R* f) {
        f[0] = n.rho;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            f[1 + d] = n.v[d];
            f[5 + d] = n.B[d];
        }
        f[4] = n.p;
        f[8] = n.psi;
        f[9] = n.irho;
        f[10] = n.ip;
    }
    template <typename R>
    __device__ static Node<R> from_rec(const R* f, double lrho, double lp) {
        Node<R> n;
        n.rho = f[0];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            n.v[d] = f[1 + d];
            n.B[d] = f[5 + d];
        }
        n.p = f[4];
        n.psi = f[8];
        n.irho = f[9];
        n.ip = f[10];
        n.lrho = lrho;
        n.lp = lp;
        return n;
    }
    template <typename R>
    __device__ static void zero(Bar<R>& b) {
        b.rho = b.p = b.psi = R(0);
#pragma unroll
        for (int d = 0; d < 3; ++d) b.v[d] = b.B[d] = R(0);
    }
    template <typename R>
    __device__ static void bar_to_array(const Bar<R>& b, R* a) {
        a[0] = b.rho;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            a[1 + d] = b.v[d];
            a[5 + d] = b.B[d];
        }
        a[4] = b.p;
        a[8] = b.psi;
    }
    template <typename R>
    __device__ static void array_to_bar(const R* a, Bar<R>& b) {
        b.rho = a[0];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            b.v[d] = a[1 + d];
            b.B[d] = a[5 + d];
        }
        b.p = a[4];
        b.psi = a[8];
    }
    template <typename R>
    __device__ static R dot3(const R* a, const R* b) {
        return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    }
    template <typename R>
    __device__ static R cons(const Node<R>& n, int q, const PhysR<R>& P) {
        if (q == 0) return n.rho;
        if (q <= 3) return n.rho * n.v[q - 1];
        if (q == 4)
            return n.p * P.inv_gm1 + R(0.5) * n.rho * dot3(n.v, n.v) + R(0.5) * dot3(n.B, n.B) +
                   R(0.5) * n.psi * n.psi;
        if (q <= 7) return n.B[q - 5];
        return n.psi;
    }
    // (d cons / d prim)^T c
    template <typename R>
    __device__ static void add_cons_bar(const Node<R>& n, const R* c, Bar<R>& b, const PhysR<R>& P) {
        const R cE = c[4];
        R t = c[0] + cE * R(0.5) * dot3(n.v, n.v);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            t += c[1 + d] * n.v[d];
            b.v[d] += (c[1 + d] + cE * n.v[d]) * n.rho;
            b.B[d] += c[5 + d] + cE * n.B[d];
        }
        b.rho += t;
        b.p += cE * P.inv_gm1;
        b.psi += c[8] + cE * n.psi;
    }
    // (d prim / d cons)^T b
    template <typename R>
    __device__ static void to_cons_bar(const Node<R>& n, const Bar<R>& b, R* ub, const PhysR<R>& P) {
        const R pg = b.p * P.gm1;
        R t = b.rho + pg * R(0.5) * dot3(n.v, n.v);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            t -= b.v[d] * n.v[d] * n.irho;
            ub[1 + d] = b.v[d] * n.irho - pg * n.v[d];
            ub[5 + d] = b.B[d] - pg * n.B[d];
        }
        ub[0] = t;
        ub[4] = pg;
        ub[8] = b.psi - pg * n.psi;
    }

    // physical flux f_o (Trixi's flux(u, orientation, ::IdealGlmMhdEquations3D))
    template <typename R>
    __device__ static void flux_vjp(const Node<R>& n, int o, const R* g, Bar<R>& b, const PhysR<R>& P) {
        const R vo = n.v[o], Bo = n.B[o];
        const R vv = dot3(n.v, n.v), BB = dot3(n.B, n.B), vB = dot3(n.v, n.B);
        const R gE = g[4];
        R gv = R(0), gB = R(0), gi = R(0), gvi = R(0);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            gv += g[1 + d] * n.v[d];
            gB += g[1 + d] * n.B[d];
            gi += g[5 + d] * n.B[d];
            gvi += g[5 + d] * n.v[d];
        }
        // f0 = rho v_o; f_{1+d} = rho v_o v_d + delta(p + |B|^2/2) - B_o B_d
        b.rho += (g[0] + gv) * vo + gE * R(0.5) * vv * vo;
        R vob = (g[0] + gv) * n.rho;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            b.v[d] += g[1 + d] * n.rho * vo + gE * n.rho * n.v[d] * vo - gE * Bo * n.B[d] - g[5 + d] * Bo;
            b.B[d] += g[1 + o] * n.B[d] - g[1 + d] * Bo + gE * (R(2) * n.B[d] * vo - Bo * n.v[d]) +
                      g[5 + d] * vo;
        }
        b.p += g[1 + o] + gE * P.gamma * P.inv_gm1 * vo;
        // f4 = (rho|v|^2/2 + gamma p/(gamma-1) + |B|^2) v_o - B_o v.B + c_h psi B_o
        vob += gE * (R(0.5) * n.rho * vv + P.gamma * n.p * P.inv_gm1 + BB) + gi;
        R Bob = -gB - gE * vB + gE * P.c_h * n.psi - gvi + g[8] * P.c_h;
        b.v[o] += vob;
        b.B[o] += Bob;
        // f_{5+d} = v_o B_d - v_d B_o (+ c_h psi for d == o); f8 = c_h B_o
        b.psi += gE * P.c_h * Bo + g[5 + o] * P.c_h;
    }

    // fast magnetosonic speed c_f(n, o) and its reverse
    template <typename R>
    __device__ static R fast_speed(const Node<R>& n, int o, const PhysR<R>& P, R& a2, R& b2, R& bo2, R& D) {
        a2 = P.gamma * n.p * n.irho;
        b2 = dot3(n.B, n.B) * n.irho;
        bo2 = n.B[o] * n.B[o] * n.irho;
        const R s = a2 + b2;
        D = rsqrt_(s * s - R(4) * a2 * bo2);
        return rsqrt_(R(0.5) * s + R(0.5) * D);
    }
    template <typename R>
    __device__ static void fast_speed_vjp(const Node<R>& n, int o, R cb, Bar<R>& b, const PhysR<R>& P) {
        R a2, b2, bo2, D;
        const R cf = fast_speed(n, o, P, a2, b2, bo2, D);
        const R q = cb * rrcp(R(4) * cf);
        const R iD = rrcp(D);
        const R sb = q * (R(1) + (a2 + b2) * iD);
        const R a2b = sb - q * R(2) * bo2 * iD;
        const R bo2b = -q * R(2) * a2 * iD;
        b.p += a2b * P.gamma * n.irho;
        b.rho -= (a2b * a2 + sb * b2 + bo2b * bo2) * n.irho;
#pragma unroll
        for (int d = 0; d < 3; ++d) b.B[d] += sb * R(2) * n.B[d] * n.irho;
        b.B[o] += bo2b * R(2) * n.B[o] * n.irho;
    }

    template <int KIND, typename R>
    __device__ static void two_point_vjp(const Node<R>& l, const Node<R>& r, int o, const R* g, Bar<R>& bl,
                                         Bar<R>& br, const PhysR<R>& P) {
        if (KIND == FLUX_HINDENLANG_GASSNER) {
            R rho_mean, dLx, dLy;
            ln_mean_grad<R>(l.rho, r.rho, l.lrho, r.lrho, l.irho, r.irho, rho_mean, dLx, dLy);
            const R X = l.rho * r.p, Y = r.rho * l.p;
            R I, dIx, dIy;
            inv_ln_mean_grad<R>(X, Y, l.lrho + r.lp, r.lrho + l.lp, l.irho * r.ip, r.irho * l.ip, I, dIx, dIy);
            const R iprm = l.p * r.p * I;
            R va[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) va[d] = R(0.5) * (l.v[d] + r.v[d]);
            const R vsa = R(0.5) * dot3(l.v, r.v);
            const R f1 = rho_mean * va[o];
            const R gE = g[4];
            const R sb = R(0.5) * gE;  // cotangent of the sum s in f4 = ... + s/2
            R f1b = g[0] + gE * (vsa + iprm * P.inv_gm1);
#pragma unroll
            for (int d = 0; d < 3; ++d) f1b += g[1 + d] * va[d];
            // rho_mean, va
            const R rmb = f1b * va[o];
            R vab[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) vab[d] = g[1 + d] * f1;
            vab[o] += f1b * rho_mean;
            // p_avg + magnetic_square_avg in f_{1+o}
            const R pab = g[1 + o];
            // velocity_square_avg and inv_rho_p_mean in f4
            const R vsab = gE * f1;
            const R iprmb = gE * f1 * P.inv_gm1;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                bl.v[d] += R(0.5) * (vab[d] + vsab * r.v[d]);
                br.v[d] += R(0.5) * (vab[d] + vsab * l.v[d]);
                // magnetic_square_avg = (Bl . Br)/2 ; -(Bl_o Br_d + Br_o Bl_d)/2
                bl.B[d] += R(0.5) * pab * r.B[d] - R(0.5) * g[1 + d] * r.B[o];
                br.B[d] += R(0.5) * pab * l.B[d] - R(0.5) * g[1 + d] * l.B[o];
                bl.B[o] -= R(0.5) * g[1 + d] * r.B[d];
                br.B[o] -= R(0.5) * g[1 + d] * l.B[d];
            }
            bl.p += R(0.5) * pab;
            br.p += R(0.5) * pab;
            // inv_rho_p_mean = p_l p_r I(rho_l p_r, rho_r p_l)
            const R Ib = iprmb * l.p * r.p;
            bl.p += iprmb * r.p * I;
            br.p += iprmb * l.p * I;
            const R Xb = Ib * dIx, Yb = Ib * dIy;
            bl.rho += Xb * r.p + rmb * dLx;
            br.p += Xb * l.rho;
            br.rho += Yb * l.p + rmb * dLy;
            bl.p += Yb * r.rho;
            // s = p_l v_r,o + p_r v_l,o + sum_{d != o} [...] + c_h (Bl_o psi_r + Br_o psi_l)
            bl.p += sb * r.v[o];
            br.v[o] += sb * l.p;
            br.p += sb * l.v[o];
            bl.v[o] += sb * r.p;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                if (d == o) continue;
                const R BB = l.B[d] * r.B[d];
                bl.v[o] += sb * BB;
                br.v[o] += sb * BB;
                bl.B[d] += sb * (l.v[o] + r.v[o]) * r.B[d];
                br.B[d] += sb * (l.v[o] + r.v[o]) * l.B[d];
                // - v_l,d Bl_o Br_d - v_r,d Br_o Bl_d
                bl.v[d] -= sb * l.B[o] * r.B[d];
                bl.B[o] -= sb * l.v[d] * r.B[d];
                br.B[d] -= sb * l.v[d] * l.B[o];
                br.v[d] -= sb * r.B[o] * l.B[d];
                br.B[o] -= sb * r.v[d] * l.B[d];
                bl.B[d] -= sb * r.v[d] * r.B[o];
                // f_{5+d} = (v_l,o Bl_d - v_l,d Bl_o + v_r,o Br_d - v_r,d Br_o)/2
                const R h = R(0.5) * g[5 + d];
                bl.v[o] += h * l.B[d];
                bl.B[d] += h * l.v[o];
                bl.v[d] -= h * l.B[o];
                bl.B[o] -= h * l.v[d];
                br.v[o] += h * r.B[d];
                br.B[d] += h * r.v[o];
                br.v[d] -= h * r.B[o];
                br.B[o] -= h * r.v[d];
            }
            bl.B[o] += sb * P.c_h * r.psi + R(0.5) * P.c_h * g[8];
            br.psi += sb * P.c_h * l.B[o];
            br.B[o] += sb * P.c_h * l.psi + R(0.5) * P.c_h * g[8];
            bl.psi += sb * P.c_h * r.B[o];
            // f_{5+o} = c_h (psi_l + psi_r)/2
            bl.psi += R(0.5) * P.c_h * g[5 + o];
            br.psi += R(0.5) * P.c_h * g[5 + o];
        } else {
            R h[NV];
#pragma unroll
            for (int q = 0; q < NV; ++q) h[q] = R(0.5) * g[q];
            flux_vjp(l, o, h, bl, P);
            flux_vjp(r, o, h, br, P);
            if (KIND == FLUX_LLF || KIND == FLUX_LLF_NAIVE) {
                R t0, t1, t2, t3;
                const R cl = fast_speed(l, o, P, t0, t1, t2, t3);
                const R cr = fast_speed(r, o, P, t0, t1, t2, t3);
                const R al = dabs(l.v[o]), ar = dabs(r.v[o]);
                R lam;
                bool right_v, right_c;
                if (KIND == FLUX_LLF_NAIVE) {
                    right_v = ar > al;
                    right_c = cr > cl;
                    lam = (right_v ? ar : al) + (right_c ? cr : cl);
                } else {
                    right_v = (ar + cr) > (al + cl);
                    right_c = right_v;
                    lam = right_v ? ar + cr : al + cl;
                }
                R cvl[NV], cvr[NV], dot = R(0);
#pragma unroll
                for (int q = 0; q < NV; ++q) {
                    const R hl = R(0.5) * lam * g[q];
                    cvl[q] = hl;
                    cvr[q] = -hl;
                    dot += g[q] * (cons(r, q, P) - cons(l, q, P));
                }
                add_cons_bar(l, cvl, bl, P);
                add_cons_bar(r, cvr, br, P);
                const R lb = R(-0.5) * dot;
                if (right_v)
                    br.v[o] += lb * (r.v[o] >= R(0) ? R(1) : R(-1));
                else
                    bl.v[o] += lb * (l.v[o] >= R(0) ? R(1) : R(-1));
                if (right_c)
                    fast_speed_vjp(r, o, lb, br, P);
                else
                    fast_speed_vjp(l, o, lb, bl, P);
            }
        }
    }

    // flux_nonconservative_powell(l, r, o) is linear in (r.B_o, r.psi):
    //   f = A(l) r.B_o + C(l) r.psi,
    //   A = (0, B_l, v_l.B_l, v_l, 0), C = (0, 0, psi_l v_l,o, 0, v_l,o).
    // Reverse of sum_j c_j f(l, r_j) for cotangent g: the own side is the
    // derivative of A, C at (beta, phi) = sum_j c_j (r_j.B_o, r_j.psi); the
    // other side gets a = g.A(l) (for B_o) and c = g.C(l) (for psi).
    template <typename R>
    __device__ static void other_vals(const Node<R>& n, int o, R& bo, R& ps) {
        bo = n.B[o];
        ps = n.psi;
    }
    template <typename R>
    __device__ static void add_other(Bar<R>& b, int o, R a, R c) {
        b.B[o] += a;
        b.psi += c;
    }
    template <typename R>
    __device__ static void noncons_other(const Node<R>& l, int o, const R* g, R& a, R& c) {
        a = g[4] * dot3(l.v, l.B);
#pragma unroll
        for (int d = 0; d < 3; ++d) a += g[1 + d] * l.B[d] + g[5 + d] * l.v[d];
        c = (g[4] * l.psi + g[8]) * l.v[o];
    }
    template <typename R>
    __device__ static void noncons_own(const Node<R>& l, int o, const R* g, R beta, R phi, Bar<R>& bl) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            bl.B[d] += (g[1 + d] + g[4] * l.v[d]) * beta;
            bl.v[d] += (g[4] * l.B[d] + g[5 + d]) * beta;
        }
        bl.psi += g[4] * l.v[o] * phi;
        bl.v[o] += (g[4] * l.psi + g[8]) * phi;
    }
};

}  // namespace trixi_hip
