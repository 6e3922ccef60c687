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
    const bool series = dxy * dxy < R(1.0e-4) * (s * s);
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
    const bool series = dxy * dxy < R(1.0e-4) * (s * s);
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
        n.lrho = log((double)n.rho);
        n.lp = log((double)n.p);
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

}  // namespace trixi_hip
