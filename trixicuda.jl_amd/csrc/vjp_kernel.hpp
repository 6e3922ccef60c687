// vjp_kernel.hpp -- adjoint (vector-Jacobian product) of the fused rhs!.
//
//   ubar = beta * ubar + (d rhs / d u)^T (alpha * vbar)
//
// rhs(u)_e = -invJ_e [ volume(u_e) + surface(u_e, u_neighbours) ], so with
// w_e = -invJ_e * alpha * vbar_e (the cotangent of the element accumulator):
//  * a volume pair (i, j) along a line, acc_i += Ds[i,j] f*, acc_j += Ds[j,i] f*,
//    has flux cotangent  g = Ds[i,j] w_i + Ds[j,i] w_j;
//  * the weak form acc_i += Dhat[i,p] f(u_p) has g_p = sum_i Dhat[i,p] w_i;
//  * an interface flux F(u_L, u_R) enters acc_L (end node) with +lhat_p and
//    acc_R (first node) with -lhat_m, so g = lhat_p w_L - lhat_m w_R.
//    Each element evaluates the interfaces of its own faces and keeps only
//    its own side's cotangent, reading the neighbour's w at the face node:
//    every element writes only its own ubar, with no atomics.
// The work decomposition (M elements per workgroup, one lane per element
// line, direction phases combined through LDS, XCD-aware tiles) is that of
// rhs_kernel; cotangents are accumulated in primitive space and converted to
// conservative space once per node.
#pragma once

#include <type_traits>

#include "physics_vjp.hpp"
#include "rhs_kernel.hpp"

namespace trixi_hip {

struct VjpArgs {
    const double* u;
    const double* vbar;
    double* ubar;
    const int32_t* nbr;
    const double* invJ;
    int64_t E;
    PhysParams phys;
    double Ds[MAXN * MAXN];
    double Dhat[MAXN * MAXN];
    double lhat_m, lhat_p;
    double alpha, beta;
    int nxcd;
    // Hennemann-Gassner shock capturing
    const double* alpha_raw;  // [E] blending factors before smoothing (from u)
    double* alphabar;         // [E] out: cotangent of the smoothed blending factor
    int32_t* winner;          // [E] out: argument of the smoothing max (-1 own, else face)
    double invw[MAXN];        // inverse LGL weights (FV subcells)
    double atol;              // pure flux-differencing threshold
    int smooth;
    // non-periodic meshes: boundary faces (nbr < 0) with Dirichlet states
    const double* x;          // node coordinates x[(d*NN + node)*E + e]
    IcParams bc_ic;
    int bc;
    double t;                 // time of the linearisation point
    int64_t pd_lanes;         // meshes with fewer element lines run the direction-group form
    // reverse RK stage: the next (earlier) stage's register cotangent
    // kbar_out = kbar_a * vbar + kbar_b * ubar (ubar as this call leaves it),
    // written with each node's ubar (with shock capturing the reverse
    // indicator then adds kbar_b times its own ubar increments); nullptr: none
    double* kbar_out;
    double kbar_a, kbar_b;
};

#ifndef TRIXI_HIP_VJP_LDS_KIB
#define TRIXI_HIP_VJP_LDS_KIB 64
#endif
#ifndef TRIXI_HIP_VJP_LDS_ONE_GROUP_KIB
#define TRIXI_HIP_VJP_LDS_ONE_GROUP_KIB 128
#endif

// Workgroup width of the adjoint kernel (FP64), a power of two, with GA LDS
// accumulators: node records with 1/rho and 1/p, FP64 logarithms, the scaled
// cotangents and the accumulators within TRIXI_HIP_VJP_LDS_KIB, or twice that
// for the direction groups of systems of more than five variables (the
// nine-variable MHD adjoint: whole waves per direction group; gfx950 gives a
// workgroup up to 160 KiB of LDS).  The one-group form (large meshes) takes
// up to 128 KiB: 3D Euler at level 5 with 8-element tiles, VJP 165 -> 159 us
// (the direction-group forms of the small meshes measured slower with wider
// tiles).
template <int DIM, int N, int NV, int NRF, int G, int GA>
__host__ __device__ constexpr int vjp_width() {
    constexpr int NL = cpow(N, DIM - 1), NN = cpow(N, DIM);
    constexpr int per_elem = NN * (NRF * 8 + 2 * 8 + (1 + GA) * NV * 8);
    constexpr int budget =
        (G == 1 ? TRIXI_HIP_VJP_LDS_ONE_GROUP_KIB : (NV > 5 ? 2 : 1) * TRIXI_HIP_VJP_LDS_KIB) * 1024;
    int m = 1;
    while (2 * m * NL * G <= 256) m *= 2;
    while (m > 1 && m * per_elem > budget) m /= 2;
    return m;
}

// Surface groups of the direction-group form: beside the direction groups'
// volume terms, more lane groups evaluate the interface fluxes, each into an
// LDS accumulator of its own.  A workgroup of the nine-variable MHD adjoint
// (three 64-lane direction groups, one workgroup per CU) then keeps all four
// SIMDs of its CU busy: one face group for every direction, 3D GLM-MHD at
// level 4, VJP 69.1 -> 54.4 us (the + faces alone in it, the direction groups
// keeping the - faces, 60.3 us).  For systems of up to five variables one
// group per direction takes that direction's faces (and with shock capturing
// its FV subcells), leaving the direction groups the volume pairs.
// TRIXI_HIP_VJP_FACE_GROUP: 2, per-direction groups for up to five
// variables where they fit, else one face group; 1, one face group; 0, none.
// They apply where the extra accumulators leave the tile width unchanged,
// and the kernel then takes the reloading form (LDS accumulators, zeroed in
// the node pass).
#ifndef TRIXI_HIP_VJP_FACE_GROUP
#define TRIXI_HIP_VJP_FACE_GROUP 2
#endif
// the number of surface groups: 0, 1 (one face group) or G (per direction)
template <int DIM, int N, int VOL, int NV, int NRF, int G>
__host__ __device__ constexpr int vjp_surface_groups() {
    if (!(G > 1 && VOL != VI_WEAK) || TRIXI_HIP_VJP_FACE_GROUP == 0) return 0;
    constexpr int w = vjp_width<DIM, N, NV, NRF, G, G>();
    if (TRIXI_HIP_VJP_FACE_GROUP >= 2 && NV <= 5 && vjp_width<DIM, N, NV, NRF, G, 2 * G>() == w) return G;
    if (vjp_width<DIM, N, NV, NRF, G, G + 1>() == w) return 1;
    return 0;
}
// lane groups of a workgroup: G, plus the surface groups
template <int DIM, int N, int VOL, int NV, int NRF, int G>
__host__ __device__ constexpr int vjp_groups() {
    return G + vjp_surface_groups<DIM, N, VOL, NV, NRF, G>();
}
template <int DIM, int N, int VOL, int NV, int NRF, int G = 1>
__host__ __device__ constexpr int pick_m_vjp() {
    return vjp_width<DIM, N, NV, NRF, G, vjp_groups<DIM, N, VOL, NV, NRF, G>()>();
}

// G direction groups: G = 1 runs the directions one after the other in every
// lane; G = DIM gives each direction its own lanes (M * N^(d-1) per group)
// and its own LDS accumulator, summed in direction order at the end, for
// meshes too small to fill the GPU with one lane per element line.
template <int DIM, int N, class EQ, int VOL, int VF, int SF, int FVF, typename R, int M, bool EXT, int G>
__global__ void __launch_bounds__((M * cpow(N, DIM - 1) * vjp_groups<DIM, N, VOL, EQ::NV, EQ::NRF, G>()))
vjp_kernel(const VjpArgs a) {
    constexpr int NL = cpow(N, DIM - 1);
    constexpr int NN = cpow(N, DIM);
    constexpr int NV = EQ::NV;
    constexpr int NRF = EQ::NRF;
    constexpr bool PD = G > 1;
    static_assert(G == 1 || G == DIM, "one group, or one group per direction");
    constexpr int NS = vjp_surface_groups<DIM, N, VOL, NV, NRF, G>();
    constexpr int GT = G + NS;  // lane groups (and LDS accumulators)
    // one face group lies within one wave (its directions' accumulations
    // into shared face nodes stay in program order)
    static_assert(NS != 1 || (M * NL <= 64 && 64 % (M * NL) == 0), "face group within one wave");
    constexpr int NT = M * NL * GT;  // workgroup size
    using Node = typename EQ::template Node<R>;
    using Bar = typename EQ::template Bar<R>;
    // flux differencing with node records of more than 10 values, shock
    // capturing (which also recomputes the forward fluxes) and the forms
    // with surface groups reload the nodes per pair and accumulate
    // cotangents in LDS
    constexpr bool RELOAD = VOL != VI_WEAK && (NRF > 10 || VOL == VI_HG || NS > 0);

    __shared__ R rec[NRF * NN * M];
    __shared__ double recd[2 * NN * M];
    __shared__ R ws[NV * NN * M];   // w = -invJ alpha vbar of the tile
    __shared__ R acc[GT * NV * NN * M];  // primitive cotangents, summed over directions (per group)
    __shared__ double abar_s[VOL == VI_HG ? NT : 1];

    const int tid = threadIdx.x;
    const int grp = tid / (M * NL);  // direction of this lane's group (PD; >= G: surface groups)
    const int gtid = tid - grp * (M * NL);
    const int el = gtid % M;
    const int line = gtid / M;
    R* const accg = acc + (PD ? grp * NV * NN * M : 0);
    const int64_t E = a.E;
    const uint32_t E32 = (uint32_t)E;
    const int64_t slab = (int64_t)NN * E;
    const int64_t e0 = xcd_tile(a.nxcd) * M;
    if (e0 >= E) return;
    const int64_t e = e0 + el;
    const int64_t ec = e < E ? e : E - 1;
    const PhysR<R> P = PhysR<R>::from(a.phys);

    int32_t nb[2 * DIM];
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) nb[f] = a.nbr[(int64_t)f * E + ec];

    // HG: the smoothed blending factor of this lane's element, as in the
    // forward kernel, and which argument of the max produced it
    constexpr bool HG = VOL == VI_HG;
    R alpha = R(0);
    bool dg_only = true;
    int win = -1;
    double abar_part = 0.0;
    if constexpr (HG) {
        double al = a.alpha_raw[ec];
        if (a.smooth) {
#pragma unroll
            for (int f = 0; f < 2 * DIM; ++f) {
                if (EXT && nb[f] < 0) continue;  // smoothing runs over interfaces only
                const double an = 0.5 * a.alpha_raw[nb[f]];
                if (an > al) {
                    al = an;
                    win = f;
                }
            }
        }
        dg_only = fabs(al) <= a.atol;
        alpha = (R)al;
    }

    for (int q = tid; q < NN * M; q += NT) {
        const int c = q % M;
        const int n = q / M;
        int64_t ee = e0 + c;
        if (ee >= E) ee = E - 1;
        R uc[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) uc[v] = (R)a.u[soa(v, n, ee, slab, E32)];
        const Node nd = EQ::node(uc, P);
        R f[NRF];
        EQ::to_rec(nd, f);
#pragma unroll
        for (int k = 0; k < NRF; ++k) rec[(k * NN + n) * M + c] = f[k];
        recd[n * M + c] = nd.lrho;
        recd[(NN + n) * M + c] = nd.lp;
        const R wsc = (R)(-a.invJ[ee] * a.alpha);
#pragma unroll
        for (int v = 0; v < NV; ++v) ws[(v * NN + n) * M + c] = wsc * (R)a.vbar[soa(v, n, ee, slab, E32)];
        if constexpr (RELOAD)
#pragma unroll
            for (int g = 0; g < GT; ++g)
#pragma unroll
                for (int v = 0; v < NV; ++v) acc[((g * NV + v) * NN + n) * M + c] = R(0);
    }
    __syncthreads();

    const R lm = (R)a.lhat_m, lp = (R)a.lhat_p;
    auto Wn = [&](int node, int v) -> R { return ws[(v * NN + node) * M + el]; };
    auto load_node = [&](int node) -> Node {
        R f[NRF];
#pragma unroll
        for (int k = 0; k < NRF; ++k) f[k] = rec[(k * NN + node) * M + el];
        return EQ::template from_rec<R>(f, recd[node * M + el], recd[(NN + node) * M + el]);
    };
    // acc (+)= b at a node of this lane's element; the first direction of
    // the group stores
    auto flush = [&](int d, int node, const Bar& b) {
        R t[NV];
        EQ::bar_to_array(b, t);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const int idx = (v * NN + node) * M + el;
            accg[idx] = (PD || d == 0) ? t[v] : accg[idx] + t[v];
        }
    };
    auto add_acc = [&](int node, const Bar& b) {
        R t[NV];
        EQ::bar_to_array(b, t);
#pragma unroll
        for (int v = 0; v < NV; ++v) accg[(v * NN + node) * M + el] += t[v];
    };
    // Interface at the -/+ face of the line: neighbour face node state and
    // cotangent, then the flux VJP keeping only this element's side.
    // A boundary face (EXT, nbr < 0) has a Dirichlet state that does not
    // depend on u, and no neighbour cotangent.
    auto face = [&](int d, int base, int s, bool plus, const Node& own, Bar& b) {
        const int32_t fnb = nb[2 * d + (plus ? 1 : 0)];
        const bool bnd = EXT && fnb < 0;
        const int64_t fn = bnd ? ec : fnb;
        const R wn = bnd ? R(0) : (R)(-a.invJ[fn] * a.alpha);
        const int nnode = plus ? base : base + (N - 1) * s;
        const int onode = plus ? base + (N - 1) * s : base;
        R un[NV], g[NV], go[NV], gn[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const int64_t r = soa(v, nnode, fn, slab, E32);
            un[v] = (R)a.u[r];
            // cotangents of this element's and the neighbour's face flux
            go[v] = (plus ? lp : -lm) * Wn(onode, v);
            gn[v] = (plus ? -lm : lp) * wn * (R)a.vbar[r];
            g[v] = go[v] + gn[v];
        }
        if constexpr (EXT) {
            if (bnd) {
                double xx[3] = {0.0, 0.0, 0.0};
#pragma unroll
                for (int dd = 0; dd < DIM; ++dd) xx[dd] = a.x[((int64_t)dd * NN + onode) * E + ec];
                double uu[5];
                boundary_value(a.bc_ic, xx, a.t, uu);
#pragma unroll
                for (int v = 0; v < NV; ++v) un[v] = (R)uu[v];
            }
        }
        const Node ndn = EQ::node(un, P);
        Bar dummy;
        EQ::zero(dummy);
        if (plus)
            EQ::template two_point_vjp<SF>(own, ndn, d, g, b, dummy, P);
        else
            EQ::template two_point_vjp<SF>(ndn, own, d, g, dummy, b, P);
        if constexpr (EQ::NONCONS) {
            // each side adds half the nonconservative flux with itself first:
            // own term noncons(own, nbr), neighbour term noncons(nbr, own)
            R bo, ps, ca, cc;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                go[v] *= R(0.5);
                gn[v] *= R(0.5);
            }
            EQ::other_vals(ndn, d, bo, ps);
            EQ::noncons_own(own, d, go, bo, ps, b);
            EQ::noncons_other(ndn, d, gn, ca, cc);
            EQ::add_other(b, d, ca, cc);
        }
    };

    // FV subcells of the shock-capturing blend along this lane's line in
    // direction d: acc_p += alpha invw_p F, acc_{p+1} -= alpha invw_{p+1} F
    auto fv_block = [&](auto dc) {
        constexpr int d = decltype(dc)::value;
        const int s = cpow(N, d);
        const int base = line_base<DIM, N>(line, d);
        auto W = [&](int p, int v) -> R { return Wn(base + p * s, v); };
        const bool blend = HG && !dg_only;
        if constexpr (HG) {
            if (blend) {
#pragma unroll
                for (int p = 0; p + 1 < N; ++p) {
                    asm volatile("" ::: "memory");
                    const Node np0 = load_node(base + p * s);
                    const Node np1 = load_node(base + (p + 1) * s);
                    R h[NV], F[NV];
                    EQ::template two_point_value<FVF>(np0, np1, d, P, F);
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        h[v] = (R)a.invw[p] * W(p, v) - (R)a.invw[p + 1] * W(p + 1, v);
                        abar_part += (double)(h[v] * F[v]);
                        h[v] *= alpha;
                    }
                    Bar b0, b1;
                    EQ::zero(b0);
                    EQ::zero(b1);
                    EQ::template two_point_vjp<FVF>(np0, np1, d, h, b0, b1, P);
                    add_acc(base + p * s, b0);
                    add_acc(base + (p + 1) * s, b1);
                }
            }
        }
    };

    // one phase per direction; the direction is a compile-time constant so
    // that every component index (v[d], B[d], g[1 + d]) stays static
    auto phase = [&](auto dc) {
        constexpr int d = decltype(dc)::value;
        const int s = cpow(N, d);
        const int base = line_base<DIM, N>(line, d);
        auto W = [&](int p, int v) -> R { return Wn(base + p * s, v); };

        if constexpr (RELOAD) {
            // Large node records: the pair's nodes are reloaded from LDS (the
            // empty asm with a memory clobber keeps the compiler from merging
            // the reloads back into N live records) and cotangents go straight
            // to the LDS accumulator, zeroed in the node pass.
            // HG: acc = (1 - alpha) V_DG + alpha V_FV + surface, so the DG
            // cotangents scale by 1 - alpha and alphabar = <w, V_FV - V_DG>
            const bool blend = HG && !dg_only;
            const R sc = blend ? R(1) - alpha : R(1);
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const Node ni = load_node(base + i * s);
                Bar bi;
                EQ::zero(bi);
#pragma unroll
                for (int j = i + 1; j < N; ++j) {
                    asm volatile("" ::: "memory");
                    const Node nj = load_node(base + j * s);
                    const R fa = (R)a.Ds[i * N + j], fb = (R)a.Ds[j * N + i];
                    R g[NV];
#pragma unroll
                    for (int v = 0; v < NV; ++v) g[v] = fa * W(i, v) + fb * W(j, v);
                    if constexpr (HG) {
                        if (blend) {
                            R f[NV];
                            EQ::template two_point_value<VF>(ni, nj, d, P, f);
#pragma unroll
                            for (int v = 0; v < NV; ++v) abar_part -= (double)(g[v] * f[v]);
                        }
#pragma unroll
                        for (int v = 0; v < NV; ++v) g[v] *= sc;
                    }
                    Bar bj;
                    EQ::zero(bj);
                    EQ::template two_point_vjp<VF>(ni, nj, d, g, bi, bj, P);
                    add_acc(base + j * s, bj);
                }
                add_acc(base + i * s, bi);
            }
            if constexpr (NS != G) fv_block(dc);
            if constexpr (EQ::NONCONS) {
                R ca[N], cc[N];
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    asm volatile("" ::: "memory");
                    R beta = R(0), phi = R(0);
#pragma unroll
                    for (int j = 0; j < N; ++j) {
                        R bo, ps;
                        EQ::other_vals(load_node(base + j * s), d, bo, ps);
                        beta += (R)a.Ds[i * N + j] * bo;
                        phi += (R)a.Ds[i * N + j] * ps;
                    }
                    const Node ni = load_node(base + i * s);
                    R g[NV];
#pragma unroll
                    for (int v = 0; v < NV; ++v) g[v] = R(0.5) * W(i, v);
                    Bar bi;
                    EQ::zero(bi);
                    EQ::noncons_own(ni, d, g, beta, phi, bi);
                    EQ::noncons_other(ni, d, g, ca[i], cc[i]);
                    add_acc(base + i * s, bi);
                }
#pragma unroll
                for (int j = 0; j < N; ++j) {
                    R sa = R(0), sc = R(0);
#pragma unroll
                    for (int i = 0; i < N; ++i) {
                        sa += (R)a.Ds[i * N + j] * ca[i];
                        sc += (R)a.Ds[i * N + j] * cc[i];
                    }
                    Bar t;
                    EQ::zero(t);
                    EQ::add_other(t, d, sa, sc);
                    add_acc(base + j * s, t);
                }
            }
            if constexpr (NS == 0) {
                asm volatile("" ::: "memory");
                Bar b;
                EQ::zero(b);
                face(d, base, s, false, load_node(base), b);
                add_acc(base, b);
            }
            if constexpr (NS == 0) {
                asm volatile("" ::: "memory");
                Bar b;
                EQ::zero(b);
                face(d, base, s, true, load_node(base + (N - 1) * s), b);
                add_acc(base + (N - 1) * s, b);
            }
        } else {
            Bar bb[N];
#pragma unroll
            for (int p = 0; p < N; ++p) EQ::zero(bb[p]);
            Node nd[N];
#pragma unroll
            for (int p = 0; p < N; ++p) nd[p] = load_node(base + p * s);
            if (VOL == VI_WEAK) {
#pragma unroll
                for (int p = 0; p < N; ++p) {
                    R g[NV];
#pragma unroll
                    for (int v = 0; v < NV; ++v) g[v] = R(0);
#pragma unroll
                    for (int i = 0; i < N; ++i) {
                        const R dh = (R)a.Dhat[i * N + p];
#pragma unroll
                        for (int v = 0; v < NV; ++v) g[v] += dh * W(i, v);
                    }
                    EQ::flux_vjp(nd[p], d, g, bb[p], P);
                }
            } else {
#pragma unroll
                for (int i = 0; i < N; ++i)
#pragma unroll
                    for (int j = i + 1; j < N; ++j) {
                        const R fa = (R)a.Ds[i * N + j], fb = (R)a.Ds[j * N + i];
                        R g[NV];
#pragma unroll
                        for (int v = 0; v < NV; ++v) g[v] = fa * W(i, v) + fb * W(j, v);
                        EQ::template two_point_vjp<VF>(nd[i], nd[j], d, g, bb[i], bb[j], P);
                    }
                if constexpr (EQ::NONCONS) {
                    // acc_i += 1/2 sum_j Ds[i,j] noncons(u_i, u_j), linear in u_j's (B_d, psi)
                    R ca[N], cc[N];
#pragma unroll
                    for (int i = 0; i < N; ++i) {
                        R beta = R(0), phi = R(0);
#pragma unroll
                        for (int j = 0; j < N; ++j) {
                            R bo, ps;
                            EQ::other_vals(nd[j], d, bo, ps);
                            beta += (R)a.Ds[i * N + j] * bo;
                            phi += (R)a.Ds[i * N + j] * ps;
                        }
                        R g[NV];
#pragma unroll
                        for (int v = 0; v < NV; ++v) g[v] = R(0.5) * W(i, v);
                        EQ::noncons_own(nd[i], d, g, beta, phi, bb[i]);
                        EQ::noncons_other(nd[i], d, g, ca[i], cc[i]);
                    }
#pragma unroll
                    for (int j = 0; j < N; ++j) {
                        R sa = R(0), sc = R(0);
#pragma unroll
                        for (int i = 0; i < N; ++i) {
                            sa += (R)a.Ds[i * N + j] * ca[i];
                            sc += (R)a.Ds[i * N + j] * cc[i];
                        }
                        EQ::add_other(bb[j], d, sa, sc);
                    }
                }
            }
#pragma unroll
            for (int p = 1; p < N - 1; ++p) flush(d, base + p * s, bb[p]);
            {
                const Node own = load_node(base);
                face(d, base, s, false, own, bb[0]);
            }
            flush(d, base, bb[0]);
            {
                const Node own = load_node(base + (N - 1) * s);
                face(d, base, s, true, own, bb[N - 1]);
            }
            flush(d, base + (N - 1) * s, bb[N - 1]);
        }
        // one group: the next direction adds to nodes other lanes stored
        if constexpr (!PD) __syncthreads();
    };
    // (with groups, each runs its own direction into its own accumulator;
    // no barrier inside a phase, one after all)
    if (!PD || grp == 0) phase(std::integral_constant<int, 0>{});
    if constexpr (DIM > 1) {
        if (!PD || grp == 1) phase(std::integral_constant<int, 1>{});
    }
    if constexpr (DIM > 2) {
        if (!PD || grp == 2) phase(std::integral_constant<int, 2>{});
    }
    if constexpr (NS != 0) {
        // surface groups: the interfaces of this lane's line (and with one
        // group per direction the FV subcells), into their own accumulators
        auto faces_of_line = [&](auto dc) {
            constexpr int d = decltype(dc)::value;
            const int s = cpow(N, d);
            const int base = line_base<DIM, N>(line, d);
            {
                asm volatile("" ::: "memory");
                Bar b;
                EQ::zero(b);
                face(d, base, s, false, load_node(base), b);
                add_acc(base, b);
            }
            {
                asm volatile("" ::: "memory");
                Bar b;
                EQ::zero(b);
                face(d, base, s, true, load_node(base + (N - 1) * s), b);
                add_acc(base + (N - 1) * s, b);
            }
        };
        if constexpr (NS == 1) {
            if (grp == G) {
                faces_of_line(std::integral_constant<int, 0>{});
                if constexpr (DIM > 1) faces_of_line(std::integral_constant<int, 1>{});
                if constexpr (DIM > 2) faces_of_line(std::integral_constant<int, 2>{});
            }
        } else {
            if (grp == G) {
                fv_block(std::integral_constant<int, 0>{});
                faces_of_line(std::integral_constant<int, 0>{});
            }
            if constexpr (DIM > 1) {
                if (grp == G + 1) {
                    fv_block(std::integral_constant<int, 1>{});
                    faces_of_line(std::integral_constant<int, 1>{});
                }
            }
            if constexpr (DIM > 2) {
                if (grp == G + 2) {
                    fv_block(std::integral_constant<int, 2>{});
                    faces_of_line(std::integral_constant<int, 2>{});
                }
            }
        }
    }
    if constexpr (PD) __syncthreads();

    if constexpr (HG) {
        // alphabar of each element: sum of its lines' (and groups') parts in
        // a fixed order
        abar_s[tid] = abar_part;
        __syncthreads();
        if (grp == 0 && line == 0 && e < E) {
            double sum = 0.0;
            for (int g = 0; g < GT; ++g)
                for (int l = 0; l < NL; ++l) sum += abar_s[g * M * NL + l * M + el];
            a.alphabar[e] = dg_only ? 0.0 : sum;
            a.winner[e] = win;
        }
    }

    // (d prim / d cons)^T per node, ubar = beta ubar + result
    for (int q = tid; q < NN * M; q += NT) {
        const int c = q % M;
        const int n = q / M;
        const int64_t ee = e0 + c;
        if (ee >= E) continue;
        R f[NRF];
#pragma unroll
        for (int k = 0; k < NRF; ++k) f[k] = rec[(k * NN + n) * M + c];
        const Node nd = EQ::template from_rec<R>(f, recd[n * M + c], recd[(NN + n) * M + c]);
        R t[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            t[v] = acc[(v * NN + n) * M + c];
#pragma unroll
            for (int g = 1; g < GT; ++g) t[v] += acc[((g * NV + v) * NN + n) * M + c];
        }
        Bar tot;
        EQ::array_to_bar(t, tot);
        R ub[NV];
        EQ::to_cons_bar(nd, tot, ub, P);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const int64_t g = soa(v, n, ee, slab, E32);
            const double prev = a.beta != 0.0 ? a.beta * a.ubar[g] : 0.0;
            const double ubn = prev + (double)ub[v];
            a.ubar[g] = ubn;
            // (vbar is not written by this kernel: the neighbours' reads of
            // it stay valid, kbar_out being another buffer)
            if (a.kbar_out != nullptr) a.kbar_out[g] = a.kbar_a * a.vbar[g] + a.kbar_b * ubn;
        }
    }
}

// M: elements per workgroup of the one-group form; MG: of the
// direction-group form, used for meshes whose one-group form would launch
// fewer lanes than a.pd_lanes (on 2D and 3D meshes).
template <int DIM, int N, class EQ, int VOL, int VF, int SF, int FVF, typename R, int M, bool EXT = false,
          int MG = M>
hipError_t launch_vjp(const VjpArgs& a, hipStream_t stream) {
    constexpr int NL = cpow(N, DIM - 1);
    if constexpr (DIM > 1) {
        if (a.E * NL < a.pd_lanes) {
            int64_t blocks = (a.E + MG - 1) / MG;
            if (a.nxcd > 1) blocks = (blocks + a.nxcd - 1) / a.nxcd * a.nxcd;
            hipLaunchKernelGGL((vjp_kernel<DIM, N, EQ, VOL, VF, SF, FVF, R, MG, EXT, DIM>), dim3((unsigned)blocks),
                               dim3(MG * NL * vjp_groups<DIM, N, VOL, EQ::NV, EQ::NRF, DIM>()), 0, stream, a);
            return hipGetLastError();
        }
    }
    int64_t blocks = (a.E + M - 1) / M;
    if (a.nxcd > 1) blocks = (blocks + a.nxcd - 1) / a.nxcd * a.nxcd;
    hipLaunchKernelGGL((vjp_kernel<DIM, N, EQ, VOL, VF, SF, FVF, R, M, EXT, 1>), dim3((unsigned)blocks), dim3(M * NL),
                       0, stream, a);
    return hipGetLastError();
}

}  // namespace trixi_hip
