// rhs_kernel.hpp -- the fused DGSEM rhs! kernel for gfx950.
//
// One launch computes, for every element of a periodic TreeMesh,
//   volume integral (weak form, flux differencing incl. nonconservative
//   terms, or Hennemann-Gassner DG/FV blending)
//   + surface integral of the numerical interface fluxes
//   + Jacobian scaling du *= -1/J,
// i.e. the reference stages cuda_volume_integral!, cuda_prolong2interfaces!,
// cuda_interface_flux!, cuda_surface_integral! and cuda_jacobian!
// (src/solvers/dg_3d.jl:57-102, 361-412, 846-873 and the 1D/2D twins) in a
// single pass over HBM.  Optionally the low-storage Runge-Kutta register
// update is fused in as an epilogue (FP64): k = a k + dt du, u_out = u + b k.
//
// Work decomposition.  A workgroup owns M consecutive elements (Trixi order,
// element-innermost SoA so that lanes of a wave read consecutive addresses).
// Lane (el, line) owns one line of N nodes through element el: for d =
// 0..DIM-1 it evaluates the N(N-1)/2 symmetric two-point fluxes along its
// line in direction d, the two interface fluxes at the line's end faces
// (recomputed on both sides of each interface, so no face buffer ever goes
// to HBM), and accumulates the line's N nodes in registers.  Directions are
// combined through an LDS accumulator: the lines of one direction partition
// the element's nodes, so each phase is race free between barriers.
#pragma once

#include <type_traits>

#include "indicator.hpp"
#include "initial_conditions.hpp"
#include "physics.hpp"

namespace trixi_hip {

typedef float f32x4 __attribute__((ext_vector_type(4)));

#ifndef TRIXI_HIP_WEAK_MFMA
#define TRIXI_HIP_WEAK_MFMA 1  // mixed-precision weak form: D-hat on MFMA
#endif

constexpr int MAXN = 8;

// Tuning knobs (compile-time): LDS row padding in elements and a minimum
// number of waves per SIMD requested from the register allocator.
#ifndef TRIXI_HIP_LDS_PAD
#define TRIXI_HIP_LDS_PAD 0
#endif
#ifndef TRIXI_HIP_MIN_WAVES
#define TRIXI_HIP_MIN_WAVES 1
#endif
// Fused RK stage: 1 keeps u and k of the node pass in registers for the
// register update in the mixed mode (FP32 arithmetic, FP64 state), where
// that measured faster, and reads them again there in FP64 and with an FP32
// state (fewer registers: 3D Euler level 5, RK step 0.588 -> 0.573 ms FP64,
// 0.340 -> 0.330 ms FP32 state, against 0.476 -> 0.498 ms mixed); 0 always
// reads them again.
#ifndef TRIXI_HIP_RK_KEEP
#define TRIXI_HIP_RK_KEEP 1
#endif
// Fused RK stage of the sequential form with an FP32 state (u and k read
// again): 1 requests them once the lane's last direction is computed, before
// the barrier ahead of the update, so that their latency overlaps the wait;
// 0 reads them in the update itself.  3D Euler level 5, FP32-state RK step
// 0.306 -> 0.260 ms (three rounds on one box).  The FP64 parallel-direction
// form measured 0.578 -> 0.588 ms with it and reads them in the update.
#ifndef TRIXI_HIP_RK_PREFETCH
#define TRIXI_HIP_RK_PREFETCH 1
#endif
// 1: the node pass's state loads are issued before the face loads (systems
// of up to five variables); 0: each (node, element) pair is loaded as the
// node pass reaches it.
#ifndef TRIXI_HIP_EARLY_LOADS
#define TRIXI_HIP_EARLY_LOADS 1
#endif
// 1: node records are computed once per node into LDS; 0: every lane
// computes the node data of its line in each direction phase (no record LDS).
#ifndef TRIXI_HIP_STAGE_NODES
#define TRIXI_HIP_STAGE_NODES 1
#endif
// Workgroup size of the node-parallel form (lanes; whole elements of
// N^d * d lanes each)
#ifndef TRIXI_HIP_NP_THREADS
#define TRIXI_HIP_NP_THREADS 256
#endif
// Direction phases in parallel: one lane group (whole waves) per direction,
// combined through the LDS accumulator in direction order.
#ifndef TRIXI_HIP_PAR_DIRS
#define TRIXI_HIP_PAR_DIRS 1
#endif
// ... also in the mixed (FP32) mode
#ifndef TRIXI_HIP_PAR_DIRS_F32
#define TRIXI_HIP_PAR_DIRS_F32 0
#endif
// ... also in the fused RK stage (FP64: 256 -> 198 VGPRs, 1 -> 2 waves per
// SIMD; 3D Euler level 5 RK step 0.663 -> 0.588 ms)
#ifndef TRIXI_HIP_PAR_DIRS_RK
#define TRIXI_HIP_PAR_DIRS_RK 1
#endif

struct RhsArgs {
    // state vectors of the kernel's state type S (double, or float for the
    // FP32-state precision)
    const void* u;            // SoA state  u[(v*NN + node)*E + e]
    void* du;                 // SoA output (rhs mode)
    void* k;                  // 2N register (RK mode)
    void* u_out;              // RK mode: u_out = u + b k
    const double* time;       // RK mode: device scalars {t, dt}
    const int32_t* nbr;       // [2*DIM][E] neighbour element per face (-x,+x,-y,+y,-z,+z)
    const double* invJ;       // [E]
    const double* alpha_raw;  // [E] HG blending factors before smoothing
    int64_t E;
    PhysParams phys;
    double Ds[MAXN * MAXN];   // derivative_split (row-major)
    double Dhat[MAXN * MAXN]; // derivative_dhat (row-major)
    double invw[MAXN];        // inverse weights
    double lhat_m, lhat_p;    // boundary_interpolation[1,1], [N,2]
    double atol;              // HG pure-DG threshold
    int smooth;               // HG smoothing
    double rk_a, rk_b;        // RK stage coefficients (rk_a unused when rk_first)
    int rk_first;
    int nxcd;                 // XCDs the dispatcher deals workgroups to round robin
    // non-periodic meshes and source terms
    const double* x;          // node coordinates x[(d*NN + node)*E + e]
    IcParams bc_ic;           // Dirichlet boundary state (faces with nbr < 0)
    int bc;                   // enum trixi_hip_boundary_condition
    int src;                  // enum trixi_hip_source_terms
    int outflow;              // bit f: boundary face direction f takes the inner state
    double t_host;            // rhs mode: time t
    // L2 mortars: a face with nbr = -2 - slot takes its surface flux from
    // mflux[(slot*NL + face node)*NV + v]; mnb[slot*mK + k] lists the elements
    // across that face (-1 padded) for the blending-factor smoothing
    const double* mflux;
    const int32_t* mnb;
    int mK;
    // node-parallel form (small periodic conforming meshes, chosen on the host)
    int np;
    // fused RK stage of the node-parallel form with shock capturing
    // (compressible Euler): the raw blending factors of u_out, the next
    // stage's input, go to alpha_next (nullptr: not computed), with the
    // indicator's inverse Vandermonde and parameters
    double* alpha_next;
    const double* Vinv;
    double ind_threshold, ind_s, ind_alpha_min, ind_alpha_max;
    int ind_var;
};

// Dirichlet state boundary_ic(x, t) at node `node` of element e (its face
// node on a boundary face), in R.
template <int DIM, int NN, int NV, typename R>
__device__ __forceinline__ void boundary_state(const RhsArgs& a, int64_t e, int node, double t, R* ub) {
    double xx[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) xx[d] = a.x[((int64_t)d * NN + node) * a.E + e];
    double uu[5];
    boundary_value(a.bc_ic, xx, t, uu);
#pragma unroll
    for (int v = 0; v < NV; ++v) ub[v] = (R)uu[v];
}

// source_terms_convergence_test at node `node` of element e, FP64
// (compressible Euler, or the inviscid Burgers equation: the only scalar
// equation built with source terms).
template <int DIM, int NN, int NV>
__device__ __forceinline__ void source_at(const RhsArgs& a, int64_t e, int node, double t, double* sv) {
    double xx[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) xx[d] = a.x[((int64_t)d * NN + node) * a.E + e];
    if constexpr (NV == 1) {
        sv[0] = source_burgers_convergence_test(xx[0], t);
        return;
    }
    if (a.src == TRIXI_HIP_SOURCE_POISSON_NONPERIODIC || a.src == TRIXI_HIP_SOURCE_HARMONIC) {
        // hyperbolic diffusion: the relaxation -q / T_r reads this node's q
        double qv[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int d = 0; d < DIM; ++d) qv[d] = static_cast<const double*>(a.u)[((int64_t)(1 + d) * NN + node) * a.E + e];
        double s4[4];
        source_hypdiff(DIM, a.src, a.phys.nu, a.phys.inv_Tr, xx, qv, s4);
#pragma unroll
        for (int v = 0; v < NV && v < 4; ++v) sv[v] = s4[v];
        return;
    }
    double s5[5];
    if (a.bc_ic.eq == EQ_SHALLOW_WATER)
        source_shallow_water_convergence_test(a.phys.grav, xx, t, s5);
    else if (a.bc_ic.eq == EQ_EULER_QUASI1D)
        source_quasi1d_convergence_test(a.phys.gamma, xx, t, s5);
    else if (a.src == TRIXI_HIP_SOURCE_EOC_TEST_EULER)
        source_eoc_test_euler(DIM, a.phys.gm1, xx, t, s5);
    else
        source_convergence_test(DIM, a.phys.gm1, xx, t, s5);
#pragma unroll
    for (int v = 0; v < NV; ++v) sv[v] = s5[v];
}

// Workgroup -> element tile.  The dispatcher deals consecutive workgroups
// to the XCDs round robin, and each XCD has its own L2; remapping so that
// XCD x processes the contiguous tiles [x*per, (x+1)*per) keeps every cache
// line of the element-innermost state, and the face neighbours of a tile
// (nearby in Morton order), inside one L2.  The grid is launched with
// nxcd*per workgroups; tiles past the end exit before any barrier.
__device__ __forceinline__ int64_t xcd_tile(int nxcd) {
    const int64_t bid = blockIdx.x, nb = gridDim.x;
    if (nxcd <= 1) return bid;
    const int64_t per = nb / nxcd;
    return (bid % nxcd) * per + bid / nxcd;
}

// Line geometry: base node of line `line` in direction d, the other
// directions enumerated in increasing order with the first one fastest
// (the face-node order of Trixi's interface arrays, containers_3d.jl:58).
template <int DIM, int N>
__device__ __forceinline__ int line_base(int line, int d) {
    int base = 0, rem = line;
#pragma unroll
    for (int dd = 0; dd < DIM; ++dd) {
        if (dd == d) continue;
        base += (rem % N) * cpow(N, dd);
        rem /= N;
    }
    return base;
}

// Lane groups of the rhs kernel: with parallel directions each direction has
// its own M * N^(d-1) lanes (whole waves), otherwise one group runs the
// directions one after the other.  Parallel directions triple the waves per
// element, which pays where LDS or registers limit the occupancy of the
// sequential form: the FP64 rhs! (3D Euler flux differencing at level 5:
// 93.8 -> 89.5 us) and the FP64 fused RK stage (fewer registers than the
// sequential form's 256).  In FP32 arithmetic (63.4 -> 76.8 us) the
// sequential form stays, with one copy of each direction's code live at a
// time, and so it does for the nine-variable MHD system, whose parallel form
// spills (3D GLM-MHD at level 4: 28.6 -> 45.5 us).
template <int DIM, int N, int M, typename R, bool RK, int NV>
__host__ __device__ constexpr bool rhs_par_dirs() {
    return TRIXI_HIP_PAR_DIRS != 0 && DIM > 1 && (M * cpow(N, DIM - 1)) % 64 == 0 &&
           (std::is_same<R, double>::value || TRIXI_HIP_PAR_DIRS_F32 != 0) && (!RK || TRIXI_HIP_PAR_DIRS_RK != 0) &&
           NV <= 5;
}
template <int DIM, int N, int M, typename R, bool RK, int NV>
__host__ __device__ constexpr int rhs_threads() {
    return M * cpow(N, DIM - 1) * (rhs_par_dirs<DIM, N, M, R, RK, NV>() ? DIM : 1);
}

// EXT: the variant with boundary faces (nbr = -1, Dirichlet states), L2
// mortar faces (nbr = -2 - slot, fluxes from the mortar pass) and source
// terms; the periodic, conforming, source-free kernels carry none of that code.
template <int DIM, int N, class EQ, int VOL, int VF, int SF, int FVF, typename R, int M, bool RK, bool EXT,
          typename L, typename S>
__global__ void __launch_bounds__((rhs_threads<DIM, N, M, R, RK, EQ::NV>()), TRIXI_HIP_MIN_WAVES)
rhs_kernel(const RhsArgs a) {
    static_assert(!EXT || std::is_same<S, double>::value, "boundaries, mortars and sources take FP64 states");
    const S* __restrict__ U = static_cast<const S*>(a.u);
    S* __restrict__ DU = static_cast<S*>(a.du);
    S* __restrict__ K = static_cast<S*>(a.k);
    S* __restrict__ UO = static_cast<S*>(a.u_out);
    constexpr int NL = cpow(N, DIM - 1);
    constexpr int NN = cpow(N, DIM);
    // PD: lane group g (whole waves) evaluates direction g; the groups meet in
    // the LDS accumulator in direction order (x stores, y adds, z adds and
    // writes), the same summation order as the sequential phases.
    constexpr bool PD = rhs_par_dirs<DIM, N, M, R, RK, EQ::NV>();
    constexpr int NG = PD ? DIM : 1;
    constexpr int NT = M * NL * NG;         // workgroup size
    constexpr int NIT = (N + NG - 1) / NG;  // node-pass (node, element) pairs per lane
    constexpr int NV = EQ::NV;
    constexpr int NRF = EQ::NRF;
    constexpr bool SF_LOGS = flux_needs_logs(SF);
    constexpr bool VF_LOGS = (VOL != VI_WEAK) && flux_needs_logs(VF);
    constexpr bool FV_LOGS = (VOL == VI_HG) && flux_needs_logs(FVF);
    constexpr bool NODE_LOGS = SF_LOGS || VF_LOGS || FV_LOGS;
    constexpr int NDF = NODE_LOGS ? 2 : 0;
    using Node = typename EQ::template Node<R, L>;
    // mixed precision weak form with N = 4 on whole waves: D-hat on MFMA
    constexpr bool WEAK_MFMA = TRIXI_HIP_WEAK_MFMA && VOL == VI_WEAK && std::is_same<R, float>::value &&
                               N == 4 && (M * NL) % 64 == 0;

    // LDS: node records (R fields and FP64 logarithms) of the M elements,
    // field-major with the element innermost, and the direction accumulator.
    constexpr int LDM = M + TRIXI_HIP_LDS_PAD;  // LDS row length (elements)
    constexpr bool STAGE = TRIXI_HIP_STAGE_NODES != 0;
    // RK stage: the node pass keeps u (and prefetches k) in registers, and the
    // register update runs over the same (node, element) pairs at the end,
    // from the accumulator, instead of re-reading u and k line by line.
    constexpr bool COOP = RK && STAGE;
    __shared__ R rec[STAGE ? NRF * NN * LDM : 1];
    __shared__ L recd[(STAGE && NDF > 0) ? NDF * NN * LDM : 1];
    __shared__ R acc[(DIM > 1 || COOP) ? NV * NN * LDM : 1];

    const int tid = threadIdx.x;
    const int grp = PD ? tid / (M * NL) : 0;  // wave-uniform
    const int gtid = tid - grp * (M * NL);
    const int el = gtid % M;
    const int line = gtid / M;
    const int64_t E = a.E;
    const uint32_t E32 = (uint32_t)E;
    const int64_t slab = (int64_t)NN * E;
    const int64_t e0 = xcd_tile(a.nxcd) * M;
    if (e0 >= E) return;  // whole workgroup: padding tile of the XCD mapping
    const int64_t e = e0 + el;
    const bool valid = e < E;
    const int64_t ec = valid ? e : E - 1;
    const PhysR<R> P = PhysR<R>::from(a.phys);

    // Face neighbours of all 2*DIM faces, and the neighbour face data of the
    // first direction, are requested before the node pass so that their
    // latency overlaps it; later directions are prefetched one phase ahead.
    // A face with nbr < 0 is a boundary face (-1, non-periodic mesh) or a
    // mortar face (<= -2): its loads read this element instead, and the state
    // is replaced by the boundary state, or the flux by the mortar flux, in
    // the direction phase.
    int32_t nb[2 * DIM], nbl[2 * DIM];
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) {
        nb[f] = a.nbr[(int64_t)f * E + ec];
        nbl[f] = (!EXT || nb[f] >= 0) ? nb[f] : (int32_t)ec;
    }
    // ---- node pass loads: where they take few registers (at most 24 VGPRs,
    // systems of up to five variables), every (node, element) pair of this
    // lane is loaded here, before anything waits on memory, so that a guarded
    // last pair is not loaded only after the others are computed (indices
    // past the tile are clamped to a valid pair and not used).  Otherwise
    // the node pass below loads each pair as it goes.
    constexpr bool RK_KEEP = COOP && TRIXI_HIP_RK_KEEP != 0 && std::is_same<R, float>::value &&
                             std::is_same<S, double>::value;
    constexpr bool EARLY =
        TRIXI_HIP_EARLY_LOADS != 0 && STAGE && NV <= 5 && NIT * NV * (int)sizeof(S) <= 96;
    S ul[EARLY ? NIT : 1][NV];
    double ukeep[RK_KEEP && !EARLY ? NIT : 1][NV], kkeep[RK_KEEP ? NIT : 1][NV];
    // RK_PRE: u and k of the update's (node, element) pairs, requested after
    // the direction work (few registers are live there; at most 40 VGPRs)
    constexpr bool RK_PRE = COOP && !RK_KEEP && !PD && TRIXI_HIP_RK_PREFETCH != 0 && std::is_same<S, float>::value &&
                            NV <= 5 && NIT * NV * (int)sizeof(S) <= 80;
    S upre[RK_PRE ? NIT : 1][NV], kpre[RK_PRE ? NIT : 1][NV];
    auto rk_prefetch = [&]() {
#pragma unroll
        for (int it = 0; it < (RK_PRE ? NIT : 0); ++it) {
            int q = tid + it * NT;
            if ((it + 1) * NT > NN * M && q >= NN * M) q = NN * M - 1;
            const int c = q % M;
            const int n = q / M;
            int64_t ee = e0 + c;
            if (ee >= E) ee = E - 1;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const int64_t g = soa(v, n, ee, slab, E32);
                upre[it][v] = U[g];
                kpre[it][v] = a.rk_first ? S(0) : K[g];
            }
        }
    };
#pragma unroll
    for (int it = 0; it < (EARLY ? NIT : 0); ++it) {
        int q = tid + it * NT;
        if ((it + 1) * NT > NN * M && q >= NN * M) q = NN * M - 1;
        const int c = q % M;
        const int n = q / M;
        int64_t ee = e0 + c;
        if (ee >= E) ee = E - 1;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const int64_t g = soa(v, n, ee, slab, E32);
            ul[it][v] = U[g];
            if constexpr (RK_KEEP) kkeep[it][v] = a.rk_first ? 0.0 : (double)K[g];
        }
    }
    // keep these loads ahead of the face loads below: loads complete in issue
    // order, so the node pass then waits for its own data only
    if constexpr (EARLY) __builtin_amdgcn_sched_barrier(0);

    double tnow = 0.0;
    if constexpr (EXT) tnow = RK ? a.time[0] : a.t_host;
    double um_next[NV], up_next[NV];
    {
        // first direction of this group: 0, or the group's own with PD
        const int base0 = line_base<DIM, N>(line, grp);
        const int s0 = grp == 0 ? 1 : (grp == 1 ? N : N * N);
        int32_t nm0 = nbl[0], np0 = nbl[1];
#pragma unroll
        for (int f = 1; f < DIM; ++f)
            if (PD && grp == f) {
                nm0 = nbl[2 * f];
                np0 = nbl[2 * f + 1];
            }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            um_next[v] = (double)U[soa(v, base0 + (N - 1) * s0, nm0, slab, E32)];
            up_next[v] = (double)U[soa(v, base0, np0, slab, E32)];
        }
    }

    // ---- node data, once per node: coalesced loads (element innermost) -----
    // Every lane visits NIT (node, element) pairs: NN * M / NT, rounded up.
    // RK_KEEP: the node pass keeps u and k in registers for the update;
    // otherwise the update reads them again with the same coalesced pattern
#pragma unroll
    for (int it = 0; it < (STAGE ? NIT : 0); ++it) {
        const int q = tid + it * NT;
        if ((it + 1) * NT > NN * M && q >= NN * M) continue;
        const int c = q % M;
        const int n = q / M;
        R uc[NV];
        if constexpr (EARLY) {
#pragma unroll
            for (int v = 0; v < NV; ++v) uc[v] = (R)(double)ul[it][v];
        } else {
            int64_t ee = e0 + c;
            if (ee >= E) ee = E - 1;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const int64_t g = soa(v, n, ee, slab, E32);
                const double ud = (double)U[g];
                uc[v] = (R)ud;
                if constexpr (RK_KEEP) {
                    ukeep[it][v] = ud;
                    kkeep[it][v] = a.rk_first ? 0.0 : (double)K[g];
                }
            }
        }
        const Node nd = EQ::template node<R, NODE_LOGS, L>(uc, P);
        R f[NRF];
        EQ::to_rec(nd, f);
#pragma unroll
        for (int k = 0; k < NRF; ++k) rec[(k * NN + n) * LDM + c] = f[k];
        if (NODE_LOGS) {
            recd[n * LDM + c] = nd.lrho;
            recd[(NN + n) * LDM + c] = nd.lp;
        }
    }
    if (STAGE) __syncthreads();

    R alpha = R(0);
    bool dg_only = true;
    if (VOL == VI_HG) {
        double al = a.alpha_raw[ec];
        if (a.smooth) {
#pragma unroll
            for (int s = 0; s < 2 * DIM; ++s) {
                if (EXT && nb[s] < 0) {
                    // boundary faces have no neighbour; a mortar face has the
                    // elements on the other side of the mortar
                    if (nb[s] == -1) continue;
                    const int32_t* mn = a.mnb + (int64_t)(-2 - nb[s]) * a.mK;
                    for (int k = 0; k < a.mK; ++k) {
                        if (mn[k] < 0) continue;
                        const double an = 0.5 * a.alpha_raw[mn[k]];
                        al = an > al ? an : al;
                    }
                    continue;
                }
                const double an = 0.5 * a.alpha_raw[nb[s]];
                al = an > al ? an : al;
            }
        }
        dg_only = fabs(al) <= a.atol;
        alpha = (R)al;
    }

    // last direction, sequential or not: -1/J (+ sources), then du or the RK update
    auto line_epilogue = [&](R (&al)[N][NV], int base, int s) {
        const double fac = -a.invJ[e];
        double t_dt = 0.0;
        if (RK) t_dt = a.time[1];
#pragma unroll
        for (int p = 0; p < N; ++p) {
            const int node = base + p * s;
            double sv[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) sv[v] = 0.0;
            if (EXT && a.src) source_at<DIM, NN, NV>(a, e, node, tnow, sv);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                R total = al[p][v];
                if (DIM > 1) total = acc[(v * NN + node) * LDM + el] + total;
                const int64_t g = soa(v, node, e, slab, E32);
                const double val = (double)total * fac + sv[v];
                if (!RK) {
                    DU[g] = (S)val;
                } else {
                    const S kk = (S)(a.rk_first ? t_dt * val : a.rk_a * (double)K[g] + t_dt * val);
                    K[g] = kk;
                    UO[g] = (S)((double)U[g] + a.rk_b * (double)kk);
                }
            }
        }
    };

    R aclp[PD ? N : 1][NV];  // this group's line sums (PD)
    int base_g = 0, s_g = 1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        if (PD && d != grp) continue;  // wave-uniform
        const int s = cpow(N, d);
        const int base = line_base<DIM, N>(line, d);
        R um[NV], up[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            um[v] = (R)um_next[v];
            up[v] = (R)up_next[v];
        }
        if (EXT && a.bc) {
            // outflow sides: the inner face state on both sides, so that the
            // (consistent) surface flux is flux(u_inner)
            if (nb[2 * d] == -1) {
                if ((a.outflow >> (2 * d)) & 1) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) um[v] = (R)U[soa(v, base, ec, slab, E32)];
                } else if (a.bc == TRIXI_HIP_BC_SLIP_WALL) {
                    // slip wall: the inner state with its normal momentum reversed
#pragma unroll
                    for (int v = 0; v < NV; ++v) um[v] = (R)U[soa(v, base, ec, slab, E32)];
                    um[1 + d] = -um[1 + d];
                } else {
                    boundary_state<DIM, NN, NV, R>(a, ec, base, tnow, um);
                }
            }
            if (nb[2 * d + 1] == -1) {
                if ((a.outflow >> (2 * d + 1)) & 1) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) up[v] = (R)U[soa(v, base + (N - 1) * s, ec, slab, E32)];
                } else if (a.bc == TRIXI_HIP_BC_SLIP_WALL) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) up[v] = (R)U[soa(v, base + (N - 1) * s, ec, slab, E32)];
                    up[1 + d] = -up[1 + d];
                } else {
                    boundary_state<DIM, NN, NV, R>(a, ec, base + (N - 1) * s, tnow, up);
                }
            }
        }
        if (!PD && d + 1 < DIM) {
            const int sn = cpow(N, d + 1);
            const int bn = line_base<DIM, N>(line, d + 1);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                um_next[v] = (double)U[soa(v, bn + (N - 1) * sn, nbl[2 * (d + 1)], slab, E32)];
                up_next[v] = (double)U[soa(v, bn, nbl[2 * (d + 1) + 1], slab, E32)];
            }
        }

        Node nd[N];
#pragma unroll
        for (int p = 0; p < N; ++p) {
            const int node = base + p * s;
            if (STAGE) {
                R f[NRF];
#pragma unroll
                for (int k = 0; k < NRF; ++k) f[k] = rec[(k * NN + node) * LDM + el];
                const L lr = NODE_LOGS ? recd[node * LDM + el] : L(0);
                const L lp = NODE_LOGS ? recd[(NN + node) * LDM + el] : L(0);
                nd[p] = EQ::from_rec(f, lr, lp, P);
            } else {
                R uc[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) uc[v] = (R)U[soa(v, node, ec, slab, E32)];
                nd[p] = EQ::template node<R, NODE_LOGS, L>(uc, P);
            }
        }

        R acl[N][NV];
#pragma unroll
        for (int p = 0; p < N; ++p)
#pragma unroll
            for (int v = 0; v < NV; ++v) acl[p][v] = R(0);

        // ---- volume integral along the line ---------------------------------
        if (VOL == VI_WEAK) {
            R fl[N][NV];
#pragma unroll
            for (int p = 0; p < N; ++p) EQ::flux(nd[p], d, P, fl[p]);
            if constexpr (WEAK_MFMA) {
                // D-hat on the matrix cores: v_mfma_f32_4x4x1_16b_f32 runs 16
                // blocks of 4x4x1 per wave, block b on lanes 4b..4b+3.  Each
                // lane gives row (lane % 4) of D-hat as A and its own line's
                // flux at node k as B, so after k = 0..3 it holds its line's
                // contraction sum_k Dhat[i][k] f_k in VGPRs i = 0..3.
                const int r = threadIdx.x & 3;
                float dh[N];
#pragma unroll
                for (int k = 0; k < N; ++k) dh[k] = (float)a.Dhat[r * N + k];
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    f32x4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int k = 0; k < N; ++k) c = __builtin_amdgcn_mfma_f32_4x4x1f32(dh[k], fl[k][v], c, 0, 0, 0);
#pragma unroll
                    for (int i = 0; i < N; ++i) acl[i][v] += c[i];
                }
            } else {
#pragma unroll
                for (int i = 0; i < N; ++i)
#pragma unroll
                    for (int p = 0; p < N; ++p) {
                        const R w = (R)a.Dhat[i * N + p];
#pragma unroll
                        for (int v = 0; v < NV; ++v) acl[i][v] += w * fl[p][v];
                    }
            }
        } else {
            const R sc = (VOL == VI_HG && !dg_only) ? R(1) - alpha : R(1);
#pragma unroll
            for (int i = 0; i < N; ++i)
#pragma unroll
                for (int j = i + 1; j < N; ++j) {
                    R f[NV];
                    EQ::template two_point<VF>(nd[i], nd[j], d, P, f);
                    const R fa = sc * (R)a.Ds[i * N + j];
                    const R fb = sc * (R)a.Ds[j * N + i];
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        acl[i][v] += fa * f[v];
                        acl[j][v] += fb * f[v];
                    }
                }
            if constexpr (EQ::NONCONS) {
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    R integral[NV];
#pragma unroll
                    for (int v = 0; v < NV; ++v) integral[v] = R(0);
#pragma unroll
                    for (int j = 0; j < N; ++j) {
                        R f[NV];
                        EQ::template noncons<-1>(nd[i], nd[j], d, P, f);
                        const R w = (R)a.Ds[i * N + j];
#pragma unroll
                        for (int v = 0; v < NV; ++v) integral[v] += w * f[v];
                    }
                    const R h = sc * R(0.5);
#pragma unroll
                    for (int v = 0; v < NV; ++v) acl[i][v] += h * integral[v];
                }
            }
            if (VOL == VI_HG && !dg_only) {
#pragma unroll
                for (int p = 0; p + 1 < N; ++p) {
                    R F[NV];
                    EQ::template two_point<FVF>(nd[p], nd[p + 1], d, P, F);
                    const R wl = alpha * (R)a.invw[p];
                    const R wr = alpha * (R)a.invw[p + 1];
                    if constexpr (EQ::NONCONS) {
                        // subcell interface with nonconservative terms: the left
                        // node takes f + nc(u_l, u_r)/2, the right f + nc(u_r, u_l)/2
                        R ncl[NV], ncr[NV];
                        EQ::template noncons<FVF>(nd[p], nd[p + 1], d, P, ncl);
                        EQ::template noncons<FVF>(nd[p + 1], nd[p], d, P, ncr);
#pragma unroll
                        for (int v = 0; v < NV; ++v) {
                            acl[p][v] += wl * (F[v] + R(0.5) * ncl[v]);
                            acl[p + 1][v] -= wr * (F[v] + R(0.5) * ncr[v]);
                        }
                    } else {
#pragma unroll
                        for (int v = 0; v < NV; ++v) {
                            acl[p][v] += wl * F[v];
                            acl[p + 1][v] -= wr * F[v];
                        }
                    }
                }
            }
        }

        // ---- interface fluxes at both ends of the line, surface integral ----
        {
            const Node ndm = EQ::template node<R, SF_LOGS, L>(um, P);
            const Node ndp = EQ::template node<R, SF_LOGS, L>(up, P);
            R fm[NV], fp[NV];
            EQ::template two_point<SF>(ndm, nd[0], d, P, fm);       // this element is "right"
            EQ::template two_point<SF>(nd[N - 1], ndp, d, P, fp);   // this element is "left"
            if constexpr (EQ::NONCONS) {
                R ncm[NV], ncp[NV];
                EQ::template noncons<SF>(nd[0], ndm, d, P, ncm);
                EQ::template noncons<SF>(nd[N - 1], ndp, d, P, ncp);
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    fm[v] = fm[v] + R(0.5) * ncm[v];
                    fp[v] = fp[v] + R(0.5) * ncp[v];
                }
            }
            if constexpr (EXT) {
                // mortar faces: the face node of this line is `line`
                if (nb[2 * d] <= -2) {
                    const double* mf = a.mflux + ((int64_t)(-2 - nb[2 * d]) * NL + line) * NV;
#pragma unroll
                    for (int v = 0; v < NV; ++v) fm[v] = (R)mf[v];
                }
                if (nb[2 * d + 1] <= -2) {
                    const double* mf = a.mflux + ((int64_t)(-2 - nb[2 * d + 1]) * NL + line) * NV;
#pragma unroll
                    for (int v = 0; v < NV; ++v) fp[v] = (R)mf[v];
                }
            }
            const R lm = (R)a.lhat_m, lp = (R)a.lhat_p;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                acl[0][v] = acl[0][v] - fm[v] * lm;
                acl[N - 1][v] = acl[N - 1][v] + fp[v] * lp;
            }
        }

        // ---- combine directions ---------------------------------------------
        if constexpr (PD) {
#pragma unroll
            for (int p = 0; p < N; ++p)
#pragma unroll
                for (int v = 0; v < NV; ++v) aclp[p][v] = acl[p][v];
            base_g = base;
            s_g = s;
        } else if (d + 1 < DIM || COOP) {
            if (d + 1 == DIM) rk_prefetch();
#pragma unroll
            for (int p = 0; p < N; ++p)
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    const int idx = (v * NN + base + p * s) * LDM + el;
                    if (d == 0)
                        acc[idx] = acl[p][v];
                    else
                        acc[idx] += acl[p][v];
                }
            __syncthreads();
        } else if (valid) {
            line_epilogue(acl, base, s);
        }
    }

    if constexpr (PD) {
        // the groups add their line sums in direction order; the last one
        // finishes (or, for a fused RK stage, every group stores and the
        // update below runs over the node-pass pairs)
#pragma unroll
        for (int ph = 0; ph < DIM; ++ph) {
            const bool last = (ph + 1 == DIM) && !COOP;
            if (grp == ph) {
                if (!last) {
#pragma unroll
                    for (int p = 0; p < N; ++p)
#pragma unroll
                        for (int v = 0; v < NV; ++v) {
                            const int idx = (v * NN + base_g + p * s_g) * LDM + el;
                            if (ph == 0)
                                acc[idx] = aclp[p][v];
                            else
                                acc[idx] += aclp[p][v];
                        }
                } else if (valid) {
                    line_epilogue(aclp, base_g, s_g);
                }
            }
            if (!last) __syncthreads();
        }
    }

    if constexpr (COOP) {
        // k = a_s k + dt du, u_out = u + b_s k over the node pass's pairs;
        // the accumulator holds every direction (same summation order as the
        // line-wise path)
        const double t_dt = a.time[1];
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int q = tid + it * NT;
            if ((NN * M) % NT != 0 && q >= NN * M) continue;
            const int c = q % M;
            const int n = q / M;
            const int64_t ee = e0 + c;
            if (ee >= E) continue;
            const double fac = -a.invJ[ee];
            double sv[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) sv[v] = 0.0;
            if (EXT && a.src) source_at<DIM, NN, NV>(a, ee, n, tnow, sv);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const int64_t g = soa(v, n, ee, slab, E32);
                const double val = (double)acc[(v * NN + n) * LDM + c] * fac + sv[v];
                const double kold = RK_KEEP  ? kkeep[it][v]
                                    : RK_PRE ? (double)kpre[it][v]
                                             : (a.rk_first ? 0.0 : (double)K[g]);
                const double uold = RK_KEEP  ? (EARLY ? (double)ul[it][v] : ukeep[it][v])
                                    : RK_PRE ? (double)upre[it][v]
                                             : (double)U[g];
                const S kk = (S)(a.rk_first ? t_dt * val : a.rk_a * kold + t_dt * val);
                K[g] = kk;
                UO[g] = (S)(uold + a.rk_b * (double)kk);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Node-parallel form for small meshes.  The line form gives each lane a whole
// line (N(N-1)/2 two-point fluxes and two interface fluxes per direction), so
// a mesh of a few thousand elements runs one or two waves per SIMD, each a
// long dependent chain.  Here lane (el, node, d) evaluates only its node's
// row of the line in direction d: the N-1 two-point fluxes with the other
// nodes of the line (each pair is evaluated by both of its nodes, twice the
// volume-flux work), the subcell fluxes on either side of the node and, at
// the line's ends, the interface flux; N^d * DIM lanes per element.
//
// Every node receives the same terms in the same order as in the line form
// (partners in increasing order, then the nonconservative sum, the subcell
// fluxes, the interface flux; directions added x, y, z); the two forms
// differ only where the compiler fuses a multiply-add in one code shape and
// not in the other.
//
// The last direction's lanes also do the node pass (node records to LDS,
// and for a fused RK stage u and k kept in registers) and the epilogue.
// ---------------------------------------------------------------------------
// Either of two nodes, chosen field by field through their node records
// (register selects; a conditional copy of the whole struct would go through
// scratch).  The record holds every field the fluxes read.
template <class EQ, typename R, typename L>
__device__ __forceinline__ typename EQ::template Node<R, L> select_node(
    bool c, const typename EQ::template Node<R, L>& a, const typename EQ::template Node<R, L>& b,
    const PhysR<R>& P) {
    R fa[EQ::NRF], fb[EQ::NRF];
    EQ::to_rec(a, fa);
    EQ::to_rec(b, fb);
#pragma unroll
    for (int k = 0; k < EQ::NRF; ++k) fa[k] = c ? fa[k] : fb[k];
    return EQ::from_rec(fa, c ? a.lrho : b.lrho, c ? a.lp : b.lp, P);
}

template <int DIM, int N>
__host__ __device__ constexpr int np_elems() {
    return (TRIXI_HIP_NP_THREADS / (cpow(N, DIM) * DIM)) > 0 ? TRIXI_HIP_NP_THREADS / (cpow(N, DIM) * DIM) : 1;
}

template <int DIM, int N, class EQ, int VOL, int VF, int SF, int FVF, typename R, bool RK, typename L, typename S>
__global__ void __launch_bounds__((np_elems<DIM, N>() * cpow(N, DIM) * DIM))
rhs_np_kernel(const RhsArgs a) {
    const S* __restrict__ U = static_cast<const S*>(a.u);
    S* __restrict__ DU = static_cast<S*>(a.du);
    S* __restrict__ K = static_cast<S*>(a.k);
    S* __restrict__ UO = static_cast<S*>(a.u_out);
    constexpr int NN = cpow(N, DIM);
    constexpr int M = np_elems<DIM, N>();
    constexpr int NV = EQ::NV;
    constexpr int NRF = EQ::NRF;
    constexpr bool SF_LOGS = flux_needs_logs(SF);
    constexpr bool VF_LOGS = flux_needs_logs(VF);
    constexpr bool FV_LOGS = (VOL == VI_HG) && flux_needs_logs(FVF);
    constexpr bool NODE_LOGS = SF_LOGS || VF_LOGS || FV_LOGS;
    constexpr int NDF = NODE_LOGS ? 2 : 0;
    using Node = typename EQ::template Node<R, L>;
    static_assert(VOL != VI_WEAK, "the node-parallel form is built for flux differencing and HG");

    __shared__ R rec[NRF * NN * M];
    __shared__ L recd[NDF > 0 ? NDF * NN * M : 1];
    __shared__ R acc[DIM > 1 ? (DIM - 1) * NV * NN * M : 1];
    // fused RK stage with shock capturing: the next stage's raw blending
    // factors from u_out (indicator variable, modal coefficients, energies)
    constexpr bool FUSE_IND = RK && VOL == VI_HG && std::is_same<EQ, Euler<DIM>>::value;
    __shared__ double ind_buf[FUSE_IND ? 2 * M * NN : 1];

    const int tid = threadIdx.x;
    const int el = tid % M;
    const int node = (tid / M) % NN;
    const int d = tid / (M * NN);
    const int64_t E = a.E;
    const uint32_t E32 = (uint32_t)E;
    const int64_t slab = (int64_t)NN * E;
    const int64_t e0 = xcd_tile(a.nxcd) * M;
    if (e0 >= E) return;  // whole workgroup: padding tile of the XCD mapping
    const int64_t e = e0 + el;
    const bool valid = e < E;
    const int64_t ec = valid ? e : E - 1;
    const PhysR<R> P = PhysR<R>::from(a.phys);
    const bool last = d == DIM - 1;  // node pass and epilogue lanes

    // this node's position along direction d, and the interface state at the
    // line's ends (requested before the node pass to overlap its latency)
    const int s = d == 0 ? 1 : (d == 1 ? N : N * N);
    const int p = (node / s) % N;
    const int base = node - p * s;
    const bool face = p == 0 || p == N - 1;
    R ub[NV];
    if (face) {
        const int32_t nbe = a.nbr[(int64_t)(2 * d + (p == 0 ? 0 : 1)) * E + ec];
        const int nn_face = p == 0 ? node + (N - 1) * s : node - (N - 1) * s;
#pragma unroll
        for (int v = 0; v < NV; ++v) ub[v] = (R)U[soa(v, nn_face, nbe, slab, E32)];
    }

    if (last) {
        R uc[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) uc[v] = (R)U[soa(v, node, ec, slab, E32)];
        const Node nd = EQ::template node<R, NODE_LOGS, L>(uc, P);
        R f[NRF];
        EQ::to_rec(nd, f);
#pragma unroll
        for (int k = 0; k < NRF; ++k) rec[(k * NN + node) * M + el] = f[k];
        if (NODE_LOGS) {
            recd[node * M + el] = nd.lrho;
            recd[(NN + node) * M + el] = nd.lp;
        }
    }
    __syncthreads();

    R alpha = R(0);
    bool dg_only = true;
    if constexpr (VOL == VI_HG) {
        double al = a.alpha_raw[ec];
        if (a.smooth) {
#pragma unroll
            for (int f = 0; f < 2 * DIM; ++f) {
                const double an = 0.5 * a.alpha_raw[a.nbr[(int64_t)f * E + ec]];
                al = an > al ? an : al;
            }
        }
        dg_only = fabs(al) <= a.atol;
        alpha = (R)al;
    }

    auto load_node = [&](int n) {
        R f[NRF];
#pragma unroll
        for (int k = 0; k < NRF; ++k) f[k] = rec[(k * NN + n) * M + el];
        const L lr = NODE_LOGS ? recd[n * M + el] : L(0);
        const L lq = NODE_LOGS ? recd[(NN + n) * M + el] : L(0);
        return EQ::from_rec(f, lr, lq, P);
    };
    const Node me = load_node(node);

    R acl[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acl[v] = R(0);

    // this node's row of D_split and its inverse weight, read with constant
    // indices (a dynamic index into the kernel arguments goes through scratch)
    R dsr[N];
    R invw_p = R(0);
#pragma unroll
    for (int j = 0; j < N; ++j) dsr[j] = R(0);
#pragma unroll
    for (int pp = 0; pp < N; ++pp) {
        if (pp != p) continue;
#pragma unroll
        for (int j = 0; j < N; ++j) dsr[j] = (R)a.Ds[pp * N + j];
        invw_p = (R)a.invw[pp];
    }

    // the direction's terms with the direction a compile-time constant (a
    // runtime direction would index the node records' vector fields
    // dynamically, which goes through scratch); d is uniform over each wave
    // of the usual shapes, so the dispatch below does not diverge
    auto dir_terms = [&](auto dc) {
        constexpr int D = decltype(dc)::value;
        constexpr int SD = cpow(N, D);
        // volume: this node's row of the split form, partners in increasing
        // order; every lane runs N-1 pairs with the lower node first (operands
        // selected, not branched, so that a wave evaluates one flux per pair)
        const R sc = (VOL == VI_HG && !dg_only) ? R(1) - alpha : R(1);
    #pragma unroll
        for (int q = 0; q < N - 1; ++q) {
            const bool below = q < p;  // partner q below this node, else q + 1 above
            const Node other = load_node(base + (below ? q : q + 1) * SD);
            const Node lo = select_node<EQ, R, L>(below, other, me, P);
            const Node hi = select_node<EQ, R, L>(below, me, other, P);
            R f[NV];
            EQ::template two_point<VF>(lo, hi, D, P, f);
            const R w = sc * (below ? dsr[q] : dsr[q + 1]);
    #pragma unroll
            for (int v = 0; v < NV; ++v) acl[v] += w * f[v];
        }
        if constexpr (EQ::NONCONS) {
            R integral[NV];
    #pragma unroll
            for (int v = 0; v < NV; ++v) integral[v] = R(0);
    #pragma unroll
            for (int j = 0; j < N; ++j) {
                const Node other = load_node(base + j * SD);
                R f[NV];
                EQ::template noncons<-1>(me, other, D, P, f);
                const R w = dsr[j];
    #pragma unroll
                for (int v = 0; v < NV; ++v) integral[v] += w * f[v];
            }
            const R h = sc * R(0.5);
    #pragma unroll
            for (int v = 0; v < NV; ++v) acl[v] += h * integral[v];
        }
        if (VOL == VI_HG && !dg_only) {
            const R w = alpha * invw_p;
            if (p > 0) {  // subcell interface p - 1/2: this node is its right node
                const Node left = load_node(node - SD);
                R F[NV];
                EQ::template two_point<FVF>(left, me, D, P, F);
                if constexpr (EQ::NONCONS) {
                    R nc[NV];
                    EQ::template noncons<FVF>(me, left, D, P, nc);
    #pragma unroll
                    for (int v = 0; v < NV; ++v) acl[v] -= w * (F[v] + R(0.5) * nc[v]);
                } else {
    #pragma unroll
                    for (int v = 0; v < NV; ++v) acl[v] -= w * F[v];
                }
            }
            if (p < N - 1) {  // subcell interface p + 1/2: this node is its left node
                const Node right = load_node(node + SD);
                R F[NV];
                EQ::template two_point<FVF>(me, right, D, P, F);
                if constexpr (EQ::NONCONS) {
                    R nc[NV];
                    EQ::template noncons<FVF>(me, right, D, P, nc);
    #pragma unroll
                    for (int v = 0; v < NV; ++v) acl[v] += w * (F[v] + R(0.5) * nc[v]);
                } else {
    #pragma unroll
                    for (int v = 0; v < NV; ++v) acl[v] += w * F[v];
                }
            }
        }

        // interface flux at the line's ends and its surface-integral weight
        // (this element is "right" at its -d face, "left" at its +d face)
        if (face) {
            const Node nb = EQ::template node<R, SF_LOGS, L>(ub, P);
            const bool minus = p == 0;
            const Node lo = select_node<EQ, R, L>(minus, nb, me, P);
            const Node hi = select_node<EQ, R, L>(minus, me, nb, P);
            R fs[NV];
            EQ::template two_point<SF>(lo, hi, D, P, fs);
            if constexpr (EQ::NONCONS) {
                R nc[NV];
                EQ::template noncons<SF>(me, nb, D, P, nc);
    #pragma unroll
                for (int v = 0; v < NV; ++v) fs[v] = fs[v] + R(0.5) * nc[v];
            }
            const R lw = minus ? -(R)a.lhat_m : (R)a.lhat_p;
    #pragma unroll
            for (int v = 0; v < NV; ++v) acl[v] = acl[v] + fs[v] * lw;
        }

    };
    if (d == 0) dir_terms(std::integral_constant<int, 0>{});
    if constexpr (DIM > 1) {
        if (d == 1) dir_terms(std::integral_constant<int, 1>{});
    }
    if constexpr (DIM > 2) {
        if (d == 2) dir_terms(std::integral_constant<int, 2>{});
    }

    // directions meet in LDS; the last direction's lane adds them in order
    if (!last) {
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[((d * NV + v) * NN + node) * M + el] = acl[v];
    }
    if (DIM > 1) __syncthreads();
    // (no early exit: the fused indicator below synchronises the workgroup)
    double uo[FUSE_IND ? NV : 1];
    if (last && valid) {
        const double fac = -a.invJ[e];
        double t_dt = 0.0;
        if (RK) t_dt = a.time[1];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            R total = acl[v];
            if constexpr (DIM > 1) {
                R sum = acc[((0 * NV + v) * NN + node) * M + el];
#pragma unroll
                for (int dd = 1; dd < DIM - 1; ++dd) sum = sum + acc[((dd * NV + v) * NN + node) * M + el];
                total = sum + acl[v];
            }
            const int64_t g = soa(v, node, e, slab, E32);
            const double val = (double)total * fac;
            if constexpr (!RK) {
                DU[g] = (S)val;
            } else {
                // u and k again (cache resident on the small meshes this form runs)
                const S kk = (S)(a.rk_first ? t_dt * val : a.rk_a * (double)K[g] + t_dt * val);
                K[g] = kk;
                const S un = (S)((double)U[g] + a.rk_b * (double)kk);
                UO[g] = un;
                if constexpr (FUSE_IND) uo[v] = (double)un;
            }
        }
    }
    if constexpr (FUSE_IND) {
        if (a.alpha_next == nullptr) return;  // uniform: the whole workgroup
        // the indicator kernel's steps on the stored (rounded) u_out, element
        // slot el of the node-pass lanes: the same values in the same order,
        // so the factors equal a separate indicator launch's bit for bit
        double* const buf = ind_buf;
        double* const e2 = ind_buf + M * NN;
        const PhysR<double> P64 = PhysR<double>::from(a.phys);
        const VinvRows<DIM, N> V = vinv_rows<DIM, N>(a.Vinv, node);
        double m = 0.0;
        if (last && valid) m = Euler<DIM>::indicator_variable(uo, P64, a.ind_var);
#pragma unroll
        for (int dd = 0; dd < DIM; ++dd) {
            const int s2 = cpow(N, dd);
            const int idx = (node / s2) % N, b2 = node - idx * s2;
            if (last) buf[el * NN + node] = m;
            __syncthreads();
            if (last) {
                double r = 0.0;
#pragma unroll
                for (int ii = 0; ii < N; ++ii) r += V.r[dd][ii] * buf[el * NN + b2 + ii * s2];
                m = r;
            }
            __syncthreads();
        }
        if (last) e2[el * NN + node] = m * m;
        __syncthreads();
        if (last && node == 0 && valid) {
            double total, clip1, clip2;
            modal_energies<DIM, N>(e2, el, total, clip1, clip2);
            a.alpha_next[e] = hg_alpha(total, clip1, clip2, a.ind_threshold, a.ind_s, a.ind_alpha_min,
                                       a.ind_alpha_max);
        }
    }
}

// Shapes the node-parallel form is built for: periodic conforming meshes in
// 2D and 3D, flux differencing or shock capturing, at most five variables
// (the nine-variable MHD system is faster in the line form), whole
// workgroups within 1024 lanes.  The host sets RhsArgs::np only for these.
template <int DIM, int N, int NV, int VOL, bool EXT>
__host__ __device__ constexpr bool np_eligible() {
    return !EXT && VOL != VI_WEAK && DIM > 1 && NV <= 5 && cpow(N, DIM) * DIM <= 1024;
}

// Launch helper used by the instantiation units: the node-parallel form for
// small periodic conforming meshes (RhsArgs::np), the line form otherwise.
template <int DIM, int N, class EQ, int VOL, int VF, int SF, int FVF, typename R, int M, bool RK, bool EXT = false,
          typename L = double, typename S = double>
hipError_t launch_rhs(const RhsArgs& a, hipStream_t stream) {
    if constexpr (np_eligible<DIM, N, EQ::NV, VOL, EXT>()) {
        if (a.np) {
            constexpr int MN = np_elems<DIM, N>();
            int64_t blocks = (a.E + MN - 1) / MN;
            if (a.nxcd > 1) blocks = (blocks + a.nxcd - 1) / a.nxcd * a.nxcd;
            hipLaunchKernelGGL((rhs_np_kernel<DIM, N, EQ, VOL, VF, SF, FVF, R, RK, L, S>), dim3((unsigned)blocks),
                               dim3(MN * cpow(N, DIM) * DIM), 0, stream, a);
            return hipGetLastError();
        }
    }
    // only the node-parallel form computes the next stage's blending factors
    if (a.alpha_next != nullptr) return hipErrorInvalidValue;
    int64_t blocks = (a.E + M - 1) / M;
    if (a.nxcd > 1) blocks = (blocks + a.nxcd - 1) / a.nxcd * a.nxcd;
    hipLaunchKernelGGL((rhs_kernel<DIM, N, EQ, VOL, VF, SF, FVF, R, M, RK, EXT, L, S>), dim3((unsigned)blocks),
                       dim3(rhs_threads<DIM, N, M, R, RK, EQ::NV>()), 0, stream, a);
    return hipGetLastError();
}

}  // namespace trixi_hip
