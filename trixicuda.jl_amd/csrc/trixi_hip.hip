// trixi_hip.hip -- C ABI, context management and auxiliary kernels of
// libtrixihip.so (see include/trixi_hip.h for the contract).
#include "../../include/trixi_hip.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "registry.hpp"
#include "initial_conditions.hpp"
#include "indicator.hpp"

namespace trixi_hip {

// ---------------------------------------------------------------------------
// registry
// ---------------------------------------------------------------------------
static std::map<KernelKey, LaunchFn>& table() {
    static std::map<KernelKey, LaunchFn> t;
    return t;
}
void register_kernel(const KernelKey& key, LaunchFn fn) { table()[key] = fn; }
LaunchFn find_kernel(const KernelKey& key) {
    auto it = table().find(key);
    return it == table().end() ? nullptr : it->second;
}
static std::map<KernelKey, VjpFn>& vjp_table() {
    static std::map<KernelKey, VjpFn> t;
    return t;
}
void register_vjp(const KernelKey& key, VjpFn fn) { vjp_table()[key] = fn; }
VjpFn find_vjp(const KernelKey& key) {
    auto it = vjp_table().find(key);
    return it == vjp_table().end() ? nullptr : it->second;
}
static std::map<KernelKey, MortarFn>& mortar_table() {
    static std::map<KernelKey, MortarFn> t;
    return t;
}
void register_mortar(const KernelKey& key, MortarFn fn) { mortar_table()[key] = fn; }
MortarFn find_mortar(const KernelKey& key) {
    auto it = mortar_table().find(key);
    return it == mortar_table().end() ? nullptr : it->second;
}

void register_linadv();
void register_euler1d();
void register_euler2d();
void register_euler3d();
void register_euler1d_hg();
void register_euler2d_hg();
void register_euler3d_hg();
void register_mhd3d();
void register_mhd2d();
void register_burgers();
void register_mhd1d();
void register_hypdiff();
void register_euler_multi();
void register_euler_quasi1d();
void register_mhd_mortar();
void register_shallow_water();
void register_vjp_linadv();
void register_vjp_euler1d();
void register_vjp_euler2d();
void register_vjp_euler3d();
void register_vjp_mhd3d();
void register_euler_ext();
void register_vjp_euler_ext();
void register_mortar_kernels();
void register_fp32();

static void ensure_registered() {
    static std::once_flag once;
    std::call_once(once, [] {
        register_linadv();
        register_euler1d();
        register_euler2d();
        register_euler3d();
        register_euler1d_hg();
        register_euler2d_hg();
        register_euler3d_hg();
        register_mhd3d();
        register_mhd2d();
        register_burgers();
        register_mhd1d();
        register_hypdiff();
        register_euler_multi();
        register_euler_quasi1d();
        register_mhd_mortar();
        register_shallow_water();
        register_vjp_linadv();
        register_vjp_euler1d();
        register_vjp_euler2d();
        register_vjp_euler3d();
        register_vjp_mhd3d();
        register_euler_ext();
        register_vjp_euler_ext();
        register_mortar_kernels();
        register_fp32();
    });
}

// ---------------------------------------------------------------------------
// error reporting
// ---------------------------------------------------------------------------
static thread_local std::string g_last_error;

static int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess)                                                          \
            return fail(TRIXI_HIP_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// ---------------------------------------------------------------------------
// Static device scratch of the *_noalloc entry points: a file-scope array that
// every GPU the library runs on holds until the process ends.  In host code
// `scratch` names a host-side stand-in, never the device memory: the device
// address is looked up once per device (hipGetSymbolAddress, with that device
// current) and only that address is handed to kernels.  All calls on one GPU
// share the array; calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
#define TRIXI_HIP_SCRATCH_DOUBLES 2048  // 16 KiB per device
static __device__ __align__(256) double scratch[TRIXI_HIP_SCRATCH_DOUBLES];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mutex;
static double* g_scratch_addr[kMaxDevices] = {};

// Device address of `scratch` on the current device.
static int scratch_ptr(double** out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(TRIXI_HIP_EDEVICE, "device ordinal out of range for scratch");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(TRIXI_HIP_EDEVICE, "hipGetSymbolAddress returned no address for scratch");
        g_scratch_addr[dev] = (double*)p;
    }
    *out = g_scratch_addr[dev];
    return TRIXI_HIP_OK;
}

// ---------------------------------------------------------------------------
// Lobatto-Legendre basis on the host (the matrices of
// LobattoLegendreBasisGPU, reference basis_lobatto_legendre.jl:60-100).
// Interior nodes: roots of P_{n-1} - P_{n+1} (proportional to (1-x^2) P_n'),
// Newton from Chebyshev-Gauss-Lobatto guesses; w_i = 2 / (n (n+1) P_n(x_i)^2).
// ---------------------------------------------------------------------------
static void legendre(int n, double x, double& P, double& dP) {
    double p0 = 1.0, p1 = x, d0 = 0.0, d1 = 1.0;
    if (n == 0) {
        P = 1.0;
        dP = 0.0;
        return;
    }
    for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        const double d2 = d0 + (2.0 * k - 1.0) * p1;
        p0 = p1;
        p1 = p2;
        d0 = d1;
        d1 = d2;
    }
    P = p1;
    dP = d1;
}

struct HostBasis {
    int N;
    std::vector<double> x, w, wbary, D, Ds, Dhat, lm, lp, Vinv;
};

static void lgl(int N, std::vector<double>& x, std::vector<double>& w) {
    const int n = N - 1;
    x.assign(N, 0.0);
    w.assign(N, 0.0);
    x[0] = -1.0;
    x[n] = 1.0;
    for (int i = 1; i < n; ++i) {
        double xi = -cos(M_PI * i / n);
        for (int it = 0; it < 100; ++it) {
            double Pm, dPm, Pp, dPp;
            legendre(n - 1, xi, Pm, dPm);
            legendre(n + 1, xi, Pp, dPp);
            const double q = Pm - Pp, dq = dPm - dPp;
            const double dx = q / dq;
            xi -= dx;
            if (fabs(dx) < 1e-17) break;
        }
        x[i] = xi;
    }
    // enforce symmetry and the centre node
    for (int i = 0; i < N / 2; ++i) {
        const double s = 0.5 * (x[n - i] - x[i]);
        x[i] = -s;
        x[n - i] = s;
    }
    if (N % 2 == 1) x[N / 2] = 0.0;
    for (int i = 0; i < N; ++i) {
        double P, dP;
        legendre(n, x[i], P, dP);
        w[i] = 2.0 / (n * (n + 1.0) * P * P);
    }
}

static bool invert(int n, const std::vector<double>& A, std::vector<double>& Ainv) {
    std::vector<double> M = A;
    Ainv.assign(n * n, 0.0);
    for (int i = 0; i < n; ++i) Ainv[i * n + i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (fabs(M[r * n + c]) > fabs(M[piv * n + c])) piv = r;
        if (M[piv * n + c] == 0.0) return false;
        for (int k = 0; k < n; ++k) {
            std::swap(M[piv * n + k], M[c * n + k]);
            std::swap(Ainv[piv * n + k], Ainv[c * n + k]);
        }
        const double d = M[c * n + c];
        for (int k = 0; k < n; ++k) {
            M[c * n + k] /= d;
            Ainv[c * n + k] /= d;
        }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const double f = M[r * n + c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; ++k) {
                M[r * n + k] -= f * M[c * n + k];
                Ainv[r * n + k] -= f * Ainv[c * n + k];
            }
        }
    }
    return true;
}

static std::vector<double> interpolation_matrix(const std::vector<double>& xin,
                                                const std::vector<double>& wbary,
                                                const std::vector<double>& xout) {
    const int nin = (int)xin.size(), nout = (int)xout.size();
    std::vector<double> V(nout * nin, 0.0);
    for (int k = 0; k < nout; ++k) {
        int hit = -1;
        for (int j = 0; j < nin; ++j)
            if (xout[k] == xin[j]) hit = j;
        if (hit >= 0) {
            V[k * nin + hit] = 1.0;
            continue;
        }
        double s = 0.0;
        for (int j = 0; j < nin; ++j) {
            V[k * nin + j] = wbary[j] / (xout[k] - xin[j]);
            s += V[k * nin + j];
        }
        for (int j = 0; j < nin; ++j) V[k * nin + j] /= s;
    }
    return V;
}

static HostBasis make_basis(int N) {
    HostBasis b;
    b.N = N;
    lgl(N, b.x, b.w);
    b.wbary.assign(N, 1.0);
    for (int j = 0; j < N; ++j)
        for (int k = 0; k < N; ++k)
            if (k != j) b.wbary[j] *= (b.x[j] - b.x[k]);
    for (int j = 0; j < N; ++j) b.wbary[j] = 1.0 / b.wbary[j];
    b.D.assign(N * N, 0.0);
    for (int i = 0; i < N; ++i) {
        double diag = 0.0;
        for (int j = 0; j < N; ++j)
            if (j != i) {
                b.D[i * N + j] = (b.wbary[j] / b.wbary[i]) / (b.x[i] - b.x[j]);
                diag -= b.D[i * N + j];
            }
        b.D[i * N + i] = diag;
    }
    b.Ds.resize(N * N);
    for (int q = 0; q < N * N; ++q) b.Ds[q] = 2.0 * b.D[q];
    b.Ds[0] += 1.0 / b.w[0];
    b.Ds[N * N - 1] -= 1.0 / b.w[N - 1];
    b.Dhat.resize(N * N);
    for (int j = 0; j < N; ++j)
        for (int n = 0; n < N; ++n) b.Dhat[j * N + n] = -b.D[n * N + j] * b.w[n] / b.w[j];
    b.lm.assign(N, 0.0);
    b.lp.assign(N, 0.0);
    b.lm[0] = 1.0 / b.w[0];
    b.lp[N - 1] = 1.0 / b.w[N - 1];
    std::vector<double> V(N * N);
    for (int i = 0; i < N; ++i)
        for (int m = 0; m < N; ++m) {
            double P, dP;
            legendre(m, b.x[i], P, dP);
            V[i * N + m] = P * sqrt(m + 0.5);
        }
    invert(N, V, b.Vinv);
    return b;
}

// ---------------------------------------------------------------------------
// auxiliary kernels
// ---------------------------------------------------------------------------
__global__ void pack_kernel(double* __restrict__ dst, const double* __restrict__ src, int nv, int nn,
                            int64_t E, int to_native) {
    const int64_t total = (int64_t)nv * nn * E;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        // native index i = (v*nn + n)*E + e  <->  trixi index v + nv*(n + nn*e)
        const int64_t e = i % E;
        const int64_t r = i / E;
        const int64_t n = r % nn, v = r / nn;
        const int64_t t = v + nv * (n + nn * e);
        if (to_native)
            dst[i] = src[t];
        else
            dst[t] = src[i];
    }
}

__global__ void coefficients_kernel(double* __restrict__ u, const double* __restrict__ xs, IcParams q,
                                    int N, int nn, int64_t E, double t) {
    const int64_t total = (int64_t)nn * E;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = i % E;
        const int n = (int)(i / E);
        double x[3] = {0, 0, 0};
        for (int d = 0; d < q.dim; ++d) x[d] = xs[((int64_t)d * nn + n) * E + e];
        if (q.dim == 1) {
            if (n == 0) x[0] = nextafter(x[0], INFINITY);
            if (n == N - 1) x[0] = nextafter(x[0], -INFINITY);
        }
        double uu[9];
        initial_condition(q, x, t, (uint64_t)(e * nn + n), uu);
        for (int v = 0; v < q.nv; ++v) u[((int64_t)v * nn + n) * E + e] = uu[v];
    }
}

// Hennemann-Gassner blending factor, one lane per node: blocks of whole
// elements, one wave (64 lanes) when an element fits, else the element's
// node count rounded up to whole waves (polydeg 4 in 3D: 125 nodes, 128
// lanes); lanes past the last whole element idle.  The dimension-wise modal
// transform goes through LDS; the element's first lane sums the modal
// energies in node order (the order of the element-per-thread formulation)
// and applies the logistic and clips.
template <int DIM, int N>
constexpr int ind_block() {
    return cpow(N, DIM) <= 64 ? 64 : (cpow(N, DIM) + 63) / 64 * 64;
}
template <int DIM, int N>
constexpr int ind_epb() {
    return ind_block<DIM, N>() / cpow(N, DIM);
}

template <int DIM, int N, typename S = double>
__global__ void __launch_bounds__(256) indicator_kernel(const S* __restrict__ u, double* __restrict__ alpha_raw,
                                                        int64_t E, PhysParams phys, const double* __restrict__ Vinv,
                                                        double threshold, double parameter_s, double alpha_min,
                                                        double alpha_max, int var, int nv) {
    constexpr int NN = cpow(N, DIM);
    constexpr int BS = ind_block<DIM, N>();
    constexpr int EPB = ind_epb<DIM, N>();
    static_assert(BS <= 256 && EPB >= 1, "one block holds whole elements");
    __shared__ double buf[BS];
    __shared__ double e2[BS];
    // idle lanes (el >= EPB) shadow the last element: their LDS slots lie
    // past EPB * NN, which no element reads
    const int el = threadIdx.x / NN, n = threadIdx.x % NN;
    const int elc = el < EPB ? el : EPB - 1;
    const int64_t e = (int64_t)blockIdx.x * EPB + elc;
    const int64_t ec = e < E ? e : E - 1;
    const PhysR<double> P = PhysR<double>::from(phys);
    const VinvRows<DIM, N> V = vinv_rows<DIM, N>(Vinv, n);
    const double m =
        modal_transform_rows<DIM, N>(indicator_value<DIM, N, S>(u, ec, E, P, n, var, nv), V, buf, elc, n);
    e2[threadIdx.x] = m * m;
    __syncthreads();
    if (n != 0 || el >= EPB || e >= E) return;
    double total, clip1, clip2;
    modal_energies<DIM, N>(e2, el, total, clip1, clip2);
    alpha_raw[e] = hg_alpha(total, clip1, clip2, threshold, parameter_s, alpha_min, alpha_max);
}

// alpha_e = max(raw_e, max over the face neighbours n of raw_n / 2), the
// neighbours being the elements across interfaces and L2 mortars (Trixi's
// apply_smoothing! over interfaces and mortars); boundary faces (nbr = -1)
// have none, a mortar face (nbr = -2 - slot) lists its in mnb[slot*mK ..].
__global__ void smooth_alpha_kernel(const double* __restrict__ raw, double* __restrict__ out,
                                    const int32_t* __restrict__ nbr, int dim, int64_t E,
                                    const int32_t* __restrict__ mnb, int mK) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= E) return;
    double a = raw[e];
    for (int s = 0; s < 2 * dim; ++s) {
        const int32_t nb = nbr[(int64_t)s * E + e];
        if (nb == -1) continue;
        if (nb <= -2) {
            for (int k = 0; k < mK; ++k) {
                const int32_t n = mnb[(int64_t)(-2 - nb) * mK + k];
                if (n < 0) continue;
                const double an = 0.5 * raw[n];
                a = an > a ? an : a;
            }
            continue;
        }
        const double an = 0.5 * raw[nb];
        a = an > a ? an : a;
    }
    out[e] = a;
}

// Reverse of the HG blending factor (smoothing max, clipping, logistic,
// modal energy, density * pressure), one lane per node like the forward:
//   ubar_e += abar_raw_e * d alpha_raw_e / d u_e,
// where abar_raw_e gathers the cotangents of the smoothed factors that took
// their value from element e: its own (winner -1) and, for each face f, the
// neighbour n = nbr(e, f) whose max was half of e's factor (winner(n) = f ^ 1,
// the face of n that points back to e).
template <int DIM, int N>
__global__ void __launch_bounds__(64) indicator_vjp_kernel(
    const double* __restrict__ u, double* __restrict__ ubar, const double* __restrict__ alphabar,
    const int32_t* __restrict__ winner, const int32_t* __restrict__ nbr, int smooth, int64_t E, PhysParams phys,
    const double* __restrict__ Vinv, double threshold, double parameter_s, double alpha_min, double alpha_max,
    int var, int nv, double* __restrict__ kbar_out, double kb) {
    constexpr int NN = cpow(N, DIM);
    constexpr int NV = DIM + 2;
    constexpr int EPB = 64 / NN;
    static_assert(NN <= 64 && 64 % NN == 0, "one block holds whole elements");
    __shared__ double buf[64];
    __shared__ double e2[64];
    __shared__ double coef[EPB][3];  // cotangents of total, clip1, clip2 (0: nothing to do)
    __shared__ double ab_s[EPB];     // cotangent of each element's raw factor
    const int el = threadIdx.x / NN, n = threadIdx.x % NN;
    const int64_t e = (int64_t)blockIdx.x * EPB + el;
    const bool valid = e < E;
    const int64_t ec = valid ? e : E - 1;
    // ubar[i] += d, and in a reverse RK stage the earlier stage's register
    // cotangent (kbar_out = a kbar + kb ubar, written by the VJP kernel
    // before this one) follows: kbar_out[i] += kb d
    auto bump = [&](int64_t i, double d) {
        ubar[i] += d;
        if (kbar_out != nullptr) kbar_out[i] += kb * d;
    };
    // the gathered cotangent first (all loads issued together, the
    // neighbours' unconditionally), so that a block whose elements all have
    // none (smoothed factors of 0: most elements of a mesh) ends here
    if (n == 0) {
        double ab = 0.0;
        if (valid) {
            if (smooth) {
                int32_t nbv[2 * DIM], wn[2 * DIM];
                double an[2 * DIM];
#pragma unroll
                for (int f = 0; f < 2 * DIM; ++f) nbv[f] = nbr[(int64_t)f * E + e];
#pragma unroll
                for (int f = 0; f < 2 * DIM; ++f) {
                    const int64_t q = nbv[f] >= 0 ? (int64_t)nbv[f] : e;
                    wn[f] = winner[q];
                    an[f] = alphabar[q];
                }
                ab = winner[e] < 0 ? alphabar[e] : 0.0;
#pragma unroll
                for (int f = 0; f < 2 * DIM; ++f)
                    if (nbv[f] >= 0 && wn[f] == (f ^ 1)) ab += 0.5 * an[f];
            } else {
                ab = alphabar[e];
            }
        }
        ab_s[el] = ab;
    }
    __syncthreads();
    bool any = false;
#pragma unroll
    for (int k = 0; k < EPB; ++k) any = any || ab_s[k] != 0.0;
    if (!any) return;  // block-uniform
    const PhysR<double> P = PhysR<double>::from(phys);
    // this node's state, kept for the reverse of the indicator variable below
    // (compressible Euler: the shock-capturing adjoint is built for it)
    double uc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) uc[v] = u[((int64_t)v * NN + n) * E + ec];
    const double m = modal_transform<DIM, N>(Euler<DIM>::indicator_variable(uc, P, var), Vinv, buf, el, n);
    e2[threadIdx.x] = m * m;
    __syncthreads();
    if (n == 0) {
        double tb = 0.0, c1b = 0.0, c2b = 0.0;
        const double ab = ab_s[el];
        if (ab != 0.0) {
            double total, clip1, clip2;
            modal_energies<DIM, N>(e2, el, total, clip1, clip2);
            const double frac1 = (total != 0.0) ? (total - clip1) / total : 0.0;
            const double frac2 = (clip1 != 0.0) ? (clip1 - clip2) / clip1 : 0.0;
            const bool second = frac2 > frac1;
            const double energy = second ? frac2 : frac1;
            const double k = parameter_s / threshold;
            const double a = 1.0 / (1.0 + exp(-k * (energy - threshold)));
            // clipped to 0, to 1 or to alpha_max: locally constant
            if (!(a < alpha_min || a > 1.0 - alpha_min || !(a < alpha_max))) {
                const double eb = ab * k * a * (1.0 - a);
                if (second) {
                    if (clip1 != 0.0) {
                        c1b = eb * clip2 / (clip1 * clip1);
                        c2b = -eb / clip1;
                    }
                } else if (total != 0.0) {
                    tb = eb * clip1 / (total * total);
                    c1b = -eb / total;
                }
            }
        }
        coef[el][0] = tb;
        coef[el][1] = c1b;
        coef[el][2] = c2b;
    }
    __syncthreads();
    const double tb = coef[el][0], c1b = coef[el][1], c2b = coef[el][2];
    const bool active = valid && (tb != 0.0 || c1b != 0.0 || c2b != 0.0);
    const int maxc = max_mode<DIM, N>(n);
    double mb = 2.0 * m * (tb + (maxc < N - 1 ? c1b : 0.0) + (maxc < N - 2 ? c2b : 0.0));
    // transpose of the dimension-wise modal transform, last direction first
#pragma unroll
    for (int d = DIM - 1; d >= 0; --d) {
        const int s = cpow(N, d);
        const int idx = (n / s) % N, base = n - idx * s;
        buf[threadIdx.x] = mb;
        __syncthreads();
        double r = 0.0;
#pragma unroll
        for (int jj = 0; jj < N; ++jj) r += Vinv[jj * N + idx] * buf[el * NN + base + jj * s];
        __syncthreads();
        mb = r;
    }
    if (!active) return;
    const int64_t r0 = (int64_t)n * E + e;
    if (var == 2) {  // density
        bump(r0, mb);
        return;
    }
    const double g = mb * P.gm1;
    const double rho = uc[0], en = uc[DIM + 1];
    if (var == 1) {
        // pressure = (gamma - 1) (rho E - |rho v|^2 / (2 rho))
        const double irho = 1.0 / rho;
        double ke = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const int64_t rd = (int64_t)(1 + d) * NN * E + r0;
            const double vd = uc[1 + d] * irho;
            ke += vd * vd;
            bump(rd, -(g * vd));
        }
        bump(r0, g * 0.5 * ke);
        bump((int64_t)(DIM + 1) * NN * E + r0, g);
        return;
    }
    // density * pressure = (gamma - 1) (rho E - |rho v|^2 / 2)
    bump(r0, g * en);
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const int64_t rd = (int64_t)(1 + d) * NN * E + r0;
        bump(rd, -(g * uc[1 + d]));
    }
    bump((int64_t)(DIM + 1) * NN * E + r0, g * rho);
}

__global__ void copy_kernel(double* __restrict__ dst, const double* __restrict__ src, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Reverse of one low-storage RK stage's register update:
// kbar <- a * kbar + b * ubar   (a is ignored, kbar overwritten, when first)
__global__ void kbar_update_kernel(double* __restrict__ kbar, const double* __restrict__ ubar, double a,
                                   double b, int first, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        kbar[i] = first ? b * ubar[i] : a * kbar[i] + b * ubar[i];
}

__global__ void set_time_kernel(double* time, double t, double dt, double c0, double c1, double c2,
                                double c3, double c4) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const double c[5] = {c0, c1, c2, c3, c4};
        for (int s = 0; s < 5; ++s) {
            time[2 * s] = t + c[s] * dt;
            time[2 * s + 1] = dt;
        }
    }
}

// Per-element max_abs_speeds scaled by invJ, reduced per block (max).
template <class EQ, int DIM>
__global__ void max_speed_kernel(const double* __restrict__ u, const double* __restrict__ invJ,
                                 double* __restrict__ partial, int nn, int64_t E, PhysParams phys) {
    __shared__ double red[256];
    const PhysR<double> P = PhysR<double>::from(phys);
    double best = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < E;
         e += (int64_t)gridDim.x * blockDim.x) {
        double lmax[3] = {0, 0, 0};
        for (int n = 0; n < nn; ++n) {
            double uc[EQ::NV], lam[3];
            for (int v = 0; v < EQ::NV; ++v) uc[v] = u[((int64_t)v * nn + n) * E + e];
            EQ::max_abs_speeds(uc, P, lam);
            for (int d = 0; d < DIM; ++d) lmax[d] = lam[d] > lmax[d] ? lam[d] : lmax[d];
        }
        double s = 0.0;
        for (int d = 0; d < DIM; ++d) s += lmax[d];
        const double sc = invJ[e] * s;
        best = sc > best ? sc : best;
    }
    red[threadIdx.x] = best;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Deterministic final reduction of `nblocks` partial vectors of length m:
// components [0, nsum) are summed, [nsum, m) take the maximum; one thread
// per component, fixed order.
__global__ void reduce_partials_kernel(const double* __restrict__ partial, double* __restrict__ out,
                                       int nblocks, int m, int nsum) {
    const int c = threadIdx.x;
    if (c >= m) return;
    const bool is_sum = c < nsum;
    double acc = is_sum ? 0.0 : partial[c];
    for (int b = is_sum ? 0 : 1; b < nblocks; ++b) {
        const double v = partial[(int64_t)b * m + c];
        acc = is_sum ? acc + v : fmax(acc, v);
    }
    out[c] = acc;
}

// Error norms: each thread handles analysis points (element, analysis node),
// interpolating u and x with the (NA x N) Vandermonde of the 2N analysis
// basis.  Block partials: [l2 (nv) | linf (nv)].
template <int DIM>
__global__ void error_norms_kernel(const double* __restrict__ u, const double* __restrict__ xs,
                                   const double* __restrict__ invJ, double* __restrict__ partial,
                                   const double* __restrict__ Vdm_g, const double* __restrict__ wa_g,
                                   int N, int NA, int64_t E, IcParams q, double t, int nv) {
    __shared__ double red[256 * 18];
    const int nn = (int)lrint(pow((double)N, (double)DIM));
    const int nna = (int)lrint(pow((double)NA, (double)DIM));
    double l2[9], li[9];
    for (int v = 0; v < nv; ++v) {
        l2[v] = 0.0;
        li[v] = 0.0;
    }
    const int64_t total = (int64_t)nna * E;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = i / nna;
        const int a = (int)(i % nna);
        int ai[3] = {0, 0, 0};
        {
            int r = a;
            for (int d = 0; d < DIM; ++d) {
                ai[d] = r % NA;
                r /= NA;
            }
        }
        double ua[9], xa[3] = {0, 0, 0};
        for (int v = 0; v < nv; ++v) ua[v] = 0.0;
        for (int n = 0; n < nn; ++n) {
            double w = 1.0;
            int r = n;
            for (int d = 0; d < DIM; ++d) {
                w *= Vdm_g[ai[d] * N + (r % N)];
                r /= N;
            }
            if (w == 0.0) continue;
            for (int v = 0; v < nv; ++v) ua[v] += w * u[((int64_t)v * nn + n) * E + e];
            for (int d = 0; d < DIM; ++d) xa[d] += w * xs[((int64_t)d * nn + n) * E + e];
        }
        double uex[9];
        initial_condition(q, xa, t, 0, uex);
        double wq = 1.0;
        for (int d = 0; d < DIM; ++d) wq *= wa_g[ai[d]];
        const double vj = pow(1.0 / invJ[e], (double)DIM);
        for (int v = 0; v < nv; ++v) {
            const double diff = uex[v] - ua[v];
            l2[v] += diff * diff * (wq * vj);
            li[v] = fmax(li[v], fabs(diff));
        }
    }
    for (int v = 0; v < nv; ++v) {
        red[threadIdx.x * 18 + v] = l2[v];
        red[threadIdx.x * 18 + 9 + v] = li[v];
    }
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int v = 0; v < nv; ++v) {
                red[threadIdx.x * 18 + v] += red[(threadIdx.x + w) * 18 + v];
                red[threadIdx.x * 18 + 9 + v] =
                    fmax(red[threadIdx.x * 18 + 9 + v], red[(threadIdx.x + w) * 18 + 9 + v]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int v = 0; v < nv; ++v) {
            partial[(int64_t)blockIdx.x * 2 * nv + v] = red[v];
            partial[(int64_t)blockIdx.x * 2 * nv + nv + v] = red[9 + v];
        }
}

__global__ void integrate_kernel(const double* __restrict__ u, const double* __restrict__ invJ,
                                 double* __restrict__ partial, const double* __restrict__ w_g, int N,
                                 int dim, int64_t E, int nv) {
    __shared__ double red[256 * 9];
    int nn = 1;
    for (int d = 0; d < dim; ++d) nn *= N;
    double acc[9];
    for (int v = 0; v < nv; ++v) acc[v] = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < E;
         e += (int64_t)gridDim.x * blockDim.x) {
        const double vj = pow(1.0 / invJ[e], (double)dim);
        for (int n = 0; n < nn; ++n) {
            double w = 1.0;
            int r = n;
            for (int d = 0; d < dim; ++d) {
                w *= w_g[r % N];
                r /= N;
            }
            for (int v = 0; v < nv; ++v) acc[v] += vj * w * u[((int64_t)v * nn + n) * E + e];
        }
    }
    for (int v = 0; v < nv; ++v) red[threadIdx.x * 9 + v] = acc[v];
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int v = 0; v < nv; ++v) red[threadIdx.x * 9 + v] += red[(threadIdx.x + w) * 9 + v];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int v = 0; v < nv; ++v) partial[(int64_t)blockIdx.x * nv + v] = red[v];
}

// Positivity / finiteness flags per element (FP64).
__global__ void check_state_kernel(const double* __restrict__ u, int32_t* __restrict__ flags,
                                   double* __restrict__ partial, int eq, int dim, int nv, int nn,
                                   int64_t E, PhysParams phys) {
    __shared__ double red[256];
    double cnt = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < E;
         e += (int64_t)gridDim.x * blockDim.x) {
        int f = 0;
        for (int n = 0; n < nn; ++n) {
            double uc[9];
            for (int v = 0; v < nv; ++v) {
                uc[v] = u[((int64_t)v * nn + n) * E + e];
                if (!isfinite(uc[v])) f |= 1;
            }
            if (eq == EQ_LINADV || eq == EQ_BURGERS || eq == EQ_HYPDIFF) continue;
            if (eq == EQ_SHALLOW_WATER) {
                if (!(uc[0] > 0.0)) f |= 2;  // water height
                continue;
            }
            if (eq == EQ_EULER_QUASI1D) {
                // rho = (a rho) / a, p = (gamma - 1)((a e) / a - (a rho v)^2 / (2 a (a rho)))
                const double a = uc[3];
                if (!(a > 0.0) || !(uc[0] / a > 0.0)) f |= 2;
                if (!(phys.gm1 * (uc[2] / a - 0.5 * uc[1] * uc[1] / (a * uc[0])) > 0.0)) f |= 4;
                continue;
            }
            if (eq == EQ_EULER_MULTI) {
                // total density and the mixture pressure
                double rho = 0.0, h1 = 0.0, h2 = 0.0, ke = 0.0;
                for (int k = 0; k < nv - dim - 1; ++k) {
                    rho += uc[dim + 1 + k];
                    h1 += uc[dim + 1 + k] * phys.cvg[k];
                    h2 += uc[dim + 1 + k] * phys.cv[k];
                }
                if (!(rho > 0.0)) f |= 2;
                for (int d = 0; d < dim; ++d) ke += uc[d] * uc[d] / rho;
                if (!((h1 / h2 - 1.0) * (uc[dim] - 0.5 * ke) > 0.0)) f |= 4;
                continue;
            }
            if (!(uc[0] > 0.0)) f |= 2;
            double ke = 0.0;
            const int nvel = (eq == EQ_EULER) ? dim : 3;
            for (int d = 0; d < nvel; ++d) ke += uc[1 + d] * uc[1 + d] / uc[0];
            double p = uc[nvel + 1] - 0.5 * ke;
            if (eq == EQ_GLM_MHD)
                p -= 0.5 * (uc[5] * uc[5] + uc[6] * uc[6] + uc[7] * uc[7] + (nv > 8 ? uc[8] * uc[8] : 0.0));
            p *= phys.gm1;
            if (!(p > 0.0)) f |= 4;
        }
        if (flags) flags[e] = f;
        if (f) cnt += 1.0;
    }
    red[threadIdx.x] = cnt;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

}  // namespace trixi_hip

using namespace trixi_hip;

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct trixi_hip_ctx {
    trixi_hip_config cfg;
    int dim, N, nv, nn;
    int64_t E, I, nvalues;
    double total_volume;
    HostBasis basis;
    // analysis basis
    int NA;
    // device buffers (all allocated in create)
    int32_t* d_nbr = nullptr;
    double* d_invJ = nullptr;
    double* d_x = nullptr;
    double* d_alpha_raw = nullptr;
    double* d_alpha = nullptr;
    double* d_time = nullptr;
    double* d_partial = nullptr;
    double* d_result = nullptr;
    double* d_Vinv = nullptr;
    double* d_Vdm = nullptr;
    double* d_wa = nullptr;
    double* d_w = nullptr;
    double* d_tmp = nullptr;  // RK stage states (two buffers of nvalues)
    int nblocks_red = 1024;
    RhsArgs args;
    LaunchFn rhs_fn = nullptr, rk_fn = nullptr;
    // adjoint
    VjpFn vjp_fn = nullptr;
    LaunchFn rk64_fn = nullptr;   // FP64 forward stage for the adjoint's recomputation
    double* d_stage = nullptr;    // 5 stage inputs of one RK step
    double* d_k = nullptr;        // forward register during recomputation
    double* d_kbar = nullptr;     // adjoint register
    double* d_alphabar = nullptr; // HG: cotangent of the smoothed blending factor
    double* d_stage_alpha = nullptr;  // fused HG adjoint: raw blending factors of the 5 stage inputs
    int32_t* d_winner = nullptr;  // HG: argument of the smoothing max
    bool adjoint = false;
    // L2 mortars
    int64_t nmortars = 0;
    int mK = 0;                   // small elements per mortar, 2^(d-1)
    int32_t* d_mids = nullptr;    // [nmortars][mK+1]
    int32_t* d_mori = nullptr;
    int32_t* d_mls = nullptr;
    int32_t* d_mnb = nullptr;     // [slots][mK] elements across each mortar face
    double* d_mflux = nullptr;    // [slots][NL][nv]
    MortarArgs margs;
    MortarFn mortar_fn = nullptr;
    // graph cache for rk steps
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    void* graph_u = nullptr;
    void* graph_k = nullptr;
    hipStream_t graph_stream = nullptr;
    bool has_graph = false;
    // the fused RK stages compute the next stage's raw blending factors
    // (node-parallel form, compressible Euler shock capturing)
    bool fuse_ind = false;
};

static const double CK_A[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                               -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
static const double CK_B[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                               1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                               2277821191437.0 / 14882151754819.0};
static const double CK_C[5] = {0.0, 1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
                               2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0};

static int nvars_of(const trixi_hip_config& c) {
    if (c.equation == TRIXI_HIP_LINEAR_SCALAR_ADVECTION || c.equation == TRIXI_HIP_INVISCID_BURGERS) return 1;
    if (c.equation == TRIXI_HIP_COMPRESSIBLE_EULER_MULTICOMPONENT) return c.ndims + 1 + c.ncomponents;
    if (c.equation == TRIXI_HIP_COMPRESSIBLE_EULER_QUASI1D) return 4;
    if (c.equation == TRIXI_HIP_SHALLOW_WATER) return c.ndims + 2;
    if (c.equation == TRIXI_HIP_COMPRESSIBLE_EULER) return c.ndims + 2;
    if (c.equation == TRIXI_HIP_HYPERBOLIC_DIFFUSION) return c.ndims + 1;
    return c.ndims == 1 ? 8 : 9;  // ideal MHD: no cleaning field in 1D
}

static KernelKey make_key(const trixi_hip_config& c, int rk, bool mortars = false) {
    KernelKey k;
    k.dim = c.ndims;
    k.N = c.polydeg + 1;
    // multicomponent Euler: one kernel family per component count
    k.eq = (c.equation == TRIXI_HIP_COMPRESSIBLE_EULER_MULTICOMPONENT) ? multi_key(c.ncomponents) : c.equation;
    k.vol = c.volume_integral;
    k.vf = (c.volume_integral == TRIXI_HIP_VOLUME_WEAK_FORM) ? FLUX_CENTRAL : c.volume_flux;
    k.sf = c.surface_flux;
    k.fvf = (c.volume_integral == TRIXI_HIP_VOLUME_SHOCK_CAPTURING_HG) ? c.fv_flux : FLUX_CENTRAL;
    k.prec = c.precision;
    k.rk = rk;
    k.ext = (c.boundary_condition != TRIXI_HIP_BC_PERIODIC || c.source_terms != TRIXI_HIP_SOURCE_NONE || mortars)
                ? 1 : 0;
    return k;
}

static void free_ctx(trixi_hip_ctx* c) {
    if (!c) return;
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    void* bufs[] = {c->d_nbr, c->d_invJ, c->d_x,     c->d_alpha_raw, c->d_alpha, c->d_time,
                    c->d_partial, c->d_result, c->d_Vinv, c->d_Vdm,   c->d_wa,    c->d_w,
                    c->d_tmp, c->d_stage, c->d_k, c->d_kbar, c->d_alphabar, c->d_winner, c->d_stage_alpha,
                    c->d_mids, c->d_mori, c->d_mls, c->d_mnb, c->d_mflux};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    delete c;
}

static bool fp32_state(const trixi_hip_ctx* c) { return c->cfg.precision == TRIXI_HIP_FP32; }

// Raw blending factors of state u (FP64, or FP32 when f32).
static int launch_indicator(trixi_hip_ctx* c, const void* u, hipStream_t s, bool f32 = false,
                            double* out = nullptr) {
    double* const araw = out ? out : c->d_alpha_raw;
    const double threshold = 0.5 * pow(10.0, -1.8 * pow((double)c->N, 0.25));
    const double parameter_s = log((1.0 - 0.0001) / 0.0001);
    const PhysParams& P = c->args.phys;
    if (c->N != 4 && c->N != 5) return fail(TRIXI_HIP_EUNSUPPORTED, "shock capturing is built for polydeg 3 and 4");
    const int var = c->cfg.equation == TRIXI_HIP_SHALLOW_WATER ? IND_WATERHEIGHT_PRESSURE : c->cfg.indicator_variable;
#define TRIXI_HIP_IND_S(D, NQ, ST)                                                                             \
    hipLaunchKernelGGL((indicator_kernel<D, NQ, ST>),                                                         \
                       dim3((unsigned)((c->E + ind_epb<D, NQ>() - 1) / ind_epb<D, NQ>())),                    \
                       dim3(ind_block<D, NQ>()), 0, s, static_cast<const ST*>(u), araw, c->E, P,             \
                       c->d_Vinv, threshold, parameter_s, c->cfg.alpha_min, c->cfg.alpha_max, var, c->nv)
#define TRIXI_HIP_IND(D, NQ)                                  \
    do {                                                      \
        if (f32) TRIXI_HIP_IND_S(D, NQ, float);                \
        else TRIXI_HIP_IND_S(D, NQ, double);                   \
    } while (0)
    if (c->N == 4) {
        if (c->dim == 1) TRIXI_HIP_IND(1, 4); else if (c->dim == 2) TRIXI_HIP_IND(2, 4); else TRIXI_HIP_IND(3, 4);
    } else {
        if (c->dim == 1) TRIXI_HIP_IND(1, 5); else if (c->dim == 2) TRIXI_HIP_IND(2, 5); else TRIXI_HIP_IND(3, 5);
    }
#undef TRIXI_HIP_IND
#undef TRIXI_HIP_IND_S
    HIP_TRY(hipGetLastError());
    return TRIXI_HIP_OK;
}

// Mortar fluxes of state u, before the rhs kernel that reads them.
static int launch_mortars(trixi_hip_ctx* c, const double* u, hipStream_t s) {
    if (c->nmortars == 0) return TRIXI_HIP_OK;
    MortarArgs m = c->margs;
    m.u = u;
    HIP_TRY(c->mortar_fn(m, s));
    return TRIXI_HIP_OK;
}

static IcParams ic_params(const trixi_hip_ctx* c, int ic, uint64_t seed);

// Mortar operators of MortarL2 (Lobatto basis of N nodes): forward_{lower,upper}
// [j][i] = l_i((x_j -+ 1)/2); reverse = G2L * P * L2G with the Gauss-basis
// projection P[i][j] = 1/2 lg_i((g_j -+ 1)/2) wg_j / wg_i (Trixi's
// calc_forward_{lower,upper}, calc_reverse_{lower,upper}(n, Val(:gauss)),
// used by the reference's MortarL2GPU, basis_lobatto_legendre.jl:155-173).
static void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < n; ++i) {
        double xi = -cos(M_PI * (i + 0.75) / (n + 0.5)), P = 0.0, dP = 1.0;
        for (int it = 0; it < 100; ++it) {
            legendre(n, xi, P, dP);
            const double dx = P / dP;
            xi -= dx;
            if (fabs(dx) < 1e-16) break;
        }
        legendre(n, xi, P, dP);
        x[i] = xi;
        w[i] = 2.0 / ((1.0 - xi * xi) * dP * dP);
    }
}

static void mortar_operators(const HostBasis& b, int N, double F[2][MAXN * MAXN], double R[2][MAXN * MAXN]) {
    std::vector<double> g, wg;
    gauss_legendre(N, g, wg);
    std::vector<double> gb(N);
    for (int j = 0; j < N; ++j) {
        double p = 1.0;
        for (int k = 0; k < N; ++k)
            if (k != j) p *= g[j] - g[k];
        gb[j] = 1.0 / p;
    }
    for (int half = 0; half < 2; ++half) {
        const double sh = half ? 1.0 : -1.0;
        std::vector<double> xs(N), gs(N);
        for (int j = 0; j < N; ++j) {
            xs[j] = 0.5 * (b.x[j] + sh);
            gs[j] = 0.5 * (g[j] + sh);
        }
        const std::vector<double> Fh = interpolation_matrix(b.x, b.wbary, xs);  // [j][i]
        const std::vector<double> Pg = interpolation_matrix(g, gb, gs);         // [j][i] = lg_i(gs_j)
        const std::vector<double> L2G = interpolation_matrix(b.x, b.wbary, g);
        const std::vector<double> G2L = interpolation_matrix(g, gb, b.x);
        std::vector<double> P(N * N), T(N * N);
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) P[i * N + j] = 0.5 * Pg[j * N + i] * wg[j] / wg[i];
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) {
                double acc = 0.0;
                for (int k = 0; k < N; ++k) acc += P[i * N + k] * L2G[k * N + j];
                T[i * N + j] = acc;
            }
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) {
                double acc = 0.0;
                for (int k = 0; k < N; ++k) acc += G2L[i * N + k] * T[k * N + j];
                R[half][i * N + j] = acc;
                F[half][i * N + j] = Fh[i * N + j];
            }
    }
}

extern "C" {

const char* trixi_hip_version(void) { return "trixi_hip 0.1.0 (gfx950)"; }

const char* trixi_hip_last_error(void) { return g_last_error.c_str(); }

int trixi_hip_basis(int32_t nnodes, double* nodes, double* weights, double* derivative_matrix,
                    double* derivative_split, double* derivative_dhat, double* boundary_interpolation,
                    double* inverse_vandermonde_legendre) {
    if (nnodes < 2 || nnodes > MAXN) return fail(TRIXI_HIP_EINVAL, "nnodes out of range");
    const HostBasis b = make_basis(nnodes);
    const int N = nnodes;
    if (nodes) memcpy(nodes, b.x.data(), N * sizeof(double));
    if (weights) memcpy(weights, b.w.data(), N * sizeof(double));
    if (derivative_matrix) memcpy(derivative_matrix, b.D.data(), N * N * sizeof(double));
    if (derivative_split) memcpy(derivative_split, b.Ds.data(), N * N * sizeof(double));
    if (derivative_dhat) memcpy(derivative_dhat, b.Dhat.data(), N * N * sizeof(double));
    if (boundary_interpolation)
        for (int j = 0; j < N; ++j) {
            boundary_interpolation[2 * j] = b.lm[j];
            boundary_interpolation[2 * j + 1] = b.lp[j];
        }
    if (inverse_vandermonde_legendre) memcpy(inverse_vandermonde_legendre, b.Vinv.data(), N * N * sizeof(double));
    return TRIXI_HIP_OK;
}

int trixi_hip_create(const trixi_hip_config* config, const trixi_hip_mesh* mesh, trixi_hip_ctx** out) {
    if (!config || !mesh || !out) return fail(TRIXI_HIP_EINVAL, "null argument");
    *out = nullptr;
    ensure_registered();
    const trixi_hip_config& cfg = *config;
    if (cfg.ndims < 1 || cfg.ndims > 3) return fail(TRIXI_HIP_EINVAL, "ndims must be 1, 2 or 3");
    if (cfg.polydeg < 1 || cfg.polydeg + 1 > MAXN) return fail(TRIXI_HIP_EINVAL, "polydeg out of range");
    if (cfg.equation < TRIXI_HIP_LINEAR_SCALAR_ADVECTION || cfg.equation > TRIXI_HIP_SHALLOW_WATER)
        return fail(TRIXI_HIP_EINVAL, "unknown equation");
    if (cfg.equation == TRIXI_HIP_SHALLOW_WATER) {
        if (cfg.ndims > 2) return fail(TRIXI_HIP_EUNSUPPORTED, "shallow water is built for 1D and 2D");
        if (!(cfg.gravity > 0.0)) return fail(TRIXI_HIP_EINVAL, "shallow water needs gravity > 0");
        if (cfg.boundary_condition == TRIXI_HIP_BC_DIRICHLET &&
            cfg.boundary_ic != TRIXI_HIP_IC_CONVERGENCE_TEST && cfg.boundary_ic != TRIXI_HIP_IC_CONSTANT)
            return fail(TRIXI_HIP_EUNSUPPORTED, "shallow water Dirichlet states: the convergence test or constant");
        if (cfg.boundary_outflow_faces != 0) return fail(TRIXI_HIP_EUNSUPPORTED, "no outflow sides for shallow water");
    }
    if (cfg.equation == TRIXI_HIP_COMPRESSIBLE_EULER_QUASI1D) {
        if (cfg.ndims != 1) return fail(TRIXI_HIP_EINVAL, "quasi-1D Euler is one-dimensional");
        if (cfg.boundary_condition != TRIXI_HIP_BC_PERIODIC ||
            (cfg.source_terms != TRIXI_HIP_SOURCE_NONE && cfg.source_terms != TRIXI_HIP_SOURCE_CONVERGENCE_TEST))
            return fail(TRIXI_HIP_EUNSUPPORTED,
                        "quasi-1D Euler is built for periodic meshes, with or without source_terms_convergence_test");
    }
    if (cfg.equation == TRIXI_HIP_INVISCID_BURGERS && cfg.ndims != 1)
        return fail(TRIXI_HIP_EUNSUPPORTED, "the inviscid Burgers equation is built for 1D");
    if (cfg.equation == TRIXI_HIP_COMPRESSIBLE_EULER_MULTICOMPONENT) {
        if (cfg.ndims > 2) return fail(TRIXI_HIP_EUNSUPPORTED, "multicomponent Euler is built for 1D and 2D");
        if (cfg.ncomponents < 1 || cfg.ncomponents > 2)
            return fail(TRIXI_HIP_EUNSUPPORTED, "multicomponent Euler is built for 1 or 2 components");
        for (int k = 0; k < cfg.ncomponents; ++k)
            if (!(cfg.gammas[k] > 1.0) || !(cfg.gas_constants[k] > 0.0))
                return fail(TRIXI_HIP_EINVAL, "multicomponent Euler needs gamma_k > 1 and R_k > 0");
        if (cfg.boundary_condition != TRIXI_HIP_BC_PERIODIC || cfg.source_terms != TRIXI_HIP_SOURCE_NONE)
            return fail(TRIXI_HIP_EUNSUPPORTED, "multicomponent Euler is built for periodic meshes without sources");
    }
    if (cfg.volume_integral == TRIXI_HIP_VOLUME_SHOCK_CAPTURING_HG && cfg.equation != TRIXI_HIP_COMPRESSIBLE_EULER &&
        cfg.equation != TRIXI_HIP_INVISCID_BURGERS && cfg.equation != TRIXI_HIP_IDEAL_GLM_MHD &&
        cfg.equation != TRIXI_HIP_SHALLOW_WATER)
        return fail(TRIXI_HIP_EUNSUPPORTED,
                    "shock capturing is built for compressible Euler, Burgers, GLM-MHD and shallow water");
    if (cfg.volume_integral == TRIXI_HIP_VOLUME_SHOCK_CAPTURING_HG && cfg.equation == TRIXI_HIP_SHALLOW_WATER &&
        cfg.indicator_variable != TRIXI_HIP_INDICATOR_DENSITY_PRESSURE)
        return fail(TRIXI_HIP_EUNSUPPORTED, "the shallow-water indicator variable is waterheight_pressure");
    if (cfg.volume_integral == TRIXI_HIP_VOLUME_SHOCK_CAPTURING_HG && cfg.equation == TRIXI_HIP_IDEAL_GLM_MHD &&
        (cfg.indicator_variable != TRIXI_HIP_INDICATOR_DENSITY_PRESSURE || cfg.ndims == 1))
        return fail(TRIXI_HIP_EUNSUPPORTED, "GLM-MHD shock capturing is built for 2D/3D with density_pressure");
    if (cfg.volume_integral == TRIXI_HIP_VOLUME_SHOCK_CAPTURING_HG && cfg.equation == TRIXI_HIP_INVISCID_BURGERS &&
        cfg.indicator_variable != TRIXI_HIP_INDICATOR_DENSITY)
        return fail(TRIXI_HIP_EUNSUPPORTED, "the Burgers indicator variable is the solution itself (first)");
    if (cfg.indicator_variable < TRIXI_HIP_INDICATOR_DENSITY_PRESSURE ||
        cfg.indicator_variable > TRIXI_HIP_INDICATOR_DENSITY)
        return fail(TRIXI_HIP_EINVAL, "unknown indicator variable");
    if (cfg.boundary_condition != TRIXI_HIP_BC_PERIODIC) {
        if (cfg.boundary_condition != TRIXI_HIP_BC_DIRICHLET && cfg.boundary_condition != TRIXI_HIP_BC_SLIP_WALL)
            return fail(TRIXI_HIP_EINVAL, "unknown boundary condition");
        if (cfg.boundary_condition == TRIXI_HIP_BC_SLIP_WALL && cfg.equation != TRIXI_HIP_SHALLOW_WATER)
            return fail(TRIXI_HIP_EUNSUPPORTED, "slip walls are built for shallow water");
        const bool swe_ok = cfg.equation == TRIXI_HIP_SHALLOW_WATER;
        const bool euler_ok = cfg.equation == TRIXI_HIP_COMPRESSIBLE_EULER &&
                              (cfg.boundary_ic == TRIXI_HIP_IC_CONVERGENCE_TEST || cfg.boundary_ic == TRIXI_HIP_IC_CONSTANT);
        const bool hypdiff_ok = cfg.equation == TRIXI_HIP_HYPERBOLIC_DIFFUSION &&
                                (cfg.boundary_ic == TRIXI_HIP_IC_POISSON_NONPERIODIC ||
                                 cfg.boundary_ic == TRIXI_HIP_IC_HARMONIC_NONPERIODIC ||
                                 cfg.boundary_ic == TRIXI_HIP_IC_CONSTANT);
        const bool burgers_ok = cfg.equation == TRIXI_HIP_INVISCID_BURGERS &&
                                (cfg.boundary_ic == TRIXI_HIP_IC_CONVERGENCE_TEST || cfg.boundary_ic == TRIXI_HIP_IC_CONSTANT ||
                                 cfg.boundary_ic == TRIXI_HIP_IC_SHOCK || cfg.boundary_ic == TRIXI_HIP_IC_RAREFACTION);
        if (!euler_ok && !burgers_ok && !hypdiff_ok && !swe_ok)
            return fail(TRIXI_HIP_EUNSUPPORTED,
                        "Dirichlet boundaries are built for compressible Euler (convergence-test or constant state) "
                        "Burgers (convergence-test wave, shock, rarefaction, constant) and hyperbolic "
                        "diffusion (Poisson steady state, harmonic state, constant)");
        if (cfg.boundary_outflow_faces < 0 || cfg.boundary_outflow_faces >= (1 << (2 * cfg.ndims)))
            return fail(TRIXI_HIP_EINVAL, "outflow face bits beyond the mesh's faces");
    } else if (cfg.boundary_outflow_faces != 0) {
        return fail(TRIXI_HIP_EINVAL, "outflow faces need a non-periodic boundary configuration");
    }
    if (cfg.equation == TRIXI_HIP_HYPERBOLIC_DIFFUSION) {
        if (!(cfg.nu > 0.0) || !(cfg.Lr > 0.0))
            return fail(TRIXI_HIP_EINVAL, "hyperbolic diffusion needs nu > 0 and Lr > 0");
        if (cfg.source_terms != TRIXI_HIP_SOURCE_NONE && cfg.source_terms != TRIXI_HIP_SOURCE_POISSON_NONPERIODIC &&
            cfg.source_terms != TRIXI_HIP_SOURCE_HARMONIC)
            return fail(TRIXI_HIP_EUNSUPPORTED, "hyperbolic diffusion takes the Poisson or harmonic source terms");
    } else if (cfg.source_terms == TRIXI_HIP_SOURCE_POISSON_NONPERIODIC ||
               cfg.source_terms == TRIXI_HIP_SOURCE_HARMONIC) {
        return fail(TRIXI_HIP_EUNSUPPORTED, "these source terms are built for hyperbolic diffusion");
    } else if (cfg.source_terms == TRIXI_HIP_SOURCE_EOC_TEST_EULER) {
        if (cfg.equation != TRIXI_HIP_COMPRESSIBLE_EULER || cfg.ndims == 1)
            return fail(TRIXI_HIP_EUNSUPPORTED, "source_terms_eoc_test_euler is built for compressible Euler 2D/3D");
    } else if (cfg.source_terms != TRIXI_HIP_SOURCE_NONE) {
        if (cfg.source_terms != TRIXI_HIP_SOURCE_CONVERGENCE_TEST) return fail(TRIXI_HIP_EINVAL, "unknown source terms");
        if (cfg.equation != TRIXI_HIP_COMPRESSIBLE_EULER && cfg.equation != TRIXI_HIP_INVISCID_BURGERS &&
            cfg.equation != TRIXI_HIP_COMPRESSIBLE_EULER_QUASI1D && cfg.equation != TRIXI_HIP_SHALLOW_WATER)
            return fail(TRIXI_HIP_EUNSUPPORTED, "source_terms_convergence_test is built for compressible Euler, "
                                                "Burgers, quasi-1D Euler and shallow water");
    }
    const int64_t NM = mesh->nmortars;
    if (NM < 0) return fail(TRIXI_HIP_EINVAL, "negative mortar count");
    if (NM > 0 && cfg.ndims == 1) return fail(TRIXI_HIP_EINVAL, "1D meshes have no mortars");
    const KernelKey key = make_key(cfg, 0, NM > 0);
    LaunchFn fn = find_kernel(key), fn_rk = find_kernel(make_key(cfg, 1, NM > 0));
    MortarFn mfn = nullptr;
    if (NM > 0) {
        mfn = find_mortar({cfg.ndims, cfg.polydeg + 1, cfg.equation, 0, 0, cfg.surface_flux, 0, 0, 0, 0});
        if (!mfn) fn = nullptr;
    }
    if (!fn || !fn_rk) {
        char buf[320];
        snprintf(buf, sizeof buf,
                 "unsupported combination: ndims=%d equation=%d polydeg=%d volume_integral=%d volume_flux=%d "
                 "surface_flux=%d fv_flux=%d precision=%d boundaries/sources/mortars=%d",
                 cfg.ndims, cfg.equation, cfg.polydeg, cfg.volume_integral, cfg.volume_flux, cfg.surface_flux,
                 cfg.fv_flux, cfg.precision, key.ext);
        return fail(TRIXI_HIP_EUNSUPPORTED, buf);
    }
    const int dim = cfg.ndims, N = cfg.polydeg + 1, nv = nvars_of(cfg);
    int nn = 1;
    for (int d = 0; d < dim; ++d) nn *= N;
    const int64_t E = mesh->nelements, I = mesh->ninterfaces;
    if (E <= 0 || I < 0 || !mesh->neighbor_ids || !mesh->orientations || !mesh->inverse_jacobian)
        return fail(TRIXI_HIP_EINVAL, "incomplete mesh");
    if (E > (int64_t)INT32_MAX / 2) return fail(TRIXI_HIP_EINVAL, "too many elements");
    // Face neighbours from the interfaces; every face must be matched at most
    // once (conforming TreeMesh), and faces without an interface are boundary
    // faces, which need a boundary condition.
    std::vector<int32_t> nbr((size_t)2 * dim * E, -1);
    for (int64_t i = 0; i < I; ++i) {
        const int32_t L = mesh->neighbor_ids[2 * i], R = mesh->neighbor_ids[2 * i + 1];
        const int o = mesh->orientations[i];
        if (L < 0 || L >= E || R < 0 || R >= E || o < 1 || o > dim)
            return fail(TRIXI_HIP_EINVAL, "interface " + std::to_string(i) + " out of range");
        int32_t& lp = nbr[(size_t)(2 * (o - 1) + 1) * E + L];
        int32_t& rm = nbr[(size_t)(2 * (o - 1)) * E + R];
        if (lp != -1 || rm != -1)
            return fail(TRIXI_HIP_EINVAL, "face matched twice (non-conforming meshes need mortars)");
        lp = R;
        rm = L;
    }
    // Mortar faces: nbr = -2 - slot, slot m*(K+1)+p for the small element at
    // position p, m*(K+1)+K for the large element; mnb lists the elements
    // across each slot's face.
    const int K = 1 << (dim - 1);
    std::vector<int32_t> mnb;
    if (NM > 0) {
        if (!mesh->mortar_neighbor_ids || !mesh->mortar_orientations || !mesh->mortar_large_sides)
            return fail(TRIXI_HIP_EINVAL, "incomplete mortar arrays");
        if (NM * (K + 1) > (int64_t)INT32_MAX / 2) return fail(TRIXI_HIP_EINVAL, "too many mortars");
        mnb.assign((size_t)NM * (K + 1) * K, -1);
        for (int64_t m = 0; m < NM; ++m) {
            const int o = mesh->mortar_orientations[m], ls = mesh->mortar_large_sides[m];
            const int32_t* ids = mesh->mortar_neighbor_ids + m * (K + 1);
            bool in_range = o >= 1 && o <= dim && (ls == 1 || ls == 2);
            for (int p = 0; p <= K; ++p) in_range = in_range && ids[p] >= 0 && ids[p] < E;
            if (!in_range) return fail(TRIXI_HIP_EINVAL, "mortar " + std::to_string(m) + " out of range");
            const int dir_large = 2 * (o - 1) + (ls == 1 ? 1 : 0), dir_small = 2 * (o - 1) + (ls == 1 ? 0 : 1);
            const int64_t slot_large = m * (K + 1) + K;
            for (int p = 0; p < K; ++p) {
                int32_t& f = nbr[(size_t)dir_small * E + ids[p]];
                if (f != -1) return fail(TRIXI_HIP_EINVAL, "face matched twice (mortar " + std::to_string(m) + ")");
                const int64_t slot = m * (K + 1) + p;
                f = (int32_t)(-2 - slot);
                mnb[(size_t)slot * K] = ids[K];
                mnb[(size_t)slot_large * K + p] = ids[p];
            }
            int32_t& f = nbr[(size_t)dir_large * E + ids[K]];
            if (f != -1) return fail(TRIXI_HIP_EINVAL, "face matched twice (mortar " + std::to_string(m) + ")");
            f = (int32_t)(-2 - slot_large);
        }
    }
    if (cfg.boundary_condition == TRIXI_HIP_BC_PERIODIC)
        for (size_t q = 0; q < nbr.size(); ++q)
            if (nbr[q] == -1)
                return fail(TRIXI_HIP_EUNSUPPORTED,
                            "element face without interface: a non-periodic mesh needs a boundary condition");
    if (cfg.boundary_condition != TRIXI_HIP_BC_PERIODIC && !mesh->node_coordinates)
        return fail(TRIXI_HIP_EINVAL, "boundary states need the node coordinates");

    trixi_hip_ctx* c = new trixi_hip_ctx();
    c->cfg = cfg;
    c->dim = dim;
    c->N = N;
    c->nv = nv;
    c->nn = nn;
    c->E = E;
    c->I = I;
    c->nvalues = (int64_t)nv * nn * E;
    c->total_volume = mesh->total_volume;
    c->basis = make_basis(N);
    c->NA = 2 * (N - 1) + 1;
    c->rhs_fn = fn;
    c->rk_fn = fn_rk;
    c->nmortars = NM;
    c->mK = K;
    c->mortar_fn = mfn;

    auto cleanup_fail = [&](int code, const std::string& msg) {
        free_ctx(c);
        return fail(code, msg);
    };
#define ALLOC(ptr, bytes)                                                      \
    do {                                                                       \
        if (hipMalloc((void**)&(ptr), (bytes)) != hipSuccess)                  \
            return cleanup_fail(TRIXI_HIP_ENOMEM, "hipMalloc failed for " #ptr); \
    } while (0)
    ALLOC(c->d_nbr, nbr.size() * sizeof(int32_t));
    ALLOC(c->d_invJ, E * sizeof(double));
    ALLOC(c->d_x, (size_t)dim * nn * E * sizeof(double));
    ALLOC(c->d_alpha_raw, E * sizeof(double));
    ALLOC(c->d_alpha, E * sizeof(double));
    ALLOC(c->d_time, 16 * sizeof(double));
    ALLOC(c->d_partial, (size_t)c->nblocks_red * 18 * sizeof(double));
    ALLOC(c->d_result, 32 * sizeof(double));
    ALLOC(c->d_Vinv, N * N * sizeof(double));
    ALLOC(c->d_Vdm, c->NA * N * sizeof(double));
    ALLOC(c->d_wa, c->NA * sizeof(double));
    ALLOC(c->d_w, N * sizeof(double));
    ALLOC(c->d_tmp, 2 * (size_t)c->nvalues * sizeof(double));
    if (NM > 0) {
        const size_t slots = (size_t)NM * (K + 1);
        ALLOC(c->d_mids, slots * sizeof(int32_t));
        ALLOC(c->d_mori, NM * sizeof(int32_t));
        ALLOC(c->d_mls, NM * sizeof(int32_t));
        ALLOC(c->d_mnb, slots * K * sizeof(int32_t));
        ALLOC(c->d_mflux, slots * (nn / N) * nv * sizeof(double));
    }
#undef ALLOC
    auto up = [&](void* dst, const void* src, size_t bytes) {
        return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    bool ok = up(c->d_nbr, nbr.data(), nbr.size() * sizeof(int32_t));
    ok = ok && up(c->d_invJ, mesh->inverse_jacobian, E * sizeof(double));
    if (mesh->node_coordinates) {
        std::vector<double> xs((size_t)dim * nn * E);
        for (int64_t e = 0; e < E; ++e)
            for (int n = 0; n < nn; ++n)
                for (int d = 0; d < dim; ++d)
                    xs[((size_t)d * nn + n) * E + e] = mesh->node_coordinates[((size_t)e * nn + n) * dim + d];
        ok = ok && up(c->d_x, xs.data(), xs.size() * sizeof(double));
    } else {
        ok = ok && hipMemset(c->d_x, 0, (size_t)dim * nn * E * sizeof(double)) == hipSuccess;
    }
    ok = ok && up(c->d_Vinv, c->basis.Vinv.data(), N * N * sizeof(double));
    {
        std::vector<double> xa, wa;
        lgl(c->NA, xa, wa);
        std::vector<double> V = interpolation_matrix(c->basis.x, c->basis.wbary, xa);
        ok = ok && up(c->d_Vdm, V.data(), V.size() * sizeof(double));
        ok = ok && up(c->d_wa, wa.data(), wa.size() * sizeof(double));
    }
    ok = ok && up(c->d_w, c->basis.w.data(), N * sizeof(double));
    if (NM > 0) {
        ok = ok && up(c->d_mids, mesh->mortar_neighbor_ids, (size_t)NM * (K + 1) * sizeof(int32_t));
        ok = ok && up(c->d_mori, mesh->mortar_orientations, NM * sizeof(int32_t));
        ok = ok && up(c->d_mls, mesh->mortar_large_sides, NM * sizeof(int32_t));
        ok = ok && up(c->d_mnb, mnb.data(), mnb.size() * sizeof(int32_t));
    }
    if (!ok) return cleanup_fail(TRIXI_HIP_EDEVICE, "host to device copy failed in create");

    RhsArgs& a = c->args;
    memset(&a, 0, sizeof(a));
    a.nbr = c->d_nbr;
    a.invJ = c->d_invJ;
    a.alpha_raw = c->d_alpha_raw;
    a.E = E;
    a.phys.gamma = cfg.gamma;
    a.phys.gm1 = cfg.gamma - 1.0;
    a.phys.inv_gm1 = (cfg.gamma != 1.0) ? 1.0 / (cfg.gamma - 1.0) : 0.0;
    a.phys.c_h = (cfg.equation == TRIXI_HIP_IDEAL_GLM_MHD && dim == 1) ? 0.0 : cfg.c_h;
    if (cfg.equation == TRIXI_HIP_HYPERBOLIC_DIFFUSION) {
        a.phys.nu = cfg.nu;
        a.phys.inv_Tr = cfg.nu / (cfg.Lr * cfg.Lr);  // T_r = Lr^2 / nu
    }
    a.phys.grav = cfg.gravity;
    if (cfg.equation == TRIXI_HIP_COMPRESSIBLE_EULER_MULTICOMPONENT) {
        // c_v,k = R_k / (gamma_k - 1) (Trixi's cv), and c_v,k gamma_k
        for (int k = 0; k < cfg.ncomponents; ++k) {
            a.phys.cv[k] = cfg.gas_constants[k] / (cfg.gammas[k] - 1.0);
            a.phys.cvg[k] = a.phys.cv[k] * cfg.gammas[k];
        }
    }
    for (int d = 0; d < 3; ++d) a.phys.a[d] = cfg.advection_velocity[d];
    a.bc_ic = ic_params(c, cfg.boundary_ic, cfg.boundary_seed);
    for (int q = 0; q < N * N; ++q) {
        a.Ds[q] = c->basis.Ds[q];
        a.Dhat[q] = c->basis.Dhat[q];
    }
    for (int q = 0; q < N; ++q) a.invw[q] = 1.0 / c->basis.w[q];
    a.lhat_m = c->basis.lm[0];
    a.lhat_p = c->basis.lp[N - 1];
    {
        const double eps = 2.220446049250313e-16;
        a.atol = std::max(100 * eps, pow(eps, 0.75));
    }
    a.smooth = cfg.alpha_smooth;
    a.x = c->d_x;
    a.bc = cfg.boundary_condition;
    a.src = cfg.source_terms;
    a.outflow = cfg.boundary_outflow_faces;
    a.mflux = c->d_mflux;
    a.mnb = c->d_mnb;
    a.mK = K;
    {
        MortarArgs& m = c->margs;
        memset(&m, 0, sizeof(m));
        m.mflux = c->d_mflux;
        m.ids = c->d_mids;
        m.ori = c->d_mori;
        m.ls = c->d_mls;
        m.nmortars = NM;
        m.E = E;
        m.phys = a.phys;
        if (NM > 0) mortar_operators(c->basis, N, m.F, m.R);
    }
    a.nxcd = 1;
    {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
            // CDNA3/CDNA4 (gfx942, gfx950) parts are built from 8 XCDs.
            if (strstr(prop.gcnArchName, "gfx950") || strstr(prop.gcnArchName, "gfx942")) a.nxcd = 8;
        }
        const char* env = getenv("TRIXI_HIP_NXCD");
        if (env) a.nxcd = atoi(env) > 0 ? atoi(env) : 1;
    }
    {
        // The node-parallel form for meshes whose line form (one lane per
        // element line and direction) has fewer lanes than two waves per SIMD
        // (256 CUs x 4 SIMDs x 2 waves x 64 lanes); TRIXI_HIP_NP_LANES
        // overrides the threshold, 0 turns the form off.  Eligible shapes as
        // np_eligible() in rhs_kernel.hpp: periodic, conforming, no sources,
        // 2D/3D, flux differencing or HG, at most five variables.
        int64_t np_lanes = 131072;
        const char* env = getenv("TRIXI_HIP_NP_LANES");
        if (env) np_lanes = atoll(env) > 0 ? atoll(env) : 0;
        const KernelKey key = make_key(cfg, 0, NM > 0);
        const bool eligible = key.ext == 0 && cfg.volume_integral != TRIXI_HIP_VOLUME_WEAK_FORM && dim > 1 &&
                              nv <= 5 && nn * dim <= 1024;
        const int64_t line_lanes = E * (nn / N) * dim;
        a.np = eligible && line_lanes < np_lanes ? 1 : 0;
        // shock capturing: the indicator's parameters (Trixi's threshold
        // 0.5 * 10^(-1.8 N^0.25) and logistic steepness log((1 - 1e-4) / 1e-4)),
        // for the fused RK stages of the node-parallel form, which write the
        // next stage's raw blending factors; TRIXI_HIP_FUSE_INDICATOR=0 keeps
        // a separate indicator launch per stage
        a.Vinv = c->d_Vinv;
        a.ind_threshold = 0.5 * pow(10.0, -1.8 * pow((double)N, 0.25));
        a.ind_s = log((1.0 - 0.0001) / 0.0001);
        a.ind_alpha_min = cfg.alpha_min;
        a.ind_alpha_max = cfg.alpha_max;
        a.ind_var = cfg.indicator_variable;
        const char* fenv = getenv("TRIXI_HIP_FUSE_INDICATOR");
        c->fuse_ind = a.np && cfg.volume_integral == TRIXI_HIP_VOLUME_SHOCK_CAPTURING_HG &&
                      cfg.equation == TRIXI_HIP_COMPRESSIBLE_EULER && (N == 4 || N == 5) &&
                      !(fenv && atoi(fenv) == 0);
    }
    {
        // look up this device's static scratch once, at set-up, with the
        // device current (used by the *_noalloc entry points)
        double* sp = nullptr;
        const int rc = scratch_ptr(&sp);
        if (rc) {
            free_ctx(c);
            return rc;
        }
    }
    *out = c;
    return TRIXI_HIP_OK;
}

int trixi_hip_destroy(trixi_hip_ctx* ctx) {
    if (!ctx) return fail(TRIXI_HIP_EINVAL, "null context");
    free_ctx(ctx);
    return TRIXI_HIP_OK;
}

int trixi_hip_sizes(const trixi_hip_ctx* ctx, int32_t* nvars, int32_t* nnodes, int64_t* nvalues) {
    if (!ctx) return fail(TRIXI_HIP_EINVAL, "null context");
    if (nvars) *nvars = ctx->nv;
    if (nnodes) *nnodes = ctx->N;
    if (nvalues) *nvalues = ctx->nvalues;
    return TRIXI_HIP_OK;
}

static const char* const kNeedF32 = "this context's state is FP32: use trixi_hip_rhs_f32 / trixi_hip_rk_step_f32";
static const char* const kNeedF64 = "this context's state is FP64: use trixi_hip_rhs / trixi_hip_rk_step";

// rhs! of state u into du (state type of the context)
static int enqueue_rhs(trixi_hip_ctx* ctx, void* du, const void* u, double t, hipStream_t s);

int trixi_hip_rhs(trixi_hip_ctx* ctx, double* du, const double* u, double t, trixi_hip_stream stream) {
    if (!ctx || !du || !u) return fail(TRIXI_HIP_EINVAL, "null argument");
    if (fp32_state(ctx)) return fail(TRIXI_HIP_EINVAL, kNeedF32);
    return enqueue_rhs(ctx, du, u, t, (hipStream_t)stream);
}

int trixi_hip_rhs_f32(trixi_hip_ctx* ctx, float* du, const float* u, double t, trixi_hip_stream stream) {
    if (!ctx || !du || !u) return fail(TRIXI_HIP_EINVAL, "null argument");
    if (!fp32_state(ctx)) return fail(TRIXI_HIP_EINVAL, kNeedF64);
    return enqueue_rhs(ctx, du, u, t, (hipStream_t)stream);
}

static int enqueue_rhs(trixi_hip_ctx* ctx, void* du, const void* u, double t, hipStream_t s) {
    if (du == u) return fail(TRIXI_HIP_EINVAL, "du and u must not alias");
    if (ctx->cfg.volume_integral == TRIXI_HIP_VOLUME_SHOCK_CAPTURING_HG) {
        int rc = launch_indicator(ctx, u, s, fp32_state(ctx));
        if (rc) return rc;
    }
    if (!fp32_state(ctx)) {
        int rc = launch_mortars(ctx, static_cast<const double*>(u), s);
        if (rc) return rc;
    }
    RhsArgs a = ctx->args;
    a.u = u;
    a.du = du;
    a.t_host = t;
    HIP_TRY(ctx->rhs_fn(a, s));
    return TRIXI_HIP_OK;
}

int trixi_hip_pack(trixi_hip_ctx* ctx, double* u_native, const double* u_trixi, trixi_hip_stream stream) {
    if (!ctx || !u_native || !u_trixi) return fail(TRIXI_HIP_EINVAL, "null argument");
    hipLaunchKernelGGL(pack_kernel, dim3(1024), dim3(256), 0, (hipStream_t)stream, u_native, u_trixi, ctx->nv,
                       ctx->nn, ctx->E, 1);
    HIP_TRY(hipGetLastError());
    return TRIXI_HIP_OK;
}

int trixi_hip_unpack(trixi_hip_ctx* ctx, double* u_trixi, const double* u_native, trixi_hip_stream stream) {
    if (!ctx || !u_native || !u_trixi) return fail(TRIXI_HIP_EINVAL, "null argument");
    hipLaunchKernelGGL(pack_kernel, dim3(1024), dim3(256), 0, (hipStream_t)stream, u_trixi, u_native, ctx->nv,
                       ctx->nn, ctx->E, 0);
    HIP_TRY(hipGetLastError());
    return TRIXI_HIP_OK;
}

static IcParams ic_params(const trixi_hip_ctx* c, int ic, uint64_t seed) {
    IcParams q;
    q.dim = c->dim;
    q.eq = c->cfg.equation;
    q.nv = c->nv;
    q.ic = ic;
    q.gamma = c->cfg.gamma;
    q.inv_gm1 = c->args.phys.inv_gm1;
    q.c_h = c->cfg.c_h;
    for (int d = 0; d < 3; ++d) q.a[d] = c->cfg.advection_velocity[d];
    q.length = c->cfg.ic_length;
    q.seed = seed;
    q.nc = c->cfg.equation == TRIXI_HIP_COMPRESSIBLE_EULER_MULTICOMPONENT ? c->cfg.ncomponents : 0;
    q.grav = c->cfg.gravity;
    for (int k = 0; k < MAXC; ++k) {
        q.cv[k] = c->args.phys.cv[k];
        q.cvg[k] = c->args.phys.cvg[k];
    }
    return q;
}

int trixi_hip_compute_coefficients(trixi_hip_ctx* ctx, double* u, int32_t ic, double t, uint64_t seed,
                                   trixi_hip_stream stream) {
    if (!ctx || !u) return fail(TRIXI_HIP_EINVAL, "null argument");
    if (ic < 0 || ic > TRIXI_HIP_IC_HARMONIC_NONPERIODIC) return fail(TRIXI_HIP_EINVAL, "unknown initial condition");
    hipLaunchKernelGGL(coefficients_kernel, dim3(1024), dim3(256), 0, (hipStream_t)stream, u, ctx->d_x,
                       ic_params(ctx, ic, seed), ctx->N, ctx->nn, ctx->E, t);
    HIP_TRY(hipGetLastError());
    return TRIXI_HIP_OK;
}

// The five stages of CarpenterKennedy2N54, each one fused rhs+update
// launch (plus the HG indicator when shock capturing).  A stage reads its
// input state while other workgroups still read that state's face nodes, so
// it cannot update in place: stage 0 reads u and writes buffer A, then the
// stages alternate A -> B -> A -> B, and stage 4 reads B and writes the
// caller's u, which no stage after 0 reads.
static int enqueue_rk_stages(trixi_hip_ctx* c, void* u, void* k, hipStream_t s) {
    // two stage buffers of the context's state type in d_tmp
    const size_t bytes = (size_t)c->nvalues * (fp32_state(c) ? sizeof(float) : sizeof(double));
    void* bufA = c->d_tmp;
    void* bufB = static_cast<char*>(static_cast<void*>(c->d_tmp)) + bytes;
    void* const in[5] = {u, bufA, bufB, bufA, bufB};
    void* const out[5] = {bufA, bufB, bufA, bufB, u};
    // fused shock-capturing stages: stage st reads the raw blending factors
    // of its input from alpha[st % 2] and writes those of its output (the
    // next stage's input) to alpha[(st + 1) % 2]; only the first stage needs
    // an indicator launch
    double* const alpha[2] = {c->d_alpha_raw, c->d_alpha};
    for (int st = 0; st < 5; ++st) {
        void* cur = in[st];
        void* nxt = out[st];
        if (c->cfg.volume_integral == TRIXI_HIP_VOLUME_SHOCK_CAPTURING_HG && (!c->fuse_ind || st == 0)) {
            int rc = launch_indicator(c, cur, s, fp32_state(c));
            if (rc) return rc;
        }
        if (!fp32_state(c)) {
            int rc = launch_mortars(c, static_cast<const double*>(cur), s);
            if (rc) return rc;
        }
        RhsArgs a = c->args;
        a.u = cur;
        a.k = k;
        a.u_out = nxt;
        a.time = c->d_time + 2 * st;
        a.rk_a = CK_A[st];
        a.rk_b = CK_B[st];
        a.rk_first = (st == 0);
        if (c->fuse_ind) {
            a.alpha_raw = alpha[st % 2];
            a.alpha_next = st < 4 ? alpha[(st + 1) % 2] : nullptr;
        }
        HIP_TRY(c->rk_fn(a, s));
    }
    return TRIXI_HIP_OK;
}

int trixi_hip_rk_stage(trixi_hip_ctx* ctx, const double* u, double* k, double* u_out, int32_t stage, double t,
                       double dt, trixi_hip_stream stream) {
    if (!ctx || !u || !k || !u_out) return fail(TRIXI_HIP_EINVAL, "null argument");
    if (stage < 0 || stage > 4) return fail(TRIXI_HIP_EINVAL, "stage must be 0..4");
    if (u_out == u || k == u || k == u_out) return fail(TRIXI_HIP_EINVAL, "u, k and u_out must be distinct");
    if (fp32_state(ctx)) return fail(TRIXI_HIP_EINVAL, kNeedF32);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(set_time_kernel, dim3(1), dim3(1), 0, s, ctx->d_time, t, dt, CK_C[0], CK_C[1], CK_C[2],
                       CK_C[3], CK_C[4]);
    HIP_TRY(hipGetLastError());
    if (ctx->cfg.volume_integral == TRIXI_HIP_VOLUME_SHOCK_CAPTURING_HG) {
        int rc = launch_indicator(ctx, u, s);
        if (rc) return rc;
    }
    {
        int rc = launch_mortars(ctx, u, s);
        if (rc) return rc;
    }
    RhsArgs a = ctx->args;
    a.u = u;
    a.k = k;
    a.u_out = u_out;
    a.time = ctx->d_time + 2 * stage;
    a.rk_a = CK_A[stage];
    a.rk_b = CK_B[stage];
    a.rk_first = (stage == 0);
    HIP_TRY(ctx->rk_fn(a, s));
    return TRIXI_HIP_OK;
}

int trixi_hip_sync(trixi_hip_stream stream) {
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    HIP_TRY(hipGetLastError());
    return TRIXI_HIP_OK;
}

static int rk_step_any(trixi_hip_ctx* ctx, void* u, void* k, double t, double dt, int32_t use_graph,
                       hipStream_t s);

int trixi_hip_rk_step(trixi_hip_ctx* ctx, double* u, double* k, double t, double dt, int32_t use_graph,
                      trixi_hip_stream stream) {
    if (!ctx || !u || !k) return fail(TRIXI_HIP_EINVAL, "null argument");
    if (fp32_state(ctx)) return fail(TRIXI_HIP_EINVAL, kNeedF32);
    return rk_step_any(ctx, u, k, t, dt, use_graph, (hipStream_t)stream);
}

int trixi_hip_rk_step_f32(trixi_hip_ctx* ctx, float* u, float* k, double t, double dt, int32_t use_graph,
                          trixi_hip_stream stream) {
    if (!ctx || !u || !k) return fail(TRIXI_HIP_EINVAL, "null argument");
    if (!fp32_state(ctx)) return fail(TRIXI_HIP_EINVAL, kNeedF64);
    return rk_step_any(ctx, u, k, t, dt, use_graph, (hipStream_t)stream);
}

static int rk_step_any(trixi_hip_ctx* ctx, void* u, void* k, double t, double dt, int32_t use_graph,
                       hipStream_t s) {
    if (u == k) return fail(TRIXI_HIP_EINVAL, "u and k must be distinct");
    hipLaunchKernelGGL(set_time_kernel, dim3(1), dim3(1), 0, s, ctx->d_time, t, dt, CK_C[0], CK_C[1], CK_C[2],
                       CK_C[3], CK_C[4]);
    HIP_TRY(hipGetLastError());
    if (!use_graph) return enqueue_rk_stages(ctx, u, k, s);
    if (!s) return fail(TRIXI_HIP_EINVAL, "graph replay needs an explicit (non-default) stream");
    if (!ctx->has_graph || ctx->graph_u != u || ctx->graph_k != k) {
        if (ctx->graph_exec) {
            HIP_TRY(hipGraphExecDestroy(ctx->graph_exec));
            ctx->graph_exec = nullptr;
        }
        if (ctx->graph) {
            HIP_TRY(hipGraphDestroy(ctx->graph));
            ctx->graph = nullptr;
        }
        ctx->has_graph = false;
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_rk_stages(ctx, u, k, s);
        hipGraph_t g = nullptr;
        hipError_t ec = hipStreamEndCapture(s, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess) return fail(TRIXI_HIP_EDEVICE, std::string("graph capture: ") + hipGetErrorString(ec));
        ctx->graph = g;
        HIP_TRY(hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0));
        ctx->graph_u = u;
        ctx->graph_k = k;
        ctx->has_graph = true;
    }
    HIP_TRY(hipGraphLaunch(ctx->graph_exec, s));
    return TRIXI_HIP_OK;
}

// max_dt with its block partials in `partial` (at least `blocks` doubles) and
// the reduced maximum in `result`.
static int max_dt_impl(trixi_hip_ctx* ctx, const double* u, double* max_dt, hipStream_t s, int blocks,
                       double* partial, double* result) {
    const int threads = 256;
    const PhysParams& P = ctx->args.phys;
    const int eq = ctx->cfg.equation, dim = ctx->dim;
#define MS(EQT, D) \
    hipLaunchKernelGGL((max_speed_kernel<EQT, D>), dim3(blocks), dim3(threads), 0, s, u, ctx->d_invJ, partial, ctx->nn, ctx->E, P)
    if (eq == TRIXI_HIP_LINEAR_SCALAR_ADVECTION) {
        if (dim == 1) MS(LinAdv<1>, 1); else if (dim == 2) MS(LinAdv<2>, 2); else MS(LinAdv<3>, 3);
    } else if (eq == TRIXI_HIP_COMPRESSIBLE_EULER) {
        if (dim == 1) MS(Euler<1>, 1); else if (dim == 2) MS(Euler<2>, 2); else MS(Euler<3>, 3);
    } else if (eq == TRIXI_HIP_INVISCID_BURGERS) {
        MS(Burgers, 1);
    } else if (eq == TRIXI_HIP_HYPERBOLIC_DIFFUSION) {
        if (dim == 1) MS(HypDiff<1>, 1); else if (dim == 2) MS(HypDiff<2>, 2); else MS(HypDiff<3>, 3);
    } else if (eq == TRIXI_HIP_COMPRESSIBLE_EULER_QUASI1D) {
        MS(EulerQuasi1D, 1);
    } else if (eq == TRIXI_HIP_SHALLOW_WATER) {
        if (dim == 1) MS(ShallowWater<1>, 1); else MS(ShallowWater<2>, 2);
    } else if (eq == TRIXI_HIP_COMPRESSIBLE_EULER_MULTICOMPONENT) {
        using M11 = EulerMulti<1, 1>;
        using M12 = EulerMulti<1, 2>;
        using M21 = EulerMulti<2, 1>;
        using M22 = EulerMulti<2, 2>;
        const int nc = ctx->cfg.ncomponents;
        if (dim == 1) {
            if (nc == 1) MS(M11, 1); else MS(M12, 1);
        } else {
            if (nc == 1) MS(M21, 2); else MS(M22, 2);
        }
    } else {
        if (dim == 1) MS(GlmMhd1D, 1); else if (dim == 2) MS(GlmMhd, 2); else MS(GlmMhd, 3);
    }
#undef MS
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(32), 0, s, partial, result, blocks, 1, 0);
    HIP_TRY(hipGetLastError());
    double m = 0.0;
    HIP_TRY(hipMemcpyAsync(&m, result, sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const double tiny = nextafter(0.0, 1.0);
    if (!(m > tiny)) m = tiny;
    *max_dt = 2.0 / (ctx->N * m);
    return TRIXI_HIP_OK;
}

int trixi_hip_max_dt(trixi_hip_ctx* ctx, const double* u, double* max_dt, trixi_hip_stream stream) {
    if (!ctx || !u || !max_dt) return fail(TRIXI_HIP_EINVAL, "null argument");
    const int blocks = (int)std::min<int64_t>((ctx->E + 255) / 256, ctx->nblocks_red);
    return max_dt_impl(ctx, u, max_dt, (hipStream_t)stream, blocks, ctx->d_partial, ctx->d_result);
}

int trixi_hip_max_dt_noalloc(trixi_hip_ctx* ctx, const double* u, double* max_dt, trixi_hip_stream stream) {
    if (!ctx || !u || !max_dt) return fail(TRIXI_HIP_EINVAL, "null argument");
    // the same grid as trixi_hip_max_dt: the same partials and maximum
    const int blocks = (int)std::min<int64_t>((ctx->E + 255) / 256, ctx->nblocks_red);
    if (blocks + 1 > TRIXI_HIP_SCRATCH_DOUBLES)
        return fail(TRIXI_HIP_EINVAL, "max_dt_noalloc: " + std::to_string(blocks) +
                                          " block partials exceed the static scratch of " +
                                          std::to_string(TRIXI_HIP_SCRATCH_DOUBLES - 1));
    double* sp = nullptr;
    const int rc = scratch_ptr(&sp);
    if (rc) return rc;
    // partials in [0, blocks), the maximum in the last slot; every partial is
    // written before it is read, so nothing needs clearing between calls
    return max_dt_impl(ctx, u, max_dt, (hipStream_t)stream, blocks, sp, sp + TRIXI_HIP_SCRATCH_DOUBLES - 1);
}

int trixi_hip_error_norms(trixi_hip_ctx* ctx, const double* u, int32_t ic, double t, uint64_t seed, double* l2,
                          double* linf, trixi_hip_stream stream) {
    if (!ctx || !u || !l2 || !linf) return fail(TRIXI_HIP_EINVAL, "null argument");
    if (ic < 0 || ic > TRIXI_HIP_IC_HARMONIC_NONPERIODIC) return fail(TRIXI_HIP_EINVAL, "unknown initial condition");
    hipStream_t s = (hipStream_t)stream;
    const int threads = 256;
    const int64_t npts = (int64_t)ctx->E * (int64_t)lrint(pow((double)ctx->NA, (double)ctx->dim));
    const int blocks = (int)std::min<int64_t>((npts + threads - 1) / threads, ctx->nblocks_red);
    const IcParams q = ic_params(ctx, ic, seed);
    const int nv = ctx->nv;
#define EN(D)                                                                                                   \
    hipLaunchKernelGGL((error_norms_kernel<D>), dim3(blocks), dim3(threads), 0, s, u, ctx->d_x, ctx->d_invJ,    \
                       ctx->d_partial, ctx->d_Vdm, ctx->d_wa, ctx->N, ctx->NA, ctx->E, q, t, nv)
    if (ctx->dim == 1) EN(1); else if (ctx->dim == 2) EN(2); else EN(3);
#undef EN
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(32), 0, s, ctx->d_partial, ctx->d_result, blocks,
                       2 * nv, nv);
    HIP_TRY(hipGetLastError());
    double res[18];
    HIP_TRY(hipMemcpyAsync(res, ctx->d_result, 2 * nv * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int v = 0; v < nv; ++v) {
        l2[v] = sqrt(res[v] / ctx->total_volume);
        linf[v] = res[nv + v];
    }
    return TRIXI_HIP_OK;
}

int trixi_hip_integrate(trixi_hip_ctx* ctx, const double* u, double* integral, trixi_hip_stream stream) {
    if (!ctx || !u || !integral) return fail(TRIXI_HIP_EINVAL, "null argument");
    hipStream_t s = (hipStream_t)stream;
    const int threads = 256;
    const int blocks = (int)std::min<int64_t>((ctx->E + threads - 1) / threads, ctx->nblocks_red);
    hipLaunchKernelGGL(integrate_kernel, dim3(blocks), dim3(threads), 0, s, u, ctx->d_invJ, ctx->d_partial,
                       ctx->d_w, ctx->N, ctx->dim, ctx->E, ctx->nv);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(32), 0, s, ctx->d_partial, ctx->d_result, blocks,
                       ctx->nv, ctx->nv);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(integral, ctx->d_result, ctx->nv * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return TRIXI_HIP_OK;
}

int trixi_hip_indicator(trixi_hip_ctx* ctx, const double* u, double* alpha, trixi_hip_stream stream) {
    if (!ctx || !u || !alpha) return fail(TRIXI_HIP_EINVAL, "null argument");
    const trixi_hip_config& cfg = ctx->cfg;
    const bool built = cfg.equation == TRIXI_HIP_COMPRESSIBLE_EULER ||
                       (cfg.equation == TRIXI_HIP_IDEAL_GLM_MHD && cfg.ndims > 1 &&
                        cfg.indicator_variable == TRIXI_HIP_INDICATOR_DENSITY_PRESSURE) ||
                       (cfg.equation == TRIXI_HIP_INVISCID_BURGERS &&
                        cfg.indicator_variable == TRIXI_HIP_INDICATOR_DENSITY) ||
                       (cfg.equation == TRIXI_HIP_SHALLOW_WATER &&
                        cfg.indicator_variable == TRIXI_HIP_INDICATOR_DENSITY_PRESSURE);
    if (!built)
        return fail(TRIXI_HIP_EUNSUPPORTED,
                    "the HG indicator is built for compressible Euler, 2D/3D GLM-MHD (density_pressure) and "
                    "Burgers (first)");
    hipStream_t s = (hipStream_t)stream;
    int rc = launch_indicator(ctx, u, s);
    if (rc) return rc;
    if (ctx->cfg.alpha_smooth) {
        hipLaunchKernelGGL(smooth_alpha_kernel, dim3((unsigned)((ctx->E + 255) / 256)), dim3(256), 0, s,
                           ctx->d_alpha_raw, alpha, ctx->d_nbr, ctx->dim, ctx->E, ctx->d_mnb, ctx->mK);
    } else {
        hipLaunchKernelGGL(copy_kernel, dim3(256), dim3(256), 0, s, alpha, ctx->d_alpha_raw, ctx->E);
    }
    HIP_TRY(hipGetLastError());
    return TRIXI_HIP_OK;
}

static VjpArgs vjp_args(const trixi_hip_ctx* c) {
    VjpArgs v;
    memset(&v, 0, sizeof(v));
    v.nbr = c->d_nbr;
    v.invJ = c->d_invJ;
    v.E = c->E;
    v.phys = c->args.phys;
    for (int q = 0; q < MAXN * MAXN; ++q) {
        v.Ds[q] = c->args.Ds[q];
        v.Dhat[q] = c->args.Dhat[q];
    }
    for (int q = 0; q < MAXN; ++q) v.invw[q] = c->args.invw[q];
    v.lhat_m = c->args.lhat_m;
    v.lhat_p = c->args.lhat_p;
    v.nxcd = c->args.nxcd;
    // the direction-group form below two waves per SIMD of the one-group
    // form (as the rhs kernel's node-parallel form); 0 turns it off
    v.pd_lanes = 131072;
    if (const char* env = getenv("TRIXI_HIP_VJP_PD_LANES")) v.pd_lanes = atoll(env) > 0 ? atoll(env) : 0;
    v.alpha_raw = c->d_alpha_raw;
    v.alphabar = c->d_alphabar;
    v.winner = c->d_winner;
    v.atol = c->args.atol;
    v.smooth = c->args.smooth;
    v.x = c->d_x;
    v.bc_ic = c->args.bc_ic;
    v.bc = c->args.bc;
    return v;
}

static bool is_hg(const trixi_hip_ctx* c) {
    return c->cfg.volume_integral == TRIXI_HIP_VOLUME_SHOCK_CAPTURING_HG;
}

static int require_vjp(trixi_hip_ctx* c) {
    if (c->vjp_fn) return TRIXI_HIP_OK;
    if (fp32_state(c)) return fail(TRIXI_HIP_EUNSUPPORTED, "the adjoint takes FP64 states (no FP32-state adjoint)");
    if (c->nmortars > 0) return fail(TRIXI_HIP_EUNSUPPORTED, "no adjoint kernel for meshes with mortars");
    if (c->cfg.boundary_outflow_faces != 0)
        return fail(TRIXI_HIP_EUNSUPPORTED, "no adjoint kernel for outflow boundaries");
    KernelKey k = make_key(c->cfg, 0);
    k.prec = 0;
    VjpFn fn = find_vjp(k);
    if (!fn)
        return fail(TRIXI_HIP_EUNSUPPORTED,
                    "no adjoint kernel for this configuration (built: linear advection, compressible "
                    "Euler (weak form, flux differencing, HG shock capturing) and GLM-MHD flux differencing, "
                    "polydeg 3)");
    // the reverse indicator (indicator_vjp_kernel) is built for polydeg 3 and
    // the compressible Euler variables only
    if (is_hg(c) && (c->N != 4 || c->cfg.equation != TRIXI_HIP_COMPRESSIBLE_EULER))
        return fail(TRIXI_HIP_EUNSUPPORTED, "the shock-capturing adjoint is built for compressible Euler, polydeg 3");
    if (is_hg(c) && !c->d_alphabar) {
        if (hipMalloc((void**)&c->d_alphabar, (size_t)c->E * sizeof(double)) != hipSuccess) {
            c->d_alphabar = nullptr;
            return fail(TRIXI_HIP_ENOMEM, "hipMalloc failed for the shock-capturing adjoint buffers");
        }
        if (hipMalloc((void**)&c->d_winner, (size_t)c->E * sizeof(int32_t)) != hipSuccess) {
            (void)hipFree(c->d_alphabar);
            c->d_alphabar = nullptr;
            c->d_winner = nullptr;
            return fail(TRIXI_HIP_ENOMEM, "hipMalloc failed for the shock-capturing adjoint buffers");
        }
    }
    c->vjp_fn = fn;
    return TRIXI_HIP_OK;
}

// ubar = beta ubar + J(v.u)^T (alpha vbar): the VJP kernel and, for HG shock
// capturing, the blending factor of v.u before it and its reverse after it.
// alpha_known: v.alpha_raw already holds the raw blending factors of v.u
static int enqueue_vjp(trixi_hip_ctx* c, const VjpArgs& v, hipStream_t s, bool alpha_known = false) {
    if (is_hg(c) && !alpha_known) {
        int rc = launch_indicator(c, v.u, s);
        if (rc) return rc;
    }
    HIP_TRY(c->vjp_fn(v, s));
    if (is_hg(c)) {
        const int epb = 64 / (c->dim == 1 ? 4 : c->dim == 2 ? 16 : 64);
        const unsigned blocks = (unsigned)((c->E + epb - 1) / epb);
        const int threads = 64;
        const double threshold = 0.5 * pow(10.0, -1.8 * pow((double)c->N, 0.25));
        const double parameter_s = log((1.0 - 0.0001) / 0.0001);
        const PhysParams& P = c->args.phys;
#define TRIXI_HIP_IND_VJP(D)                                                                                    \
    hipLaunchKernelGGL((indicator_vjp_kernel<D, 4>), dim3(blocks), dim3(threads), 0, s, v.u, v.ubar,            \
                       (const double*)c->d_alphabar, (const int32_t*)c->d_winner, (const int32_t*)c->d_nbr,    \
                       c->cfg.alpha_smooth, c->E, P, (const double*)c->d_Vinv, threshold, parameter_s,       \
                       c->cfg.alpha_min, c->cfg.alpha_max, c->cfg.indicator_variable, c->nv, v.kbar_out,     \
                       v.kbar_b)
        if (c->dim == 1)
            TRIXI_HIP_IND_VJP(1);
        else if (c->dim == 2)
            TRIXI_HIP_IND_VJP(2);
        else
            TRIXI_HIP_IND_VJP(3);
#undef TRIXI_HIP_IND_VJP
        HIP_TRY(hipGetLastError());
    }
    return TRIXI_HIP_OK;
}

int trixi_hip_rhs_vjp(trixi_hip_ctx* ctx, double* ubar, const double* u, const double* vbar, double t,
                      double alpha, double beta, trixi_hip_stream stream) {
    if (!ctx || !ubar || !u || !vbar) return fail(TRIXI_HIP_EINVAL, "null argument");
    if (ubar == vbar || ubar == u) return fail(TRIXI_HIP_EINVAL, "ubar must not alias u or vbar");
    int rc = require_vjp(ctx);
    if (rc) return rc;
    VjpArgs v = vjp_args(ctx);
    v.u = u;
    v.vbar = vbar;
    v.ubar = ubar;
    v.alpha = alpha;
    v.beta = beta;
    v.t = t;
    return enqueue_vjp(ctx, v, (hipStream_t)stream);
}

int trixi_hip_enable_adjoint(trixi_hip_ctx* ctx) {
    if (!ctx) return fail(TRIXI_HIP_EINVAL, "null context");
    int rc = require_vjp(ctx);
    if (rc) return rc;
    if (ctx->adjoint) return TRIXI_HIP_OK;
    trixi_hip_config c64 = ctx->cfg;
    c64.precision = TRIXI_HIP_FP64;
    ctx->rk64_fn = find_kernel(make_key(c64, 1));
    if (!ctx->rk64_fn) return fail(TRIXI_HIP_EUNSUPPORTED, "no FP64 stage kernel for the adjoint");
    const size_t bytes = (size_t)ctx->nvalues * sizeof(double);
    double** bufs[4] = {&ctx->d_stage, &ctx->d_k, &ctx->d_kbar, &ctx->d_stage_alpha};
    // the stage inputs' blending factors only where the stages compute them
    const size_t sizes[4] = {5 * bytes, bytes, bytes,
                             is_hg(ctx) && ctx->fuse_ind ? 5 * (size_t)ctx->E * sizeof(double) : 0};
    for (int b = 0; b < 4; ++b) {
        if (*bufs[b] || sizes[b] == 0) continue;
        if (hipMalloc((void**)bufs[b], sizes[b]) != hipSuccess) {
            // leave no partial set behind: a retry starts from scratch
            *bufs[b] = nullptr;
            for (int q = 0; q < b; ++q) {
                (void)hipFree(*bufs[q]);
                *bufs[q] = nullptr;
            }
            return fail(TRIXI_HIP_ENOMEM, "hipMalloc failed for the adjoint buffers");
        }
    }
    ctx->adjoint = true;
    return TRIXI_HIP_OK;
}

int trixi_hip_rk_step_vjp(trixi_hip_ctx* ctx, const double* u_n, double* ubar, double t, double dt,
                          trixi_hip_stream stream) {
    if (!ctx || !u_n || !ubar) return fail(TRIXI_HIP_EINVAL, "null argument");
    if (!ctx->adjoint) return fail(TRIXI_HIP_EINVAL, "call trixi_hip_enable_adjoint first");
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = ctx->nvalues;
    // stage inputs: U_0 is the checkpoint itself (only read), U_1 .. U_4 are
    // recomputed in FP64 into the stage buffers
    const double* U[5];
    double* Uw[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    U[0] = u_n;
    for (int st = 1; st < 5; ++st) U[st] = Uw[st] = ctx->d_stage + (size_t)st * n;
    hipLaunchKernelGGL(set_time_kernel, dim3(1), dim3(1), 0, s, ctx->d_time, t, dt, CK_C[0], CK_C[1], CK_C[2],
                       CK_C[3], CK_C[4]);
    HIP_TRY(hipGetLastError());
    // fused shock-capturing stages: one indicator launch for U_0, and every
    // stage writes the raw blending factors of the next input; they are kept
    // per stage for the reverse sweep
    const bool fused = is_hg(ctx) && ctx->fuse_ind && ctx->d_stage_alpha;
    double* A[5];
    for (int st = 0; st < 5; ++st) A[st] = fused ? ctx->d_stage_alpha + (size_t)st * ctx->E : nullptr;
    for (int st = 0; st < 4; ++st) {
        if (is_hg(ctx) && (!fused || st == 0)) {
            int rc = launch_indicator(ctx, U[st], s, false, A[st]);
            if (rc) return rc;
        }
        RhsArgs a = ctx->args;
        a.u = U[st];
        a.k = ctx->d_k;
        a.u_out = Uw[st + 1];
        a.time = ctx->d_time + 2 * st;
        a.rk_a = CK_A[st];
        a.rk_b = CK_B[st];
        a.rk_first = (st == 0);
        if (fused) {
            a.alpha_raw = A[st];
            a.alpha_next = A[st + 1];
        }
        HIP_TRY(ctx->rk64_fn(a, s));
    }
    // reverse the stages: kbar_4 = B_4 ubar here, then each stage's VJP
    // leaves kbar_{s-1} = A_s kbar_s + B_{s-1} ubar in the other register
    // buffer as it finishes each node (the stage buffers' first slot is free:
    // U_0 is u_n)
    double* Kb[2] = {ctx->d_kbar, ctx->d_stage};
    int cur = 0;
    hipLaunchKernelGGL(kbar_update_kernel, dim3(1024), dim3(256), 0, s, Kb[cur], (const double*)ubar, 0.0, CK_B[4],
                       1, n);
    HIP_TRY(hipGetLastError());
    for (int st = 4; st >= 0; --st) {
        VjpArgs v = vjp_args(ctx);
        v.u = U[st];
        v.vbar = Kb[cur];
        if (st > 0) {
            v.kbar_out = Kb[cur ^ 1];
            v.kbar_a = CK_A[st];
            v.kbar_b = CK_B[st - 1];
        }
        v.ubar = ubar;
        v.alpha = dt;
        v.beta = 1.0;
        v.t = t + CK_C[st] * dt;
        if (fused) v.alpha_raw = A[st];
        int rc = enqueue_vjp(ctx, v, s, fused);
        if (rc) return rc;
        cur ^= 1;
    }
    return TRIXI_HIP_OK;
}

int trixi_hip_check_state(trixi_hip_ctx* ctx, const double* u, int32_t* flags, int64_t* nbad,
                          trixi_hip_stream stream) {
    if (!ctx || !u || !nbad) return fail(TRIXI_HIP_EINVAL, "null argument");
    hipStream_t s = (hipStream_t)stream;
    const int threads = 256;
    const int blocks = (int)std::min<int64_t>((ctx->E + threads - 1) / threads, ctx->nblocks_red);
    hipLaunchKernelGGL(check_state_kernel, dim3(blocks), dim3(threads), 0, s, u, flags, ctx->d_partial,
                       ctx->cfg.equation, ctx->dim, ctx->nv, ctx->nn, ctx->E, ctx->args.phys);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(32), 0, s, ctx->d_partial, ctx->d_result, blocks, 1, 1);
    HIP_TRY(hipGetLastError());
    double cnt = 0.0;
    HIP_TRY(hipMemcpyAsync(&cnt, ctx->d_result, sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *nbad = (int64_t)cnt;
    if (*nbad > 0) return fail(TRIXI_HIP_ESTATE, std::to_string(*nbad) + " element(s) with invalid state");
    return TRIXI_HIP_OK;
}

}  // extern "C"
