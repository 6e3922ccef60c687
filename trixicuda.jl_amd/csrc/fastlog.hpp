// fastlog.hpp -- natural logarithm in FP64 for the logarithmic means.
//
// The node pass takes log(rho) and log(p) at every node; the device library's
// log is a double-double evaluation of about a hundred instructions.  This
// one reduces x = m 2^e with m in [sqrt(1/2), sqrt(2)), so that
// s = (m - 1)/(m + 1) lies in [-0.1716, 0.1716], and sums the series
//     log(m) = 2 atanh(s) = 2 s (1 + s^2/3 + s^4/5 + ... + s^18/19),
// whose first omitted term is below 2^-53 of the sum (s^20/21 < 5e-17), then
// adds e log(2) with log(2) split into a 32-bit-exact high part and a low
// part.  About 30 instructions; within 2 ulp of the correctly rounded value
// for every positive finite x (pinned by a CPU test against long double),
// exact at x = 1; log(0) = -inf, log(x < 0) = NaN, log(+inf) = +inf,
// log(NaN) = NaN.
//
// Plain C++ (host and device) so that the CPU test compiles the same source.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

#ifdef __HIPCC__
#define TRIXI_FASTLOG_QUAL __host__ __device__ __forceinline__
#else
#define TRIXI_FASTLOG_QUAL inline
#endif

namespace trixi_hip {

TRIXI_FASTLOG_QUAL double fast_log(double x) {
    int e;
    double m = std::frexp(x, &e);  // x = m 2^e, m in [1/2, 1) for positive finite x
    // m in [sqrt(1/2), sqrt(2))
    const bool lo = m < 0.70710678118654752440;
    m = lo ? 2.0 * m : m;
    e = lo ? e - 1 : e;
    const double num = m - 1.0;  // exact (Sterbenz)
    const double den = m + 1.0;
    const double s = num / den;  // IEEE division (host and device alike)
    const double z = s * s;
    // 1 + z/3 + z^2/5 + ... + z^9/19 (Horner, explicit fused multiply-adds so
    // that host and device round alike)
    double q = 1.0 / 19.0;
    q = std::fma(q, z, 1.0 / 17.0);
    q = std::fma(q, z, 1.0 / 15.0);
    q = std::fma(q, z, 1.0 / 13.0);
    q = std::fma(q, z, 1.0 / 11.0);
    q = std::fma(q, z, 1.0 / 9.0);
    q = std::fma(q, z, 1.0 / 7.0);
    q = std::fma(q, z, 1.0 / 5.0);
    q = std::fma(q, z, 1.0 / 3.0);
    // log(m) = 2 s + 2 s z q
    const double two_s = 2.0 * s;
    const double logm = std::fma(two_s, z * q, two_s);
    // e log 2: high part with 32 significant bits (e * hi exact for |e| < 2^21)
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    const double de = (double)e;
    double r = std::fma(de, ln2_hi, std::fma(de, ln2_lo, logm));
    // special values: frexp leaves 0, inf and NaN unchanged
    if (!(x > 0.0)) r = (x == 0.0) ? -HUGE_VAL : __builtin_nan("");
    if (x == HUGE_VAL) r = HUGE_VAL;
    return r;
}

}  // namespace trixi_hip
