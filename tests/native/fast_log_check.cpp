// Accuracy check of trixi_hip::fast_log (csrc/fastlog.hpp, the FP64
// logarithm of the logarithmic means) against the x87 extended-precision
// logl: the largest error in units in the last place over random positive
// doubles (uniform bit patterns, values near 1, near powers of two,
// subnormals), and the special values.  Prints "max_ulp <v>" and
// "special ok" / "special FAIL <x>".
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>

#include "fastlog.hpp"

static double ulp_error(double x) {
    const long double ref = logl((long double)x);
    const double got = trixi_hip::fast_log(x);
    const double rd = (double)ref;  // correctly rounded reference (up to double rounding)
    if (rd == 0.0) return got == 0.0 ? 0.0 : INFINITY;
    const double ulp = std::nextafter(std::fabs(rd), INFINITY) - std::fabs(rd);
    return (double)(std::fabs((long double)got - ref) / ulp);
}

int main(int argc, char** argv) {
    const long n = argc > 1 ? atol(argv[1]) : 2000000;
    std::mt19937_64 rng(20261202);
    double worst = 0.0, worst_x = 0.0;
    auto take = [&](double x) {
        const double e = ulp_error(x);
        if (e > worst) {
            worst = e;
            worst_x = x;
        }
    };
    for (long i = 0; i < n; ++i) {
        // positive finite doubles with uniform bit patterns (subnormals included)
        uint64_t b = rng() & 0x7fefffffffffffffULL;
        double x;
        std::memcpy(&x, &b, 8);
        if (x > 0.0) take(x);
        // near 1 and near the reduction boundaries sqrt(1/2) 2^k
        const double u = std::ldexp((double)(rng() >> 11), -53);  // [0, 1)
        take(1.0 + (u - 0.5) * 1e-3);
        take(1.0 + (u - 0.5) * 0.7);
        take(std::ldexp(0.70710678118654752440 * (1.0 + (u - 0.5) * 1e-6), (int)(rng() % 200) - 100));
        // the states the rhs sees: densities and pressures of order one
        take(0.01 + 10.0 * u);
    }
    for (int k = -1074; k <= 1023; ++k) take(std::ldexp(1.0, k));
    std::printf("max_ulp %.4f at %.17g\n", worst, worst_x);
    bool ok = trixi_hip::fast_log(1.0) == 0.0 && std::isinf(trixi_hip::fast_log(0.0)) &&
              trixi_hip::fast_log(0.0) < 0 && std::isnan(trixi_hip::fast_log(-1.0)) &&
              std::isnan(trixi_hip::fast_log(NAN)) && std::isinf(trixi_hip::fast_log(INFINITY)) &&
              trixi_hip::fast_log(INFINITY) > 0 && std::isnan(trixi_hip::fast_log(-INFINITY));
    std::printf(ok ? "special ok\n" : "special FAIL\n");
    return 0;
}
