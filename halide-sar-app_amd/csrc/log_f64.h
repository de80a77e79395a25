// Natural logarithm in f64 for the dB image output, the same code on the host and the device.
//
// x = m 2^e with m in [sqrt(1/2), sqrt(2)) (frexp, then one exact doubling), and
//   log x = e ln 2 + log m,   log m = 2 atanh(s) = 2 s + 2 s z P(z),   s = (m - 1) / (m + 1), z = s^2,
//   P(z) = sum_{k=0..9} z^k / (2k + 3)
// (the Taylor series of atanh). |s| < 0.1716 and z < 0.0295, so the first omitted term, 2 s z^11 / 23,
// is below 1e-18 of 2 s. f = m - 1 is exact (Sterbenz) and 2 s = f - f s, so the leading term is f
// itself and s (one IEEE division) only enters a correction below f^2 / 2. ln 2 is split so that
// e ln2_hi is exact for every exponent of an f64. Measured against logl over (0, 1], all binades,
// near 1 and subnormals: within 1 ulp (tests/img_math/log_check.cpp).
// Special values follow IEEE log: log 0 = -inf, log of a negative number or NaN is NaN,
// log +inf = +inf. Every operation is an explicit IEEE operation or fma, with contraction off, so
// the host build (tests) and the device build give the same bits.
#pragma once

#include <cmath>
#include <limits>

#ifndef SARBP_HD
#if defined(__HIP__) || defined(__HIPCC__)
#define SARBP_HD __host__ __device__
#else
#define SARBP_HD
#endif
#endif

namespace sarbp {

// ln 2 = kLn2Hi + kLn2Lo: kLn2Hi has 42 significant bits, so e * kLn2Hi is exact for |e| < 2^11
constexpr double kLn2Hi = 0x1.62e42fefa3000p-1;
constexpr double kLn2Lo = 0x1.3de6af278ece6p-42;

SARBP_HD inline double log_f64(double x) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    if (!(x > 0.0)) return x == 0.0 ? -HUGE_VAL : std::numeric_limits<double>::quiet_NaN();
    if (x == HUGE_VAL) return x;
    int e = 0;
    double m = frexp(x, &e);  // x = m 2^e, m in [1/2, 1)
    if (m < 0.70710678118654752440) {
        m = m + m;
        e = e - 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double P = 1.0 / 21.0;
    P = fma(P, z, 1.0 / 19.0);
    P = fma(P, z, 1.0 / 17.0);
    P = fma(P, z, 1.0 / 15.0);
    P = fma(P, z, 1.0 / 13.0);
    P = fma(P, z, 1.0 / 11.0);
    P = fma(P, z, 1.0 / 9.0);
    P = fma(P, z, 1.0 / 7.0);
    P = fma(P, z, 1.0 / 5.0);
    P = fma(P, z, 1.0 / 3.0);
    // 2 s = f - f s exactly, so log m = f - (f s - 2 s z P(z)): f is exact and the rounding of s
    // only enters the correction, whose magnitude is below f^2 / 2
    const double c = fma(f, s, -2.0 * ((s * z) * P));
    const double de = (double)e;
    const double r = fma(de, kLn2Lo, -c) + f;
    return fma(de, kLn2Hi, r);
}

}  // namespace sarbp
