// FP64 power x^a for the viscosity law mu = (x)^a / Re (x > 0), built from
// textbook series instead of the library pow, which costs about four times
// as many instructions:
//   log x = e ln2 + 2 atanh(s),  x = 2^e m,  m in [1/sqrt2, sqrt2),
//           s = (m - 1)/(m + 1), |s| <= 0.1716, atanh series to s^21;
//   exp y = 2^k exp(r), k = rint(y/ln2), r = y - k ln2 with ln2 split into a
//           32-bit head and a tail (Cody and Waite, "Software Manual for the
//           Elementary Functions", 1980), |r| <= 0.347, Taylor series to r^13.
// Series truncation is below 2^-56 relative in both; the result is within a
// few units in the last place of the correctly rounded power.  Arguments
// that are not positive and finite go to the library pow.
//
// Plain C++ apart from the qualifier, so that host code (and the tests) can
// compile exactly the device arithmetic.
#pragma once

#include <cmath>

#ifdef __HIPCC__
#define ST_HD __host__ __device__ __forceinline__
#else
#define ST_HD inline
#endif

namespace st {

ST_HD double series_log(double x)
{
    int e;
    double m = std::frexp(x, &e);                 // m in [0.5, 1)
    if (m < 0.70710678118654752440) {
        m *= 2.0;
        e -= 1;
    }
    const double s = (m - 1.0) / (m + 1.0), z = s * s;
    // atanh(s)/s - 1 = z/3 + z^2/5 + ... + z^10/21
    double p = 1.0 / 21.0;
    p = std::fma(p, z, 1.0 / 19.0);
    p = std::fma(p, z, 1.0 / 17.0);
    p = std::fma(p, z, 1.0 / 15.0);
    p = std::fma(p, z, 1.0 / 13.0);
    p = std::fma(p, z, 1.0 / 11.0);
    p = std::fma(p, z, 1.0 / 9.0);
    p = std::fma(p, z, 1.0 / 7.0);
    p = std::fma(p, z, 1.0 / 5.0);
    p = std::fma(p, z, 1.0 / 3.0);
    p *= z;
    const double ln2_hi = 0x1.62e42ff000000p-1, ln2_lo = -0x1.718432a1b0e26p-35;
    const double t = 2.0 * s;
    return std::fma(double(e), ln2_hi, std::fma(t, p, std::fma(double(e), ln2_lo, t)));
}

ST_HD double series_exp(double y)
{
    const double ln2_hi = 0x1.62e42ff000000p-1, ln2_lo = -0x1.718432a1b0e26p-35;
    const double k = std::rint(y * 1.4426950408889634);
    const double r = std::fma(-k, ln2_lo, std::fma(-k, ln2_hi, y));
    // sum_{n=0}^{13} r^n / n!
    double p = 1.0 / 6227020800.0;
    p = std::fma(p, r, 1.0 / 479001600.0);
    p = std::fma(p, r, 1.0 / 39916800.0);
    p = std::fma(p, r, 1.0 / 3628800.0);
    p = std::fma(p, r, 1.0 / 362880.0);
    p = std::fma(p, r, 1.0 / 40320.0);
    p = std::fma(p, r, 1.0 / 5040.0);
    p = std::fma(p, r, 1.0 / 720.0);
    p = std::fma(p, r, 1.0 / 120.0);
    p = std::fma(p, r, 1.0 / 24.0);
    p = std::fma(p, r, 1.0 / 6.0);
    p = std::fma(p, r, 0.5);
    p = std::fma(p, r, 1.0);
    p = std::fma(p, r, 1.0);
    return std::ldexp(p, int(k));
}

// x^a for x > 0 and |a log x| < 700.
ST_HD double series_pow(double x, double a)
{
    if (!(x > 0.0) || !(x < 1.0e300)) return std::pow(x, a);
    return series_exp(a * series_log(x));
}

}  // namespace st
