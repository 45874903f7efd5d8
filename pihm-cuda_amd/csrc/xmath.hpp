// xmath.hpp - x^y in FP64 as exp(y ln x) with short polynomials.
//
// The flux functions raise positive reals to non-integer powers (van
// Genuchten retention and conductivity, Manning's depth^(2/3), channel
// geometry) about seven times per element per call.  The device library's
// pow carries a double-double logarithm and full special-case handling
// (some 215 VALU instructions); this one is about 65:
//
//   ln x:   x = m 2^e with m in [sqrt(1/2), sqrt(2)),  f = (m-1)/(m+1),
//           ln m = 2 atanh f = 2f (1 + f^2/3 + f^4/5 + ... + f^18/21)
//           (|f| <= 0.1716, the first omitted term is below 2^-56), and
//           ln x = e ln2_hi + (e ln2_lo + ln m) kept as a hi + lo pair
//           (ln2 split as in Cody and Waite, ln2_hi exact in 32 bits).
//   t:      y ln x = thi + tlo with the product's rounding error from fma.
//   exp t:  k = rint(thi / ln2),  r = (thi - k ln2_hi) + (tlo - k ln2_lo)
//           (|r| <= ln2 / 2), exp r by its Taylor series to r^13/13!
//           (remainder below 2^-57), scaled by 2^k.
//
// Error: t carries about |y| 1.2e-16 absolute error, exp about 1 ulp, so x^y
// is within a few ulp of the exact power for the exponents used here
// (|y| <= 3); the unit tests measure it against the host libm's pow.
// Exact where the reference's results are exact and the formulas depend on
// it: 1^y = 1, 0^y = +0 for y > 0, and x^y >= 1 for x >= 1, y >= 0 (so
// pow(1/s, m) - 1 never goes negative).  Arguments outside finite x > 0 with
// |y ln x| < 708 go to the library pow.
//
// Host and device compile the same operations (explicit fma, no
// contraction), so both return the same bits.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace pihm {

#define PIHM_HD __host__ __device__ __forceinline__

constexpr double kLn2Hi = 6.93147180369123816490e-01;  // 0x3FE62E42FEE00000: 32 significant bits
constexpr double kLn2Lo = 1.90821492927058770002e-10;  // ln 2 - kLn2Hi

// ln x = hi + lo for finite x > 0 (normal or subnormal); the pair carries
// ln x to about 1.2e-16 absolute whatever the exponent of x.
PIHM_HD double ln_pos(double x, double *lo) {
  int e;
  double m = frexp(x, &e);                       // x = m 2^e, m in [1/2, 1)
  if (m < 0.70710678118654752440) {
    m = m + m;
    e -= 1;
  }
  const double f = (m - 1.0) / (m + 1.0);        // m - 1 is exact
  const double s = f * f;
  double p = 1.0 / 21.0;
  p = fma(p, s, 1.0 / 19.0);
  p = fma(p, s, 1.0 / 17.0);
  p = fma(p, s, 1.0 / 15.0);
  p = fma(p, s, 1.0 / 13.0);
  p = fma(p, s, 1.0 / 11.0);
  p = fma(p, s, 1.0 / 9.0);
  p = fma(p, s, 1.0 / 7.0);
  p = fma(p, s, 1.0 / 5.0);
  p = fma(p, s, 1.0 / 3.0);
  const double f2 = f + f;
  const double lnm = fma(f2 * s, p, f2);         // |lnm| <= 0.35
  const double de = (double)e;
  const double a = de * kLn2Hi;                  // exact: |e| < 2^11
  const double b = fma(de, kLn2Lo, lnm);
  const double hi = a + b;                       // |a| >= |b| unless a == 0
  *lo = (a - hi) + b;
  return hi;
}

// e^(hi + lo) for |hi| < 708 and |lo| <= ulp(hi) (normal results).
PIHM_HD double exp_core(double hi, double lo) {
  const double k = rint(hi * 1.44269504088896340736);
  double r = fma(-k, kLn2Hi, hi);                // exact for |k| < 2^21
  r = r + fma(-k, kLn2Lo, lo);                   // |r| <= ln2 / 2 + 1e-10
  double p = 1.0 / 6227020800.0;                 // 1/13!
  p = fma(p, r, 1.0 / 479001600.0);
  p = fma(p, r, 1.0 / 39916800.0);
  p = fma(p, r, 1.0 / 3628800.0);
  p = fma(p, r, 1.0 / 362880.0);
  p = fma(p, r, 1.0 / 40320.0);
  p = fma(p, r, 1.0 / 5040.0);
  p = fma(p, r, 1.0 / 720.0);
  p = fma(p, r, 1.0 / 120.0);
  p = fma(p, r, 1.0 / 24.0);
  p = fma(p, r, 1.0 / 6.0);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)k);
}

// The library pow, kept out of line: its call sites are rarely taken and
// inlining it at each would multiply the kernels' code size.
__host__ __device__ inline __attribute__((noinline)) double pow_lib(double x, double y) { return pow(x, y); }

// x^y: the fast path for finite x > 0 and finite y; (+0)^y = +0 for y > 0;
// everything else (x < 0, -0, non-finite arguments, |y ln x| >= 708) from
// the library pow.
PIHM_HD double xpow(double x, double y) {
  if (x > 0.0 && x < HUGE_VAL && fabs(y) < HUGE_VAL) {
    double lnlo;
    const double lnhi = ln_pos(x, &lnlo);
    const double thi = y * lnhi;
    const double tlo = fma(y, lnhi, -thi) + y * lnlo;   // y ln x = thi + tlo
    if (fabs(thi) < 708.0) return exp_core(thi, tlo);
  } else if (x == 0.0 && y > 0.0 && !signbit(x)) {
    return 0.0;
  }
  return pow_lib(x, y);
}

}  // namespace pihm
