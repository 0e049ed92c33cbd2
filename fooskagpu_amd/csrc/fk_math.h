// fk_math.h -- portable FP64 logarithm and exponential shared by host and
// device code.
//
// The relaxation Newton iterations take finite-difference Jacobians with
// steps of 1e-3 .. 1e-10 relative, which magnify a one-ulp difference in a
// logarithm by up to 1e10; their converged states then differ in the 10th
// digit between two libm implementations.  fk_log below is a fixed sequence
// of IEEE-754 double operations (frexp, +, -, *, / only, each correctly
// rounded on the CPU and on gfx950), so every compiler that keeps the
// operation order (-ffp-contract=off) gives the same bits for it.
//
// Method: x = m 2^e with m in [sqrt(1/2), sqrt(2)), f = m - 1 (exact),
// s = f / (2 + f), so that log(1 + f) = 2 atanh(s) = f - s (f - R) with
// R = sum_{k>=1} 2 s^(2k) / (2k + 1)  (since f - 2 s = s f).
// |s| <= 0.1716, s^2 <= 0.02944: nine terms leave a truncation below
// 3e-17 relative.  The correction s (f - R) is O(f^2), so its rounding is far
// below ulp(f) and the result is within about one ulp of log(x).
// log(x) = e ln2_hi + (log(m) + e ln2_lo), ln2_hi carrying 29 significant
// bits so that e ln2_hi is exact.
#ifndef FK_MATH_H
#define FK_MATH_H
#include <cmath>

#if defined(__HIPCC__)
#define FK_HD __host__ __device__
#else
#define FK_HD
#endif

namespace fkmath {

FK_HD inline double log(double x) {
  if (!(x > 0)) return x == 0 ? -HUGE_VAL : NAN;  // also NaN in -> NaN out
  if (x == HUGE_VAL) return x;
  int e;
  double m = frexp(x, &e);  // m in [0.5, 1)
  if (m < 0.70710678118654752440) {
    m = m + m;
    e -= 1;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double q = 2.0 / 19.0;
  q = q * z + 2.0 / 17.0;
  q = q * z + 2.0 / 15.0;
  q = q * z + 2.0 / 13.0;
  q = q * z + 2.0 / 11.0;
  q = q * z + 2.0 / 9.0;
  q = q * z + 2.0 / 7.0;
  q = q * z + 2.0 / 5.0;
  q = q * z + 2.0 / 3.0;
  const double R = z * q;
  const double lm = f - s * (f - R);
  const double de = (double)e;
  const double ln2_hi = 0x1.62e42ffp-1;             // 0.69314718060195446
  const double ln2_lo = -0x1.718432a1b0e26p-35;    // ln 2 - ln2_hi
  return de * ln2_hi + (lm + de * ln2_lo);
}

FK_HD inline double log10(double x) {
  const double inv_ln10 = 0x1.bcb7b1526e50ep-2;  // nearest double to 1 / ln 10
  return fkmath::log(x) * inv_ln10;
}

// exp(x) for the finite-rate relaxation factor, the same bits on the CPU and
// on gfx950 for the same reason as log.  Method: k = nearest integer to
// x / ln 2, r = (x - k ln2_hi) - k ln2_lo (k ln2_hi exact, |r| <= ln2 / 2 up to
// rounding), exp(r) by its Taylor series through r^13 / 13! (truncation below
// 4e-18 relative for |r| <= 0.347) in Horner form, and exp(x) = 2^k exp(r) by
// an exact ldexp.  x < -708 (exp(x) < 2^-1021) returns 0, so no result is
// subnormal; x > 709.78 (beyond the double range) returns +inf.
FK_HD inline double exp(double x) {
  if (x != x) return x;
  if (x > 709.78) return HUGE_VAL;
  if (x < -708.0) return 0.0;
  const double inv_ln2 = 0x1.71547652b82fep0;  // nearest double to 1 / ln 2
  const double ln2_hi = 0x1.62e42ffp-1;
  const double ln2_lo = -0x1.718432a1b0e26p-35;
  const double kd = rint(x * inv_ln2);
  const double r = (x - kd * ln2_hi) - kd * ln2_lo;
  double q = 1.0 / 6227020800.0;  // 1 / 13!
  q = q * r + 1.0 / 479001600.0;
  q = q * r + 1.0 / 39916800.0;
  q = q * r + 1.0 / 3628800.0;
  q = q * r + 1.0 / 362880.0;
  q = q * r + 1.0 / 40320.0;
  q = q * r + 1.0 / 5040.0;
  q = q * r + 1.0 / 720.0;
  q = q * r + 1.0 / 120.0;
  q = q * r + 1.0 / 24.0;
  q = q * r + 1.0 / 6.0;
  q = q * r + 0.5;
  q = q * r + 1.0;
  q = q * r + 1.0;
  return ldexp(q, (int)kd);
}

}  // namespace fkmath
#endif
