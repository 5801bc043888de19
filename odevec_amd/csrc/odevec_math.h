// Deterministic elementary functions for the integrator and the rate code.
//
// Each function is a fixed sequence of IEEE operations that every conforming
// platform rounds identically (+, -, *, /, floor, frexp, ldexp), evaluated
// without contraction, so the device and a host build return the same bits.
// Accuracy is a few ulp:
//   exp: k = round(x/ln2), r = x - k*ln2 (two-part ln2), degree-13 Taylor
//        polynomial in Horner form, scaled by 2^k;
//   log: x = m*2^e with m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1),
//        log m = 2 s (1 + s^2/3 + ... + s^22/23), plus e*ln2 (two-part);
//   pow: exp(y*log(x)) for x > 0.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#define ODV_MATH_HD __host__ __device__ __attribute__((always_inline))

namespace odv {

ODV_MATH_HD inline double m_exp(double x) {
#pragma clang fp contract(off)
  if (x != x) return x;
  if (x > 709.782712893384) return INFINITY;
  if (x < -745.1332191019412) return 0.0;
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  const double inv_ln2 = 1.44269504088896338700e+00;
  double kd = x * inv_ln2;
  kd = (kd >= 0.0) ? floor(kd + 0.5) : -floor(0.5 - kd);
  const int k = (int)kd;
  const double r = (x - kd * ln2_hi) - kd * ln2_lo;
  double p = 1.0 / 6227020800.0;
  p = p * r + 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  return ldexp(p, k);
}

ODV_MATH_HD inline double m_log(double x) {
#pragma clang fp contract(off)
  if (x != x) return x;
  if (x < 0.0) return NAN;
  if (x == 0.0) return -INFINITY;
  if (x == INFINITY) return x;
  int e = 0;
  double m = frexp(x, &e);
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e = e - 1;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double s2 = s * s;
  double p = 1.0 / 23.0;
  p = p * s2 + 1.0 / 21.0;
  p = p * s2 + 1.0 / 19.0;
  p = p * s2 + 1.0 / 17.0;
  p = p * s2 + 1.0 / 15.0;
  p = p * s2 + 1.0 / 13.0;
  p = p * s2 + 1.0 / 11.0;
  p = p * s2 + 1.0 / 9.0;
  p = p * s2 + 1.0 / 7.0;
  p = p * s2 + 1.0 / 5.0;
  p = p * s2 + 1.0 / 3.0;
  const double logm = 2.0 * s + (2.0 * s) * (s2 * p);
  const double ed = (double)e;
  return ed * 6.93147180369123816490e-01 + (logm + ed * 1.90821492927058770002e-10);
}

ODV_MATH_HD inline double m_log10(double x) {
#pragma clang fp contract(off)
  return m_log(x) * 0.43429448190325182765;
}

ODV_MATH_HD inline double m_pow(double x, double y) {
#pragma clang fp contract(off)
  if (y == 0.0) return 1.0;
  if (x != x || y != y) return x + y;
  if (x == 0.0) return (y > 0.0) ? 0.0 : INFINITY;
  if (x < 0.0) return NAN;
  return m_exp(y * m_log(x));
}

}  // namespace odv
