// pf_math.h -- FP64 exp / log / log1p for the van Genuchten closures.
//
// The closures spend most of the residual's and the Jacobian's arithmetic
// in these three functions.  The general-purpose device library versions
// handle every special case (NaN, infinities, zeros, negative arguments)
// and carry extra-precision intermediates; the closures only ever pass
// finite positive arguments of moderate size, so these versions skip that
// work and keep the error within 1.5 ulp (checked on the CPU against long
// double references over the ranges the closures use).
//
// Methods (textbook, no tables):
//   exp(x):   x = k ln2 + r, |r| <= ln2/2, with ln2 split in a 32-bit head
//             and its remainder (Cody & Waite) so k*head is exact; e^r by
//             its Taylor series to degree 13 (truncation < 5e-18); 2^k by
//             ldexp.
//   log(x):   x = 2^e m, m in [sqrt(1/2), sqrt(2)), f = m - 1; log m =
//             2 atanh(s) with s = f/(2+f), |s| <= 0.1716, series to s^19
//             (truncation < 3e-17 relative) arranged as f - s (f - R) so
//             the exact f leads; e ln2 added with the split constant.
//   log1p(t): t >= 0: u = 1+t, log u plus the exact rounding error of the
//             sum divided by u.
//
// The header compiles for the device (HIP) and for the host (the CPU
// accuracy test); pf_div / pf_rcp use the hardware reciprocal with Newton
// steps on the device and plain division on the host.
#ifndef PF_MATH_H
#define PF_MATH_H

#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PF_HD __host__ __device__ __forceinline__
#else
#define PF_HD static inline
#endif

// 1/b and a/b: on the device the hardware reciprocal estimate plus two
// Newton steps and one residual correction (the compiler's FP64 division
// without its range scaling and special-value fix-up; operands here are
// finite, normal and far from overflow), within 1 ulp.
PF_HD double pf_rcp(double b)
{
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(b);
  double e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  e = fma(-b, r, 1.0);
  return fma(r, e, r);
#else
  return 1.0 / b;
#endif
}

PF_HD double pf_div(double a, double b)
{
#if defined(__HIP_DEVICE_COMPILE__)
  double r = pf_rcp(b);
  double q = a * r;
  double e = fma(-b, q, a);
  return fma(e, r, q);
#else
  return a / b;
#endif
}

// ln 2 = PF_LN2_HI + PF_LN2_LO; PF_LN2_HI has a 32-bit significand
#define PF_LN2_HI 0.6931471803691238
#define PF_LN2_LO 1.9082149292705877e-10
#define PF_LOG2E 1.4426950408889634
#define PF_SQRT_HALF 0.7071067811865476

// e^x for |x| < 700
PF_HD double pf_exp(double x)
{
  double k = rint(x * PF_LOG2E);
  double r = fma(-k, PF_LN2_HI, x);
  r = fma(-k, PF_LN2_LO, r);
  double p = 1.0 / 6227020800.0;            // 1/13!
  p = fma(p, r, 1.0 / 479001600.0);         // 1/12!
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

// natural log for finite x > 0
PF_HD double pf_log(double x)
{
  int e;
  double m = frexp(x, &e);                  // [0.5, 1)
  bool lo = m < PF_SQRT_HALF;
  m = lo ? m + m : m;                       // [sqrt(1/2), sqrt(2))
  e = lo ? e - 1 : e;
  double f = m - 1.0;                       // exact
  double s = pf_div(f, m + 1.0);
  double s2 = s * s;
  // log m = 2 atanh(s) = 2s + s R, R = s^2 (2/3 + s^2 (2/5 + ...)); with
  // 2s = f - s f this is f - s (f - R): the exact f leads and the rounded
  // part is only ~f/2 of the result
  double P = 2.0 / 19.0;
  P = fma(P, s2, 2.0 / 17.0);
  P = fma(P, s2, 2.0 / 15.0);
  P = fma(P, s2, 2.0 / 13.0);
  P = fma(P, s2, 2.0 / 11.0);
  P = fma(P, s2, 2.0 / 9.0);
  P = fma(P, s2, 2.0 / 7.0);
  P = fma(P, s2, 2.0 / 5.0);
  P = fma(P, s2, 2.0 / 3.0);
  double lm = f - s * fma(-s2, P, f);
  double de = (double)e;
  return fma(de, PF_LN2_HI, fma(de, PF_LN2_LO, lm));
}

// log(1 + t) for finite t >= 0
PF_HD double pf_log1p(double t)
{
  double u = 1.0 + t;
  double c = t - (u - 1.0);                 // rounding error of 1 + t, exact
  return pf_log(u) + c * pf_rcp(u);
}

#endif
