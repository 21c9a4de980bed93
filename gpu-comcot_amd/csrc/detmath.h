// Deterministic float pow / exp for the nonlinear momentum terms.
//
// COMCOT's nonlinear momentum evaluates Manning friction as
//   FF = 0.5*dt*g * n^2 * sqrt(P^2 + Q^2) / D^(2.3333333)
// (moment.f90:459-460, 945-946, 656-657, 1139-1140, 207-208, 250-251) and the
// breaking damping as EXP(-REF) (moment.f90:1323).  The Fortran calls libm
// powf/expf on the host; a GPU math library gives different last bits.  To
// make the device result reproducible on any host, both functions here are
// computed in IEEE double using only +, -, *, / (each correctly rounded on
// CPU and GPU alike, with FP contraction disabled) and bit manipulation, then
// rounded once to float.  The double result is within a few 1e-16 of the
// exact value, so the float result is the correctly rounded one except when
// the exact value lies within ~1e-15 (relative) of a float rounding midpoint.
//
// Method (textbook):
//   ln x = e*ln2 + 2*atanh(s),  s = (m-1)/(m+1),  m in [sqrt(1/2), sqrt(2))
//   exp t = 2^k * exp(r),  k = round(t/ln2),  r = t - k*ln2 (Cody-Waite split)
// with Taylor series for atanh (12 odd terms) and exp (degree 14).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define GC_HD __host__ __device__ __forceinline__
#else
#define GC_HD static inline
#endif

namespace gcdet {

GC_HD double bits_to_double(uint64_t u) {
  union { uint64_t u; double d; } v; v.u = u; return v.d;
}
GC_HD uint64_t double_to_bits(double d) {
  union { uint64_t u; double d; } v; v.d = d; return v.u;
}

// natural log of a positive, finite, normal double
GC_HD double ln_pos(double x) {
  uint64_t b = double_to_bits(x);
  int e = (int)((b >> 52) & 0x7ff) - 1023;
  double m = bits_to_double((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
  if (m > 1.4142135623730951) { m = m * 0.5; e = e + 1; }
  double s = (m - 1.0) / (m + 1.0);
  double s2 = s * s;
  // atanh(s)/s = sum_{k>=0} s^(2k) / (2k+1)
  double p = 1.0 / 25.0;
  p = p * s2 + 1.0 / 23.0;
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
  p = p * s2 + 1.0;
  const double ln2_hi = 6.93147180369123816490e-01;  // 0x3FE62E42FEE00000
  const double ln2_lo = 1.90821492927058770002e-10;  // 0x3DEA39EF35793C76
  double de = (double)e;
  return (de * ln2_hi + 2.0 * s * p) + de * ln2_lo;
}

// e^t; only ever rounded to float by the callers, so beyond |t| = 200 the
// result is 0 or +inf there, as libm's expf/powf give (WAVE_BREAKING meets
// t = -inf where the cell is dry: COEF = EXP(-(+inf)) = 0).  NaN passes.
GC_HD double exp_core(double t) {
  if (t != t) return t;
  if (t < -200.0) return 0.0;
  if (t > 200.0) return bits_to_double(0x7ff0000000000000ULL);
  const double inv_ln2 = 1.44269504088896338700e+00;
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  double kd = t * inv_ln2;
  // round half away from zero, done with truncation so that both sides agree
  long long k = (long long)(kd + (kd >= 0.0 ? 0.5 : -0.5));
  kd = (double)k;
  double r = (t - kd * ln2_hi) - kd * ln2_lo;
  double p = 1.0 / 87178291200.0;          // 1/14!
  p = p * r + 1.0 / 6227020800.0;          // 1/13!
  p = p * r + 1.0 / 479001600.0;           // 1/12!
  p = p * r + 1.0 / 39916800.0;            // 1/11!
  p = p * r + 1.0 / 3628800.0;             // 1/10!
  p = p * r + 1.0 / 362880.0;              // 1/9!
  p = p * r + 1.0 / 40320.0;               // 1/8!
  p = p * r + 1.0 / 5040.0;                // 1/7!
  p = p * r + 1.0 / 720.0;                 // 1/6!
  p = p * r + 1.0 / 120.0;                 // 1/5!
  p = p * r + 1.0 / 24.0;                  // 1/4!
  p = p * r + 1.0 / 6.0;                   // 1/3!
  p = p * r + 0.5;                         // 1/2!
  p = p * r + 1.0;
  p = p * r + 1.0;
  uint64_t scale = (uint64_t)(k + 1023) << 52;
  return p * bits_to_double(scale);
}

// x^y for the depths the momentum routines raise to 7/3: x > 0 finite
// (they bound it below by 1e-3 or 1e-5).  NaN, zero, negative and infinite
// x follow C's powf for y > 0: NaN, 0, NaN, +inf.
GC_HD float powf_det(float x, float y) {
  if (x != x) return x;
  if (!(x > 0.0f)) return x == 0.0f ? 0.0f : (float)bits_to_double(0x7ff8000000000000ULL);
  if (x > 3.0e38f) return (float)bits_to_double(0x7ff0000000000000ULL);
  double t = (double)y * ln_pos((double)x);
  return (float)exp_core(t);
}

// e^x for |x| < 80
GC_HD float expf_det(float x) {
  return (float)exp_core((double)x);
}

}  // namespace gcdet
