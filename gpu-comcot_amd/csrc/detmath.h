// Deterministic float pow / exp for the nonlinear momentum terms.
//
// COMCOT's nonlinear momentum evaluates Manning friction as
//   FF = 0.5*dt*g * n^2 * sqrt(P^2 + Q^2) / D^(2.3333333)
// (moment.f90:459-460, 945-946, 656-657, 1139-1140, 207-208, 250-251) and the
// breaking damping as EXP(-REF) (moment.f90:1323).  The Fortran calls libm
// powf/expf on the host; a GPU math library gives different last bits.  To
// make the device result reproducible on any host, both functions here are
// computed in IEEE double using only +, -, *, explicit fma (each correctly
// rounded on CPU and GPU alike; FP contraction is disabled around them) and
// bit manipulation, then rounded once to float.  The double result is
// within ~1e-16 (relative) of the exact value, so the float result is the
// correctly rounded one except when the exact value lies that close to a
// float rounding midpoint.
//
// Method (table-driven, as in the usual libm designs):
//   log2 x = e + log2(c_k) + log2(1 + r),  x = 2^e m, m in [1,2),
//            c_k the centre of m's 1/128 interval, r = m/c_k - 1 computed
//            exactly as fma(m, 1/c_k, -1) with 1/c_k rounded to 28 bits;
//   2^t    = 2^(k/32) * 2^g,  k = round(32 t), |g| <= 1/64;
// with degree-6 polynomials for log2(1+r) and 2^g.  The tables are built by
// constexpr code from series (atanh for ln, Taylor for exp) at compile time,
// so host and device hold identical values.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define GC_HD __host__ __device__ __forceinline__
#define GC_TABLE __device__ __constant__
#else
#define GC_HD static inline
#define GC_TABLE static
#endif

namespace gcdet {

GC_HD double bits_to_double(uint64_t u) {
  union { uint64_t u; double d; } v; v.u = u; return v.d;
}
GC_HD uint64_t double_to_bits(double d) {
  union { uint64_t u; double d; } v; v.d = d; return v.u;
}

constexpr double kLn2 = 0.69314718055994530942;

// ---- compile-time table construction (plain IEEE double arithmetic)
constexpr double cx_ln(double x) {  // x in [0.5, 2]: 2*atanh((x-1)/(x+1))
  double s = (x - 1.0) / (x + 1.0), s2 = s * s, p = 0.0;
  for (int k = 40; k >= 0; --k) p = p * s2 + 1.0 / (2.0 * k + 1.0);
  return 2.0 * s * p;
}
constexpr double cx_exp(double r) {  // |r| <= 1
  double p = 0.0;
  for (int k = 30; k >= 1; --k) p = 1.0 + p * r / k;
  return p;
}
constexpr double cx_round28(double v) {  // v in (0.5, 1]: keep 28 significant bits
  double s = v * 268435456.0;               // 2^28
  long long q = (long long)(s + 0.5);
  return (double)q / 268435456.0;
}

struct Tables {
  double invc[128];
  double logc[128];  // log2(1/invc), of the stored invc
  double exp2j[32];  // 2^(j/32)
};

constexpr Tables make_tables() {
  Tables t{};
  for (int k = 0; k < 128; ++k) {
    double c = 1.0 + (k + 0.5) / 128.0;
    double ic = cx_round28(1.0 / c);
    t.invc[k] = ic;
    t.logc[k] = -cx_ln(ic) / kLn2;
  }
  for (int j = 0; j < 32; ++j) t.exp2j[j] = cx_exp(j * (kLn2 / 32.0));
  return t;
}

GC_TABLE constexpr Tables kTab = make_tables();

// log2(1+r)/r and 2^g coefficients
constexpr double kL1 = 1.0 / kLn2, kL2 = -1.0 / (2.0 * kLn2), kL3 = 1.0 / (3.0 * kLn2),
                 kL4 = -1.0 / (4.0 * kLn2), kL5 = 1.0 / (5.0 * kLn2),
                 kL6 = -1.0 / (6.0 * kLn2);
constexpr double kE1 = kLn2, kE2 = kLn2 * kLn2 / 2.0, kE3 = kLn2 * kLn2 * kLn2 / 6.0,
                 kE4 = kLn2 * kLn2 * kLn2 * kLn2 / 24.0,
                 kE5 = kLn2 * kLn2 * kLn2 * kLn2 * kLn2 / 120.0,
                 kE6 = kLn2 * kLn2 * kLn2 * kLn2 * kLn2 * kLn2 / 720.0;

GC_HD double dfma(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_fma(a, b, c);
#else
  return __builtin_fma(a, b, c);  // correctly rounded (libm fma if no FMA unit)
#endif
}

// fma(a, b, c) for a constant addend c.  On the device it is one VOP3
// v_fma_f64 that reads c from a scalar register pair; left to itself the
// compiler writes the two-address v_fmac_f64 instead, which needs c copied
// into the accumulator first (an extra vector move per polynomial step) and
// keeps every coefficient in vector registers for the whole kernel.  The
// operation, and so the result, is the same IEEE double fma either way.
GC_HD double dfma_k(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r;
  __asm__("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
  return r;
#else
  return __builtin_fma(a, b, c);
#endif
}

// x * 2^e for a result in the normal range, where it is exact: one
// instruction on the device (the same value as multiplying by the power).
GC_HD double scale2(double x, int e) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_ldexp(x, e);
#else
  return x * bits_to_double((uint64_t)(uint32_t)(e + 1023) << 52);
#endif
}

GC_HD float bits_to_float(uint32_t u) {
  union { uint32_t u; float f; } v; v.u = u; return v.f;
}
GC_HD uint32_t float_to_bits(float f) {
  union { uint32_t u; float f; } v; v.f = f; return v.u;
}

// log2 of a positive, finite, normal float (32-bit decomposition)
template <class TAB>
GC_HD double log2_pos(float x, const TAB &T) {
  const uint32_t b = float_to_bits(x);
  const int e = (int)(b >> 23) - 127;
  const int k = (int)((b >> 16) & 127);
  const double m = (double)bits_to_float((b & 0x007fffffu) | 0x3f800000u);
  const double r = dfma(m, T.invc[k], -1.0);
  double p = dfma(r, kL6, kL5);
  p = dfma_k(r, p, kL4);
  p = dfma_k(r, p, kL3);
  p = dfma_k(r, p, kL2);
  p = dfma_k(r, p, kL1);
  return (double)e + (T.logc[k] + r * p);
}

// 2^t; only ever rounded to float by the callers, so beyond |t| = 300 the
// result is 0 or +inf there, as libm's expf/powf give (WAVE_BREAKING meets
// t = -inf where the cell is dry: COEF = EXP(-(+inf)) = 0).  NaN passes.
template <class TAB>
GC_HD double exp2_core(double t, const TAB &T) {
  if (t != t) return t;
  if (t < -300.0) return 0.0;
  if (t > 300.0) return bits_to_double(0x7ff0000000000000ULL);
  const double s = t * 32.0;
  const int k = (int)(s + (s >= 0.0 ? 0.5 : -0.5));
  const double g = t - (double)k * (1.0 / 32.0);
  const int j = k & 31;
  const int e2 = k >> 5;  // floor(k / 32)
  double q = dfma(g, kE6, kE5);
  q = dfma_k(g, q, kE4);
  q = dfma_k(g, q, kE3);
  q = dfma_k(g, q, kE2);
  q = dfma_k(g, q, kE1);
  q = dfma(g, q, 1.0);
  return scale2(T.exp2j[j] * q, e2);
}

// x^y for the depths the momentum routines raise to 7/3: x > 0 finite
// (they bound it below by 1e-3 or 1e-5).  NaN, zero, negative and infinite
// x follow C's powf for y > 0: NaN, 0, NaN, +inf; subnormal x (whose power
// underflows for y > 1) gives 0.
template <class TAB>
GC_HD float powf_det(float x, float y, const TAB &T) {
  if (x != x) return x;
  if (!(x > 0.0f)) return x == 0.0f ? 0.0f : (float)bits_to_double(0x7ff8000000000000ULL);
  if (x > 3.0e38f) return (float)bits_to_double(0x7ff0000000000000ULL);
  if (x < 1.17549435e-38f) return 0.0f;
  return (float)exp2_core((double)y * log2_pos(x, T), T);
}

// x^y for the depth terms of Manning friction, whose callers guarantee
// x >= 1e-5 (the momentum routines clamp or test the depth first) and
// y = 2.3333333: there t = y*log2(x) lies in [-39, 299], so none of the
// range cases of exp2_core or powf_det can occur and the same arithmetic
// runs without them.  Equal to powf_det(x, y) for every x >= 1e-5 and for
// NaN and +inf (NaN passes, +inf and the largest floats give +inf).
template <class TAB>
GC_HD float powf_manning(float x, float y, const TAB &T) {
  const double t = (double)y * log2_pos(x, T);
  const double s = t * 32.0;
  const int k = (int)(s + (s >= 0.0 ? 0.5 : -0.5));
  const double g = t - (double)k * (1.0 / 32.0);
  const int j = k & 31;
  const int e2 = k >> 5;
  double q = dfma(g, kE6, kE5);
  q = dfma_k(g, q, kE4);
  q = dfma_k(g, q, kE3);
  q = dfma_k(g, q, kE2);
  q = dfma_k(g, q, kE1);
  q = dfma(g, q, 1.0);
  const float r = (float)scale2(T.exp2j[j] * q, e2);
  return x != x ? x : r;
}

// e^x
template <class TAB>
GC_HD float expf_det(float x, const TAB &T) {
  return (float)exp2_core((double)x * (1.0 / kLn2), T);
}

#if !defined(__HIP_DEVICE_COMPILE__)
// host entry points on the compile-time tables
static inline float powf_det(float x, float y) { return powf_det(x, y, kTab); }
static inline float expf_det(float x) { return expf_det(x, kTab); }
static inline float powf_manning(float x, float y) { return powf_manning(x, y, kTab); }
#endif

}  // namespace gcdet
