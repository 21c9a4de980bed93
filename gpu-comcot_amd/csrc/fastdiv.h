// Shorter division and square root sequences for the nonlinear momentum
// terms, equal to the IEEE operations wherever their results can matter.
//
// The IEEE single-precision division on gfx950 is eleven VALU operations:
// two operand scalings (V_DIV_SCALE), a reciprocal with one Newton step,
// the quotient with two residual corrections (the last one V_DIV_FMAS) and
// a special-case fix-up (V_DIV_FIXUP).  The momentum kernels spend about a
// fifth of their instructions on fourteen divisions per cell.
//
// div_inrange(a, b) is the same sequence without the scalings and the
// fix-up.  V_DIV_SCALE leaves both operands unchanged, and V_DIV_FIXUP
// passes the quotient through, unless b is 0, denormal or so large that
// 1/b is, a is 0 or below 2^-103 in magnitude, a/b is below 2^-126 or the
// exponents differ by 96 or more; for the other (a, b) the result is the
// IEEE quotient bit for bit (div_inrange_applies states the range that the
// device self-check, gcomcot_division_check_, covers).
// At the call sites below the divisor is a total depth (>= 1e-5 where the
// result is used), a friction power of one (>= 1e-7) or 1 + FF (>= 1), so
// only a numerator below 2^-103 leaves that range.  The quotient is then
// below 2^-86, and every path from it ends in the flux, which is set to 0
// when below EPS = 1e-10 (moment.f90:565, 759, 1054, 1247): a change
// of that size cannot move a sum that survives the test, since it is far
// below half an ulp of 1e-10.  A zero numerator gives +0 for -0/b (IEEE:
// -0), which the same test turns into the same +0.
//
// div12(x) is x / 12 as RN(1/12) times x with one residual correction.
// Checked against IEEE division for every float (tests/test_fastdiv.py):
// equal for all |x| >= 2^-124 and for zeros; the only differences are
// denormal quotients.  Its argument at the call sites is a positive
// coefficient times a depth of at least 1e-5.
#pragma once

#if defined(__HIPCC__) || defined(__HIP__)
#define GC_FD_HD __host__ __device__ __forceinline__
#else
#define GC_FD_HD static inline
#include <math.h>
#endif

namespace gc {

// RN(1/12)
constexpr float kRcp12 = 1.0f / 12.0f;

GC_FD_HD float div12(float x) {
  const float q = x * kRcp12;
  const float r = fmaf(-12.0f, q, x);
  const float d = fmaf(r, kRcp12, q);
  return x == 0.0f ? q : d;  // q keeps the sign of a zero x
}

// The operands div_inrange is claimed for (with q = a / b): both normal,
// |b| within 2^+-kDivRange, |a| >= 2^-103, an exponent gap below 96 and a
// normal quotient.  The call sites stay far inside it.
constexpr int kDivRange = 120;
GC_FD_HD bool div_inrange_applies(float a, float b, float q) {
  const float lo = 0x1p-120f, hi = 0x1p120f;
  return fabsf(b) >= lo && fabsf(b) <= hi && fabsf(a) >= 0x1p-103f && fabsf(a) <= 0x1p127f &&
         fabsf(q) >= 0x1p-126f && fabsf(q) <= 0x1p95f;
}

#if defined(__HIPCC__) || defined(__HIP__)
// div_inrange split in two: the refined reciprocal of b (three operations
// that depend on b alone), and the quotient from it.  div_by(a, b,
// recip_refined(b)) is div_inrange(a, b) bit for bit, so quotients with one
// divisor can share the first part.
__device__ __forceinline__ float recip_refined(float b) {
  const float y = __builtin_amdgcn_rcpf(b);
  const float e = fmaf(-b, y, 1.0f);
  return fmaf(e, y, y);
}
__device__ __forceinline__ float div_by(float a, float b, float y) {
  float q = a * y;
  const float r0 = fmaf(-b, q, a);
  q = fmaf(r0, y, q);
  const float r1 = fmaf(-b, q, a);
  return fmaf(r1, y, q);
}
__device__ __forceinline__ float div_inrange(float a, float b) {
  return div_by(a, b, recip_refined(b));
}

// sqrt(x) as the correctly rounded square root sequence computes it for
// x >= 2^-96: the hardware estimate, then the neighbour one ulp below or
// above where the residual says so.  That sequence first scales smaller x
// by 2^32 (and selects x itself for +-0 and +inf, which the steps below
// already return).  At the call sites x is g*D (D = 0 or >= 1e-5) or
// P^2 + Q^2 of the friction term; there a smaller x only feeds a friction
// numerator below 2^-48 into the quotient argument above.  Equal to IEEE
// sqrt for every float >= 2^-96 and for +-0 and +inf (device self-check
// over all floats, gcomcot_division_check_).
__device__ __forceinline__ float sqrt_inrange(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const float sm = __uint_as_float(__float_as_uint(s) - 1u);
  const float rm = fmaf(-sm, s, x);
  const float sp = __uint_as_float(__float_as_uint(s) + 1u);
  const float rp = fmaf(-sp, s, x);
  const float t = rm <= 0.0f ? sm : s;
  return rp > 0.0f ? sp : t;
}
#endif

}  // namespace gc
