// FP32 definitions of the OpenCL C built-ins that the trax seeding kernels
// call: atan2, atan, acos, sign, the common-function max, and the float->int
// conversions behind `(uint)floor(f)` / `(int)floor(f)`.
//
// OpenCL leaves the results of these built-ins implementation-defined within
// ulp bounds (atan2 <= 6 ulp, atan <= 5 ulp, acos <= 4 ulp; OpenCL 1.2
// section 7.4), so the original device's exact bits cannot be reproduced.
// This header fixes one definition.  It uses only IEEE-754 binary32
// +, -, *, / and sqrt (correctly rounded, no fused multiply-add), so every
// compiler that honours `-ffp-contract=off` and IEEE division/sqrt produces
// the same bits for the same inputs.  The CPU restatement of the chain
// carries its own copy of the same definitions; the parity tests require
// identical outputs from both.
//
// Method: S. L. Moshier, Cephes Math Library (single precision `atanf`,
// `asinf`), minimax polynomials on |u| <= tan(pi/8) and |a| <= 0.5 with the
// usual range reductions.  The pi, pi/2, pi/4 constants are split into a
// float head and a float tail to keep the reductions within ~2 ulp.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define TRAX_FP __host__ __device__ __forceinline__
#else
#define TRAX_FP static inline
#endif

#define TRAX_PI_F 3.14159274101257324f      /* float(pi) == OpenCL M_PI_F */
#define TRAX_PI_LO -8.74227765734758577e-08f
#define TRAX_PIO2_F 1.57079637050628662f
#define TRAX_PIO2_LO -4.37113882867379289e-08f
#define TRAX_PIO4_F 0.785398185253143311f
#define TRAX_PIO4_LO -2.18556941433689644e-08f
#define TRAX_TAN_PI8 0.414213562373095049f

// atan(u) for |u| <= tan(pi/8): u + u^3 P(u^2)
TRAX_FP float trax_atan_poly(float u) {
  float z = u * u;
  float p = ((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z -
            3.33329491539e-1f;
  return (p * z) * u + u;
}

// atan(t) for t in [0, 1].  Both reductions are evaluated with one division
// and one polynomial (operands chosen by selects), so a wave whose lanes take
// different branches pays for one of each.
TRAX_FP float trax_atan_unit(float t) {
  const int big = t > TRAX_TAN_PI8;
  const float u = big ? (t - 1.0f) / (t + 1.0f) : t;  // the division is evaluated either way
  const float p = trax_atan_poly(u);
  return big ? TRAX_PIO4_F + (p + TRAX_PIO4_LO) : p;
}

TRAX_FP float trax_copysign(float mag, float sgn) { return __builtin_copysignf(mag, sgn); }

TRAX_FP float trax_atanf(float x) {
  if (x != x) return x;
  float a = __builtin_fabsf(x);
  // atan(1/a) reduction for a > 1, through one call (operand by select)
  const int inv = a > 1.0f;
  const float t = inv ? 1.0f / a : a;  // the division is evaluated either way
  const float u = trax_atan_unit(t);
  const float r = inv ? (TRAX_PIO2_F - u) + TRAX_PIO2_LO : u;
  return trax_copysign(r, x);
}

// atan2 with the C99 Annex F / OpenCL special cases (signed zeros, infinities).
TRAX_FP float trax_atan2f(float y, float x) {
  if (x != x || y != y) return x + y;
  float ax = __builtin_fabsf(x), ay = __builtin_fabsf(y);
  int xneg = __builtin_signbit(x) != 0;
  float r;
  if (ay == 0.0f) {
    r = xneg ? TRAX_PI_F : 0.0f;
  } else if (ax == 0.0f) {
    r = TRAX_PIO2_F;
  } else if (ax == __builtin_inff() || ay == __builtin_inff()) {
    if (ax == ay)
      r = xneg ? 2.35619449019234492f : TRAX_PIO4_F;
    else if (ay == __builtin_inff())
      r = TRAX_PIO2_F;
    else
      r = xneg ? TRAX_PI_F : 0.0f;
  } else {
    // atan of the smaller over the larger magnitude: one division
    const int swap = ay > ax;
    const float num = swap ? ax : ay, den = swap ? ay : ax;
    const float a = trax_atan_unit(num / den);
    r = swap ? (TRAX_PIO2_F - a) + TRAX_PIO2_LO : a;
    if (xneg) r = (TRAX_PI_F - r) + TRAX_PI_LO;
  }
  return trax_copysign(r, y);
}

// asin(a) for |a| <= 0.5: a + a^3 Q(a^2)
TRAX_FP float trax_asin_poly(float a) {
  float z = a * a;
  float p = (((4.2163199048e-2f * z + 2.4181311049e-2f) * z + 4.5470025998e-2f) * z +
             7.4953002686e-2f) * z +
            1.6666752422e-1f;
  return (p * z) * a + a;
}

TRAX_FP float trax_acosf(float x) {
  if (x != x) return x;
  if (x > 1.0f || x < -1.0f) return __builtin_nanf("");
  // x > 0.5: 2 asin(sqrt((1 - x) / 2)); x < -0.5: pi - 2 asin(sqrt((1 + x) / 2));
  // else pi/2 - asin(x).  One square root and one polynomial, operands by
  // select (1 + x == 1 - |x| for negative x, exactly).
  const int hi = x > 0.5f, lo = x < -0.5f;
  const float s = __builtin_sqrtf(0.5f * (1.0f - __builtin_fabsf(x)));
  const float p = trax_asin_poly(hi || lo ? s : x);
  return hi ? 2.0f * p : lo ? (TRAX_PI_F - 2.0f * p) + TRAX_PI_LO : (TRAX_PIO2_F - p) + TRAX_PIO2_LO;
}

// OpenCL sign(): 1, -1, the zero itself, or 0 for NaN.
TRAX_FP float trax_sign(float x) {
  if (x > 0.0f) return 1.0f;
  if (x < 0.0f) return -1.0f;
  if (x == 0.0f) return x;
  return 0.0f;
}

// OpenCL common-function max(x, y): y if x < y, otherwise x.
TRAX_FP float trax_maxf(float x, float y) { return x < y ? y : x; }

// bool promoted to float, as `(a < b) * v` does in OpenCL C.
TRAX_FP float trax_b2f(int b) { return b ? 1.0f : 0.0f; }

// (int)f with the AMDGPU v_cvt_i32_f32 saturation (NaN -> 0).
TRAX_FP int32_t trax_f2i(float f) {
  if (f != f) return 0;
  if (f >= 2147483648.0f) return 2147483647;
  if (f <= -2147483648.0f) return (int32_t)(-2147483647 - 1);
  return (int32_t)f;
}

// (uint)f with the AMDGPU v_cvt_u32_f32 saturation (NaN and negatives -> 0).
TRAX_FP uint32_t trax_f2u(float f) {
  if (!(f > -1.0f)) return 0u;
  if (f >= 4294967296.0f) return 4294967295u;
  return (uint32_t)f;
}
