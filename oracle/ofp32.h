// Oracle copy of the FP32 built-in definitions used by the CPU restatement.
//
// TEST INFRASTRUCTURE: this file belongs to the CPU oracle.  It defines the
// OpenCL C built-ins that the reference kernels call (atan2, atan, acos,
// sign, max, float->int conversions) with IEEE binary32 +,-,*,/,sqrt only,
// following S. L. Moshier, Cephes Math Library (single precision atanf and
// asinf).  OpenCL makes these built-ins implementation-defined within ulp
// bounds, so this file is the definition the oracle uses.  It is pinned by
// tests/test_oracle_math.py against double-precision libm (ulp bounds,
// special values) and must be compiled with -ffp-contract=off.
#pragma once

#include <stdint.h>

#define OFP static inline

#define O_PI_F 3.14159274101257324f      /* float(pi) == OpenCL M_PI_F */
#define O_PI_LO -8.74227765734758577e-08f
#define O_PIO2_F 1.57079637050628662f
#define O_PIO2_LO -4.37113882867379289e-08f
#define O_PIO4_F 0.785398185253143311f
#define O_PIO4_LO -2.18556941433689644e-08f
#define O_TAN_PI8 0.414213562373095049f

// atan(u) for |u| <= tan(pi/8): u + u^3 P(u^2)
OFP float o_atan_poly(float u) {
  float z = u * u;
  float p = ((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z -
            3.33329491539e-1f;
  return (p * z) * u + u;
}

// atan(t) for t in [0, 1]
OFP float o_atan_unit(float t) {
  if (t > O_TAN_PI8) {
    float u = (t - 1.0f) / (t + 1.0f);
    return O_PIO4_F + (o_atan_poly(u) + O_PIO4_LO);
  }
  return o_atan_poly(t);
}

OFP float o_copysign(float mag, float sgn) { return __builtin_copysignf(mag, sgn); }

OFP float o_atanf(float x) {
  if (x != x) return x;
  float a = __builtin_fabsf(x);
  float r;
  if (a > 1.0f)
    r = (O_PIO2_F - o_atan_unit(1.0f / a)) + O_PIO2_LO;
  else
    r = o_atan_unit(a);
  return o_copysign(r, x);
}

// atan2 with the C99 Annex F / OpenCL special cases (signed zeros, infinities).
OFP float o_atan2f(float y, float x) {
  if (x != x || y != y) return x + y;
  float ax = __builtin_fabsf(x), ay = __builtin_fabsf(y);
  int xneg = __builtin_signbit(x) != 0;
  float r;
  if (ay == 0.0f) {
    r = xneg ? O_PI_F : 0.0f;
  } else if (ax == 0.0f) {
    r = O_PIO2_F;
  } else if (ax == __builtin_inff() || ay == __builtin_inff()) {
    if (ax == ay)
      r = xneg ? 2.35619449019234492f : O_PIO4_F;
    else if (ay == __builtin_inff())
      r = O_PIO2_F;
    else
      r = xneg ? O_PI_F : 0.0f;
  } else {
    if (ay > ax)
      r = (O_PIO2_F - o_atan_unit(ax / ay)) + O_PIO2_LO;
    else
      r = o_atan_unit(ay / ax);
    if (xneg) r = (O_PI_F - r) + O_PI_LO;
  }
  return o_copysign(r, y);
}

// asin(a) for |a| <= 0.5: a + a^3 Q(a^2)
OFP float o_asin_poly(float a) {
  float z = a * a;
  float p = (((4.2163199048e-2f * z + 2.4181311049e-2f) * z + 4.5470025998e-2f) * z +
             7.4953002686e-2f) * z +
            1.6666752422e-1f;
  return (p * z) * a + a;
}

OFP float o_acosf(float x) {
  if (x != x) return x;
  if (x > 1.0f || x < -1.0f) return __builtin_nanf("");
  if (x > 0.5f) {
    float s = __builtin_sqrtf(0.5f * (1.0f - x));
    return 2.0f * o_asin_poly(s);
  }
  if (x < -0.5f) {
    float s = __builtin_sqrtf(0.5f * (1.0f + x));
    return (O_PI_F - 2.0f * o_asin_poly(s)) + O_PI_LO;
  }
  return (O_PIO2_F - o_asin_poly(x)) + O_PIO2_LO;
}

// OpenCL sign(): 1, -1, the zero itself, or 0 for NaN.
OFP float o_sign(float x) {
  if (x > 0.0f) return 1.0f;
  if (x < 0.0f) return -1.0f;
  if (x == 0.0f) return x;
  return 0.0f;
}

// OpenCL common-function max(x, y): y if x < y, otherwise x.
OFP float o_maxf(float x, float y) { return x < y ? y : x; }

// bool promoted to float, as `(a < b) * v` does in OpenCL C.
OFP float o_b2f(int b) { return b ? 1.0f : 0.0f; }

// (int)f with the AMDGPU v_cvt_i32_f32 saturation (NaN -> 0).
OFP int32_t o_f2i(float f) {
  if (f != f) return 0;
  if (f >= 2147483648.0f) return 2147483647;
  if (f <= -2147483648.0f) return (int32_t)(-2147483647 - 1);
  return (int32_t)f;
}

// (uint)f with the AMDGPU v_cvt_u32_f32 saturation (NaN and negatives -> 0).
OFP uint32_t o_f2u(float f) {
  if (!(f > -1.0f)) return 0u;
  if (f >= 4294967296.0f) return 4294967295u;
  return (uint32_t)f;
}
