/* The device FP32 definitions (trax_amd/csrc/fp32_ocl.h, compiled here for the
   host with IEEE single-precision SSE arithmetic and no contraction) against
   the oracle's (oracle/ofp32.h), bit for bit: atan2, atan and acos on special
   values and on argv[1] random samples of four kinds (any bit pattern,
   detector-scale coordinates, [-1, 1], small values).  Exit status 1 on any
   mismatch. */
#include <stdio.h>
#include <stdint.h>
#include <string.h>
#include <math.h>
#include <stdlib.h>
#include "fp32_ocl.h"
#include "ofp32.h"
static uint64_t st = 0x9E3779B97F4A7C15ull;
static uint32_t rnd(void) { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return (uint32_t)(st >> 16); }
static float fbits(uint32_t b) { float f; memcpy(&f, &b, 4); return f; }
static uint32_t bitsf(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }
static float sample(void) {
  uint32_t k = rnd() % 4;
  if (k == 0) return fbits(rnd());                       /* any bit pattern */
  if (k == 1) return ((float)(int32_t)rnd()) / 2147483648.0f * 200.0f;  /* detector scale */
  if (k == 2) return ((float)(int32_t)rnd()) / 2147483648.0f;           /* [-1, 1] */
  return ((float)(int32_t)rnd()) / 2147483648.0f * 1e-3f;
}
int main(int argc, char** argv) {
  const long samples = argc > 1 ? atol(argv[1]) : 1000000L;
  const float sp[] = {0.0f, -0.0f, 1.0f, -1.0f, 0.5f, -0.5f, 0.41421356f, 2.0f, INFINITY, -INFINITY, NAN,
                      1e-38f, -1e-38f, 1e-45f, 3.4e38f, -3.4e38f, 0.49999997f, 0.50000006f, -0.50000006f};
  const int ns = sizeof(sp) / sizeof(sp[0]);
  long bad = 0, n = 0;
  for (int i = 0; i < ns; ++i) {
    for (int j = 0; j < ns; ++j, ++n) bad += bitsf(trax_atan2f(sp[i], sp[j])) != bitsf(o_atan2f(sp[i], sp[j]));
    n += 2;
    bad += bitsf(trax_atanf(sp[i])) != bitsf(o_atanf(sp[i]));
    bad += bitsf(trax_acosf(sp[i])) != bitsf(o_acosf(sp[i]));
  }
  for (long k = 0; k < samples; ++k, n += 3) {
    float a = sample(), b = sample();
    bad += bitsf(trax_atan2f(a, b)) != bitsf(o_atan2f(a, b));
    bad += bitsf(trax_atanf(a)) != bitsf(o_atanf(a));
    float c = ((float)(int32_t)rnd()) / 2147483648.0f;
    bad += bitsf(trax_acosf(c)) != bitsf(o_acosf(c));
  }
  printf("%ld mismatches in %ld evaluations\n", bad, n);
  return bad != 0;
}
