// The triplet filter's cuts (TripletThetaPhiFilter.h:68-147), shared by the
// filter pass and the fused prediction + filter walk so both evaluate the
// same expressions in the same order.
#pragma once

#include "common.h"

namespace trax {

// filterCount predicate, TripletThetaPhiFilter.h:68-147, split in three:
// the phi-difference cut (two atan2; about one candidate in seven passes),
// the theta-ratio cut (two atan2, two square roots, a division) and the
// circle-fit TIP and radius cuts.  The reference evaluates all of them and
// ANDs the results; the cuts are pure functions, so evaluating them in any
// order, each only on the previous one's survivors, gives the same product.
// The phi cut with the pair's own angle atan2(y2 - y1, x2 - x1) given.
__device__ __forceinline__ bool filter_phi_a1(const SeedParams& sp, uint32_t lt, float angle1, float x2,
                                              float y2, float x3, float y3) {
  const float dPhiCut = sp.dphi_cut[lt];
  const float PI = TRAX_PI_F;
  const float angle2 = trax_atan2f((y3 - y2), (x3 - x2));
  float delta = angle2 - angle1;
  delta += (delta > PI) ? -2 * PI : (delta < -PI) ? 2 * PI : 0;
  return __builtin_fabsf(delta) <= dPhiCut;
}

__device__ __forceinline__ bool filter_phi(const SeedParams& sp, uint32_t lt, float x1, float y1,
                                           float x2, float y2, float x3, float y3) {
  return filter_phi_a1(sp, lt, trax_atan2f((y2 - y1), (x2 - x1)), x2, y2, x3, y3);
}

__device__ __forceinline__ bool filter_theta(const SeedParams& sp, uint32_t lt, float x1, float y1,
                                             float z1, float x2, float y2, float z2, float x3,
                                             float y3, float z3) {
  const float dThetaCut = sp.dtheta_cut[lt];
  const float angle1 = trax_atan2f(__builtin_sqrtf((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1)), (z2 - z1));
  const float angle2 = trax_atan2f(__builtin_sqrtf((x3 - x2) * (x3 - x2) + (y3 - y2) * (y3 - y2)), (z3 - z2));
  const float delta = __builtin_fabsf(angle2 / angle1);
  return 1 - dThetaCut <= delta && delta <= 1 + dThetaCut;
}

// Circle through the three hits from the paraboloid (x, y, x^2 + y^2)
// plane: transverse impact parameter and radius cuts (:105-147).
__device__ __forceinline__ bool filter_circle(const SeedParams& sp, uint32_t lt, float x1, float y1,
                                              float x2, float y2, float x3, float y3) {
  const float tipCut = sp.tip_cut[lt];
  const float minRadius = sp.min_radius_curvature;
  const float p1x = x1, p1y = y1, p1z = x1 * x1 + y1 * y1;
  const float p2x = x2, p2y = y2, p2z = x2 * x2 + y2 * y2;
  const float p3x = x3, p3y = y3, p3z = x3 * x3 + y3 * y3;
  const float ax = p2x - p1x, ay = p2y - p1y, az = p2z - p1z;
  const float bx = p3x - p1x, by = p3y - p1y, bz = p3z - p1z;
  float nx = ay * bz - az * by;
  float ny = az * bx - ax * bz;
  float nz = ax * by - ay * bx;
  const float nlen = __builtin_sqrtf(nx * nx + ny * ny + nz * nz);
  if (nlen != 0.0f) {
    nx = nx / nlen;
    ny = ny / nlen;
    nz = nz / nlen;
  }
  const float cOx = (-nx) / (2 * nz);
  const float cOy = (-ny) / (2 * nz);
  const float c = -(nx * p1x + ny * p1y + nz * p1z);
  const float cR = __builtin_sqrtf((1 - nz * nz - 4 * c * nz) / (4 * nz * nz));
  float vx = -cOx, vy = -cOy;
  const float vlen = __builtin_sqrtf(vx * vx + vy * vy);
  if (vlen != 0.0f) {
    vx = vx / vlen;
    vy = vy / vlen;
  }
  const float pCAx = cOx + cR * vx;
  const float pCAy = cOy + cR * vy;
  const float tip = pCAx * pCAx + pCAy * pCAy;
  return (tip <= (tipCut * tipCut)) && (cR >= minRadius);
}

// All three cuts: the filter's accept bit for candidate (hit1, hit2, hit3).
__device__ __forceinline__ bool filter_accept(const SeedParams& sp, uint32_t lt, float x1, float y1,
                                              float z1, float x2, float y2, float z2, float x3,
                                              float y3, float z3) {
  return filter_phi(sp, lt, x1, y1, x2, y2, x3, y3) &&
         filter_theta(sp, lt, x1, y1, z1, x2, y2, z2, x3, y3, z3) &&
         filter_circle(sp, lt, x1, y1, x2, y2, x3, y3);
}

}  // namespace trax
