// Triplet prediction: extrapolate each hit pair to the third layer and
// collect the third-layer hits inside the predicted (z, phi) window.
//
// Reference: src/algorithms/TripletThetaPhiPredictor.h:46-497 and
// TripletThetaPhiPredictor.cpp:10-131 (one work-item per pair: predictCount,
// PrefixSum, blocking read of the total, predictStore, and a racy offset
// "monotonize" pass).
//
// Here count and store are the same flat, grid-strided walk over all pairs of
// the batch; the pair's (event, layer triplet) is the segment of its index in
// the pair offsets.  The reference's count and store windows differ on
// purpose-or-not (rOrigin :117-118 vs :326-327, zHigh margin sigmaPhi :130 vs
// sigmaZ :339) and its store stops at the count's slot budget (:394); both
// are reproduced exactly, and budget slots the store leaves empty are filled
// with 0xFFFFFFFF (the reference leaves them at whatever the allocator held).
// Offsets are gathered from the scanned counts, which is what the racy
// monotonize pass intends.
#include "common.h"
#include "kernels.h"

namespace trax {

struct PredWindow {
  float zLow, zHigh, phiLow, phiHigh;
  bool wrap;
  uint32_t zLowSector, zHighSector, phiLowSector, phiHighSector;
};

template <bool STORE>
__device__ __forceinline__ PredWindow predict_window(const SeedParams& sp, uint32_t lt, float p1x,
                                                     float p1y, float p1z, float phi1, float p2x,
                                                     float p2y, float p2z, float phi2, float minR,
                                                     float maxR) {
  const float sigmaZ = sp.sigma_z[lt], sigmaPhi = sp.sigma_phi[lt];
  const float R = sp.min_radius_curvature;
  const float PI = TRAX_PI_F;
  PredWindow w;
  float tip2 = p1y * (p2x - p1x) - p1x * (p2y - p1y);
  tip2 *= tip2;
  tip2 /= ((p2x - p1x) * (p2x - p1x) + (p2y - p1y) * (p2y - p1y));
  float rOrigin;
  if (!STORE) {  // :117-118
    rOrigin = p1x * p1x + p1y * p1y;
    rOrigin = __builtin_sqrtf(rOrigin - tip2);
  } else {  // :326-327
    rOrigin = __builtin_sqrtf(p1x * p1x + p1y * p1y);
    rOrigin = __builtin_sqrtf(rOrigin * rOrigin - tip2);
  }
  float cotTheta = (p2z - p1z) / (__builtin_sqrtf(p2x * p2x + p2y * p2y - tip2) - rOrigin);
  float tmp = p1z + (__builtin_sqrtf(maxR * maxR - tip2) - rOrigin) * cotTheta;
  float zLow = p1z + (__builtin_sqrtf(minR * minR - tip2) - rOrigin) * cotTheta;
  float zHigh = trax_b2f(tmp > zLow) * tmp + trax_b2f(tmp < zLow) * zLow;
  zLow = trax_b2f(tmp > zLow) * zLow + trax_b2f(tmp < zLow) * tmp;
  zLow -= sigmaZ;
  zHigh += STORE ? sigmaZ : sigmaPhi;  // :339 vs :130
  int32_t zl = trax_f2i(__builtin_floorf((zLow - sp.min_z) / sp.sector_size_z));
  w.zLowSector = (uint32_t)(zl > 0 ? zl : 0);
  uint32_t zh = trax_f2u(__builtin_floorf((zHigh - sp.min_z) / sp.sector_size_z)) + 1u;
  w.zHighSector = zh < sp.n_z ? zh : sp.n_z;

  float phi = phi2;  // atan2(p2.y, p2.x)
  tmp = phi1;        // atan2(p1.y, p1.x)
  float dPhi = phi - tmp;
  dPhi += (dPhi > PI) ? -2 * PI : (dPhi < -PI) ? 2 * PI : 0;
  dPhi = __builtin_fabsf(dPhi);
  float dHits = __builtin_sqrtf((p2x - p1x) * (p2x - p1x) + (p2y - p1y) * (p2y - p1y));
  tmp = __builtin_fabsf(trax_acosf(dHits / (2 * R)) - trax_acosf(maxR / (2 * R)));
  float phiLow = __builtin_fabsf(trax_acosf(dHits / (2 * R)) - trax_acosf(minR / (2 * R)));
  dPhi = trax_maxf(dPhi, trax_maxf(tmp, phiLow));
  float phiHigh = phi + dPhi + sigmaPhi;
  phiLow = phi - dPhi - sigmaPhi;
  w.wrap = phiLow < -PI || phiHigh > PI || phiLow > PI || phiHigh < -PI;
  phiLow -= (float)((phiLow > PI) * 2) * PI;
  phiLow += (float)((phiLow < -PI) * 2) * PI;
  phiHigh -= (float)((phiHigh > PI) * 2) * PI;
  phiHigh += (float)((phiHigh < -PI) * 2) * PI;
  uint32_t pl = trax_f2u(__builtin_floorf((phiLow - sp.min_phi) / sp.sector_size_phi));
  uint32_t ph = trax_f2u(__builtin_floorf((phiHigh - sp.min_phi) / sp.sector_size_phi)) + 1u;
  w.phiLowSector = pl > sp.n_phi ? sp.n_phi : pl;
  w.phiHighSector = ph < sp.n_phi ? ph : sp.n_phi;
  w.zLow = zLow;
  w.zHigh = zHigh;
  w.phiLow = phiLow;
  w.phiHigh = phiHigh;
  return w;
}

template <bool STORE>
__global__ __launch_bounds__(512) void k_predict(SeedParams sp, SortedHits h,
                                                 const uint32_t* __restrict__ grid,
                                                 const uint2* __restrict__ pairs,
                                                 const uint32_t* __restrict__ pair_off,
                                                 const uint32_t* __restrict__ d_npairs,
                                                 uint64_t pair_capacity,
                                                 uint32_t* __restrict__ counts,  // count: out; store: scanned
                                                 uint2* __restrict__ cands, uint64_t cand_capacity,
                                                 uint32_t* __restrict__ flags) {
  const uint32_t P = sp.n_triplets, K = sp.n_events * P, C = sp.cells, L = sp.n_layers;
  const uint32_t nPhi = sp.n_phi;
  uint64_t nP64 = *d_npairs;
  const uint32_t nP = (uint32_t)(nP64 < pair_capacity ? nP64 : pair_capacity);
  const uint32_t stride = gridDim.x * blockDim.x;
  const float PI = TRAX_PI_F;
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < nP; q += stride) {
    const uint32_t k = segment_of(pair_off, K + 1, q);
    const uint32_t e = k / P, lt = k - e * P;
    const uint2 pr = pairs[q];
    const uint32_t h1 = pr.x, h2 = pr.y;
    const uint32_t layer = sp.layer3[lt] - 1;
    const uint32_t* g3 = grid + ((uint64_t)e * L + layer) * C;
    const PredWindow win = predict_window<STORE>(sp, lt, h.x[h1], h.y[h1], h.z[h1], h.phi[h1],
                                                 h.x[h2], h.y[h2], h.z[h2], h.phi[h2],
                                                 sp.min_radius[layer], sp.max_radius[layer]);
    uint32_t n = 0;
    uint64_t pos = 0, next = 0;
    if (STORE) {
      pos = counts[q];
      next = counts[q + 1];
    }
    bool overflow = false;
    auto visit = [&](uint32_t a, uint32_t b) {
      for (uint32_t idx = a; idx < b && (!STORE || pos < next); ++idx) {
        const float zi = h.z[idx];
        bool valid = win.zLow <= zi && zi <= win.zHigh;
        const float actPhi = h.phi[idx];
        valid = valid && ((!win.wrap && (win.phiLow <= actPhi && actPhi <= win.phiHigh)) ||
                          (win.wrap && ((win.phiLow <= actPhi && actPhi <= PI) ||
                                        (-PI <= actPhi && actPhi <= win.phiHigh))));
        if (valid) {
          if (STORE) {
            if (pos < cand_capacity)
              cands[pos] = make_uint2(q, idx);
            else
              overflow = true;
            ++pos;
          } else {
            ++n;
          }
        }
      }
    };
    for (uint32_t zs = win.zLowSector; zs < win.zHighSector; ++zs) {
      const uint32_t row = zs * (nPhi + 1);
      const uint32_t zSectorStart = g3[row];
      const uint32_t zSectorEnd = g3[row + nPhi + 1];  // (zs+1)(nPhi+1): padding cell included
      const uint32_t A = g3[row + win.phiLowSector];
      const uint32_t H = g3[row + win.phiHighSector];
      if (!win.wrap) {
        visit(A, H);
      } else {  // j in [A, zSectorEnd + H - zSectorStart), wrapped at zSectorEnd
        visit(A, zSectorEnd);
        visit(zSectorStart, H);
      }
    }
    if (!STORE) {
      counts[q] = n;
    } else {
      for (; pos < next; ++pos) {
        if (pos < cand_capacity)
          cands[pos] = make_uint2(TRAX_UNWRITTEN, TRAX_UNWRITTEN);
        else
          overflow = true;
      }
      if (overflow) atomicOr(flags, TRAX_FLAG_CAND_OVERFLOW);
    }
  }
}

hipError_t launch_predict(const SeedParams& sp, const SortedHits& h, const uint32_t* grid,
                          const uint2* pairs, const uint32_t* pair_off, const uint32_t* d_npairs,
                          uint64_t pair_capacity, uint32_t* counts, uint2* cands,
                          uint64_t cand_capacity, bool store, uint32_t* flags, uint32_t block,
                          hipStream_t stream) {
  uint64_t b = (pair_capacity + block - 1) / block;
  uint32_t blocks = (uint32_t)(b == 0 ? 1 : (b > 32768u ? 32768u : b));
  if (store)
    hipLaunchKernelGGL(k_predict<true>, dim3(blocks), dim3(block), 0, stream, sp, h, grid, pairs,
                       pair_off, d_npairs, pair_capacity, counts, cands, cand_capacity, flags);
  else
    hipLaunchKernelGGL(k_predict<false>, dim3(blocks), dim3(block), 0, stream, sp, h, grid, pairs,
                       pair_off, d_npairs, pair_capacity, counts, cands, cand_capacity, flags);
  return hipGetLastError();
}

}  // namespace trax
