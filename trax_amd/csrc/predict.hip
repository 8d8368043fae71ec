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

// Third-layer hits of one pair that pass its window, in the reference's walk
// order (TripletThetaPhiPredictor.h:170-233, :379-456): per z row, phi cells
// [phiLowSector, phiHighSector), wrapping at the row end; repeated visits of
// the wrap-around walk are kept, as the reference keeps them.
template <typename F>
__device__ __forceinline__ void predict_walk(const SeedParams& sp, const uint32_t* __restrict__ g3,
                                             const PredWindow& win, const SortedHits& h, F&& f) {
  const uint32_t nPhi = sp.n_phi;
  const float PI = TRAX_PI_F;
  auto in_window = [&](float zi, float actPhi) {
    return (win.zLow <= zi && zi <= win.zHigh) &&
           ((!win.wrap && (win.phiLow <= actPhi && actPhi <= win.phiHigh)) ||
            (win.wrap && ((win.phiLow <= actPhi && actPhi <= PI) || (-PI <= actPhi && actPhi <= win.phiHigh))));
  };
  for (uint32_t zs = win.zLowSector; zs < win.zHighSector; ++zs) {
    const uint32_t row = zs * (nPhi + 1);
    const uint32_t zSectorStart = g3[row];
    const uint32_t zSectorEnd = g3[row + nPhi + 1];  // (zs+1)(nPhi+1): padding cell included
    const uint32_t A = g3[row + win.phiLowSector];
    const uint32_t H = g3[row + win.phiHighSector];
    if (!win.wrap) {
      if (scan_range(A, H, h.z, h.phi, in_window, f)) return;
    } else {  // j in [A, zSectorEnd + H - zSectorStart), wrapped at zSectorEnd
      if (scan_range(A, zSectorEnd, h.z, h.phi, in_window, f)) return;
      if (scan_range(zSectorStart, H, h.z, h.phi, in_window, f)) return;
    }
  }
}

struct PredItem {
  uint32_t k, lt;
  const uint32_t* g3;
  float p1x, p1y, p1z, phi1, p2x, p2y, p2z, phi2, minR, maxR;
};

__device__ __forceinline__ PredItem pred_item(const SeedParams& sp, const SortedHits& h,
                                              const uint32_t* __restrict__ grid,
                                              const uint2* __restrict__ pairs,
                                              const uint32_t* offs, uint32_t q) {
  const uint32_t P = sp.n_triplets, K = sp.n_events * P, C = sp.cells, L = sp.n_layers;
  PredItem it;
  it.k = segment_of(offs, K + 1, q);
  const uint32_t e = it.k / P;
  it.lt = it.k - e * P;
  const uint2 pr = pairs[q];
  const uint32_t layer = sp.layer3[it.lt] - 1;
  it.g3 = grid + ((uint64_t)e * L + layer) * C;
  it.p1x = h.x[pr.x], it.p1y = h.y[pr.x], it.p1z = h.z[pr.x], it.phi1 = h.phi[pr.x];
  it.p2x = h.x[pr.y], it.p2y = h.y[pr.y], it.p2z = h.z[pr.y], it.phi2 = h.phi[pr.y];
  it.minR = sp.min_radius[layer];
  it.maxR = sp.max_radius[layer];
  return it;
}

template <bool STORE>
__device__ __forceinline__ PredWindow item_window(const SeedParams& sp, const PredItem& it) {
  return predict_window<STORE>(sp, it.lt, it.p1x, it.p1y, it.p1z, it.phi1, it.p2x, it.p2y, it.p2z,
                               it.phi2, it.minR, it.maxR);
}

// Store walk of pair q into slots [pos, next): the store window's hits until
// the count's budget is used up (:394), then 0xFFFFFFFF in the unused slots.
__device__ __forceinline__ bool predict_store_walk(const SeedParams& sp, const SortedHits& h,
                                                   const PredItem& it, uint32_t q, uint64_t pos,
                                                   uint64_t next, uint2* __restrict__ cands,
                                                   uint64_t cand_capacity) {
  bool overflow = false;
  if (pos < next) {
    const PredWindow win = item_window<true>(sp, it);
    predict_walk(sp, it.g3, win, h, [&](uint32_t idx) {
      if (pos < cand_capacity)
        cands[pos] = make_uint2(q, idx);
      else
        overflow = true;
      ++pos;
      return pos >= next;
    });
  }
  for (; pos < next; ++pos) {
    if (pos < cand_capacity)
      cands[pos] = make_uint2(TRAX_UNWRITTEN, TRAX_UNWRITTEN);
    else
      overflow = true;
  }
  return overflow;
}

// Two-phase form (count -> prefix sum -> store), used by the plain method.
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
  extern __shared__ uint32_t s_offs[];  // K + 1 offsets when they fit (offsets_lds_bytes)
  const uint32_t K = sp.n_events * sp.n_triplets;
  const uint32_t* offs = stage_offsets(pair_off, K + 1, s_offs);
  const uint64_t nP64 = *d_npairs;
  const uint32_t nP = (uint32_t)(nP64 < pair_capacity ? nP64 : pair_capacity);
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < nP; q += stride) {
    const PredItem it = pred_item(sp, h, grid, pairs, offs, q);
    if (!STORE) {
      uint32_t n = 0;
      predict_walk(sp, it.g3, item_window<false>(sp, it), h, [&](uint32_t) { ++n; return false; });
      counts[q] = n;
    } else {
      if (predict_store_walk(sp, h, it, q, counts[q], counts[q + 1], cands, cand_capacity))
        atomicOr(flags, TRAX_FLAG_CAND_OVERFLOW);
    }
  }
}

// Single-pass form, ticketed wave tiles of 64 pairs.  One walk over the union of
// the count and store windows (their phi parts are identical; the z bounds
// differ by the rOrigin rounding and the zHigh margin) counts the count
// window's hits and keeps the first kPredBuf store-window hits, in walk
// order, in LDS.  After the block scan + look-back gives the tile's slot
// base, each pair writes min(count, store hits) candidates from the buffer
// (re-walking only if more than kPredBuf are needed) and fills the rest of
// its budget with 0xFFFFFFFF, exactly as count + scan + store would.
constexpr uint32_t kPredBuf = 8;

__global__ __launch_bounds__(256) void k_predict_fused(SeedParams sp, SortedHits h,
                                                       const uint32_t* __restrict__ grid,
                                                       const uint2* __restrict__ pairs,
                                                       const uint32_t* __restrict__ pair_off,
                                                       const uint32_t* __restrict__ d_npairs,
                                                       uint64_t pair_capacity,
                                                       uint2* __restrict__ cands,
                                                       uint64_t cand_capacity,
                                                       uint32_t* __restrict__ cand_off,
                                                       uint32_t* __restrict__ d_total, ScanCtrl* ctrl,
                                                       unsigned long long* status,
                                                       uint32_t* __restrict__ flags) {
  extern __shared__ uint32_t s_offs[];
  __shared__ uint32_t s_buf[kPredBuf][256];
  const uint32_t K = sp.n_events * sp.n_triplets, nPhi = sp.n_phi, tid = threadIdx.x;
  const uint32_t epoch = launch_epoch(ctrl);
  const uint32_t* offs = stage_offsets(pair_off, K + 1, s_offs);
  const uint64_t nP64 = *d_npairs;
  const uint32_t nP = (uint32_t)(nP64 < pair_capacity ? nP64 : pair_capacity);
  uint32_t ntiles = (nP + 63u) / 64u;
  if (ntiles == 0) ntiles = 1;
  const float PI = TRAX_PI_F;
  bool overflow = false;
  while (true) {  // wave tiles of 64 pairs
    const uint32_t tile = wave_ticket(ctrl);
    if (tile >= ntiles) break;
    const uint32_t q = tile * 64u + lane_id();
    const bool valid = q < nP;
    PredItem it;
    uint32_t n_count = 0, n_store = 0;
    if (valid) {
      it = pred_item(sp, h, grid, pairs, offs, q);
      const PredWindow wc = item_window<false>(sp, it);
      const PredWindow ws = item_window<true>(sp, it);
      const bool wrap = wc.wrap;  // same for both windows
      const float phiLow = wc.phiLow, phiHigh = wc.phiHigh;
      const uint32_t zlo = wc.zLowSector < ws.zLowSector ? wc.zLowSector : ws.zLowSector;
      const uint32_t zhi = wc.zHighSector > ws.zHighSector ? wc.zHighSector : ws.zHighSector;
      for (uint32_t zs = zlo; zs < zhi; ++zs) {
        const bool rc = zs >= wc.zLowSector && zs < wc.zHighSector;
        const bool rs = zs >= ws.zLowSector && zs < ws.zHighSector;
        const float czLo = wc.zLow, czHi = wc.zHigh, szLo = ws.zLow, szHi = ws.zHigh;
        auto visit = [&](uint32_t a, uint32_t b) {
          auto one = [&](uint32_t idx, float zi, float actPhi) {
            const bool ph = (!wrap && (phiLow <= actPhi && actPhi <= phiHigh)) ||
                            (wrap && ((phiLow <= actPhi && actPhi <= PI) || (-PI <= actPhi && actPhi <= phiHigh)));
            n_count += (rc && ph && czLo <= zi && zi <= czHi) ? 1u : 0u;
            if (rs && ph && szLo <= zi && zi <= szHi) {
              if (n_store < kPredBuf) s_buf[n_store][tid] = idx;
              ++n_store;
            }
          };
          uint32_t idx = a;
          for (; idx + 4 <= b; idx += 4) {
            float zz[4], pp[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              zz[u] = h.z[idx + u];
              pp[u] = h.phi[idx + u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) one(idx + u, zz[u], pp[u]);
          }
          for (; idx < b; ++idx) one(idx, h.z[idx], h.phi[idx]);
        };
        const uint32_t row = zs * (nPhi + 1);
        const uint32_t A = it.g3[row + wc.phiLowSector];
        const uint32_t H = it.g3[row + wc.phiHighSector];
        if (!wrap) {
          visit(A, H);
        } else {
          visit(A, it.g3[row + nPhi + 1]);
          visit(it.g3[row], H);
        }
      }
    }
    const TileScan ts = wave_scan_lookback(n_count, tile, epoch, status);
    if (valid) {
      const uint64_t pos = ts.pos;
      const uint32_t m = n_count < n_store ? n_count : n_store;
      if (m <= kPredBuf) {
        for (uint32_t j = 0; j < n_count; ++j) {
          const uint64_t d = pos + j;
          if (d < cand_capacity)
            cands[d] = j < m ? make_uint2(q, s_buf[j][tid]) : make_uint2(TRAX_UNWRITTEN, TRAX_UNWRITTEN);
          else
            overflow = true;
        }
      } else {
        overflow |= predict_store_walk(sp, h, it, q, pos, pos + n_count, cands, cand_capacity);
      }
      write_segment_starts(offs, K, it.k, q, ts.pos, cand_off);
    }
    if (tile == ntiles - 1 && lane_id() == 0) {
      const uint32_t total = ts.base + ts.total;
      for (int64_t k = K; k >= 0 && offs[k] >= nP; --k) cand_off[k] = total;
      *d_total = total;
    }
  }
  if (overflow) atomicOr(flags, TRAX_FLAG_CAND_OVERFLOW);
  __syncthreads();
  launch_done(ctrl, epoch);
}

hipError_t launch_predict(const SeedParams& sp, const SortedHits& h, const uint32_t* grid,
                          const uint2* pairs, const uint32_t* pair_off, const uint32_t* d_npairs,
                          uint64_t pair_capacity, uint32_t* counts, uint2* cands,
                          uint64_t cand_capacity, bool store, uint32_t* flags, uint32_t block,
                          hipStream_t stream) {
  uint64_t b = (pair_capacity + block - 1) / block;
  uint32_t blocks = (uint32_t)(b == 0 ? 1 : (b > 32768u ? 32768u : b));
  if (store)
    hipLaunchKernelGGL(k_predict<true>, dim3(blocks), dim3(block), offsets_lds_bytes(sp), stream, sp, h, grid, pairs,
                       pair_off, d_npairs, pair_capacity, counts, cands, cand_capacity, flags);
  else
    hipLaunchKernelGGL(k_predict<false>, dim3(blocks), dim3(block), offsets_lds_bytes(sp), stream, sp, h, grid, pairs,
                       pair_off, d_npairs, pair_capacity, counts, cands, cand_capacity, flags);
  return hipGetLastError();
}

hipError_t launch_predict_fused(const SeedParams& sp, const SortedHits& h, const uint32_t* grid,
                                const uint2* pairs, const uint32_t* pair_off,
                                const uint32_t* d_npairs, uint64_t pair_capacity, uint2* cands,
                                uint64_t cand_capacity, uint32_t* cand_off, uint32_t* d_total,
                                ScanCtrl* ctrl, unsigned long long* status, uint32_t* flags,
                                hipStream_t stream) {
  uint64_t b = (pair_capacity + 255u) / 256u;
  if (b == 0) b = 1;
  const uint32_t blocks = (uint32_t)(b < kPersistentBlocks ? b : kPersistentBlocks);
  hipLaunchKernelGGL(k_predict_fused, dim3(blocks), dim3(256), offsets_lds_bytes(sp), stream, sp, h,
                     grid, pairs, pair_off, d_npairs, pair_capacity, cands, cand_capacity, cand_off,
                     d_total, ctrl, status, flags);
  return hipGetLastError();
}

}  // namespace trax
