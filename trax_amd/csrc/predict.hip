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

// Windows of the 256 pairs of a tile, one per thread, parked in LDS so that
// lane groups can walk them: [0] count window, [1] store window; the phi part
// is the same for both.  Sector bounds are packed lo | hi << 16 (the ABI
// limits n_z to 32767 and n_phi to 16383).
struct PredTile {
  float zlo[2][256], zhi[2][256];
  uint32_t zs[2][256];
  float plo[256], phi_hi[256];
  uint32_t ps[256];  // phiLowSector | phiHighSector << 16 | wrap << 31
  uint32_t g3[256];  // grid offset of the third layer, 0xFFFFFFFF = no pair
  uint32_t cnt[256], pos[256];
};

__device__ __forceinline__ void pred_setup(const SeedParams& sp, const SortedHits& h,
                                           const uint32_t* __restrict__ grid,
                                           const uint2* __restrict__ pairs, const uint32_t* offs,
                                           uint32_t q, bool valid, bool need_count,
                                           bool need_store, PredTile& T) {
  const uint32_t t = threadIdx.x;
  if (!valid) {
    T.g3[t] = 0xFFFFFFFFu;
    return;
  }
  const PredItem it = pred_item(sp, h, grid, pairs, offs, q);
  T.g3[t] = (uint32_t)(it.g3 - grid);
  PredWindow w;
  if (need_count) {
    w = predict_window<false>(sp, it.lt, it.p1x, it.p1y, it.p1z, it.phi1, it.p2x, it.p2y, it.p2z,
                              it.phi2, it.minR, it.maxR);
    T.zlo[0][t] = w.zLow;
    T.zhi[0][t] = w.zHigh;
    T.zs[0][t] = (w.zLowSector < sp.n_z ? w.zLowSector : sp.n_z) | (w.zHighSector << 16);
  }
  if (need_store) {
    w = predict_window<true>(sp, it.lt, it.p1x, it.p1y, it.p1z, it.phi1, it.p2x, it.p2y, it.p2z,
                             it.phi2, it.minR, it.maxR);
    T.zlo[1][t] = w.zLow;
    T.zhi[1][t] = w.zHigh;
    T.zs[1][t] = (w.zLowSector < sp.n_z ? w.zLowSector : sp.n_z) | (w.zHighSector << 16);
  }
  T.plo[t] = w.phiLow;
  T.phi_hi[t] = w.phiHigh;
  T.ps[t] = w.phiLowSector | (w.phiHighSector << 16) | ((uint32_t)w.wrap << 31);
}

// The reference's walk over one pair's window (TripletThetaPhiPredictor.h:
// 170-233, :379-456), done by a lane group: per z row, phi cells
// [phiLowSector, phiHighSector), wrapping at the row end; repeated visits of
// the wrap-around walk are kept as the reference keeps them.  W selects the
// count (0) or store (1) window; f as in group_scan_range.
template <int GS, int W, typename F>
__device__ __forceinline__ void pred_group_walk(const SeedParams& sp, const SortedHits& h,
                                                const uint32_t* __restrict__ grid,
                                                const PredTile& T, uint32_t t,
                                                const LaneGroup<GS>& lg, F&& f) {
  const uint32_t g3off = T.g3[t];
  if (g3off == 0xFFFFFFFFu) return;
  const uint32_t* g3 = grid + g3off;
  const uint32_t nPhi = sp.n_phi;
  const float zLow = T.zlo[W][t], zHigh = T.zhi[W][t];
  const uint32_t zs_ = T.zs[W][t], ps = T.ps[t];
  const float phiLow = T.plo[t], phiHigh = T.phi_hi[t];
  const bool wrap = (ps >> 31) != 0;
  const uint32_t pl = ps & 0xFFFFu, ph = (ps >> 16) & 0x7FFFu;
  const float PI = TRAX_PI_F;
  auto in_window = [&](float zi, float actPhi) {
    return (zLow <= zi && zi <= zHigh) &&
           ((!wrap && (phiLow <= actPhi && actPhi <= phiHigh)) ||
            (wrap && ((phiLow <= actPhi && actPhi <= PI) || (-PI <= actPhi && actPhi <= phiHigh))));
  };
  for (uint32_t zs = zs_ & 0xFFFFu; zs < (zs_ >> 16); ++zs) {
    const uint32_t row = zs * (nPhi + 1);
    const uint32_t A = g3[row + pl];
    const uint32_t H = g3[row + ph];
    if (!wrap) {
      if (group_scan_range<GS>(lg, A, H, h.z, h.phi, in_window, f)) return;
    } else {  // j in [A, zSectorEnd + H - zSectorStart), wrapped at zSectorEnd
      const uint32_t zSectorStart = g3[row];
      const uint32_t zSectorEnd = g3[row + nPhi + 1];  // (zs+1)(nPhi+1): padding cell included
      if (group_scan_range<GS>(lg, A, zSectorEnd, h.z, h.phi, in_window, f)) return;
      if (group_scan_range<GS>(lg, zSectorStart, H, h.z, h.phi, in_window, f)) return;
    }
  }
}

// Count every slot of this wave's 64 pairs (group g takes slots g, g + 64/GS, ...).
template <int GS>
__device__ __forceinline__ void pred_count_slots(const SeedParams& sp, const SortedHits& h,
                                                 const uint32_t* __restrict__ grid, PredTile& T,
                                                 const LaneGroup<GS>& lg) {
  constexpr uint32_t NG = 64 / GS;
  const uint32_t wbase = (threadIdx.x >> 6) * 64u;
  for (uint32_t i = 0; i < (uint32_t)GS; ++i) {
    const uint32_t t = wbase + lg.g + NG * i;
    uint32_t n = 0;
    pred_group_walk<GS, 0>(sp, h, grid, T, t, lg, [&](uint32_t, bool, uint32_t, uint32_t c) {
      n += c;
      return false;
    });
    if (lg.gl == 0) T.cnt[t] = n;
  }
}

// Store the hits of every slot of this wave into [pos, pos + cnt): the store
// window's hits in walk order until the count's budget is used (:394), then
// 0xFFFFFFFF in the unused slots.  Returns true on capacity overflow.
template <int GS>
__device__ __forceinline__ bool pred_store_slots(const SeedParams& sp, const SortedHits& h,
                                                 const uint32_t* __restrict__ grid, PredTile& T,
                                                 uint32_t q0, uint2* __restrict__ cands,
                                                 uint64_t cap, const LaneGroup<GS>& lg) {
  constexpr uint32_t NG = 64 / GS;
  const uint32_t wbase = (threadIdx.x >> 6) * 64u;
  bool overflow = false;
  for (uint32_t i = 0; i < (uint32_t)GS; ++i) {
    const uint32_t t = wbase + lg.g + NG * i;
    if (T.g3[t] == 0xFFFFFFFFu) continue;
    const uint64_t pos = T.pos[t];
    const uint32_t budget = T.cnt[t], q = q0 + t;
    uint32_t written = 0;
    if (budget > 0)
      pred_group_walk<GS, 1>(sp, h, grid, T, t, lg, [&](uint32_t idx, bool p, uint32_t r, uint32_t c) {
        if (p && written + r < budget) {
          const uint64_t d = pos + written + r;
          if (d < cap)
            cands[d] = make_uint2(q, idx);
          else
            overflow = true;
        }
        written += c;
        return written >= budget;
      });
    for (uint32_t sl = written + lg.gl; sl < budget; sl += GS) {
      const uint64_t d = pos + sl;
      if (d < cap)
        cands[d] = make_uint2(TRAX_UNWRITTEN, TRAX_UNWRITTEN);
      else
        overflow = true;
    }
  }
  return overflow;
}

// Two-phase form (count -> prefix sum -> store), used by the plain method.
// Tiles of 256 pairs, block-strided.
template <int GS, bool STORE>
__global__ __launch_bounds__(256) void k_predict(SeedParams sp, SortedHits h,
                                                 const uint32_t* __restrict__ grid,
                                                 const uint2* __restrict__ pairs,
                                                 const uint32_t* __restrict__ pair_off,
                                                 const uint32_t* __restrict__ d_npairs,
                                                 uint64_t pair_capacity,
                                                 uint32_t* __restrict__ counts,  // count: out; store: scanned
                                                 uint2* __restrict__ cands, uint64_t cand_capacity,
                                                 uint32_t* __restrict__ flags) {
  extern __shared__ uint32_t s_offs[];  // K + 1 offsets when they fit (offsets_lds_bytes)
  __shared__ PredTile T;
  const uint32_t K = sp.n_events * sp.n_triplets;
  const uint32_t* offs = stage_offsets(pair_off, K + 1, s_offs);
  const uint64_t nP64 = *d_npairs;
  const uint32_t nP = (uint32_t)(nP64 < pair_capacity ? nP64 : pair_capacity);
  const LaneGroup<GS> lg;
  bool overflow = false;
  for (uint32_t tile = blockIdx.x; tile * 256u < nP; tile += gridDim.x) {
    const uint32_t q = tile * 256u + threadIdx.x;
    const bool valid = q < nP;
    pred_setup(sp, h, grid, pairs, offs, q, valid, !STORE, STORE, T);
    if (STORE && valid) {
      T.pos[threadIdx.x] = counts[q];
      T.cnt[threadIdx.x] = counts[q + 1] - counts[q];
    }
    __syncthreads();
    if (!STORE) {
      pred_count_slots<GS>(sp, h, grid, T, lg);
      __syncthreads();
      if (valid) counts[q] = T.cnt[threadIdx.x];
    } else {
      overflow |= pred_store_slots<GS>(sp, h, grid, T, tile * 256u, cands, cand_capacity, lg);
    }
    __syncthreads();
  }
  if (overflow) atomicOr(flags, TRAX_FLAG_CAND_OVERFLOW);
}

// Single-pass form: count walk, block scan + look-back for the tile's slot
// base, store walk into the count's slots, per-segment candidate offsets and
// the total, for ticketed tiles of 256 pairs.
template <int GS>
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
  __shared__ PredTile T;
  __shared__ uint32_t s_tile, s_excl, s_wave[4];
  const uint32_t K = sp.n_events * sp.n_triplets;
  const uint32_t epoch = launch_epoch(ctrl);
  const uint32_t* offs = stage_offsets(pair_off, K + 1, s_offs);
  const uint64_t nP64 = *d_npairs;
  const uint32_t nP = (uint32_t)(nP64 < pair_capacity ? nP64 : pair_capacity);
  uint32_t ntiles = (nP + 255u) / 256u;
  if (ntiles == 0) ntiles = 1;
  const LaneGroup<GS> lg;
  bool overflow = false;
  while (true) {
    const uint32_t tile = next_tile(ctrl, &s_tile);
    if (tile >= ntiles) break;
    const uint32_t q = tile * 256u + threadIdx.x;
    const bool valid = q < nP;
    pred_setup(sp, h, grid, pairs, offs, q, valid, true, true, T);
    __syncthreads();
    pred_count_slots<GS>(sp, h, grid, T, lg);
    __syncthreads();
    const uint32_t n = T.cnt[threadIdx.x];
    const TileScan ts = tile_scan_lookback(valid ? n : 0u, tile, epoch, status, s_wave, &s_excl);
    T.pos[threadIdx.x] = ts.pos;
    __syncthreads();
    overflow |= pred_store_slots<GS>(sp, h, grid, T, tile * 256u, cands, cand_capacity, lg);
    if (valid) write_segment_starts(offs, K, segment_of(offs, K + 1, q), q, ts.pos, cand_off);
    if (tile == ntiles - 1 && threadIdx.x == 0) {
      const uint32_t total = ts.base + ts.total;
      for (int64_t k = K; k >= 0 && offs[k] >= nP; --k) cand_off[k] = total;
      *d_total = total;
    }
    __syncthreads();
  }
  if (overflow) atomicOr(flags, TRAX_FLAG_CAND_OVERFLOW);
  launch_done(ctrl, epoch);
}

template <int GS>
static hipError_t predict_launch(const SeedParams& sp, const SortedHits& h, const uint32_t* grid,
                                 const uint2* pairs, const uint32_t* pair_off,
                                 const uint32_t* d_npairs, uint64_t pair_capacity, uint32_t* counts,
                                 uint2* cands, uint64_t cand_capacity, bool store, uint32_t* flags,
                                 hipStream_t stream) {
  uint64_t b = (pair_capacity + 255u) / 256u;
  const uint32_t blocks = (uint32_t)(b == 0 ? 1 : (b > 65535u ? 65535u : b));
  if (store)
    hipLaunchKernelGGL((k_predict<GS, true>), dim3(blocks), dim3(256), offsets_lds_bytes(sp), stream, sp, h,
                       grid, pairs, pair_off, d_npairs, pair_capacity, counts, cands, cand_capacity, flags);
  else
    hipLaunchKernelGGL((k_predict<GS, false>), dim3(blocks), dim3(256), offsets_lds_bytes(sp), stream, sp, h,
                       grid, pairs, pair_off, d_npairs, pair_capacity, counts, cands, cand_capacity, flags);
  return hipGetLastError();
}

// group: lanes walking one pair's window together (16, 32 or 64)
hipError_t launch_predict(const SeedParams& sp, const SortedHits& h, const uint32_t* grid,
                          const uint2* pairs, const uint32_t* pair_off, const uint32_t* d_npairs,
                          uint64_t pair_capacity, uint32_t* counts, uint2* cands,
                          uint64_t cand_capacity, bool store, uint32_t* flags, uint32_t group,
                          hipStream_t stream) {
  switch (group) {
    case 32: return predict_launch<32>(sp, h, grid, pairs, pair_off, d_npairs, pair_capacity, counts, cands,
                                       cand_capacity, store, flags, stream);
    case 64: return predict_launch<64>(sp, h, grid, pairs, pair_off, d_npairs, pair_capacity, counts, cands,
                                       cand_capacity, store, flags, stream);
    default: return predict_launch<16>(sp, h, grid, pairs, pair_off, d_npairs, pair_capacity, counts, cands,
                                       cand_capacity, store, flags, stream);
  }
}

template <int GS>
static hipError_t predict_fused_launch(const SeedParams& sp, const SortedHits& h, const uint32_t* grid,
                                       const uint2* pairs, const uint32_t* pair_off,
                                       const uint32_t* d_npairs, uint64_t pair_capacity, uint2* cands,
                                       uint64_t cand_capacity, uint32_t* cand_off, uint32_t* d_total,
                                       ScanCtrl* ctrl, unsigned long long* status, uint32_t* flags,
                                       hipStream_t stream) {
  // one tile per block; the hardware overlaps blocks waiting in look-back
  uint64_t b = (pair_capacity + 255u) / 256u;
  if (b == 0) b = 1;
  const uint32_t blocks = (uint32_t)(b < kStatusWords ? b : kStatusWords);
  hipLaunchKernelGGL((k_predict_fused<GS>), dim3(blocks), dim3(256), offsets_lds_bytes(sp), stream, sp, h,
                     grid, pairs, pair_off, d_npairs, pair_capacity, cands, cand_capacity, cand_off,
                     d_total, ctrl, status, flags);
  return hipGetLastError();
}

hipError_t launch_predict_fused(const SeedParams& sp, const SortedHits& h, const uint32_t* grid,
                                const uint2* pairs, const uint32_t* pair_off,
                                const uint32_t* d_npairs, uint64_t pair_capacity, uint2* cands,
                                uint64_t cand_capacity, uint32_t* cand_off, uint32_t* d_total,
                                ScanCtrl* ctrl, unsigned long long* status, uint32_t* flags,
                                uint32_t group, hipStream_t stream) {
  switch (group) {
    case 32: return predict_fused_launch<32>(sp, h, grid, pairs, pair_off, d_npairs, pair_capacity, cands,
                                             cand_capacity, cand_off, d_total, ctrl, status, flags, stream);
    case 64: return predict_fused_launch<64>(sp, h, grid, pairs, pair_off, d_npairs, pair_capacity, cands,
                                             cand_capacity, cand_off, d_total, ctrl, status, flags, stream);
    default: return predict_fused_launch<16>(sp, h, grid, pairs, pair_off, d_npairs, pair_capacity, cands,
                                             cand_capacity, cand_off, d_total, ctrl, status, flags, stream);
  }
}

}  // namespace trax
