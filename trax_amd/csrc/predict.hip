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
#include "cuts.h"
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
                                                     float maxR, float acosMinR, float acosMaxR) {
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
  // acos(maxR / (2 * R)) and acos(minR / (2 * R)) come precomputed per layer
  const float acosHits = trax_acosf(dHits / (2 * R));
  tmp = __builtin_fabsf(acosHits - acosMaxR);
  float phiLow = __builtin_fabsf(acosHits - acosMinR);
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
      if (scan_range(A, H, h.zphi, in_window, f)) return;
    } else {  // j in [A, zSectorEnd + H - zSectorStart), wrapped at zSectorEnd
      if (scan_range(A, zSectorEnd, h.zphi, in_window, f)) return;
      if (scan_range(zSectorStart, H, h.zphi, in_window, f)) return;
    }
  }
}

struct PredItem {
  uint32_t k, lt;
  const uint32_t* g3;
  float p1x, p1y, p1z, phi1, p2x, p2y, p2z, phi2, minR, maxR, acosMinR, acosMaxR;
};

// Pair q's item from its hits' (x, y, z, phi), a = inner, b = outer.
__device__ __forceinline__ PredItem pred_item_from(const SeedParams& sp, const uint32_t* __restrict__ grid,
                                                   const uint32_t* offs, uint32_t q, const float4& a,
                                                   const float4& b) {
  const uint32_t P = sp.n_triplets, K = sp.n_events * P, C = sp.cells, L = sp.n_layers;
  PredItem it;
  it.k = segment_of(offs, K + 1, q);
  const uint32_t e = it.k / P;
  it.lt = it.k - e * P;
  const uint32_t layer = sp.layer3[it.lt] - 1;
  it.g3 = grid + ((uint64_t)e * L + layer) * C;
  it.p1x = a.x, it.p1y = a.y, it.p1z = a.z, it.phi1 = a.w;
  it.p2x = b.x, it.p2y = b.y, it.p2z = b.z, it.phi2 = b.w;
  it.minR = sp.min_radius[layer];
  it.maxR = sp.max_radius[layer];
  it.acosMinR = sp.acos_min_r[layer];
  it.acosMaxR = sp.acos_max_r[layer];
  return it;
}

__device__ __forceinline__ PredItem pred_item(const SeedParams& sp, const SortedHits& h,
                                              const uint32_t* __restrict__ grid,
                                              const uint2* __restrict__ pairs,
                                              const uint32_t* offs, uint32_t q) {
  const uint2 pr = pairs[q];
  return pred_item_from(sp, grid, offs, q, load_xyzphi(h.xyzphi, pr.x), load_xyzphi(h.xyzphi, pr.y));
}

template <bool STORE>
__device__ __forceinline__ PredWindow item_window(const SeedParams& sp, const PredItem& it) {
  return predict_window<STORE>(sp, it.lt, it.p1x, it.p1y, it.p1z, it.phi1, it.p2x, it.p2y, it.p2z,
                               it.phi2, it.minR, it.maxR, it.acosMinR, it.acosMaxR);
}

// Store walk of pair q into slots [pos, next): the store window's hits until
// the count's budget is used up (:394), then 0xFFFFFFFF in the unused slots.
// The first `skip` slots are left alone (already written by the caller).
__device__ __forceinline__ bool predict_store_walk(const SeedParams& sp, const SortedHits& h,
                                                   const PredItem& it, uint32_t q, uint64_t pos,
                                                   uint64_t next, uint2* __restrict__ cands,
                                                   uint64_t cand_capacity, uint32_t skip = 0) {
  bool overflow = false;
  uint64_t seen = 0;  // store-window hits visited
  if (pos + skip < next) {
    const PredWindow win = item_window<true>(sp, it);
    predict_walk(sp, it.g3, win, h, [&](uint32_t idx) {
      const uint64_t d = pos + seen;
      if (seen >= skip) {
        if (d < cand_capacity)
          cands[d] = make_uint2(q, idx);
        else
          overflow = true;
      }
      ++seen;
      return pos + seen >= next;
    });
  }
  for (uint64_t d = pos + (seen > skip ? seen : skip); d < next; ++d) {
    if (d < cand_capacity)
      cands[d] = make_uint2(TRAX_UNWRITTEN, TRAX_UNWRITTEN);
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

// Tiled form, used by the batched chain.  A wave owns a tile of 64
// consecutive pairs.  Pass 1 walks each pair's union of the count and store
// windows once (their phi parts are identical; the z bounds differ by the
// rOrigin rounding and the zHigh margin): it counts the count window's hits
// and buffers the first kPredBuf store-window hits in the lane's own LDS
// column.  A wave prefix sum of the counts then lays the tile's slots out
// densely in its stash slab (kTileSlots uint32): the store list, 0xFFFFFFFF
// for budget slots the store list does not reach, and kSlotRewalk for slots
// past the buffer.  After the global prefix sum of the counts, pass 2 copies
// each slab to the tile's final position (slots and pairs of a tile are
// contiguous in the output) with coalesced reads and writes; only pairs with
// more than kPredBuf stored hits, or tiles whose slots overflow the slab,
// walk again.  The result is what count + scan + store give, slot for slot.
constexpr uint32_t kPredBuf = 16;
constexpr uint32_t kSlotRewalk = 0xFFFFFFFEu;

// The first BUF store-window hits go to s_col[j * STRIDE]; later ones to
// spill(j, idx).  With SCRATCH_ROW (the staged walk: 0.93 -> 0.90 ms on
// PU200; in the fused walk the LDS write per visit costs more than the
// branch it removes, 1.12 against 1.03 ms), s_col has BUF + 1 rows and row
// BUF is scratch.
struct NoSpill {
  __device__ void operator()(uint32_t, uint32_t) const {}
};

template <uint32_t BUF = kPredBuf, uint32_t STRIDE = 256, bool SCRATCH_ROW = false, typename Spill = NoSpill>
__device__ __forceinline__ uint32_t predict_count_buffer(const SeedParams& sp, const SortedHits& h,
                                                         const PredItem& it, uint32_t* s_col,
                                                         uint32_t* n_store_out, Spill&& spill = Spill{}) {
  const uint32_t nPhi = sp.n_phi;
  const float PI = TRAX_PI_F;
  const PredWindow wc = item_window<false>(sp, it);
  const PredWindow ws = item_window<true>(sp, it);
  const bool wrap = wc.wrap;  // same for both windows
  const float phiLow = wc.phiLow, phiHigh = wc.phiHigh;
  const float czLo = wc.zLow, czHi = wc.zHigh, szLo = ws.zLow, szHi = ws.zHigh;
  const uint32_t zlo = wc.zLowSector < ws.zLowSector ? wc.zLowSector : ws.zLowSector;
  const uint32_t zhi = wc.zHighSector > ws.zHighSector ? wc.zHighSector : ws.zHighSector;
  uint32_t n_count = 0, n_store = 0;
  for (uint32_t zs = zlo; zs < zhi; ++zs) {
    const bool rc = zs >= wc.zLowSector && zs < wc.zHighSector;
    const bool rs = zs >= ws.zLowSector && zs < ws.zHighSector;
    // count-window and store-window verdicts of one hit
    auto tests = [&](float zi, float actPhi, bool& in_s) {
      const bool ph = (!wrap && (phiLow <= actPhi && actPhi <= phiHigh)) ||
                      (wrap && ((phiLow <= actPhi && actPhi <= PI) || (-PI <= actPhi && actPhi <= phiHigh)));
      in_s = rs && ph && szLo <= zi && zi <= szHi;
      return (rc && ph && czLo <= zi && zi <= czHi) ? 1u : 0u;
    };
    auto one = [&](uint32_t idx, float zi, float actPhi) {
      bool in_s;
      n_count += tests(zi, actPhi, in_s);
      if (SCRATCH_ROW) {
        // every visited hit is written to the lane's next free slot (row
        // BUF, past the buffer, takes the writes beyond it), which only a
        // store-window hit claims: no branch per hit while the buffer lasts
        s_col[(n_store < BUF ? n_store : BUF) * STRIDE] = idx;
        if (in_s) {
          if (n_store >= BUF) spill(n_store, idx);
          ++n_store;
        }
      } else if (in_s) {
        if (n_store < BUF)
          s_col[n_store * STRIDE] = idx;
        else
          spill(n_store, idx);
        ++n_store;
      }
    };
    auto visit = [&](uint32_t a, uint32_t b) {
      uint32_t idx = a;
      for (; idx + 4 <= b; idx += 4) {
        float2 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = load_zphi(h.zphi, idx + u);
        if (SCRATCH_ROW) {
          // the four hits' slots; the hits past the buffer, if any, go to
          // spill() in slot order afterwards
          uint32_t sm = 0, p = n_store;  // sm: store-window hits of the four, one bit each
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            bool in_s;
            n_count += tests(v[u].x, v[u].y, in_s);
            s_col[(p < BUF ? p : BUF) * STRIDE] = idx + u;
            p += in_s ? 1u : 0u;
            sm |= in_s ? 1u << u : 0u;
          }
          if (p > BUF) {
            uint32_t j = n_store;
            for (; sm; sm &= sm - 1u, ++j)
              if (j >= BUF) spill(j, idx + (uint32_t)__builtin_ctz(sm));
          }
          n_store = p;
        } else {
#pragma unroll
          for (int u = 0; u < 4; ++u) one(idx + u, v[u].x, v[u].y);
        }
      }
      if (idx < b) {  // the last one to three hits, loaded together
        float2 v[3];
#pragma unroll
        for (uint32_t u = 0; u < 3; ++u) v[u] = load_zphi(h.zphi, idx + u < b ? idx + u : idx);
#pragma unroll
        for (uint32_t u = 0; u < 3; ++u)
          if (idx + u < b) one(idx + u, v[u].x, v[u].y);
      }
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
  *n_store_out = n_store;
  return n_count;
}

__global__ __launch_bounds__(256) void k_predict_walk(SeedParams sp, SortedHits h,
                                                      const uint32_t* __restrict__ grid,
                                                      const uint2* __restrict__ pairs,
                                                      const uint32_t* __restrict__ pair_off,
                                                      const uint32_t* __restrict__ d_npairs,
                                                      uint64_t pair_capacity,
                                                      uint32_t* __restrict__ counts,
                                                      uint32_t* __restrict__ slab,
                                                      uint32_t* __restrict__ tile_counts) {
  extern __shared__ uint32_t s_offs[];
  __shared__ uint32_t s_buf[(kPredBuf + 1) * 256];  // column per thread: s_buf[j * 256 + tid]
  const uint32_t K = sp.n_events * sp.n_triplets, tid = threadIdx.x, lane = lane_id();
  const uint32_t* offs = stage_offsets(pair_off, K + 1, s_offs);
  const uint64_t nP64 = *d_npairs;
  const uint32_t nP = (uint32_t)(nP64 < pair_capacity ? nP64 : pair_capacity);
  const uint32_t ntiles = (nP + kPredTile - 1) / kPredTile;
  const uint32_t nwaves = gridDim.x * (blockDim.x / 64u);
  uint32_t* s_col = s_buf + tid;
  for (uint32_t t = blockIdx.x * (blockDim.x / 64u) + (tid >> 6); t < ntiles; t += nwaves) {
    const uint32_t q = t * kPredTile + lane;
    uint32_t n_count = 0, n_store = 0;
    if (q < nP) {
      const PredItem it = pred_item(sp, h, grid, pairs, offs, q);
      n_count = predict_count_buffer<kPredBuf, 256, true>(sp, h, it, s_col, &n_store);
      counts[q] = n_count;
    }
    const uint32_t incl = wave_inclusive_scan(n_count);
    const uint32_t total = wave_last(incl);
    if (lane == 0) tile_counts[t] = total;
    if (total <= kTileSlots) {
      uint32_t* out = slab + (uint64_t)t * kTileSlots + (incl - n_count);
      for (uint32_t j = 0; j < n_count; ++j)
        out[j] = j < n_store ? (j < kPredBuf ? s_col[j * 256u] : kSlotRewalk) : TRAX_UNWRITTEN;
    }
  }
  // a zero after the last tile and the tile count (the scan's length) past
  // the largest tile count the workspace holds
  if (blockIdx.x == 0 && tid == 0) {
    const uint32_t tiles_max = (uint32_t)((pair_capacity + kPredTile - 1) / kPredTile);
    tile_counts[ntiles] = 0u;
    tile_counts[tiles_max + 1u] = ntiles;
  }
}

// Candidate offsets per (event, layer triplet) from the tile scan: the
// scanned slots of the tile holding the segment's first pair plus the counts
// of that pair's tile-mates before it.  The last entry is the total.
__global__ __launch_bounds__(256) void k_pred_offsets(const uint32_t* __restrict__ tile_scanned,
                                                      const uint32_t* __restrict__ counts,
                                                      const uint32_t* __restrict__ pair_off,
                                                      const uint32_t* __restrict__ d_npairs,
                                                      uint64_t pair_capacity, uint32_t n,
                                                      uint32_t* __restrict__ cand_off,
                                                      uint32_t* __restrict__ d_total) {
  const uint64_t nP64 = *d_npairs;
  const uint32_t nP = (uint32_t)(nP64 < pair_capacity ? nP64 : pair_capacity);
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    uint32_t q = pair_off[k];
    q = q < nP ? q : nP;
    const uint32_t t = q / kPredTile;
    uint32_t s = tile_scanned[t];
    for (uint32_t j = t * kPredTile; j < q; ++j) s += counts[j];
    cand_off[k] = s;
    if (k == n - 1u) *d_total = s;
  }
}

// Pass 2: slab -> final slots.  `tile_scanned` is the exclusive scan of the
// tiles' slot totals; a pair's first slot is its tile's entry plus the counts
// of its tile-mates before it (a wave scan of `counts`).
__global__ __launch_bounds__(256) void k_predict_place(SeedParams sp, SortedHits h,
                                                       const uint32_t* __restrict__ grid,
                                                       const uint2* __restrict__ pairs,
                                                       const uint32_t* __restrict__ pair_off,
                                                       const uint32_t* __restrict__ d_npairs,
                                                       uint64_t pair_capacity,
                                                       const uint32_t* __restrict__ tile_scanned,
                                                       const uint32_t* __restrict__ counts,
                                                       const uint32_t* __restrict__ slab,
                                                       uint2* __restrict__ cands,
                                                       uint64_t cand_capacity,
                                                       uint32_t* __restrict__ flags) {
  // the pair offsets are only needed by pairs that walk again (rare): read
  // them from global memory instead of staging them for every block
  __shared__ uint16_t s_own[4][kTileSlots];
  const uint32_t lane = lane_id();
  const uint32_t* offs = pair_off;
  const uint64_t nP64 = *d_npairs;
  const uint32_t nP = (uint32_t)(nP64 < pair_capacity ? nP64 : pair_capacity);
  const uint32_t ntiles = (nP + kPredTile - 1) / kPredTile;
  const uint32_t nwaves = gridDim.x * (blockDim.x / 64u);
  bool overflow = false;
  for (uint32_t t = blockIdx.x * (blockDim.x / 64u) + (threadIdx.x >> 6); t < ntiles; t += nwaves) {
    const uint32_t q = t * kPredTile + lane;
    const uint32_t n_own = q < nP ? counts[q] : 0u;  // slots of pair q
    const uint32_t incl = wave_inclusive_scan(n_own);
    const uint32_t base = tile_scanned[t];
    const uint32_t start = base + incl - n_own;  // first slot of pair q
    const uint32_t total = wave_last(incl);
    if (total > kTileSlots) {  // slab overflowed: every pair of the tile walks again
      if (q < nP) {
        const uint32_t n = n_own;
        if (n) {
          const PredItem it = pred_item(sp, h, grid, pairs, offs, q);
          overflow |= predict_store_walk(sp, h, it, q, start, (uint64_t)start + n, cands, cand_capacity);
        }
      }
      continue;
    }
    // slot -> (owner lane, slot index within the owner's budget), through a
    // per-wave LDS map: each lane marks its own slots
    const uint32_t rel = start - base;  // this lane's first slot within the tile
    uint16_t* own = s_own[threadIdx.x >> 6];
    for (uint32_t j = 0; j < n_own; ++j) own[rel + j] = (uint16_t)((j << 6) | lane);
    wave_lds_sync();
    for (uint32_t sl = lane; sl < total; sl += 64u) {
      const uint32_t o = own[sl];
      const uint32_t l = o & 63u, j = o >> 6;
      const uint32_t qo = t * kPredTile + l;
      const uint32_t v = slab[(uint64_t)t * kTileSlots + sl];
      const uint64_t d = (uint64_t)base + sl;
      if (v == kSlotRewalk) {
        if (j == kPredBuf) {  // this pair's slots past the buffer: walk again
          const uint32_t n = counts[qo];
          const PredItem it = pred_item(sp, h, grid, pairs, offs, qo);
          overflow |= predict_store_walk(sp, h, it, qo, d - j, d - j + n, cands, cand_capacity,
                                         kPredBuf);
        }
      } else if (d < cand_capacity) {
        cands[d] = v != TRAX_UNWRITTEN ? make_uint2(qo, v) : make_uint2(TRAX_UNWRITTEN, TRAX_UNWRITTEN);
      } else {
        overflow = true;
      }
    }
    wave_lds_sync();  // the map is rewritten for the wave's next tile
  }
  if (overflow) atomicOr(flags, TRAX_FLAG_CAND_OVERFLOW);
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

static uint32_t tile_blocks(uint64_t pair_capacity) {
  const uint64_t tiles = (pair_capacity + kPredTile - 1) / kPredTile;
  uint64_t b = (tiles + 3) / 4;  // 4 waves (tiles) per 256-thread block
  return (uint32_t)(b == 0 ? 1 : (b > 16384u ? 16384u : b));
}

hipError_t launch_predict_walk(const SeedParams& sp, const SortedHits& h, const uint32_t* grid,
                               const uint2* pairs, const uint32_t* pair_off, const uint32_t* d_npairs,
                               uint64_t pair_capacity, uint32_t* counts, uint32_t* slab,
                               uint32_t* tile_counts, hipStream_t stream) {
  hipLaunchKernelGGL(k_predict_walk, dim3(tile_blocks(pair_capacity)), dim3(256), offsets_lds_bytes(sp),
                     stream, sp, h, grid, pairs, pair_off, d_npairs, pair_capacity, counts, slab, tile_counts);
  return hipGetLastError();
}

// Prediction and filter in one pass (trax_triplet_predict_filter).  The
// filter's accept bit of a candidate is a function of its three hits and its
// layer triplet alone, and the filter keeps candidate order; so the triplets
// are the accepted store-window hits within each pair's slot budget, in pair
// order and walk order, and the candidate records need not be written.
// Pass 1 walks each tile of 64 pairs as k_predict_walk does (the count
// budget and the store-window hits, in LDS), queues the tile's
// buffered candidates densely in the wave's LDS and runs the phi cut over
// the queue.  The tile's phi survivors go to its slab as (lane | layer
// triplet << 6, hit1, hit2, hit3) records in pair and slot order, their
// number to the tile's accepted total.  A tile with more phi survivors
// than its slab holds (11 % of the tiles of the real ttbar events' coarse
// 8 x 50 grids, none on PU200) runs the theta and circle cuts in the walk,
// on its queue, and writes its accepted records as (pair, hit1, hit2, hit3)
// (or the per-pair accepted counts when more than kAccSlots) and
// kTileDone + its accepted total.  The theta and circle cuts of the other
// tiles run over their survivors, 32 tiles at a time (k_acc_cut), on full
// waves: a tile has ~14 phi
// survivors and ~7 theta survivors on PU200, so in the walk each of those
// cuts was a wave pass of mostly idle lanes (1.21 ms for the walk with all
// three cuts, 1.06 ms with the phi cut, plus 0.09 ms for k_acc_cut).
// k_acc_cut leaves each tile's accepted records in its slab as (pair, hit1,
// hit2, hit3), in order, and their number as the tile's accepted total.
// The candidate budgets still give the candidate offsets and total.  After
// the prefix sum of the accepted tile totals, pass 2 writes the triplets.
constexpr uint32_t kAccSlots = kTileSlots / 4;  // uint4 records per tile slab
constexpr uint32_t kTileRedo = 0xFFFFFFFFu;     // accepted total not known yet
constexpr uint32_t kTileDone = 0x80000000u;     // + accepted total: finished in the walk

// Accepted store-window hits of pair `it` among its slots [skip, budget),
// in walk order (slots past the store window's hits hold no hit).  The rare
// path: the same visit order as predict_walk, one hit at a time, so the cuts
// are inlined once rather than once per unrolled hit.
template <typename F>
__device__ __forceinline__ void accept_walk(const SeedParams& sp, const SortedHits& h, const PredItem& it,
                                            uint32_t skip, uint32_t budget, F&& emit) {
  if (skip >= budget) return;
  const PredWindow win = item_window<true>(sp, it);
  const uint32_t nPhi = sp.n_phi;
  const float PI = TRAX_PI_F;
  uint32_t seen = 0;
  for (uint32_t zs = win.zLowSector; zs < win.zHighSector; ++zs) {
    const uint32_t row = zs * (nPhi + 1);
    const uint32_t A = it.g3[row + win.phiLowSector];
    const uint32_t H = it.g3[row + win.phiHighSector];
    const uint32_t E = win.wrap ? it.g3[row + nPhi + 1] : H;  // first range [A, E)
    const uint32_t S2 = win.wrap ? it.g3[row] : 0u;            // wrapped range [S2, H)
    for (int part = 0; part < (win.wrap ? 2 : 1); ++part) {
      const uint32_t lo = part ? S2 : A, hi = part ? H : E;
      for (uint32_t idx = lo; idx < hi; ++idx) {
        const float2 v = load_zphi(h.zphi, idx);
        const float zi = v.x, actPhi = v.y;
        const bool in = (win.zLow <= zi && zi <= win.zHigh) &&
                        ((!win.wrap && (win.phiLow <= actPhi && actPhi <= win.phiHigh)) ||
                         (win.wrap && ((win.phiLow <= actPhi && actPhi <= PI) ||
                                       (-PI <= actPhi && actPhi <= win.phiHigh))));
        if (!in) continue;
        if (seen >= skip) {
          const float4 c = load_xyzphi(h.xyzphi, idx);
          if (filter_accept(sp, it.lt, it.p1x, it.p1y, it.p1z, it.p2x, it.p2y, it.p2z, c.x, c.y, c.z)) emit(idx);
        }
        if (++seen >= budget) return;
      }
    }
  }
}

// One cut over a wave's queue of 16-bit candidate entries; survivors are
// compacted in place (a survivor's new position is at most its old one, and
// the chunk is read before it is written).  With KEEP false the survivors
// are passed to mark(entry) instead.  Idle lanes see entry `lane`, live false.
template <bool KEEP, typename Cut, typename Mark>
__device__ __forceinline__ uint32_t queue_cut(uint16_t* queue, uint32_t nq, Cut&& cut, Mark&& mark) {
  const uint32_t lane = lane_id();
  uint32_t n_out = 0;
  for (uint32_t c0 = 0; c0 < nq; c0 += 64u) {
    const uint32_t e = c0 + lane;
    const uint32_t ent = e < nq ? (uint32_t)queue[e] : lane;
    const bool ok = cut(ent, e < nq);
    if (KEEP) {
      const uint64_t m = __ballot(ok);
      if (ok) queue[n_out + lanes_below(m)] = (uint16_t)ent;
      n_out += (uint32_t)__popcll(m);
    } else if (ok) {
      mark(ent);
    }
  }
  wave_lds_sync();
  return n_out;
}

// Per-wave LDS of the fused walk.  A lane buffers its first BUF
// store-window hits in its own column; later ones go to the wave's spill
// pool (pool order is each lane's slot order).  Queue entries are
// (j << 6 | lane) for buffered slot j, kSpillTag + k for pool entry k.
// BUF (the lane buffer depth), the waves per block and the waves per SIMD
// the register allocation targets are tuning knobs (trax_launch_config);
// the compiled combinations are listed in kFusedVariants.
constexpr uint32_t kSpill = 192;
constexpr uint32_t kSpillTag = 1024;
template <uint32_t BUF>
struct FuseLds {
  static_assert(BUF * 64 <= kSpillTag && kSpillTag + kSpill <= 65536 && BUF <= 32, "queue entries are 16-bit");
  uint32_t buf[BUF * 64];  // buf[j * 64 + lane]
  uint16_t queue[BUF * 64 + kSpill];
  uint32_t acc[64];             // accept bits of buffered slots, per lane
  uint32_t spill_hit[kSpill];
  uint16_t spill_tag[kSpill];   // j << 6 | lane
  uint8_t spill_acc[kSpill];
  uint32_t spill_n;
};

// One tile of pass 1 for the wave: pair q = t * 64 + lane.  One walk fills
// the buffers and the slot budgets, the phi cut runs over the dense queue of
// filled slots, and the tile's survivor total and survivor records are
// written (with more than kAccSlots survivors, the other two cuts run here
// and the tile's accepted records are).  Returns false, leaving the tile to
// k_acc_redo, when the spill pool overflows.
template <uint32_t BUF, typename Next>
__device__ __forceinline__ bool fused_tile(const SeedParams& sp, const SortedHits& h, const uint32_t* grid,
                                           const uint2* pairs, const uint32_t* offs, uint32_t nP, uint32_t t,
                                           FuseLds<BUF>& L, uint32_t* counts, uint32_t* tile_counts, uint4* acc_slab,
                                           uint32_t* acc_tiles, const float4& hit1, const float4& hit2,
                                           Next&& next) {
  const uint32_t lane = lane_id();
  const uint32_t q = t * kPredTile + lane;
  uint32_t* s_col = L.buf + lane;
  if (lane == 0) L.spill_n = 0u;
  wave_lds_sync();
  uint32_t n_count = 0, n_store = 0;
  PredItem it{};
  if (q < nP) {
    it = pred_item_from(sp, grid, offs, q, hit1, hit2);
    n_count = predict_count_buffer<BUF, 64>(sp, h, it, s_col, &n_store, [&](uint32_t j, uint32_t idx) {
      const uint32_t k = atomicAdd(&L.spill_n, 1u);
      if (k < kSpill) {
        L.spill_hit[k] = idx;
        L.spill_tag[k] = (uint16_t)((j << 6) | lane);
      }
    });
  }
  next();  // the next tile's loads start here, behind this tile's cuts
  const uint2 pr_q = q < nP ? pairs[q] : make_uint2(0u, 0u);  // for the slab records, behind the cuts
  // The per-pair slot budgets are read back only for tiles holding a
  // segment's first pair (the candidate offsets), the last tile (segments
  // that start at the pair total) and tiles walked again (spill pool or slab
  // overflow): the other tiles skip the 4 B per pair.
  const bool seg_start = q < nP && offs[it.k] == q;
  const bool keep = __ballot(seg_start) != 0ull || t == (nP - 1u) / kPredTile;
  if (keep && q < nP) counts[q] = n_count;
  const uint32_t incl = wave_inclusive_scan(n_count);
  const uint32_t total = wave_last(incl);
  if (lane == 0) tile_counts[t] = total;
  const uint32_t n_eff = n_count < n_store ? n_count : n_store;  // slots [0, n_eff) hold a hit
  wave_lds_sync();
  const uint32_t ns = L.spill_n;
  if (ns > kSpill) {
    if (!keep && q < nP) counts[q] = n_count;
    return false;
  }
  // the queue: buffered slots, then pool entries within their pair's budget
  const uint32_t nb = n_eff < BUF ? n_eff : BUF;
  const uint32_t qincl = wave_inclusive_scan(nb);
  uint32_t nq = wave_last(qincl);
  for (uint32_t j = 0; j < nb; ++j) L.queue[qincl - nb + j] = (uint16_t)((j << 6) | lane);
  for (uint32_t k0 = 0; k0 < ns; k0 += 64u) {
    const uint32_t k = k0 + lane;
    const uint32_t tag = k < ns ? (uint32_t)L.spill_tag[k] : 0u;
    const uint32_t ne = __shfl(n_eff, (int)(tag & 63u), 64);
    const bool ok = k < ns && (tag >> 6) < ne;
    const uint64_t m = __ballot(ok);
    if (ok) L.queue[nq + lanes_below(m)] = (uint16_t)(kSpillTag + k);
    nq += (uint32_t)__popcll(m);
    if (k < ns) L.spill_acc[k] = 0;
  }
  L.acc[lane] = 0u;
  wave_lds_sync();
  auto mark = [&](uint32_t ent) {
    if (ent >= kSpillTag)
      L.spill_acc[ent - kSpillTag] = 1;
    else
      atomicOr(&L.acc[ent & 63u], 1u << (ent >> 6));
  };
  // the phi cut's pair angle, once per pair
  const float a1 = nq ? trax_atan2f((it.p2y - it.p1y), (it.p2x - it.p1x)) : 0.0f;
  queue_cut<false>(L.queue, nq, [&](uint32_t ent, bool live) {
    const bool spilled = ent >= kSpillTag;
    const uint32_t tag = spilled ? (uint32_t)L.spill_tag[ent - kSpillTag] : ent;
    const uint32_t l = tag & 63u;
    const float ang = __shfl(a1, (int)l, 64);
    const float x2 = __shfl(it.p2x, (int)l, 64), y2 = __shfl(it.p2y, (int)l, 64);
    const uint32_t lt = (uint32_t)__shfl(it.lt, (int)l, 64);
    if (!live) return false;
    const uint32_t i3 = spilled ? L.spill_hit[ent - kSpillTag] : L.buf[(tag >> 6) * 64u + l];
    const float4 c = load_xyzphi(h.xyzphi, i3);
    return filter_phi_a1(sp, lt, ang, x2, y2, c.x, c.y);
  }, mark);
  auto marked_here = [&]() {  // marked slots of this lane
    uint32_t n = (uint32_t)__popc(L.acc[lane]);
    for (uint32_t k = 0; k < ns; ++k)
      n += (L.spill_acc[k] && (L.spill_tag[k] & 63u) == lane) ? 1u : 0u;
    return n;
  };
  uint32_t a = marked_here();
  uint32_t aincl = wave_inclusive_scan(a);
  uint32_t A = wave_last(aincl);
  const bool full = A > kAccSlots;
  if (full && !keep && q < nP) counts[q] = n_count;
  if (__builtin_expect(full, 0)) {
    // the theta and circle cuts here; the pair's hits come from memory (its
    // registers are not kept past the phi cut)
    auto hits = [&](uint32_t ent, bool live, float4& p1, float4& p2, float4& p3) {
      const bool spilled = ent >= kSpillTag;
      const uint32_t tag = spilled ? (uint32_t)L.spill_tag[ent - kSpillTag] : ent;
      const uint32_t l = tag & 63u;
      const uint32_t lt = (uint32_t)__shfl(it.lt, (int)l, 64);
      if (live) {
        const uint32_t i3 = spilled ? L.spill_hit[ent - kSpillTag] : L.buf[(tag >> 6) * 64u + l];
        const uint2 pr = pairs[t * kPredTile + l];
        p1 = load_xyzphi(h.xyzphi, pr.x);
        p2 = load_xyzphi(h.xyzphi, pr.y);
        p3 = load_xyzphi(h.xyzphi, i3);
      }
      return lt;
    };
    // theta over the phi survivors (the marked queue entries), kept in the queue
    const uint32_t n2 = queue_cut<true>(L.queue, nq, [&](uint32_t ent, bool live) {
      const bool ok = live && (ent >= kSpillTag ? L.spill_acc[ent - kSpillTag] != 0
                                                : ((L.acc[ent & 63u] >> (ent >> 6)) & 1u) != 0);
      float4 p1, p2, p3;
      const uint32_t lt = hits(ent, ok, p1, p2, p3);
      return ok && filter_theta(sp, lt, p1.x, p1.y, p1.z, p2.x, p2.y, p2.z, p3.x, p3.y, p3.z);
    }, [](uint32_t) {});
    L.acc[lane] = 0u;
    for (uint32_t k = lane; k < ns; k += 64u) L.spill_acc[k] = 0;
    wave_lds_sync();
    queue_cut<false>(L.queue, n2, [&](uint32_t ent, bool live) {
      float4 p1, p2, p3;
      const uint32_t lt = hits(ent, live, p1, p2, p3);
      return live && filter_circle(sp, lt, p1.x, p1.y, p2.x, p2.y, p3.x, p3.y);
    }, mark);
    a = marked_here();
    aincl = wave_inclusive_scan(a);
    A = wave_last(aincl);
  }
  const uint32_t mask = L.acc[lane];
  if (lane == 0) acc_tiles[t] = full ? kTileDone + A : A;
  uint4* slab = acc_slab + (uint64_t)t * kAccSlots;
  if (A <= kAccSlots) {
    if (a) {
      // phi survivors as (lane | layer triplet << 6, hit1, hit2, hit3), a
      // finished tile's accepted records as (pair, hit1, hit2, hit3)
      uint32_t pos = aincl - a;
      const uint2 pr = pr_q;
      const uint32_t tag = full ? q : (lane | (it.lt << 6));
      for (uint32_t m = mask; m; m &= m - 1u)
        slab[pos++] = make_uint4(tag, pr.x, pr.y, s_col[(uint32_t)__builtin_ctz(m) * 64u]);
      for (uint32_t k = 0; k < ns; ++k)
        if (L.spill_acc[k] && (L.spill_tag[k] & 63u) == lane) slab[pos++] = make_uint4(tag, pr.x, pr.y, L.spill_hit[k]);
    }
  } else {
    reinterpret_cast<uint32_t*>(slab)[lane] = a;  // pass 2 walks the tile's pairs again
  }
  wave_lds_sync();  // the buffers, queue and pool are reused by the wave's next tile
  return true;
}

// Default: an 8-deep lane buffer, four waves per block, seven waves per
// SIMD (at most 72 VGPRs, under 22.8 KB of LDS per block).  With all three
// cuts in the walk this shape ran PU200 in 1.22 ms against 1.31 ms at the
// six waves the unconstrained register allocation (79 VGPRs) allowed and
// 1.26 ms at eight.
template <uint32_t BUF, uint32_t NW, int WPE>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(WPE))) void k_predict_filter_walk(SeedParams sp, SortedHits h,
                                                             const uint32_t* __restrict__ grid,
                                                             const uint2* __restrict__ pairs,
                                                             const uint32_t* __restrict__ pair_off,
                                                             const uint32_t* __restrict__ d_npairs,
                                                             uint64_t pair_capacity,
                                                             uint32_t* __restrict__ counts,
                                                             uint32_t* __restrict__ tile_counts,
                                                             uint4* __restrict__ acc_slab,
                                                             uint32_t* __restrict__ acc_tiles) {
  extern __shared__ uint32_t s_offs[];
  __shared__ FuseLds<BUF> s_lds[NW];
  const uint32_t K = sp.n_events * sp.n_triplets, tid = threadIdx.x, wave = tid >> 6;
  const uint32_t* offs = stage_offsets(pair_off, K + 1, s_offs);
  const uint64_t nP64 = *d_npairs;
  const uint32_t nP = (uint32_t)(nP64 < pair_capacity ? nP64 : pair_capacity);
  const uint32_t ntiles = (nP + kPredTile - 1) / kPredTile;
  const XcdShare sh = xcd_share(ntiles, NW, wave);
  // Each tile's pair records are loaded a tile ahead and their hits during
  // the tile before (after its walk), so a tile's walk starts on resident
  // data (fused walk 1.005 -> 0.994 ms on PU200).
  const uint32_t lane = lane_id();
  auto has_pair = [&](uint32_t tt) { return tt < sh.end && tt * kPredTile + lane < nP; };
  uint2 pr_next = make_uint2(0u, 0u);
  float4 a_next = make_float4(0.0f, 0.0f, 0.0f, 0.0f), b_next = a_next;
  if (has_pair(sh.begin)) {
    pr_next = pairs[sh.begin * kPredTile + lane];
    a_next = load_xyzphi(h.xyzphi, pr_next.x);
    b_next = load_xyzphi(h.xyzphi, pr_next.y);
  }
  for (uint32_t t = sh.begin; t < sh.end; t += sh.stride) {
    const float4 a = a_next, b = b_next;
    const uint32_t t2 = t + sh.stride;
    const bool more = has_pair(t2);
    if (more) pr_next = pairs[t2 * kPredTile + lane];
    if (!fused_tile<BUF>(sp, h, grid, pairs, offs, nP, t, s_lds[wave], counts, tile_counts, acc_slab, acc_tiles, a, b,
                         [&]() {
                           if (more) {
                             a_next = load_xyzphi(h.xyzphi, pr_next.x);
                             b_next = load_xyzphi(h.xyzphi, pr_next.y);
                           }
                         }) &&
        lane == 0)
      acc_tiles[t] = kTileRedo;  // left to k_acc_redo
  }
  // a zero after the last tile and the tile count (the scans' length) past
  // the largest tile count the workspace holds
  if (blockIdx.x == 0 && tid == 0) {
    const uint32_t tiles_max = (uint32_t)((pair_capacity + kPredTile - 1) / kPredTile);
    tile_counts[ntiles] = 0u;
    tile_counts[tiles_max + 1u] = ntiles;
    acc_tiles[ntiles] = 0u;
    acc_tiles[tiles_max + 1u] = ntiles;
  }
}

// Tiles the walk marked kTileRedo (spill pool overflowed: more than kSpill
// slots past the lane buffers; none in the PU200 and ttbar workloads): each
// lane walks its pair's slots one by one
// with all three cuts, once to count and once to write the records (or the
// per-pair counts).  It runs after k_acc_cut, which leaves these tiles alone.  One lane per tile finds the marked
// tiles; the wave then takes them in turn.
__global__ __launch_bounds__(64) void k_acc_redo(SeedParams sp, SortedHits h, const uint32_t* __restrict__ grid,
                                                 const uint2* __restrict__ pairs,
                                                 const uint32_t* __restrict__ pair_off,
                                                 const uint32_t* __restrict__ d_npairs, uint64_t pair_capacity,
                                                 const uint32_t* __restrict__ counts,
                                                 uint4* __restrict__ acc_slab, uint32_t* __restrict__ acc_tiles) {
  const uint32_t lane = lane_id();
  const uint64_t nP64 = *d_npairs;
  const uint32_t nP = (uint32_t)(nP64 < pair_capacity ? nP64 : pair_capacity);
  const uint32_t ntiles = (nP + kPredTile - 1) / kPredTile;
  for (uint32_t t0 = blockIdx.x * 64u; t0 < ntiles; t0 += gridDim.x * 64u) {
    const uint32_t tl = t0 + lane;
    uint64_t m = __ballot(tl < ntiles && acc_tiles[tl] == kTileRedo);
    while (m) {
      const uint32_t t = t0 + (uint32_t)__builtin_ctzll(m);
      m &= m - 1u;
      const uint32_t q = t * kPredTile + lane;
      const uint32_t n = q < nP ? counts[q] : 0u;
      PredItem it{};
      uint32_t a = 0;
      if (n) {
        it = pred_item(sp, h, grid, pairs, pair_off, q);
        accept_walk(sp, h, it, 0u, n, [&](uint32_t) { ++a; });
      }
      const uint32_t aincl = wave_inclusive_scan(a);
      const uint32_t A = wave_last(aincl);
      uint4* slab = acc_slab + (uint64_t)t * kAccSlots;
      if (A <= kAccSlots) {
        uint32_t pos = aincl - a;
        if (a) {
          const uint2 pr = pairs[q];
          accept_walk(sp, h, it, 0u, n, [&](uint32_t idx) { slab[pos++] = make_uint4(q, pr.x, pr.y, idx); });
        }
      } else {
        reinterpret_cast<uint32_t*>(slab)[lane] = a;
      }
      if (lane == 0) acc_tiles[t] = A;
    }
  }
}

// Theta and circle cuts over the phi survivors of cut_group(tiles)
// consecutive tiles per block and round.  The tiles' survivor totals (written by the
// walk) are scanned in wave 0.  The survivors are then taken as one flat
// range, in windows of whole tiles holding at most 256 * kCutPer records:
// each thread reads kCutPer records (record r belongs to the tile whose
// running total first exceeds r) and evaluates the cuts, the accept bits go
// to an LDS bitmap, and after a barrier every accepted record is written back
// to its tile's slab as (pair, hit1, hit2, hit3) at the number of accepted
// records before it in the tile (all of the window's records are read before
// any is overwritten, and a record only moves down), with the tiles'
// accepted totals.  Tiles the walk finished (kTileDone + accepted total) get
// their accepted total; tiles it marked kTileRedo are left to k_acc_redo.
constexpr uint32_t kCutPer = 4;
constexpr uint32_t kCutWindow = 256u * kCutPer;

// Tiles per block and round of k_acc_cut: 32, or 16 below 65536 tiles, so
// that a small batch still spreads over more than a thousand blocks (ttbar,
// 19.5k tiles: 22 us at 16 against 35 us at 64; PU200, 485k tiles: 97 us at
// 32, 100 us at 64, 108 us at 16).
__host__ __device__ __forceinline__ uint32_t cut_group(uint64_t ntiles) {
  return ntiles >= 65536u ? 32u : 16u;
}

// Set bits of bitmap w in [lo, hi).
__device__ __forceinline__ uint32_t bits_in(const uint32_t* w, uint32_t lo, uint32_t hi) {
  uint32_t n = 0;
  for (uint32_t k = lo >> 5; k < ((hi + 31u) >> 5); ++k) {
    uint32_t m = w[k];
    const uint32_t b0 = k << 5;
    if (lo > b0) m &= ~0u << (lo - b0);
    if (hi < b0 + 32u) m &= (1u << (hi - b0)) - 1u;
    n += (uint32_t)__popc(m);
  }
  return n;
}

__global__ __launch_bounds__(256) void k_acc_cut(SeedParams sp, SortedHits h, const uint32_t* __restrict__ d_npairs,
                                                 uint64_t pair_capacity, uint4* __restrict__ acc_slab,
                                                 uint32_t* __restrict__ acc_tiles) {
  __shared__ uint32_t s_incl[64];
  __shared__ uint32_t s_ok[kCutWindow / 32u];
  __shared__ unsigned long long s_redo;  // tiles whose total is not the cut's (finished or redo)
  static_assert(kAccSlots <= kCutWindow, "a window holds a whole tile");
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6, tid = threadIdx.x;
  const uint64_t nP64 = *d_npairs;
  const uint32_t nP = (uint32_t)(nP64 < pair_capacity ? nP64 : pair_capacity);
  const uint32_t ntiles = (nP + kPredTile - 1) / kPredTile;
  const uint32_t G = cut_group(ntiles);
  for (uint32_t t0 = blockIdx.x * G; t0 < ntiles; t0 += gridDim.x * G) {
    if (wave == 0) {
      const uint32_t t = t0 + lane;
      const uint32_t v = lane < G && t < ntiles ? acc_tiles[t] : 0u;
      const bool redo = v >= kTileDone;  // kTileRedo, or finished in the walk
      if (redo && v != kTileRedo) acc_tiles[t] = v - kTileDone;
      const uint32_t incl = wave_inclusive_scan(redo ? 0u : v);
      s_incl[lane] = incl;
      const uint64_t m = __ballot(redo);
      if (lane == 0) s_redo = m;
    }
    __syncthreads();
    const uint32_t nt = ntiles - t0 < G ? ntiles - t0 : G;
    for (uint32_t tb = 0; tb < nt;) {
      // window: tiles [tb, te] whose records fit after R0
      const uint32_t R0 = tb ? s_incl[tb - 1u] : 0u;
      const uint64_t fit = __ballot(lane < nt && s_incl[lane] <= R0 + kCutWindow);
      const uint32_t te = 63u - (uint32_t)__builtin_clzll(fit);  // >= tb: a tile holds at most kAccSlots
      const uint32_t R1 = s_incl[te];
      uint4 rec[kCutPer];
      uint32_t lo[kCutPer];
      bool ok[kCutPer];
#pragma unroll
      for (uint32_t u = 0; u < kCutPer; ++u) {
        const uint32_t r = R0 + u * 256u + tid;
        rec[u] = make_uint4(0u, 0u, 0u, 0u);
        lo[u] = 0u;
        if (r < R1) {
          uint32_t a = tb, z = te;  // first tile whose running total exceeds r
          while (a < z) {
            const uint32_t mid = (a + z) >> 1;
            if (s_incl[mid] > r) z = mid; else a = mid + 1u;
          }
          lo[u] = a;
          rec[u] = acc_slab[(uint64_t)(t0 + a) * kAccSlots + (r - (a ? s_incl[a - 1u] : 0u))];
        }
      }
#pragma unroll
      for (uint32_t u = 0; u < kCutPer; ++u) {
        const uint32_t r = R0 + u * 256u + tid;
        ok[u] = false;
        if (r < R1) {
          const uint32_t lt = rec[u].x >> 6;
          const float4 p1 = load_xyzphi(h.xyzphi, rec[u].y), p2 = load_xyzphi(h.xyzphi, rec[u].z),
                       p3 = load_xyzphi(h.xyzphi, rec[u].w);
          ok[u] = filter_theta(sp, lt, p1.x, p1.y, p1.z, p2.x, p2.y, p2.z, p3.x, p3.y, p3.z) &&
                  filter_circle(sp, lt, p1.x, p1.y, p2.x, p2.y, p3.x, p3.y);
        }
        const uint64_t m = __ballot(ok[u]);
        if (lane == 0) {
          s_ok[u * 8u + 2u * wave] = (uint32_t)m;
          s_ok[u * 8u + 2u * wave + 1u] = (uint32_t)(m >> 32);
        }
      }
      __syncthreads();
#pragma unroll
      for (uint32_t u = 0; u < kCutPer; ++u) {
        if (!ok[u]) continue;
        const uint32_t r = R0 + u * 256u + tid;
        const uint32_t a = lo[u];
        const uint32_t tb0 = (a ? s_incl[a - 1u] : 0u) - R0;  // the tile's first record in the window
        const uint32_t dst = bits_in(s_ok, tb0, r - R0);
        acc_slab[(uint64_t)(t0 + a) * kAccSlots + dst] =
            make_uint4((t0 + a) * kPredTile + (rec[u].x & 63u), rec[u].y, rec[u].z, rec[u].w);
      }
      if (tid <= te - tb) {
        const uint32_t a = tb + tid;
        if (!((s_redo >> a) & 1ull)) {
          const uint32_t f = (a ? s_incl[a - 1u] : 0u) - R0;
          acc_tiles[t0 + a] = bits_in(s_ok, f, s_incl[a] - R0);
        }
      }
      __syncthreads();  // the bitmap is rewritten for the next window
      tb = te + 1u;
    }
  }
}


// Candidate and triplet offsets of the fused pass per (event, layer
// triplet), one wave per segment: the scanned tile totals at the tile
// holding the segment's first pair q, plus, within that tile, the slot
// budgets of the pairs before q (one coalesced load and a reduction) and the
// accepted records of pairs before q (their slab records, or the per-pair
// counts of a tile with more than kAccSlots).  The last entries are the
// totals.
__global__ __launch_bounds__(256) void k_fused_offsets(const uint32_t* __restrict__ tile_scanned,
                                                       const uint32_t* __restrict__ counts,
                                                       const uint32_t* __restrict__ acc_scanned,
                                                       const uint4* __restrict__ acc_slab,
                                                       const uint32_t* __restrict__ pair_off,
                                                       const uint32_t* __restrict__ d_npairs,
                                                       uint64_t pair_capacity, uint32_t n,
                                                       uint32_t* __restrict__ cand_off, uint32_t* __restrict__ trk_off,
                                                       uint32_t* __restrict__ d_totals) {
  const uint32_t lane = lane_id();
  const uint64_t nP64 = *d_npairs;
  const uint32_t nP = (uint32_t)(nP64 < pair_capacity ? nP64 : pair_capacity);
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  for (uint32_t k = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); k < n; k += waves) {
    uint32_t q = pair_off[k];
    q = q < nP ? q : nP;
    const uint32_t t = q / kPredTile, lq = q - t * kPredTile;
    // candidate slots of the tile-mates before q
    const uint32_t j = t * kPredTile + lane;
    uint32_t c = lane < lq ? counts[j] : 0u;
    c = wave_reduce_sum(c);
    // accepted records of the pairs before q
    uint32_t a = 0;
    if (lq) {
      const uint32_t s0 = acc_scanned[t], A = acc_scanned[t + 1u] - s0;
      const uint4* slab = acc_slab + (uint64_t)t * kAccSlots;
      if (A <= kAccSlots) {
        for (uint32_t r0 = 0; r0 < A; r0 += 64u) {
          const uint32_t r = r0 + lane;
          a += (uint32_t)__popcll(__ballot(r < A && slab[r].x < q));
        }
      } else {
        const uint32_t v = lane < lq ? reinterpret_cast<const uint32_t*>(slab)[lane] : 0u;
        a = wave_reduce_sum(v);
      }
    }
    if (lane == 0) {
      const uint32_t co = tile_scanned[t] + c, to = acc_scanned[t] + a;
      cand_off[k] = co;
      trk_off[k] = to;
      if (k == n - 1u) {
        d_totals[0] = co;
        d_totals[1] = to;
      }
    }
  }
}

// Pass 2: a block takes kPlaceTiles consecutive tiles.  Each wave reads the
// tiles' scanned positions and accepted totals (one lane per tile) and
// scans them; the block's threads then copy the tiles' records as one flat
// range (record r belongs to the tile whose running total first exceeds r),
// writing (hit1, hit2, hit3) per record.  A tile whose slab overflowed walks
// its pairs again, one wave on that tile.
constexpr uint32_t kPlaceTiles = 64;

__global__ __launch_bounds__(256) void k_acc_place(SeedParams sp, SortedHits h, const uint32_t* __restrict__ grid,
                                                   const uint2* __restrict__ pairs,
                                                   const uint32_t* __restrict__ pair_off,
                                                   const uint32_t* __restrict__ d_npairs, uint64_t pair_capacity,
                                                   const uint32_t* __restrict__ acc_scanned,
                                                   const uint32_t* __restrict__ counts,
                                                   const uint4* __restrict__ acc_slab, uint32_t* __restrict__ t1,
                                                   uint32_t* __restrict__ t2, uint32_t* __restrict__ t3,
                                                   uint64_t trk_capacity, uint32_t* __restrict__ flags) {
  __shared__ uint32_t s_incl[4][kPlaceTiles];
  __shared__ uint32_t s_base[4][kPlaceTiles];
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  const uint64_t nP64 = *d_npairs;
  const uint32_t nP = (uint32_t)(nP64 < pair_capacity ? nP64 : pair_capacity);
  const uint32_t ntiles = (nP + kPredTile - 1) / kPredTile;
  uint32_t* incl_w = s_incl[wave];
  uint32_t* base_w = s_base[wave];
  bool overflow = false;
  for (uint32_t t0 = blockIdx.x * kPlaceTiles; t0 < ntiles; t0 += gridDim.x * kPlaceTiles) {
    const uint32_t t = t0 + lane;
    uint32_t base = 0, A = 0;
    if (lane < kPlaceTiles && t < ntiles) {
      base = acc_scanned[t];
      A = acc_scanned[t + 1u] - base;
    }
    const bool big = A > kAccSlots;
    const uint32_t n = big ? 0u : A;
    const uint32_t incl = wave_inclusive_scan(n);
    const uint32_t R = wave_last(incl);
    if (lane < kPlaceTiles) {
      incl_w[lane] = incl;
      base_w[lane] = base;
    }
    wave_lds_sync();
    for (uint32_t r = wave * 64u + lane; r < R; r += 256u) {
      uint32_t lo = 0, hi = kPlaceTiles - 1u;  // first tile whose running total exceeds r
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (incl_w[mid] > r) hi = mid; else lo = mid + 1u;
      }
      const uint32_t j = r - (lo ? incl_w[lo - 1u] : 0u);
      const uint4 rec = acc_slab[(uint64_t)(t0 + lo) * kAccSlots + j];  // (pair, hit1, hit2, hit3)
      const uint64_t d = (uint64_t)base_w[lo] + j;
      if (d < trk_capacity) {
        t1[d] = rec.y;
        t2[d] = rec.z;
        t3[d] = rec.w;
      } else {
        overflow = true;
      }
    }
    uint32_t i = 0;
    for (uint64_t m = __ballot(big); m; m &= m - 1u, ++i) {
      if ((i & 3u) != wave) continue;  // big tiles go round the block's waves
      const uint32_t tt = t0 + (uint32_t)__builtin_ctzll(m);
      const uint32_t q = tt * kPredTile + lane;
      const uint32_t a = reinterpret_cast<const uint32_t*>(acc_slab + (uint64_t)tt * kAccSlots)[lane];
      const uint32_t ai = wave_inclusive_scan(a);
      uint64_t d = (uint64_t)base_w[tt - t0] + ai - a;
      if (a && q < nP) {
        const PredItem it = pred_item(sp, h, grid, pairs, pair_off, q);
        const uint2 pr = pairs[q];
        accept_walk(sp, h, it, 0u, counts[q], [&](uint32_t idx) {
          if (d < trk_capacity) {
            t1[d] = pr.x;
            t2[d] = pr.y;
            t3[d] = idx;
          } else {
            overflow = true;
          }
          ++d;
        });
      }
    }
    wave_lds_sync();  // the running totals are rewritten for the next tiles
  }
  if (overflow) atomicOr(flags, TRAX_FLAG_TRACKLET_OVERFLOW);
}

// The compiled (lane buffer, waves per block, waves per SIMD) combinations
// of the fused walk; entry 0 is the default.
const FusedVariant kFusedVariants[kNumFusedVariants] = {
    {8, 4, 7}, {8, 4, 6}, {8, 4, 8}, {4, 4, 8}, {4, 4, 7}, {8, 2, 7}, {12, 4, 6}};

template <uint32_t BUF, uint32_t NW, int WPE>
static void fused_launch(const SeedParams& sp, const SortedHits& h, const uint32_t* grid, const uint2* pairs,
                         const uint32_t* pair_off, const uint32_t* d_npairs, uint64_t pair_capacity,
                         uint32_t* counts, uint32_t* tile_counts, uint4* acc_slab, uint32_t* acc_tiles,
                         hipStream_t stream) {
  const uint64_t tiles = (pair_capacity + kPredTile - 1) / kPredTile;
  uint64_t b = (tiles + NW - 1) / NW;  // NW waves (tiles) per block and round
  const uint32_t blocks = (uint32_t)(b == 0 ? 1 : (b > 65536u / NW ? 65536u / NW : b));
  hipLaunchKernelGGL((k_predict_filter_walk<BUF, NW, WPE>), dim3(xcd_grid(blocks)), dim3(64 * NW),
                     offsets_lds_bytes(sp), stream, sp, h, grid, pairs, pair_off, d_npairs, pair_capacity, counts,
                     tile_counts, acc_slab, acc_tiles);
}

hipError_t launch_predict_filter_walk(const SeedParams& sp, const SortedHits& h, const uint32_t* grid,
                                      const uint2* pairs, const uint32_t* pair_off, const uint32_t* d_npairs,
                                      uint64_t pair_capacity, uint32_t* counts, uint32_t* tile_counts,
                                      uint4* acc_slab, uint32_t* acc_tiles, hipStream_t stream, uint32_t variant) {
#define TRAX_FUSED_CASE(i, B, W, E) \
  case i: fused_launch<B, W, E>(sp, h, grid, pairs, pair_off, d_npairs, pair_capacity, counts, tile_counts, acc_slab, acc_tiles, stream); break;
  switch (variant) {
    TRAX_FUSED_CASE(1, 8, 4, 6)
    TRAX_FUSED_CASE(2, 8, 4, 8)
    TRAX_FUSED_CASE(3, 4, 4, 8)
    TRAX_FUSED_CASE(4, 4, 4, 7)
    TRAX_FUSED_CASE(5, 8, 2, 7)
    TRAX_FUSED_CASE(6, 12, 4, 6)
    default: fused_launch<8, 4, 7>(sp, h, grid, pairs, pair_off, d_npairs, pair_capacity, counts, tile_counts,
                                   acc_slab, acc_tiles, stream);
  }
#undef TRAX_FUSED_CASE
  return hipGetLastError();
}

hipError_t launch_acc_redo(const SeedParams& sp, const SortedHits& h, const uint32_t* grid, const uint2* pairs,
                           const uint32_t* pair_off, const uint32_t* d_npairs, uint64_t pair_capacity,
                           const uint32_t* counts, uint4* acc_slab, uint32_t* acc_tiles, hipStream_t stream) {
  // one wave per block, one lane per tile: 64 tiles per block and round
  const uint64_t tiles = (pair_capacity + kPredTile - 1) / kPredTile;
  uint64_t b = (tiles + 63u) / 64u;
  const uint32_t blocks = (uint32_t)(b == 0 ? 1 : (b > 4096u ? 4096u : b));
  hipLaunchKernelGGL(k_acc_redo, dim3(blocks), dim3(64), 0, stream, sp, h, grid, pairs, pair_off, d_npairs,
                     pair_capacity, counts, acc_slab, acc_tiles);
  return hipGetLastError();
}

hipError_t launch_acc_cut(const SeedParams& sp, const SortedHits& h, const uint32_t* d_npairs,
                          uint64_t pair_capacity, uint4* acc_slab, uint32_t* acc_tiles, hipStream_t stream) {
  // cut_group(tiles) tiles per block and round (the kernel's group, from the
  // actual tile count, is no larger)
  const uint64_t tiles = (pair_capacity + kPredTile - 1) / kPredTile;
  const uint32_t g = cut_group(tiles);
  uint64_t b = (tiles + g - 1u) / g;
  const uint32_t blocks = (uint32_t)(b == 0 ? 1 : (b > 16384u ? 16384u : b));
  hipLaunchKernelGGL(k_acc_cut, dim3(blocks), dim3(256), 0, stream, sp, h, d_npairs, pair_capacity, acc_slab,
                     acc_tiles);
  return hipGetLastError();
}

hipError_t launch_fused_offsets(const uint32_t* tile_scanned, const uint32_t* counts, const uint32_t* acc_scanned,
                                const uint4* acc_slab, const uint32_t* pair_off, const uint32_t* d_npairs,
                                uint64_t pair_capacity, uint32_t n, uint32_t* cand_off, uint32_t* trk_off,
                                uint32_t* d_totals, hipStream_t stream) {
  uint32_t blocks = (n + 3u) / 4u;  // a wave per segment
  blocks = blocks == 0 ? 1 : (blocks > 16384u ? 16384u : blocks);
  hipLaunchKernelGGL(k_fused_offsets, dim3(blocks), dim3(256), 0, stream, tile_scanned, counts, acc_scanned,
                     acc_slab, pair_off, d_npairs, pair_capacity, n, cand_off, trk_off, d_totals);
  return hipGetLastError();
}

hipError_t launch_acc_place(const SeedParams& sp, const SortedHits& h, const uint32_t* grid, const uint2* pairs,
                            const uint32_t* pair_off, const uint32_t* d_npairs, uint64_t pair_capacity,
                            const uint32_t* acc_scanned, const uint32_t* counts, const uint4* acc_slab,
                            uint32_t* t1, uint32_t* t2, uint32_t* t3, uint64_t trk_capacity, uint32_t* flags,
                            hipStream_t stream) {
  // kPlaceTiles tiles per block and round
  const uint64_t tiles = (pair_capacity + kPredTile - 1) / kPredTile;
  uint64_t b = (tiles + kPlaceTiles - 1u) / kPlaceTiles;
  const uint32_t blocks = (uint32_t)(b == 0 ? 1 : (b > 16384u ? 16384u : b));
  hipLaunchKernelGGL(k_acc_place, dim3(blocks), dim3(256), 0, stream, sp, h, grid, pairs,
                     pair_off, d_npairs, pair_capacity, acc_scanned, counts, acc_slab, t1, t2, t3, trk_capacity,
                     flags);
  return hipGetLastError();
}

hipError_t launch_pred_offsets(const uint32_t* tile_scanned, const uint32_t* counts, const uint32_t* pair_off,
                               const uint32_t* d_npairs, uint64_t pair_capacity, uint32_t n,
                               uint32_t* cand_off, uint32_t* d_total, hipStream_t stream) {
  uint32_t blocks = (n + 255u) / 256u;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_pred_offsets, dim3(blocks), dim3(256), 0, stream, tile_scanned, counts, pair_off,
                     d_npairs, pair_capacity, n, cand_off, d_total);
  return hipGetLastError();
}

hipError_t launch_predict_place(const SeedParams& sp, const SortedHits& h, const uint32_t* grid,
                                const uint2* pairs, const uint32_t* pair_off, const uint32_t* d_npairs,
                                uint64_t pair_capacity, const uint32_t* tile_scanned, const uint32_t* counts,
                                const uint32_t* slab, uint2* cands, uint64_t cand_capacity, uint32_t* flags,
                                hipStream_t stream) {
  // one tile per wave, eight rounds of a full device at most: the grid is
  // sized from the pair capacity, which can exceed the pairs found many
  // times over, and surplus blocks only cost launch slots
  const uint64_t tiles = (pair_capacity + kPredTile - 1) / kPredTile;
  uint64_t b = (tiles + 3) / 4;
  const uint32_t blocks = (uint32_t)(b == 0 ? 1 : (b > 16384u ? 16384u : b));
  hipLaunchKernelGGL(k_predict_place, dim3(blocks), dim3(256), 0, stream, sp, h, grid, pairs, pair_off,
                     d_npairs, pair_capacity, tile_scanned, counts, slab, cands, cand_capacity, flags);
  return hipGetLastError();
}

}  // namespace trax
