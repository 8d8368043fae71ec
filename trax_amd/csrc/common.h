// Shared device-side definitions for the gfx950 seeding kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fp32_ocl.h"

#define TRAX_MAX_LAYERS 16
#define TRAX_MAX_TRIPLETS 8
#define TRAX_UNWRITTEN 0xFFFFFFFFu

// Status bits raised on the device (atomicOr into the context's flag word).
#define TRAX_FLAG_PAIR_OVERFLOW 0x1u
#define TRAX_FLAG_CAND_OVERFLOW 0x2u
#define TRAX_FLAG_TRACKLET_OVERFLOW 0x4u
#define TRAX_FLAG_LAYOUT 0x8u

namespace trax {

// Everything the kernels need from the configuration, passed by value as a
// kernel argument (no device-side table, nothing to upload, graph-safe).
struct SeedParams {
  uint32_t n_events, n_layers, n_z, n_phi, n_triplets, cells, cell_bits;
  float min_z, sector_size_z, min_phi, sector_size_phi, min_radius_curvature;
  float min_radius[TRAX_MAX_LAYERS];  // index: layer - 1
  float max_radius[TRAX_MAX_LAYERS];
  uint32_t layer1[TRAX_MAX_TRIPLETS], layer2[TRAX_MAX_TRIPLETS], layer3[TRAX_MAX_TRIPLETS];
  float dtheta_cut[TRAX_MAX_TRIPLETS], sigma_z[TRAX_MAX_TRIPLETS];
  float dphi_cut[TRAX_MAX_TRIPLETS], sigma_phi[TRAX_MAX_TRIPLETS];
  float tip_cut[TRAX_MAX_TRIPLETS], z0[TRAX_MAX_TRIPLETS];
  float pair_sigma_phi[TRAX_MAX_TRIPLETS];  // pairSpreadPhi bits read as float
};

// Sorted (grid-ordered) hit columns, as left behind by the grid build.
struct SortedHits {
  const float* __restrict__ x;
  const float* __restrict__ y;
  const float* __restrict__ z;
  const float* __restrict__ phi;  // atan2(y, x), computed once by the grid build
  const uint32_t* __restrict__ layer;
  const uint32_t* __restrict__ evt;
};

// Decoupled look-back control block shared by the scan and the filter
// compaction.  `ticket` hands out tiles in launch order, `done` counts
// finished blocks, `epoch` tags status words so that no clearing pass is
// needed between launches: the last block of a launch resets the counters
// and advances the epoch.
struct ScanCtrl {
  uint32_t ticket;
  uint32_t done;
  uint32_t epoch;
  uint32_t pad;
};

#define TRAX_AGENT __HIP_MEMORY_SCOPE_AGENT

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t t = __shfl_up(v, d, 64);
    if (lane >= (uint32_t)d) v += t;
  }
  return v;
}

__device__ __forceinline__ uint32_t wave_reduce_sum(uint32_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// Index k of the segment containing v: the last k with offs[k] <= v, for a
// non-decreasing offs[0..n) with offs[0] <= v.
__device__ __forceinline__ uint32_t segment_of(const uint32_t* __restrict__ offs, uint32_t n,
                                               uint32_t v) {
  uint32_t lo = 0, hi = n;  // answer in [lo, hi)
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (offs[mid] <= v)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Status word of the look-back: hi 32 bits = (epoch << 2) | state,
// state 1 = tile aggregate, 2 = inclusive prefix; lo 32 bits = value.
__device__ __forceinline__ void lookback_publish(unsigned long long* status, uint32_t tile,
                                                 uint32_t epoch, uint32_t state, uint32_t value) {
  unsigned long long w = ((unsigned long long)((epoch << 2) | state) << 32) | value;
  __hip_atomic_store(&status[tile], w, __ATOMIC_RELEASE, TRAX_AGENT);
}

// Exclusive prefix of `tile` (tile > 0), computed by one full wave.
__device__ __forceinline__ uint32_t lookback_wave(const unsigned long long* status, uint32_t tile,
                                                  uint32_t epoch) {
  const uint32_t lane = lane_id();
  uint32_t excl = 0;
  int64_t pred = (int64_t)tile - 1;
  while (true) {
    const int64_t idx = pred - (int64_t)lane;
    uint32_t state, value;
    uint32_t first_p;
    while (true) {
      if (idx >= 0) {
        unsigned long long w = __hip_atomic_load(&status[idx], __ATOMIC_ACQUIRE, TRAX_AGENT);
        uint32_t hi = (uint32_t)(w >> 32);
        state = ((hi >> 2) == epoch) ? (hi & 3u) : 0u;
        value = (uint32_t)w;
      } else {
        state = 2u;  // before tile 0: an inclusive prefix of 0
        value = 0u;
      }
      uint64_t pmask = __ballot(state == 2u);
      first_p = pmask ? (uint32_t)__builtin_ctzll(pmask) : 64u;
      uint64_t upto = first_p >= 63u ? ~0ull : ((2ull << first_p) - 1ull);
      uint64_t notready = __ballot(state == 0u);
      if ((notready & upto) == 0ull) break;
      __builtin_amdgcn_s_sleep(1);
    }
    excl += wave_reduce_sum(lane <= first_p ? value : 0u);
    if (first_p < 64u) break;
    pred -= 64;
  }
  return excl;
}

// Grid cell of a hit, GridBuilder.h:80-84 (clamped to the padding row/column).
__device__ __forceinline__ uint32_t grid_cell(const SeedParams& sp, float x, float y, float z,
                                              float* phi_out) {
  uint32_t zs = trax_f2u(__builtin_floorf((z - sp.min_z) / sp.sector_size_z));
  float phi = trax_atan2f(y, x);
  uint32_t ps = trax_f2u(__builtin_floorf((phi - sp.min_phi) / sp.sector_size_phi));
  zs = zs > sp.n_z ? sp.n_z : zs;
  ps = ps > sp.n_phi ? sp.n_phi : ps;
  *phi_out = phi;
  return zs * (sp.n_phi + 1) + ps;
}

}  // namespace trax
