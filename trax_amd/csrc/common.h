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
#define TRAX_FLAG_CONNECT_TABLE 0x10u

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
  uint32_t spread_z[TRAX_MAX_TRIPLETS], spread_phi[TRAX_MAX_TRIPLETS];  // sector pairs
  // acos(radius / (2 minRadiusCurvature)) per layer for the min and max
  // radius: window terms that depend on the layer only, evaluated once on
  // the host with the same FP32 definition the kernels use
  float acos_min_r[TRAX_MAX_LAYERS], acos_max_r[TRAX_MAX_LAYERS];
  // beam-spot pair work items per outer hit, 1 or 2 (runs of z rows)
  uint32_t pair_parts;
};

// Sorted (grid-ordered) hit columns, as left behind by the grid build.
struct SortedHits {
  const float* __restrict__ x;
  const float* __restrict__ y;
  const float* __restrict__ z;
  const float* __restrict__ phi;  // atan2(y, x), computed once by the grid build
  const uint32_t* __restrict__ layer;
  const uint32_t* __restrict__ evt;
  const float2* __restrict__ zphi;  // (z, phi) per hit, for the window walks
  const float4* __restrict__ xyzphi;  // (x, y, z, phi) per hit, one 16-byte load
};

// (z, phi) of sorted hit idx, addressed with a 32-bit byte offset (the C ABI
// caps hit counts below 2^29) so the load takes the scalar-base form.
__device__ __forceinline__ float2 load_zphi(const float2* __restrict__ zphi, uint32_t idx) {
  return *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(zphi) + (idx << 3));
}

// (x, y, z, phi) of sorted hit idx, the same addressing.
__device__ __forceinline__ float4 load_xyzphi(const float4* __restrict__ xyzphi, uint32_t idx) {
  return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(xyzphi) + (idx << 4));
}

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

// Inclusive prefix sum over the wave's 64 lanes (all lanes active), with
// DPP moves instead of LDS permutes: a Kogge-Stone scan inside each row of
// 16 lanes (row_shr 1, 2, 4, 8; lanes shifted in from outside the row read
// 0), then rows 1 and 3 add lane 15 of the row before (row_bcast:15) and
// rows 2 and 3 add lane 31 (row_bcast:31); rows a move does not write keep 0.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true);  // row_shr:1
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true);  // row_shr:2
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true);  // row_shr:4
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true);  // row_shr:8
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return v;
}

// Lane 63's value, read into a scalar register (the wave's total after an
// inclusive scan).
__device__ __forceinline__ uint32_t wave_last(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// Sum over the wave's 64 lanes (all lanes active), the same in every lane.
__device__ __forceinline__ uint32_t wave_reduce_sum(uint32_t v) { return wave_last(wave_inclusive_scan(v)); }

// Max (MAX true) or min over the wave's 64 lanes (all lanes active), the
// same in every lane: the scan's moves, with lanes shifted in from outside
// a row (and rows a broadcast does not write) reading the identity.
template <bool MAX>
__device__ __forceinline__ uint32_t wave_reduce_minmax(uint32_t v) {
  const int id = MAX ? 0 : -1;
  auto op = [](uint32_t a, int b) { const uint32_t u = (uint32_t)b; return MAX ? (u > a ? u : a) : (u < a ? u : a); };
  v = op(v, __builtin_amdgcn_update_dpp(id, (int)v, 0x111, 0xf, 0xf, false));  // row_shr:1
  v = op(v, __builtin_amdgcn_update_dpp(id, (int)v, 0x112, 0xf, 0xf, false));  // row_shr:2
  v = op(v, __builtin_amdgcn_update_dpp(id, (int)v, 0x114, 0xf, 0xf, false));  // row_shr:4
  v = op(v, __builtin_amdgcn_update_dpp(id, (int)v, 0x118, 0xf, 0xf, false));  // row_shr:8
  v = op(v, __builtin_amdgcn_update_dpp(id, (int)v, 0x142, 0xa, 0xf, false));  // row_bcast:15
  v = op(v, __builtin_amdgcn_update_dpp(id, (int)v, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return wave_last(v);
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

// Block-wide copy of a small offsets table into LDS (n <= cap), for
// segment_of lookups without dependent global loads.  Call from all threads.
// Larger tables stay in global memory (L2-resident): every block copying a
// table of thousands of entries, and the LDS it holds, cost more than the
// binary search's dependent loads (prediction walk over 2000 ttbar events:
// 121 us staged at 8001 entries, 41 us for 2048 events searched in place).
constexpr uint32_t kLdsOffsets = 1024;

// Dynamic LDS for the per-(event, layer triplet) offsets table of a launch.
inline size_t offsets_lds_bytes(const SeedParams& sp) {
  const uint32_t n = sp.n_events * sp.n_triplets + 1;
  return n <= kLdsOffsets ? (size_t)n * 4u : 0u;
}

__device__ __forceinline__ const uint32_t* stage_offsets(const uint32_t* __restrict__ offs,
                                                         uint32_t n, uint32_t* s_offs) {
  if (n > kLdsOffsets) return offs;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) s_offs[i] = offs[i];
  __syncthreads();
  return s_offs;
}

// Status word of the look-back: hi 32 bits = (epoch << 2) | state,
// state 1 = tile aggregate, 2 = inclusive prefix; lo 32 bits = value.
// The word carries its own payload, so relaxed agent-scope atomics suffice:
// they are coherent across XCDs (sc1) without the L2 write-back / invalidate
// that release / acquire ordering adds to every access on gfx950.
__device__ __forceinline__ void lookback_publish(unsigned long long* status, uint32_t tile,
                                                 uint32_t epoch, uint32_t state, uint32_t value) {
  unsigned long long w = ((unsigned long long)((epoch << 2) | state) << 32) | value;
  __hip_atomic_store(&status[tile], w, __ATOMIC_RELAXED, TRAX_AGENT);
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
        unsigned long long w = __hip_atomic_load(&status[idx], __ATOMIC_RELAXED, TRAX_AGENT);
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

// One tile of a single-pass compaction (256 threads = 4 waves): block-wide
// exclusive scan of one count per thread, then the tile's global base by
// decoupled look-back.  Returns the thread's global exclusive position; all
// threads get the tile total and base.  Contains two block barriers.
struct TileScan {
  uint32_t pos, total, base;
};

__device__ __forceinline__ TileScan tile_scan_lookback(uint32_t v, uint32_t tile, uint32_t epoch,
                                                       unsigned long long* status,
                                                       uint32_t* s_wave, uint32_t* s_excl) {
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  const uint32_t incl = wave_inclusive_scan(v);
  if (lane == 63) s_wave[wave] = incl;
  __syncthreads();
  uint32_t wp = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const uint32_t c = s_wave[w];
    wp += (w < (int)wave) ? c : 0u;
    tot += c;
  }
  if (wave == 0) {
    uint32_t excl = 0;
    if (tile == 0) {
      if (lane == 0) lookback_publish(status, tile, epoch, 2u, tot);
    } else {
      if (lane == 0) lookback_publish(status, tile, epoch, 1u, tot);
      excl = lookback_wave(status, tile, epoch);
      if (lane == 0) lookback_publish(status, tile, epoch, 2u, excl + tot);
    }
    if (lane == 0) *s_excl = excl;
  }
  __syncthreads();
  TileScan r;
  r.base = *s_excl;
  r.total = tot;
  r.pos = r.base + wp + incl - v;
  return r;
}

// Makes a wave's LDS writes visible to the other lanes of the same wave.
// Blocks go to the eight XCDs round-robin (block b on XCD b % 8), each XCD
// with its own L2.  xcd_share splits n work items into eight contiguous
// shares, share x for the blocks of XCD x, which stride through it: the
// items a round of an XCD's waves takes are neighbours (one event's hits and
// grid cells stay in that XCD's L2), and the shares stay balanced however
// far the capacity-sized grid exceeds the work.  `unit` is the block's
// first item of a round (items per block and round: `per_block`).  For a
// grid that is not a multiple of 8 it degenerates to the plain grid stride.
struct XcdShare {
  uint32_t begin, end, stride;
};
__device__ __forceinline__ XcdShare xcd_share(uint32_t n, uint32_t per_block, uint32_t local) {
  const uint32_t nb = gridDim.x;
  if (nb & 7u) return XcdShare{blockIdx.x * per_block + local, n, nb * per_block};
  const uint32_t x = blockIdx.x & 7u;
  const uint32_t lo = (uint32_t)((uint64_t)n * x / 8u), hi = (uint32_t)((uint64_t)n * (x + 1u) / 8u);
  return XcdShare{lo + (blockIdx.x >> 3) * per_block + local, hi, (nb >> 3) * per_block};
}

// Block count rounded up to a multiple of 8 (see xcd_share).
inline uint32_t xcd_grid(uint32_t blocks) { return blocks > 8u ? (blocks + 7u) & ~7u : blocks; }

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Ticket for the next tile of a persistent launch (block-uniform).
__device__ __forceinline__ uint32_t next_tile(ScanCtrl* ctrl, uint32_t* s_tile) {
  if (threadIdx.x == 0) *s_tile = __hip_atomic_fetch_add(&ctrl->ticket, 1u, __ATOMIC_RELAXED, TRAX_AGENT);
  __syncthreads();
  return *s_tile;
}

__device__ __forceinline__ uint32_t launch_epoch(ScanCtrl* ctrl) {
  return __hip_atomic_load(&ctrl->epoch, __ATOMIC_RELAXED, TRAX_AGENT) & 0x3FFFFFFFu;
}

// End of a persistent launch: the last block resets the tickets and advances
// the epoch for the next launch.  Call from all threads after the tile loop.
__device__ __forceinline__ void launch_done(ScanCtrl* ctrl, uint32_t epoch) {
  if (threadIdx.x == 0) {
    const uint32_t d = __hip_atomic_fetch_add(&ctrl->done, 1u, __ATOMIC_RELAXED, TRAX_AGENT);
    if (d == gridDim.x - 1) {
      __hip_atomic_store(&ctrl->done, 0u, __ATOMIC_RELAXED, TRAX_AGENT);
      __hip_atomic_store(&ctrl->ticket, 0u, __ATOMIC_RELAXED, TRAX_AGENT);
      __hip_atomic_store(&ctrl->epoch, (epoch + 1u) & 0x3FFFFFFFu, __ATOMIC_RELAXED, TRAX_AGENT);
    }
  }
}

// Write out[k] = pos for every segment k that starts at item `item` (empty
// segments share their start with the next one).
__device__ __forceinline__ void write_segment_starts(const uint32_t* offs, uint32_t K, uint32_t k,
                                                     uint32_t item, uint32_t pos, uint32_t* out) {
  (void)K;
  while (offs[k] == item) {
    out[k] = pos;
    if (k == 0) break;
    --k;
  }
}

// Test hits [a, b) in index order against a (z, phi) window; loads are
// issued four at a time so their latencies overlap.  f(idx) is called for
// every hit inside the window, in order, and returns true to stop the walk
// (the prediction store's slot budget).  Returns true if stopped.
template <typename InWindow, typename F>
__device__ __forceinline__ bool scan_range(uint32_t a, uint32_t b, const float2* __restrict__ zphi,
                                           InWindow in_window, F&& f) {
  uint32_t idx = a;
  for (; idx + 4 <= b; idx += 4) {
    float2 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = load_zphi(zphi, idx + u);
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (in_window(v[u].x, v[u].y) && f(idx + u)) return true;
  }
  if (idx < b) {  // the last one to three hits, loaded together
    float2 v[3];
#pragma unroll
    for (uint32_t u = 0; u < 3; ++u) v[u] = load_zphi(zphi, idx + u < b ? idx + u : idx);
#pragma unroll
    for (uint32_t u = 0; u < 3; ++u)
      if (idx + u < b && in_window(v[u].x, v[u].y) && f(idx + u)) return true;
  }
  return false;
}

// Grid cell of a hit with phi = atan2(y, x) already known, GridBuilder.h:80-84
// (clamped to the padding row/column).
__device__ __forceinline__ uint32_t grid_cell_zphi(const SeedParams& sp, float z, float phi) {
  uint32_t zs = trax_f2u(__builtin_floorf((z - sp.min_z) / sp.sector_size_z));
  uint32_t ps = trax_f2u(__builtin_floorf((phi - sp.min_phi) / sp.sector_size_phi));
  zs = zs > sp.n_z ? sp.n_z : zs;
  ps = ps > sp.n_phi ? sp.n_phi : ps;
  return zs * (sp.n_phi + 1) + ps;
}

// Grid cell of a hit, GridBuilder.h:80-84; phi = atan2(y, x) goes to phi_out.
__device__ __forceinline__ uint32_t grid_cell(const SeedParams& sp, float x, float y, float z,
                                              float* phi_out) {
  const float phi = trax_atan2f(y, x);
  *phi_out = phi;
  return grid_cell_zphi(sp, z, phi);
}

}  // namespace trax
