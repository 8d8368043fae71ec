// Exclusive uint32 prefix sum.
//
// Replaces the reference's three-kernel recursive Blelloch scan
// (src/algorithms/PrefixSum.h:55-245, PrefixSum.cpp:60-126), which reads and
// writes every element about twice per level.  Below 2^21 elements: a single
// pass, where each 256-thread block owns one tile of SCAN_TILE elements,
// loaded and stored as coalesced uint4, reduces it with wave64 shuffles,
// publishes the tile aggregate, and resolves its exclusive prefix by looking
// back over the status words of its predecessors (one wave, 64 predecessors
// per probe); one read and one write per element (8 B).  From 2^21 on: a
// two-pass form (below).  The result is the exclusive scan mod 2^32,
// identical to the reference's for every n (the reference's float launch
// sizing, PrefixSum.cpp:37,66, can drop tail elements above 2^24; this one
// does not).
#include "common.h"
#include "kernels.h"

namespace trax {

__device__ __forceinline__ uint64_t scan_length(uint64_t n_max, const uint32_t* d_n, uint32_t n_add) {
  if (!d_n) return n_max;
  const uint64_t dn = (uint64_t)d_n[0] + n_add;
  return dn < n_max ? dn : n_max;
}

__device__ __forceinline__ uint4 load4(const uint32_t* in, uint64_t idx, uint64_t n) {
  uint4 v;
  if (idx + 4 <= n) {
    v = *reinterpret_cast<const uint4*>(in + idx);
  } else {
    v.x = idx < n ? in[idx] : 0u;
    v.y = idx + 1 < n ? in[idx + 1] : 0u;
    v.z = idx + 2 < n ? in[idx + 2] : 0u;
    v.w = idx + 3 < n ? in[idx + 3] : 0u;
  }
  return v;
}

template <int SUB, bool PERSISTENT>
__global__ __launch_bounds__(256) void k_scan_u32(const uint32_t* in,
                                                  uint32_t* out, uint64_t n_max,
                                                  const uint32_t* __restrict__ d_n, uint32_t n_add,
                                                  ScanCtrl* ctrl, unsigned long long* status) {
  constexpr uint32_t TILE = 256u * 4u * SUB;
  __shared__ uint32_t s_tile, s_epoch, s_excl;
  __shared__ uint32_t s_wave[SUB][4];
  const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  if (tid == 0) s_epoch = __hip_atomic_load(&ctrl->epoch, __ATOMIC_RELAXED, TRAX_AGENT) & 0x3FFFFFFFu;
  uint64_t n = n_max;
  if (d_n) {
    const uint64_t dn = (uint64_t)d_n[0] + n_add;
    n = dn < n_max ? dn : n_max;
  }
  const uint64_t ntiles = (n + TILE - 1) / TILE;
  // Persistent: each block takes tiles in ticket order until none are left.
  while (true) {
    if (tid == 0) s_tile = __hip_atomic_fetch_add(&ctrl->ticket, 1u, __ATOMIC_RELAXED, TRAX_AGENT);
    __syncthreads();
    const uint32_t tile = s_tile, epoch = s_epoch;
    if (tile >= ntiles) break;
    const uint64_t base = (uint64_t)tile * TILE;
    const bool whole = base + TILE <= n;
    uint4 v[SUB];
    uint32_t t[SUB], incl[SUB];
    if (whole) {  // every load issued before the first use
#pragma unroll
      for (int s = 0; s < SUB; ++s) v[s] = *reinterpret_cast<const uint4*>(in + base + (uint64_t)s * 1024u + tid * 4u);
    } else {
#pragma unroll
      for (int s = 0; s < SUB; ++s) v[s] = load4(in, base + (uint64_t)s * 1024u + tid * 4u, n);
    }
#pragma unroll
    for (int s = 0; s < SUB; ++s) {
      t[s] = v[s].x + v[s].y + v[s].z + v[s].w;
      incl[s] = wave_inclusive_scan(t[s]);
      if (lane == 63) s_wave[s][wave] = incl[s];
    }
    __syncthreads();
    uint32_t running = 0, thread_excl[SUB];
#pragma unroll
    for (int s = 0; s < SUB; ++s) {
      uint32_t wp = 0, tot = 0;
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        const uint32_t c = s_wave[s][w];
        wp += (w < (int)wave) ? c : 0u;
        tot += c;
      }
      thread_excl[s] = running + wp + incl[s] - t[s];
      running += tot;
    }
    const uint32_t aggregate = running;
    if (wave == 0) {
      uint32_t excl = 0;
      if (tile == 0) {
        if (lane == 0) lookback_publish(status, tile, epoch, 2u, aggregate);
      } else {
        if (lane == 0) lookback_publish(status, tile, epoch, 1u, aggregate);
        excl = lookback_wave(status, tile, epoch);
        if (lane == 0) lookback_publish(status, tile, epoch, 2u, excl + aggregate);
      }
      if (lane == 0) s_excl = excl;
    }
    __syncthreads();
    const uint32_t excl_tile = s_excl;
#pragma unroll
    for (int s = 0; s < SUB; ++s) {
      const uint64_t idx = base + (uint64_t)s * 1024u + tid * 4u;
      uint4 o;
      o.x = excl_tile + thread_excl[s];
      o.y = o.x + v[s].x;
      o.z = o.y + v[s].y;
      o.w = o.z + v[s].z;
      if (whole || idx + 4 <= n) {
        *reinterpret_cast<uint4*>(out + idx) = o;
      } else {
        if (idx < n) out[idx] = o.x;
        if (idx + 1 < n) out[idx + 1] = o.y;
        if (idx + 2 < n) out[idx + 2] = o.z;
        if (idx + 3 < n) out[idx + 3] = o.w;
      }
    }
    if (!PERSISTENT) break;
    __syncthreads();  // s_tile, s_wave and s_excl are reused by the next tile
  }
  if (tid == 0) {
    const uint32_t epoch = s_epoch;
    const uint32_t d = __hip_atomic_fetch_add(&ctrl->done, 1u, __ATOMIC_RELAXED, TRAX_AGENT);
    if (d == gridDim.x - 1) {  // last block of the launch: reset for the next one
      __hip_atomic_store(&ctrl->done, 0u, __ATOMIC_RELAXED, TRAX_AGENT);
      __hip_atomic_store(&ctrl->ticket, 0u, __ATOMIC_RELAXED, TRAX_AGENT);
      __hip_atomic_store(&ctrl->epoch, (epoch + 1u) & 0x3FFFFFFFu, __ATOMIC_RELAXED, TRAX_AGENT);
    }
  }
}

// Two-pass form for large arrays (no atomics, no look-back): the reduce pass
// writes each chunk's sum; the scan pass re-reads each chunk, adds up the
// sums of the chunks before it (every block of an XCD reads them from that
// XCD's L2) and writes the chunk's exclusive scan.  12 B of HBM traffic per
// element instead of 8, but every block runs independently at streaming
// bandwidth, and the scan pass visits the chunks in reverse order so that
// the ones the reduce pass read last are still in the Infinity Cache.  The
// sums live in the look-back status words, one 64-bit word each with a zero
// upper half: a later single-pass launch then reads them as "not published"
// (state 0) and never as a tile of its own epoch, whatever the sum.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
constexpr uint32_t kChunkSub = 16;                     // uint4 per thread
constexpr uint32_t kChunk = 256u * 4u * kChunkSub;    // elements per chunk

__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t* in, uint64_t n_max,
                                                      const uint32_t* __restrict__ d_n, uint32_t n_add,
                                                      unsigned long long* __restrict__ partials) {
  __shared__ uint32_t s_wave[4];
  const uint64_t n = scan_length(n_max, d_n, n_add);
  const uint32_t chunk = blockIdx.x;
  const uint64_t base = (uint64_t)chunk * kChunk;
  if (base >= n) return;
  const uint32_t tid = threadIdx.x;
  uint32_t t = 0;
  if (base + kChunk <= n) {  // every load issued before the first use
#pragma unroll
    for (uint32_t s = 0; s < kChunkSub; ++s) {
      const uint4 v = *reinterpret_cast<const uint4*>(in + base + s * 1024u + tid * 4u);
      t += v.x + v.y + v.z + v.w;
    }
  } else {
#pragma unroll
    for (uint32_t s = 0; s < kChunkSub; ++s) {
      const uint4 v = load4(in, base + s * 1024u + tid * 4u, n);
      t += v.x + v.y + v.z + v.w;
    }
  }
  t = wave_reduce_sum(t);
  if (lane_id() == 0) s_wave[tid >> 6] = t;
  __syncthreads();
  if (tid == 0) partials[chunk] = (unsigned long long)(s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3]);
}

// NT: non-temporal stores, for a separate output (the input's lines then
// stay in the Infinity Cache for the blocks still to come); a scan in place
// writes the lines it has just read with ordinary stores (measured 121
// against 128 us at 2^26 with non-temporal ones).
template <bool NT>
__global__ __launch_bounds__(256) void k_scan_down(const uint32_t* in, uint32_t* out, uint64_t n_max,
                                                    const uint32_t* __restrict__ d_n, uint32_t n_add,
                                                    const unsigned long long* __restrict__ partials) {
  __shared__ uint32_t s_wave[kChunkSub][4];
  __shared__ uint32_t s_pre[4];
  const uint64_t n = scan_length(n_max, d_n, n_add);
  const uint32_t chunk = gridDim.x - 1u - blockIdx.x;
  const uint64_t base = (uint64_t)chunk * kChunk;
  if (base >= n) return;
  const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  const bool whole = base + kChunk <= n;
  uint4 v[kChunkSub];
  if (whole) {
#pragma unroll
    for (uint32_t s = 0; s < kChunkSub; ++s) v[s] = *reinterpret_cast<const uint4*>(in + base + s * 1024u + tid * 4u);
  } else {
#pragma unroll
    for (uint32_t s = 0; s < kChunkSub; ++s) v[s] = load4(in, base + s * 1024u + tid * 4u, n);
  }
  // the sums of chunks 0 .. chunk-1, 4096 per round
  uint32_t pre = 0;
  for (uint32_t b0 = 0; b0 < chunk; b0 += 4096u) {
#pragma unroll
    for (uint32_t r = 0; r < 16u; ++r) {
      const uint32_t i = b0 + r * 256u + tid;
      pre += i < chunk ? (uint32_t)partials[i] : 0u;
    }
  }
  pre = wave_reduce_sum(pre);
  if (lane == 0) s_pre[wave] = pre;
  uint32_t t[kChunkSub], incl[kChunkSub];
#pragma unroll
  for (uint32_t s = 0; s < kChunkSub; ++s) {
    t[s] = v[s].x + v[s].y + v[s].z + v[s].w;
    incl[s] = wave_inclusive_scan(t[s]);
    if (lane == 63) s_wave[s][wave] = incl[s];
  }
  __syncthreads();
  uint32_t running = s_pre[0] + s_pre[1] + s_pre[2] + s_pre[3];
#pragma unroll
  for (uint32_t s = 0; s < kChunkSub; ++s) {
    uint32_t wp = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      const uint32_t c = s_wave[s][w];
      wp += (w < (int)wave) ? c : 0u;
      tot += c;
    }
    const uint64_t idx = base + s * 1024u + tid * 4u;
    uint4 o;
    o.x = running + wp + incl[s] - t[s];
    o.y = o.x + v[s].x;
    o.z = o.y + v[s].y;
    o.w = o.z + v[s].z;
    if (whole || idx + 4 <= n) {
      u32x4 ov = {o.x, o.y, o.z, o.w};
      if (NT)
        __builtin_nontemporal_store(ov, reinterpret_cast<u32x4*>(out + idx));
      else
        *reinterpret_cast<u32x4*>(out + idx) = ov;
    } else {
      if (idx < n) out[idx] = o.x;
      if (idx + 1 < n) out[idx + 1] = o.y;
      if (idx + 2 < n) out[idx + 2] = o.z;
      if (idx + 3 < n) out[idx + 3] = o.w;
    }
    running += tot;
  }
}

// partials: ceil(n_max / kChunk) 64-bit words of scratch.
static hipError_t scan_two_pass(const uint32_t* in, uint32_t* out, uint64_t n_max, const uint32_t* d_n,
                                uint32_t n_add, unsigned long long* partials, hipStream_t stream) {
  const uint64_t blocks = (n_max + kChunk - 1) / kChunk;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_scan_reduce, dim3((uint32_t)blocks), dim3(256), 0, stream, in, n_max, d_n, n_add, partials);
  if (in == out)
    hipLaunchKernelGGL(k_scan_down<false>, dim3((uint32_t)blocks), dim3(256), 0, stream, in, out, n_max, d_n, n_add,
                       partials);
  else
    hipLaunchKernelGGL(k_scan_down<true>, dim3((uint32_t)blocks), dim3(256), 0, stream, in, out, n_max, d_n, n_add,
                       partials);
  return hipGetLastError();
}

static_assert(kChunk == kScanChunk, "two-pass chunk");

uint64_t scan_tile_elems() { return 256ull * 4ull * 8ull; }

// Segmented scan, step 2: the global exclusive scan's value at each segment
// start (the sum of everything before the segment), one 64-bit word per
// segment with a zero upper half (see the partials above).
__global__ __launch_bounds__(256) void k_segment_bases(const uint32_t* __restrict__ data, uint64_t n,
                                                       const uint32_t* __restrict__ off, uint32_t n_segments,
                                                       unsigned long long* __restrict__ bases) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n_segments; k += gridDim.x * blockDim.x) {
    const uint32_t o = off[k];
    bases[k] = o < n ? (unsigned long long)data[o] : 0ull;
  }
}

// Step 3: every element minus its segment's base.  A thread takes four
// consecutive elements: one binary search for the first, then the segment
// index only moves forward (empty segments share their start with the next
// one; the last segment starting at or before i holds i).
__global__ __launch_bounds__(256) void k_segment_fix(uint32_t* __restrict__ data, uint64_t n,
                                                     const uint32_t* __restrict__ off, uint32_t n_segments,
                                                     const unsigned long long* __restrict__ bases) {
  for (uint64_t i0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4u; i0 < n;
       i0 += (uint64_t)gridDim.x * blockDim.x * 4u) {
    uint32_t k = segment_of(off, n_segments, (uint32_t)i0);
    for (uint64_t i = i0; i < i0 + 4u && i < n; ++i) {
      while (k + 1u < n_segments && off[k + 1u] <= i) ++k;
      data[i] -= (uint32_t)bases[k];
    }
  }
}

__global__ void k_scan_ctrl_clear(ScanCtrl* ctrl) {
  if (threadIdx.x == 0) {
    ctrl->ticket = 0u;
    ctrl->done = 0u;
    ctrl->epoch = 0u;
    ctrl->pad = 0u;
  }
}

hipError_t launch_segment_bases(const uint32_t* data, uint64_t n, const uint32_t* off, uint32_t n_segments,
                                unsigned long long* bases, hipStream_t stream) {
  uint32_t blocks = (n_segments + 255u) / 256u;
  blocks = blocks == 0 ? 1u : (blocks > 4096u ? 4096u : blocks);
  hipLaunchKernelGGL(k_segment_bases, dim3(blocks), dim3(256), 0, stream, data, n, off, n_segments, bases);
  return hipGetLastError();
}

hipError_t launch_segment_fix(uint32_t* data, uint64_t n, const uint32_t* off, uint32_t n_segments,
                              const unsigned long long* bases, hipStream_t stream) {
  uint64_t b = (n + 1023u) / 1024u;
  const uint32_t blocks = (uint32_t)(b == 0 ? 1 : (b > 16384u ? 16384u : b));
  hipLaunchKernelGGL(k_segment_fix, dim3(blocks), dim3(256), 0, stream, data, n, off, n_segments, bases);
  return hipGetLastError();
}

hipError_t launch_scan_ctrl_clear(ScanCtrl* ctrl, hipStream_t stream) {
  hipLaunchKernelGGL(k_scan_ctrl_clear, dim3(1), dim3(64), 0, stream, ctrl);
  return hipGetLastError();
}

// Persistent blocks (256 CUs x 8) take tiles in ticket order; measured on
// MI355X this beats one tile per block by 2.4x (the look-back of a block's
// tile is overlapped by its resident neighbours instead of new dispatches).
template <int SUB>
static hipError_t scan_launch(const uint32_t* in, uint32_t* out, uint64_t n_max, const uint32_t* d_n,
                              uint32_t n_add, ScanCtrl* ctrl, unsigned long long* status,
                              hipStream_t stream) {
  constexpr uint64_t tile = 256ull * 4ull * SUB;
  uint64_t blocks = (n_max + tile - 1) / tile;
  if (blocks == 0) blocks = 1;
  if (blocks > kPersistentBlocks) blocks = kPersistentBlocks;
  hipLaunchKernelGGL((k_scan_u32<SUB, true>), dim3((uint32_t)blocks), dim3(256), 0, stream, in, out,
                     n_max, d_n, n_add, ctrl, status);
  return hipGetLastError();
}

// sub: uint4 loads per thread per tile (tile = 1024 * sub elements).
hipError_t launch_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n_max, const uint32_t* d_n,
                           uint32_t n_add, ScanCtrl* ctrl, unsigned long long* status,
                           hipStream_t stream, int sub, uint64_t status_words) {
  if (n_max >= kReduceScanMin) {
    // the status words hold the chunk sums of the two-pass form
    if ((n_max + kChunk - 1) / kChunk > status_words) return hipErrorInvalidValue;
    return scan_two_pass(in, out, n_max, d_n, n_add, status, stream);
  }
  // single pass: one status word per tile
  if ((n_max + 256ull * 4ull * sub - 1) / (256ull * 4ull * sub) > status_words) return hipErrorInvalidValue;
  switch (sub) {
    case 4: return scan_launch<4>(in, out, n_max, d_n, n_add, ctrl, status, stream);
    case 16: return scan_launch<16>(in, out, n_max, d_n, n_add, ctrl, status, stream);
    default: return scan_launch<8>(in, out, n_max, d_n, n_add, ctrl, status, stream);
  }
}

}  // namespace trax
