// Exclusive uint32 prefix sum: single pass, decoupled look-back.
//
// Replaces the reference's three-kernel recursive Blelloch scan
// (src/algorithms/PrefixSum.h:55-245, PrefixSum.cpp:60-126), which reads and
// writes every element about twice per level.  Here each 256-thread block
// owns one tile of SCAN_TILE elements, loaded and stored as coalesced uint4,
// reduces it with wave64 shuffles, publishes the tile aggregate, and resolves
// its exclusive prefix by looking back over the status words of its
// predecessors (one wave, 64 predecessors per probe).  Traffic is one read
// and one write per element (8 B).  The result is the exclusive scan mod 2^32,
// identical to the reference's for every n (the reference's float launch
// sizing, PrefixSum.cpp:37,66, can drop tail elements above 2^24; this one
// does not).
#include "common.h"
#include "kernels.h"

namespace trax {

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
    uint4 v[SUB];
    uint32_t t[SUB], incl[SUB];
#pragma unroll
    for (int s = 0; s < SUB; ++s) {
      const uint64_t idx = base + (uint64_t)s * 1024u + tid * 4u;
      if (idx + 4 <= n) {
        v[s] = *reinterpret_cast<const uint4*>(in + idx);
      } else {
        v[s].x = idx < n ? in[idx] : 0u;
        v[s].y = idx + 1 < n ? in[idx + 1] : 0u;
        v[s].z = idx + 2 < n ? in[idx + 2] : 0u;
        v[s].w = idx + 3 < n ? in[idx + 3] : 0u;
      }
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
      if (idx + 4 <= n) {
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

// Reduce-then-scan form for large arrays (no atomics, no look-back): pass 1
// reduces each chunk of kChunk elements to a partial sum, pass 2 scans the
// partials in one block, pass 3 re-reads each chunk and writes its
// exclusive scan.  12 B of HBM traffic per element instead of 8, but every
// block runs independently at streaming bandwidth, and pass 3 visits the
// chunks in reverse order so that the ones pass 1 read last are still in
// the Infinity Cache.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
constexpr uint32_t kChunkSub = 16;                     // uint4 per thread
constexpr uint32_t kChunk = 256u * 4u * kChunkSub;    // elements per chunk

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

__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t* in, uint64_t n_max,
                                                     const uint32_t* __restrict__ d_n, uint32_t n_add,
                                                     uint32_t* __restrict__ partials) {
  __shared__ uint32_t s_wave[4];
  const uint64_t n = scan_length(n_max, d_n, n_add);
  const uint64_t base = (uint64_t)blockIdx.x * kChunk;
  if (base >= n) return;
  const uint32_t tid = threadIdx.x;
  uint32_t t = 0;
#pragma unroll
  for (uint32_t s = 0; s < kChunkSub; ++s) {
    const uint4 v = load4(in, base + s * 1024u + tid * 4u, n);
    t += v.x + v.y + v.z + v.w;
  }
  t = wave_reduce_sum(t);
  if (lane_id() == 0) s_wave[tid >> 6] = t;
  __syncthreads();
  if (tid == 0) partials[blockIdx.x] = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
}

// Exclusive scan of the chunk partials in place (one block of 1024 threads).
__global__ __launch_bounds__(1024) void k_scan_partials(uint64_t n_max, const uint32_t* __restrict__ d_n,
                                                        uint32_t n_add, uint32_t* __restrict__ partials) {
  __shared__ uint32_t s_wave[16];
  const uint64_t n = scan_length(n_max, d_n, n_add);
  const uint32_t np = (uint32_t)((n + kChunk - 1) / kChunk);
  const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  uint32_t carry = 0;
  for (uint32_t b0 = 0; b0 < np; b0 += 1024u) {
    const uint32_t i = b0 + tid;
    const uint32_t v = i < np ? partials[i] : 0u;
    const uint32_t incl = wave_inclusive_scan(v);
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t wp = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < 16; ++w) {
      const uint32_t c = s_wave[w];
      wp += (w < (int)wave) ? c : 0u;
      tot += c;
    }
    if (i < np) partials[i] = carry + wp + incl - v;
    carry += tot;
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_scan_down(const uint32_t* in, uint32_t* out, uint64_t n_max,
                                                   const uint32_t* __restrict__ d_n, uint32_t n_add,
                                                   const uint32_t* __restrict__ partials) {
  __shared__ uint32_t s_wave[kChunkSub][4];
  const uint64_t n = scan_length(n_max, d_n, n_add);
  const uint32_t chunk = gridDim.x - 1u - blockIdx.x;
  const uint64_t base = (uint64_t)chunk * kChunk;
  if (base >= n) return;
  const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  uint4 v[kChunkSub];
  uint32_t t[kChunkSub], incl[kChunkSub];
#pragma unroll
  for (uint32_t s = 0; s < kChunkSub; ++s) {
    v[s] = load4(in, base + s * 1024u + tid * 4u, n);
    t[s] = v[s].x + v[s].y + v[s].z + v[s].w;
    incl[s] = wave_inclusive_scan(t[s]);
    if (lane == 63) s_wave[s][wave] = incl[s];
  }
  __syncthreads();
  uint32_t running = partials[chunk];
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
    if (idx + 4 <= n) {
      // streaming store: the result is not read again by this scan
      u32x4 ov = {o.x, o.y, o.z, o.w};
      __builtin_nontemporal_store(ov, reinterpret_cast<u32x4*>(out + idx));
    } else {
      if (idx < n) out[idx] = o.x;
      if (idx + 1 < n) out[idx + 1] = o.y;
      if (idx + 2 < n) out[idx + 2] = o.z;
      if (idx + 3 < n) out[idx + 3] = o.w;
    }
    running += tot;
  }
}

// partials: ceil(n_max / kChunk) uint32 of scratch.
static hipError_t scan_reduce_then_scan(const uint32_t* in, uint32_t* out, uint64_t n_max,
                                        const uint32_t* d_n, uint32_t n_add, uint32_t* partials,
                                        hipStream_t stream) {
  const uint64_t blocks = (n_max + kChunk - 1) / kChunk;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_scan_reduce, dim3((uint32_t)blocks), dim3(256), 0, stream, in, n_max, d_n, n_add,
                     partials);
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, stream, n_max, d_n, n_add, partials);
  hipLaunchKernelGGL(k_scan_down, dim3((uint32_t)blocks), dim3(256), 0, stream, in, out, n_max, d_n, n_add,
                     partials);
  return hipGetLastError();
}

uint64_t scan_tile_elems() { return 256ull * 4ull * 8ull; }

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
                           hipStream_t stream, int sub) {
  if (n_max >= kReduceScanMin) {
    // the status words double as the partials of the three-pass form
    if ((n_max + kChunk - 1) / kChunk > 2 * kStatusWords) return hipErrorInvalidValue;
    return scan_reduce_then_scan(in, out, n_max, d_n, n_add, reinterpret_cast<uint32_t*>(status), stream);
  }
  switch (sub) {
    case 4: return scan_launch<4>(in, out, n_max, d_n, n_add, ctrl, status, stream);
    case 16: return scan_launch<16>(in, out, n_max, d_n, n_add, ctrl, status, stream);
    default: return scan_launch<8>(in, out, n_max, d_n, n_add, ctrl, status, stream);
  }
}

}  // namespace trax
