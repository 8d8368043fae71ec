// Grid build: per-(event, layer) (z, phi) counting sort of the hit columns.
//
// Reference: src/algorithms/GridBuilder.h:48-427 and GridBuilder.cpp:6-160 run
// three passes with a host synchronisation and a full grid download between
// them: gridCount (LDS histogram, global atomics), PrefixSum over the whole
// E*L*C grid, and gridStore (scatter with atomic_inc ranks, so the order
// inside a cell is non-deterministic).
//
// Here one 256-thread block owns one (event, layer) segment and does all of
// it in LDS: histogram, in-block exclusive scan of the C cells, grid write,
// and a *stable* scatter.  The global scan of the reference is not needed:
// hits are stored event-major and layer-ascending, so the cells of all
// earlier segments hold exactly the hits before this segment, and the
// segment's global base is its hit offset.  The stable rank inside a cell is
// the count of earlier same-cell hits: across 256-hit chunks it comes from
// running per-cell counters in LDS, inside a wave from a 64-lane match built
// with one ballot per cell-id bit, and across the four waves of a chunk from
// taking the waves in order.  The block also writes the input index
// (`orig_index`) and phi = atan2(y, x) of every sorted hit, which the pair
// and prediction stages read instead of recomputing atan2.
#include "common.h"
#include "kernels.h"

namespace trax {

struct GridArgs {
  const float* __restrict__ ix;
  const float* __restrict__ iy;
  const float* __restrict__ iz;
  const uint32_t* __restrict__ ilayer;
  const uint32_t* __restrict__ idet;
  const uint32_t* __restrict__ ihid;
  const uint32_t* __restrict__ ievt;
  float* __restrict__ ox;
  float* __restrict__ oy;
  float* __restrict__ oz;
  uint32_t* __restrict__ olayer;
  uint32_t* __restrict__ odet;
  uint32_t* __restrict__ ohid;
  uint32_t* __restrict__ oevt;
  uint32_t* __restrict__ oorig;
  float* __restrict__ ophi;
  float2* __restrict__ ozphi;
  const uint32_t* __restrict__ event_offset;
  const uint32_t* __restrict__ layer_nhits;
  const uint32_t* __restrict__ layer_offset;
  uint32_t n_hits;
  uint32_t* __restrict__ grid;
  uint32_t* __restrict__ counters;  // global fallback: E*L*C running counters (else unused)
  uint32_t* __restrict__ flags;
};

// In-block exclusive scan of cnt[0..C) (64*NW threads, 4 consecutive cells
// per thread per step).  `cnt` is LDS or global memory.
template <int NW, typename Ptr>
__device__ __forceinline__ void block_exclusive_scan_cells(Ptr cnt, uint32_t C, uint32_t* s_wave) {
  const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  uint32_t carry = 0;
  for (uint32_t base = 0; base < C; base += 256u * NW) {
    uint32_t c0 = base + tid * 4u;
    uint32_t v0 = c0 < C ? cnt[c0] : 0u;
    uint32_t v1 = c0 + 1 < C ? cnt[c0 + 1] : 0u;
    uint32_t v2 = c0 + 2 < C ? cnt[c0 + 2] : 0u;
    uint32_t v3 = c0 + 3 < C ? cnt[c0 + 3] : 0u;
    uint32_t t = v0 + v1 + v2 + v3;
    uint32_t incl = wave_inclusive_scan(t);
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t wp = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      uint32_t c = s_wave[w];
      wp += (w < (int)wave) ? c : 0u;
      tot += c;
    }
    uint32_t e = carry + wp + incl - t;
    if (c0 < C) cnt[c0] = e;
    if (c0 + 1 < C) cnt[c0 + 1] = e + v0;
    if (c0 + 2 < C) cnt[c0 + 2] = e + v0 + v1;
    if (c0 + 3 < C) cnt[c0 + 3] = e + v0 + v1 + v2;
    carry += tot;
    __syncthreads();
  }
}

template <bool LDS_CELLS, int NW>
__global__ __launch_bounds__(64 * NW) void k_grid_build(SeedParams sp, GridArgs a, uint32_t cache_cap) {
  constexpr uint32_t BLOCK = 64u * NW;
  extern __shared__ uint32_t s_dyn[];
  __shared__ uint32_t s_wave[NW];
  const uint32_t tid = threadIdx.x, wave = tid >> 6;
  const uint32_t C = sp.cells, L = sp.n_layers;
  const uint32_t nseg = sp.n_events * L;
  for (uint32_t seg = blockIdx.x; seg < nseg; seg += gridDim.x) {
    const uint32_t e = seg / L;
    uint32_t base = a.event_offset[e] + a.layer_offset[seg];
    uint32_t nh = a.layer_nhits[seg];
    if (base > a.n_hits || nh > a.n_hits - base) {  // inconsistent layout: flag, emit empty cells
      if (tid == 0) atomicOr(a.flags, TRAX_FLAG_LAYOUT);
      base = base > a.n_hits ? a.n_hits : base;
      nh = 0;
    }
    uint32_t* cnt = LDS_CELLS ? s_dyn : (a.counters + (uint64_t)seg * C);
    // cached path: each hit's cell and phi kept in LDS after the histogram
    const bool cached = LDS_CELLS && nh <= cache_cap;
    uint32_t* s_cell = s_dyn + C;
    float* s_phi = reinterpret_cast<float*>(s_cell + cache_cap);
    uint32_t* s_idx = s_cell + 2 * cache_cap;
    for (uint32_t c = tid; c < C; c += BLOCK) cnt[c] = 0u;
    __syncthreads();
    // histogram (gridCount, GridBuilder.h:79-85)
    for (uint32_t i = tid; i < nh; i += BLOCK) {
      float phi;
      uint32_t cell = grid_cell(sp, a.ix[base + i], a.iy[base + i], a.iz[base + i], &phi);
      atomicAdd(&cnt[cell], 1u);
      if (cached) {
        s_cell[i] = cell;
        s_phi[i] = phi;
      }
    }
    __syncthreads();
    block_exclusive_scan_cells<NW>(cnt, C, s_wave);
    // cell starts in the global sorted array (the reference's scanned grid)
    uint32_t* g = a.grid + (uint64_t)seg * C;
    for (uint32_t c = tid; c < C; c += BLOCK) g[c] = base + cnt[c];
    __syncthreads();
    if (cached) {
      // Stable scatter without ordering the waves: every hit takes a slot of
      // its cell (in any order), then its rank is the number of same-cell
      // slots holding a smaller input index.  After the slot pass cnt[c] is
      // the end of cell c, i.e. cnt[c - 1] its start.
      for (uint32_t i = tid; i < nh; i += BLOCK) s_idx[atomicAdd(&cnt[s_cell[i]], 1u)] = i;
      __syncthreads();
      for (uint32_t i = tid; i < nh; i += BLOCK) {
        const uint32_t c = s_cell[i];
        const uint32_t lo = c == 0 ? 0u : cnt[c - 1], hi = cnt[c];
        uint32_t rank = 0;
        for (uint32_t t = lo; t < hi; ++t) rank += s_idx[t] < i ? 1u : 0u;
        const uint32_t dst = base + lo + rank;
        a.ox[dst] = a.ix[base + i];
        a.oy[dst] = a.iy[base + i];
        a.oz[dst] = a.iz[base + i];
        a.olayer[dst] = a.ilayer[base + i];
        a.odet[dst] = a.idet[base + i];
        a.ohid[dst] = a.ihid[base + i];
        a.oevt[dst] = a.ievt[base + i];
        a.oorig[dst] = base + i;
        a.ophi[dst] = s_phi[i];
        a.ozphi[dst] = make_float2(a.iz[base + i], s_phi[i]);
      }
      __syncthreads();
      continue;
    }
    // stable scatter (gridStore, GridBuilder.h:238-253, with a stable rank)
    for (uint32_t s0 = 0; s0 < nh; s0 += BLOCK) {
      const uint32_t i = s0 + tid;
      const bool valid = i < nh;
      float x = 0.f, y = 0.f, z = 0.f, phi = 0.f;
      uint32_t cell = 0;
      if (valid) {
        x = a.ix[base + i];
        y = a.iy[base + i];
        z = a.iz[base + i];
        cell = grid_cell(sp, x, y, z, &phi);
      }
      uint64_t match = __ballot(valid);
      for (uint32_t b = 0; b < sp.cell_bits; ++b) {
        const bool bit = (cell >> b) & 1u;
        const uint64_t m = __ballot(bit);
        match &= bit ? m : ~m;
      }
      const uint32_t rank = lanes_below(match);
      const uint32_t count = (uint32_t)__popcll(match);
      uint32_t pos = 0;
      for (uint32_t w = 0; w < (uint32_t)NW; ++w) {
        if (wave == w && valid) {
          pos = cnt[cell] + rank;
          if (rank == 0) cnt[cell] = pos + count;
        }
        __syncthreads();
      }
      if (valid) {
        const uint32_t dst = base + pos;
        a.ox[dst] = x;
        a.oy[dst] = y;
        a.oz[dst] = z;
        a.olayer[dst] = a.ilayer[base + i];
        a.odet[dst] = a.idet[base + i];
        a.ohid[dst] = a.ihid[base + i];
        a.oevt[dst] = a.ievt[base + i];
        a.oorig[dst] = base + i;
        a.ophi[dst] = phi;
        a.ozphi[dst] = make_float2(z, phi);
      }
    }
    __syncthreads();
  }
}

template <bool L>
static hipError_t grid_launch_nw(int nw, uint32_t blocks, size_t lds, uint32_t cap, hipStream_t stream,
                                 const SeedParams& sp, const GridArgs& a) {
  switch (nw) {
    case 1: hipLaunchKernelGGL((k_grid_build<L, 1>), dim3(blocks), dim3(64), lds, stream, sp, a, cap); break;
    case 2: hipLaunchKernelGGL((k_grid_build<L, 2>), dim3(blocks), dim3(128), lds, stream, sp, a, cap); break;
    case 8: hipLaunchKernelGGL((k_grid_build<L, 8>), dim3(blocks), dim3(512), lds, stream, sp, a, cap); break;
    default: hipLaunchKernelGGL((k_grid_build<L, 4>), dim3(blocks), dim3(256), lds, stream, sp, a, cap); break;
  }
  return hipGetLastError();
}

hipError_t launch_grid_build(const SeedParams& sp, const GridLaunch& gl, hipStream_t stream) {
  GridArgs a;
  a.ix = gl.ix; a.iy = gl.iy; a.iz = gl.iz;
  a.ilayer = gl.ilayer; a.idet = gl.idet; a.ihid = gl.ihid; a.ievt = gl.ievt;
  a.ox = gl.ox; a.oy = gl.oy; a.oz = gl.oz;
  a.olayer = gl.olayer; a.odet = gl.odet; a.ohid = gl.ohid; a.oevt = gl.oevt;
  a.oorig = gl.oorig; a.ophi = gl.ophi; a.ozphi = gl.ozphi;
  a.event_offset = gl.event_offset; a.layer_nhits = gl.layer_nhits; a.layer_offset = gl.layer_offset;
  a.n_hits = gl.n_hits;
  a.grid = gl.grid;
  a.counters = gl.counters;
  a.flags = gl.flags;
  const uint32_t nseg = sp.n_events * sp.n_layers;
  if (nseg == 0) return hipSuccess;
  const uint32_t blocks = nseg < 65535u ? nseg : 65535u;
  if (gl.use_lds) {
    // LDS: the cell counts, then (cell, phi, slot) per hit for segments of
    // up to `cap` hits (larger segments take the recomputing path)
    const size_t cells = (size_t)sp.cells * 4u;
    size_t room = gl.lds_max > cells ? gl.lds_max - cells : 0;
    uint32_t cap = (uint32_t)(room / 12u);
    if (cap > 16384u) cap = 16384u;
    if (cap < 256u) cap = 0;
    return grid_launch_nw<true>(gl.waves, blocks, cells + (size_t)cap * 12u, cap, stream, sp, a);
  }
  return grid_launch_nw<false>(gl.waves, blocks, 0, 0, stream, sp, a);
}

hipError_t grid_build_set_lds_limit(size_t bytes) {
  hipError_t e = hipSuccess;
  const void* fns[] = {reinterpret_cast<const void*>(&k_grid_build<true, 1>),
                       reinterpret_cast<const void*>(&k_grid_build<true, 2>),
                       reinterpret_cast<const void*>(&k_grid_build<true, 4>),
                       reinterpret_cast<const void*>(&k_grid_build<true, 8>)};
  for (const void* f : fns) {
    hipError_t r = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (r != hipSuccess) e = r;
  }
  return e;
}

}  // namespace trax
