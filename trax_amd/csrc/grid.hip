// Grid build: per-(event, layer) (z, phi) counting sort of the hit columns.
//
// Reference: src/algorithms/GridBuilder.h:48-427 and GridBuilder.cpp:6-160 run
// three passes with a host synchronisation and a full grid download between
// them: gridCount (LDS histogram, global atomics), PrefixSum over the whole
// E*L*C grid, and gridStore (scatter with atomic_inc ranks, so the order
// inside a cell is non-deterministic).
//
// Here two kernels, no host read and no global scan:
//
// k_grid_sort, one block per (event, layer) segment: computes every hit's
// cell, sorts the segment's hits by cell with a stable LSD radix sort (8-bit
// digits; ranks from 64-lane ballot matches and per-(digit, wave) counters,
// waves owning consecutive input ranges, so equal cells keep input order),
// and writes the sorted columns with coalesced stores (the input columns of
// a segment are gathered from L2).  Segments of up to 8192 hits sort packed
// (cell << 13 | index) keys in LDS; larger segments, or grids of more than
// 2^19 cells, sort (cell, index) pairs in place in four of the output
// columns before those are overwritten.  The block also writes the grid
// entry of the first cell of every 4096-cell chunk of its segment: the
// global sorted index of the first hit whose cell is not below that cell.
//
// k_grid_fill, one block per 4096-cell chunk of a segment: the hits of the
// chunk are the sorted range between its first-cell entry and the next
// chunk's; their cells (recomputed from the sorted z and phi) go into an LDS
// histogram whose exclusive scan, offset by the range start, gives every
// cell's start.  Most cells are empty and the grid is most of the bytes this
// stage writes, so the fill runs over thousands of blocks with coalesced
// stores.
//
// A segment's global base is its hit offset: hits are stored event-major and
// layer-ascending, so the cells of all earlier segments hold exactly the hits
// before this segment (the reference's scanned grid).
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
  // outputs; odet/ohid/oevt/olayer double as sort scratch for large segments
  float* ox;
  float* oy;
  float* oz;
  uint32_t* olayer;
  uint32_t* odet;
  uint32_t* ohid;
  uint32_t* oevt;
  uint32_t* oorig;
  float* ophi;
  float2* ozphi;
  const uint32_t* __restrict__ event_offset;
  const uint32_t* __restrict__ layer_nhits;
  const uint32_t* __restrict__ layer_offset;
  uint32_t n_hits;
  uint32_t* grid;
  uint32_t* flags;
};

constexpr uint32_t kGridChunk = 4096;  // cells per fill block
constexpr uint32_t kKeyIdxBits = 13;   // index bits of a packed LDS key
constexpr uint32_t kLdsSortMax = 1u << kKeyIdxBits;
constexpr uint32_t kDigitBits = 8;
constexpr uint32_t kDigits = 1u << kDigitBits;
constexpr uint32_t kFillThreads = 256;

// Hit range of segment `seg` (clamped, with a flag, if the layout does not
// fit the batch).
__device__ __forceinline__ void segment_extent(const GridArgs& a, uint32_t seg, uint32_t L,
                                               uint32_t* base_out, uint32_t* nh_out, bool raise) {
  uint32_t base = a.event_offset[seg / L] + a.layer_offset[seg];
  uint32_t nh = a.layer_nhits[seg];
  if (base > a.n_hits || nh > a.n_hits - base) {
    if (raise && threadIdx.x == 0) atomicOr(a.flags, TRAX_FLAG_LAYOUT);
    base = base > a.n_hits ? a.n_hits : base;
    nh = 0;
  }
  *base_out = base;
  *nh_out = nh;
}

// In-block exclusive scan of cnt[0..C) (64*NW threads, 4 consecutive cells
// per thread per step).  `cnt` is LDS.
template <int NW>
__device__ __forceinline__ void block_exclusive_scan_cells(uint32_t* cnt, uint32_t C, uint32_t* s_wave) {
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

// Lanes of this wave holding the same 8-bit digit (among `valid` lanes).
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid) {
  uint64_t m = __ballot(valid);
#pragma unroll
  for (uint32_t b = 0; b < kDigitBits; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t bb = __ballot(bit);
    m &= bit ? bb : ~bb;
  }
  return m;
}

// One stable counting pass of the radix sort over elements 0..n-1.
// load(i) -> element, digit(e) -> 0..255, store(pos, e).  Wave w owns the
// consecutive input range [w*S, (w+1)*S); per-(digit, wave) counters laid
// out digit-major make the output order (digit, wave, input order).
template <int NW, typename Load, typename Digit, typename Store>
__device__ __forceinline__ void radix_pass(uint32_t n, Load load, Digit digit, Store store,
                                           uint32_t* s_cnt, uint32_t* s_wave) {
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  const uint32_t S = (n + 64u * NW - 1u) / (64u * NW) * 64u;
  const uint32_t lo = wave * S;
  const uint32_t hi = lo + S < n ? lo + S : n;
  for (uint32_t c = threadIdx.x; c < kDigits * NW; c += 64u * NW) s_cnt[c] = 0u;
  __syncthreads();
  for (uint32_t i0 = lo; i0 < hi; i0 += 64u) {
    const uint32_t i = i0 + lane;
    const bool valid = i < hi;
    const uint32_t d = valid ? digit(load(i)) : 0u;
    const uint64_t m = match_digit(d, valid);
    if (valid && lanes_below(m) == 0u) s_cnt[d * NW + wave] += (uint32_t)__popcll(m);
  }
  __syncthreads();
  block_exclusive_scan_cells<NW>(s_cnt, kDigits * NW, s_wave);
  for (uint32_t i0 = lo; i0 < hi; i0 += 64u) {
    const uint32_t i = i0 + lane;
    const bool valid = i < hi;
    const auto e = load(valid ? i : lo);
    const uint32_t d = valid ? digit(e) : 0u;
    const uint64_t m = match_digit(d, valid);
    const uint32_t r = lanes_below(m);
    uint32_t pos = 0;
    if (valid) pos = s_cnt[d * NW + wave] + r;
    // every lane of the wave has read its digit's base before the leader
    // (rank 0) advances it: one instruction stream, LDS ops in order
    if (valid) store(pos, e);
    if (valid && r == 0u) s_cnt[d * NW + wave] = pos + (uint32_t)__popcll(m);
  }
  __syncthreads();
}

// Grid entries of the chunk-start cells k*kGridChunk (k < nch) that the
// sorted hit at position p opens: those with prev_cell < k*kGridChunk <= cell
// (prev_cell = -1 for p == 0).
__device__ __forceinline__ void chunk_starts(uint32_t* g, uint32_t nch, int64_t prev_cell,
                                             uint32_t cell, uint32_t value) {
  const uint32_t k_lo = prev_cell < 0 ? 0u : (uint32_t)(prev_cell / kGridChunk) + 1u;
  const uint32_t k_hi = cell / kGridChunk;  // inclusive
  for (uint32_t k = k_lo; k <= k_hi && k < nch; ++k) g[(uint64_t)k * kGridChunk] = value;
}

__device__ __forceinline__ void write_sorted(const GridArgs& a, uint32_t dst, uint32_t src) {
  const float x = a.ix[src], y = a.iy[src], z = a.iz[src];
  const float phi = trax_atan2f(y, x);
  a.ox[dst] = x;
  a.oy[dst] = y;
  a.oz[dst] = z;
  a.olayer[dst] = a.ilayer[src];
  a.odet[dst] = a.idet[src];
  a.ohid[dst] = a.ihid[src];
  a.oevt[dst] = a.ievt[src];
  a.oorig[dst] = src;
  a.ophi[dst] = phi;
  a.ozphi[dst] = make_float2(z, phi);
}

template <int NW>
__global__ __launch_bounds__(64 * NW) void k_grid_sort(SeedParams sp, GridArgs a, uint32_t lds_cap) {
  constexpr uint32_t BLOCK = 64u * NW;
  extern __shared__ uint32_t s_keys[];  // 2 * lds_cap packed keys (ping-pong)
  __shared__ uint32_t s_cnt[kDigits * NW];
  __shared__ uint32_t s_wave[NW];
  __shared__ uint32_t s_carry;
  const uint32_t tid = threadIdx.x;
  const uint32_t C = sp.cells, L = sp.n_layers;
  const uint32_t nseg = sp.n_events * L;
  const uint32_t nch = (C + kGridChunk - 1u) / kGridChunk;
  const uint32_t passes = (sp.cell_bits + kDigitBits - 1u) / kDigitBits;
  const bool packed_ok = sp.cell_bits + kKeyIdxBits <= 32u;
  for (uint32_t seg = blockIdx.x; seg < nseg; seg += gridDim.x) {
    uint32_t base, nh;
    segment_extent(a, seg, L, &base, &nh, true);
    uint32_t* g = a.grid + (uint64_t)seg * C;
    if (packed_ok && nh <= lds_cap) {
      uint32_t* src = s_keys;
      uint32_t* dst = s_keys + lds_cap;
      for (uint32_t i = tid; i < nh; i += BLOCK) {
        float phi;
        const uint32_t cell = grid_cell(sp, a.ix[base + i], a.iy[base + i], a.iz[base + i], &phi);
        src[i] = (cell << kKeyIdxBits) | i;
      }
      __syncthreads();
      for (uint32_t p = 0; p < passes; ++p) {
        const uint32_t shift = kKeyIdxBits + p * kDigitBits;
        const uint32_t* s = src;
        uint32_t* d = dst;
        radix_pass<NW>(
            nh, [s](uint32_t i) { return s[i]; },
            [shift](uint32_t e) { return (e >> shift) & (kDigits - 1u); },
            [d](uint32_t pos, uint32_t e) { d[pos] = e; }, s_cnt, s_wave);
        uint32_t* t = src;
        src = dst;
        dst = t;
      }
      for (uint32_t p = tid; p < nh; p += BLOCK) {
        const uint32_t v = src[p];
        const uint32_t cell = v >> kKeyIdxBits;
        const int64_t prev = p == 0 ? -1 : (int64_t)(src[p - 1] >> kKeyIdxBits);
        write_sorted(a, base + p, base + (v & (kLdsSortMax - 1u)));
        chunk_starts(g, nch, prev, cell, base + p);
      }
      if (tid == 0) {
        const int64_t last = nh == 0 ? -1 : (int64_t)(src[nh - 1] >> kKeyIdxBits);
        chunk_starts(g, nch, last, (nch - 1u) * kGridChunk, base + nh);
      }
      __syncthreads();
      continue;
    }
    // Large segment: (cell, index) pairs sorted in place in the output
    // columns odet/ohid and oevt/olayer, which are written last.
    uint32_t* cell_a = a.odet + base;
    uint32_t* idx_a = a.ohid + base;
    uint32_t* cell_b = a.oevt + base;
    uint32_t* idx_b = a.olayer + base;
    for (uint32_t i = tid; i < nh; i += BLOCK) {
      float phi;
      cell_a[i] = grid_cell(sp, a.ix[base + i], a.iy[base + i], a.iz[base + i], &phi);
      idx_a[i] = i;
    }
    __syncthreads();
    for (uint32_t p = 0; p < passes; ++p) {
      const uint32_t shift = p * kDigitBits;
      const uint32_t *sc = cell_a, *si = idx_a;
      uint32_t *dc = cell_b, *di = idx_b;
      radix_pass<NW>(
          nh, [sc, si](uint32_t i) { return make_uint2(sc[i], si[i]); },
          [shift](uint2 e) { return (e.x >> shift) & (kDigits - 1u); },
          [dc, di](uint32_t pos, uint2 e) {
            dc[pos] = e.x;
            di[pos] = e.y;
          },
          s_cnt, s_wave);
      uint32_t* t = cell_a;
      cell_a = cell_b;
      cell_b = t;
      t = idx_a;
      idx_a = idx_b;
      idx_b = t;
    }
    if (tid == 0) s_carry = 0xFFFFFFFFu;
    __syncthreads();
    for (uint32_t p0 = 0; p0 < nh; p0 += BLOCK) {
      // read this round's keys (and the previous round's last cell) before
      // any thread overwrites the scratch columns
      const uint32_t p = p0 + tid;
      const bool valid = p < nh;
      uint32_t cell = 0, idx = 0;
      int64_t prev = -1;
      if (valid) {
        cell = cell_a[p];
        idx = idx_a[p];
        if (p > 0) prev = tid == 0 ? (int64_t)s_carry : (int64_t)cell_a[p - 1];
      }
      __syncthreads();
      if (valid && (tid == BLOCK - 1u || p == nh - 1u)) s_carry = cell;
      if (valid) {
        write_sorted(a, base + p, base + idx);
        chunk_starts(g, nch, prev, cell, base + p);
      }
      __syncthreads();
    }
    if (tid == 0) {
      const int64_t last = nh == 0 ? -1 : (int64_t)s_carry;
      chunk_starts(g, nch, last, (nch - 1u) * kGridChunk, base + nh);
    }
    __syncthreads();
  }
}

// One block per (segment, 4096-cell chunk): LDS histogram of the chunk's
// hits, exclusive scan, coalesced grid stores.
__global__ __launch_bounds__(kFillThreads) void k_grid_fill(SeedParams sp, GridArgs a) {
  __shared__ uint32_t s_cnt[kGridChunk];
  __shared__ uint32_t s_wave[kFillThreads / 64];
  const uint32_t tid = threadIdx.x;
  const uint32_t C = sp.cells, L = sp.n_layers;
  const uint32_t nseg = sp.n_events * L;
  const uint32_t nch = (C + kGridChunk - 1u) / kGridChunk;
  const uint64_t nblk = (uint64_t)nseg * nch;
  for (uint64_t b = blockIdx.x; b < nblk; b += gridDim.x) {
    const uint32_t seg = (uint32_t)(b / nch), k = (uint32_t)(b % nch);
    uint32_t base, nh;
    segment_extent(a, seg, L, &base, &nh, false);
    const uint32_t c0 = k * kGridChunk;
    const uint32_t len = C - c0 < kGridChunk ? C - c0 : kGridChunk;
    uint32_t* g = a.grid + (uint64_t)seg * C + c0;
    // chunk-start entries written by k_grid_sort (in [base, base + nh])
    uint32_t lo = g[0] - base;
    uint32_t hi = k + 1u < nch ? g[kGridChunk] - base : nh;
    hi = hi > nh ? nh : hi;
    lo = lo > hi ? hi : lo;
    for (uint32_t c = tid; c < kGridChunk; c += kFillThreads) s_cnt[c] = 0u;
    __syncthreads();
    for (uint32_t j = lo + tid; j < hi; j += kFillThreads) {
      const float2 zp = a.ozphi[base + j];
      const uint32_t rel = grid_cell_zphi(sp, zp.x, zp.y) - c0;
      if (rel < len) atomicAdd(&s_cnt[rel], 1u);
    }
    __syncthreads();
    block_exclusive_scan_cells<kFillThreads / 64>(s_cnt, len, s_wave);
    const uint32_t start = base + lo;
    for (uint32_t c = tid; c < len; c += kFillThreads) g[c] = start + s_cnt[c];
    __syncthreads();
  }
}

template <int NW>
static void grid_sort_launch(uint32_t blocks, size_t lds, uint32_t cap, hipStream_t stream,
                             const SeedParams& sp, const GridArgs& a) {
  hipLaunchKernelGGL((k_grid_sort<NW>), dim3(blocks), dim3(64 * NW), lds, stream, sp, a, cap);
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
  a.flags = gl.flags;
  const uint32_t nseg = sp.n_events * sp.n_layers;
  if (nseg == 0) return hipSuccess;
  // packed keys in LDS: two buffers of `cap` keys next to the static
  // counters (at most 8 KiB + 64 B)
  const size_t room = gl.lds_max > 8448u ? gl.lds_max - 8448u : 0u;
  uint32_t cap = (uint32_t)(room / 8u);
  cap = cap > kLdsSortMax ? kLdsSortMax : cap & ~63u;
  const size_t lds = (size_t)cap * 8u;
  const uint32_t blocks = nseg < 65535u ? nseg : 65535u;
  switch (gl.waves) {
    case 1: grid_sort_launch<1>(blocks, lds, cap, stream, sp, a); break;
    case 2: grid_sort_launch<2>(blocks, lds, cap, stream, sp, a); break;
    case 4: grid_sort_launch<4>(blocks, lds, cap, stream, sp, a); break;
    default: grid_sort_launch<8>(blocks, lds, cap, stream, sp, a); break;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const uint64_t nblk = (uint64_t)nseg * ((sp.cells + kGridChunk - 1u) / kGridChunk);
  const uint32_t fill_blocks = nblk < (1u << 20) ? (uint32_t)nblk : (1u << 20);
  hipLaunchKernelGGL(k_grid_fill, dim3(fill_blocks), dim3(kFillThreads), 0, stream, sp, a);
  return hipGetLastError();
}

hipError_t grid_build_set_lds_limit(size_t bytes) {
  hipError_t e = hipSuccess;
  const void* fns[] = {reinterpret_cast<const void*>(&k_grid_sort<1>),
                       reinterpret_cast<const void*>(&k_grid_sort<2>),
                       reinterpret_cast<const void*>(&k_grid_sort<4>),
                       reinterpret_cast<const void*>(&k_grid_sort<8>)};
  const size_t dyn = bytes > 8448u ? bytes - 8448u : 0u;
  const size_t want = (size_t)kLdsSortMax * 8u;
  for (const void* f : fns) {
    hipError_t r = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)(dyn < want ? dyn : want));
    if (r != hipSuccess) e = r;
  }
  return e;
}

}  // namespace trax
