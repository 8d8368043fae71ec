// Grid build: per-(event, layer) (z, phi) counting sort of the hit columns.
//
// Reference: src/algorithms/GridBuilder.h:48-427 and GridBuilder.cpp:6-160 run
// three passes with a host synchronisation and a full grid download between
// them: gridCount (LDS histogram, global atomics), PrefixSum over the whole
// E*L*C grid, and gridStore (scatter with atomic_inc ranks, so the order
// inside a cell is non-deterministic).
//
// Here one kernel, no host read and no global scan.  k_grid_sort takes one
// (event, layer) segment per block, everything in LDS:
// * each thread loads its hits' columns (segments of up to 64 * waves * HPT
//   hits keep all of them in registers from the first load to the last
//   store, so the block waits for memory once), computes phi = atan2(y, x)
//   and the cell, and the cells go to LDS in input order;
// * the hits' cells are counted over the segment's occupied cell range
//   [cmin, cmax] (a few thousand cells for a barrel layer); the exclusive
//   scan of the counts is the segment's grid over that range, and below /
//   above it the grid is the segment base / base + hit count, filled with
//   16-byte stores that drain while the sort works;
// * each hit takes a slot of its cell (LDS atomics) and its final position is
//   the cell start plus the number of same-cell hits with a smaller input
//   index, so equal cells keep input order;
// * a cell holding more than 32 hits, or a range wider than the LDS buffer,
//   switches to a stable LSD radix sort of packed ((cell - cmin) << 13 |
//   index) keys (5-bit digits; ranks from 64-lane ballot matches and
//   per-(digit, wave) counters over consecutive input ranges) and the grid
//   from binary searches in the sorted keys;
// * the columns are permuted through LDS two per round: written in input
//   order (from registers, or read coalesced for larger segments), gathered
//   by sorted position and stored coalesced.
// Segments of more than 8192 hits sort (cell, index) pairs with the same
// radix passes in place in four of the output columns before those are
// overwritten.
//
// A segment's global base is its hit offset: hits are stored event-major and
// layer-ascending, so the cells of all earlier segments hold exactly the hits
// before this segment (the reference's scanned grid).  A layout whose
// segments are not contiguous from hit 0 raises TRAX_FLAG_LAYOUT.
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
  // outputs; odet/ohid/oevt/olayer are sort scratch for large segments
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
  float4* oxyzphi;
  const uint32_t* __restrict__ event_offset;
  const uint32_t* __restrict__ layer_nhits;
  const uint32_t* __restrict__ layer_offset;
  uint32_t n_hits;
  uint32_t* grid;
  uint32_t* flags;
};

constexpr uint32_t kKeyIdxBits = 13;   // index bits of a packed LDS key
constexpr uint32_t kLdsSortMax = 1u << kKeyIdxBits;
constexpr uint32_t kDigitBits = 5;
constexpr uint32_t kDigits = 1u << kDigitBits;
constexpr uint32_t kOccupancyMax = 32;  // counting sort: most hits in one cell

__device__ __forceinline__ void segment_range(const GridArgs& a, uint32_t seg, uint32_t L,
                                              uint32_t* base, uint32_t* nh) {
  *base = a.event_offset[seg / L] + a.layer_offset[seg];
  *nh = a.layer_nhits[seg];
}

// Hit range of segment `seg`, clamped to the batch.  With `check`, thread 0
// raises TRAX_FLAG_LAYOUT unless the segment starts where the previous one
// ends (segment 0 at hit 0) and fits the batch.
__device__ __forceinline__ void segment_extent(const GridArgs& a, uint32_t seg, uint32_t L,
                                               uint32_t* base_out, uint32_t* nh_out, bool check) {
  uint32_t base, nh;
  segment_range(a, seg, L, &base, &nh);
  const bool fits = base <= a.n_hits && nh <= a.n_hits - base;
  if (check && threadIdx.x == 0) {
    uint32_t pend = 0;
    if (seg > 0) {
      uint32_t pb, pn;
      segment_range(a, seg - 1, L, &pb, &pn);
      pend = pb + pn;
    }
    if (!fits || base != pend) atomicOr(a.flags, TRAX_FLAG_LAYOUT);
  }
  if (!fits) {
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

// Lanes of this wave holding the same digit (among `valid` lanes).
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
// load(i) -> element, digit(e) -> 0..kDigits-1, store(pos, e).  Wave w owns
// the consecutive input range [w*S, (w+1)*S); per-(digit, wave) counters
// laid out digit-major make the output order (digit, wave, input order).
template <int NW, typename Load, typename Digit, typename Store>
__device__ __forceinline__ void radix_pass(uint32_t n, Load load, Digit digit, Store store,
                                           uint32_t* s_cnt, uint32_t* s_wave) {
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  const uint32_t S = (n + 64u * NW - 1u) / (64u * NW) * 64u;
  const uint32_t lo = wave * S;
  const uint32_t hi = lo + S < n ? lo + S : n;
  for (uint32_t c = threadIdx.x; c < kDigits * NW; c += 64u * NW) s_cnt[c] = 0u;
  __syncthreads();
  for (uint32_t i = lo + lane; i < hi; i += 64u) atomicAdd(&s_cnt[digit(load(i)) * NW + wave], 1u);
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

// Block-wide min of lo and max of hi (every thread gets both).
template <int NW>
__device__ __forceinline__ void block_min_max(uint32_t lo, uint32_t hi, uint32_t* s_wave,
                                              uint32_t* out_lo, uint32_t* out_hi) {
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  lo = wave_reduce_minmax<false>(lo);
  hi = wave_reduce_minmax<true>(hi);
  if (lane == 0) {
    s_wave[wave] = lo;
    s_wave[NW + wave] = hi;
  }
  __syncthreads();
  lo = s_wave[0];
  hi = s_wave[NW];
#pragma unroll
  for (int w = 1; w < NW; ++w) {
    lo = s_wave[w] < lo ? s_wave[w] : lo;
    hi = s_wave[NW + w] > hi ? s_wave[NW + w] : hi;
  }
  __syncthreads();
  *out_lo = lo;
  *out_hi = hi;
}

// Number of keys in S[0..n) below t (S sorted ascending).
__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t* S, uint32_t n, uint32_t t) {
  uint32_t lo = 0, len = n;
  while (len > 0) {
    const uint32_t half = len >> 1;
    if (S[lo + half] < t) {
      lo += half + 1u;
      len -= half + 1u;
    } else {
      len = half;
    }
  }
  return lo;
}

// p[0..n) = v by the block: 16-byte stores between a head and a tail of at
// most three words each.
template <uint32_t BLOCK>
__device__ __forceinline__ void fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
  static_assert(BLOCK >= 4, "head and tail take one store per thread");
  const uint32_t mis = (uint32_t)((reinterpret_cast<uintptr_t>(p) >> 2) & 3u);
  uint32_t head = mis ? 4u - mis : 0u;
  head = head < n ? head : n;
  if (threadIdx.x < head) p[threadIdx.x] = v;
  uint4* q = reinterpret_cast<uint4*>(p + head);
  const uint32_t nq = (n - head) >> 2;
  const uint4 vv = make_uint4(v, v, v, v);
  for (uint32_t c = threadIdx.x; c < nq; c += BLOCK) q[c] = vv;
  const uint32_t done = head + nq * 4u;
  if (threadIdx.x < n - done) p[done + threadIdx.x] = v;
}

// Grid entries of a segment outside its occupied cell range [cmin, cmax]:
// the segment base below it, base + nh above it.
template <uint32_t BLOCK>
__device__ __forceinline__ void grid_constant_parts(uint32_t* g, uint32_t C, uint32_t cmin,
                                                    uint32_t cmax, uint32_t base, uint32_t nh) {
  fill_u32<BLOCK>(g, cmin < C ? cmin : C, base);
  if (cmax + 1u < C) fill_u32<BLOCK>(g + cmax + 1u, C - (cmax + 1u), base + nh);
}

// In-block exclusive scan of cnt[0..n) that also returns the largest count.
template <int NW>
__device__ __forceinline__ uint32_t block_scan_occupancy(uint32_t* cnt, uint32_t n, uint32_t* s_wave) {
  const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  uint32_t carry = 0, occ = 0;
  for (uint32_t base = 0; base < n; base += 256u * NW) {
    const uint32_t c0 = base + tid * 4u;
    const uint32_t v0 = c0 < n ? cnt[c0] : 0u;
    const uint32_t v1 = c0 + 1 < n ? cnt[c0 + 1] : 0u;
    const uint32_t v2 = c0 + 2 < n ? cnt[c0 + 2] : 0u;
    const uint32_t v3 = c0 + 3 < n ? cnt[c0 + 3] : 0u;
    const uint32_t m01 = v0 > v1 ? v0 : v1, m23 = v2 > v3 ? v2 : v3;
    const uint32_t m = m01 > m23 ? m01 : m23;
    occ = m > occ ? m : occ;
    const uint32_t t = v0 + v1 + v2 + v3;
    const uint32_t incl = wave_inclusive_scan(t);
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t wp = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const uint32_t c = s_wave[w];
      wp += (w < (int)wave) ? c : 0u;
      tot += c;
    }
    const uint32_t e = carry + wp + incl - t;
    if (c0 < n) cnt[c0] = e;
    if (c0 + 1 < n) cnt[c0 + 1] = e + v0;
    if (c0 + 2 < n) cnt[c0 + 2] = e + v0 + v1;
    if (c0 + 3 < n) cnt[c0 + 3] = e + v0 + v1 + v2;
    carry += tot;
    __syncthreads();
  }
  occ = wave_reduce_minmax<true>(occ);
  if (lane == 0) s_wave[NW + wave] = occ;
  __syncthreads();
  occ = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) occ = s_wave[NW + w] > occ ? s_wave[NW + w] : occ;
  __syncthreads();
  return occ;
}

// The hit columns of a segment, HPT hits per thread (hit tid + k * BLOCK),
// held in registers from the block's first load to its last store.
template <int HPT>
struct HitRegs {
  float x[HPT], y[HPT], z[HPT], phi[HPT];
  uint32_t layer[HPT], det[HPT], hid[HPT], evt[HPT];
};

// One block per (event, layer) segment at a time.  Segments of at most
// 64 * NW * HPT hits keep every column in registers: their loads are issued
// once, at the start, and the permutation rounds only move data through LDS.
// Larger segments read the columns round by round (up to the LDS cap) or
// sort in place in four output columns.
template <int NW, int HPT, int WPE>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(WPE))) void k_grid_sort(SeedParams sp, GridArgs a, uint32_t lds_cap) {
  constexpr uint32_t BLOCK = 64u * NW;
  extern __shared__ uint32_t s_lds[];  // R0, R1: lds_cap words each; R2: lds_cap uint16
  __shared__ uint32_t s_cnt[kDigits * NW];
  __shared__ uint32_t s_wave[2 * NW];
  const uint32_t tid = threadIdx.x;
  const uint32_t C = sp.cells, L = sp.n_layers;
  const uint32_t nseg = sp.n_events * L;
  const uint32_t cap = lds_cap;
  uint32_t* R0 = s_lds;
  uint32_t* R1 = s_lds + lds_cap;
  uint16_t* R2 = reinterpret_cast<uint16_t*>(s_lds + 2u * lds_cap);
  for (uint32_t seg = blockIdx.x; seg < nseg; seg += gridDim.x) {
    uint32_t base, nh;
    segment_extent(a, seg, L, &base, &nh, true);
    uint32_t* g = a.grid + (uint64_t)seg * C;
    if (nh == 0) {
      fill_u32<BLOCK>(g, C, base);
      continue;
    }
    if (nh <= cap) {
      // every hit's phi = atan2(y, x) and cell; cells in input order into R1;
      // the segment's occupied cell range
      const bool regs = nh <= BLOCK * (uint32_t)HPT;  // block-uniform
      HitRegs<HPT> v;
      uint32_t cmin = 0xFFFFFFFFu, cmax = 0u;
      if (regs) {
#pragma unroll
        for (int k = 0; k < HPT; ++k) {
          const uint32_t i = tid + (uint32_t)k * BLOCK;
          if (i < nh) {
            v.x[k] = a.ix[base + i];
            v.y[k] = a.iy[base + i];
            v.z[k] = a.iz[base + i];
            v.layer[k] = a.ilayer[base + i];
            v.det[k] = a.idet[base + i];
            v.hid[k] = a.ihid[base + i];
            v.evt[k] = a.ievt[base + i];
          }
        }
#pragma unroll
        for (int k = 0; k < HPT; ++k) {
          const uint32_t i = tid + (uint32_t)k * BLOCK;
          if (i < nh) {
            v.phi[k] = trax_atan2f(v.y[k], v.x[k]);
            const uint32_t c = grid_cell_zphi(sp, v.z[k], v.phi[k]);
            cmin = c < cmin ? c : cmin;
            cmax = c > cmax ? c : cmax;
            R1[i] = c;
          }
        }
      } else {
        for (uint32_t i = tid; i < nh; i += BLOCK) {
          const float phi = trax_atan2f(a.iy[base + i], a.ix[base + i]);
          const uint32_t c = grid_cell_zphi(sp, a.iz[base + i], phi);
          cmin = c < cmin ? c : cmin;
          cmax = c > cmax ? c : cmax;
          R1[i] = c;
        }
      }
      block_min_max<NW>(cmin, cmax, s_wave, &cmin, &cmax);
      const uint32_t span = cmax - cmin + 1u;  // cells <= 2^24
      uint32_t span_bits = 0;
      while (((span - 1u) >> span_bits) != 0u) ++span_bits;
      if (span_bits + kKeyIdxBits <= 32u) {  // block-uniform
        // most of the grid is outside the occupied range: those stores drain
        // while the sort works in LDS
        grid_constant_parts<BLOCK>(g, C, cmin, cmax, base, nh);
        const uint32_t* S = nullptr;  // sorted packed keys ((cell - cmin) << 13 | index)
        if (span + nh <= lds_cap) {
          // Counting sort over the occupied range: R0 holds the cell counts
          // (then starts, then ends) followed by the slot -> hit map.
          uint32_t* slot_hit = R0 + span;
          for (uint32_t x = tid; x < span; x += BLOCK) R0[x] = 0u;
          __syncthreads();
          for (uint32_t i = tid; i < nh; i += BLOCK) atomicAdd(&R0[R1[i] - cmin], 1u);
          __syncthreads();
          const uint32_t occ = block_scan_occupancy<NW>(R0, span, s_wave);
          if (occ <= kOccupancyMax) {  // block-uniform
            // the scanned counts are the grid over the occupied range
            for (uint32_t x = tid; x < span; x += BLOCK) g[cmin + x] = base + R0[x];
            __syncthreads();
            // each hit takes a slot of its cell: R1[i] = (cell - cmin) << 13 | slot
            for (uint32_t i = tid; i < nh; i += BLOCK) {
              const uint32_t x = R1[i] - cmin;
              const uint32_t slot = atomicAdd(&R0[x], 1u);
              slot_hit[slot] = i;
              R1[i] = (x << kKeyIdxBits) | slot;
            }
            __syncthreads();
            for (uint32_t i = tid; i < nh; i += BLOCK) {
              const uint32_t x = R1[i] >> kKeyIdxBits;
              const uint32_t st = x ? R0[x - 1u] : 0u, en = R0[x];
              uint32_t rank = 0;
              for (uint32_t t = st; t < en; ++t) rank += slot_hit[t] < i ? 1u : 0u;
              R1[i] = (x << kKeyIdxBits) | (st + rank);
            }
            __syncthreads();
            for (uint32_t i = tid; i < nh; i += BLOCK) {
              const uint32_t w = R1[i];
              R0[w & (kLdsSortMax - 1u)] = (w & ~(kLdsSortMax - 1u)) | i;
            }
            __syncthreads();
            S = R0;
          }
        }
        if (S == nullptr) {
          // radix sort of the packed keys (wide range or a crowded cell);
          // R1 still holds the cells
          for (uint32_t i = tid; i < nh; i += BLOCK) R0[i] = ((R1[i] - cmin) << kKeyIdxBits) | i;
          __syncthreads();
          uint32_t* src = R0;
          uint32_t* dst = R1;
          const uint32_t passes = (span_bits + kDigitBits - 1u) / kDigitBits;
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
          S = src;
          for (uint32_t x = tid; x < span; x += BLOCK)
            g[cmin + x] = base + lower_bound_u32(S, nh, x << kKeyIdxBits);
        }
        // Columns in sorted order, two per round: the sorted input indexes
        // move to the 16-bit region R2, the columns go in input order into
        // R0/R1 (from registers, or read coalesced), are gathered by sorted
        // position and stored coalesced.
        for (uint32_t p = tid; p < nh; p += BLOCK) R2[p] = (uint16_t)(S[p] & (kLdsSortMax - 1u));
        __syncthreads();
        for (uint32_t r = 0; r < 4; ++r) {
          if (regs) {
#pragma unroll
            for (int k = 0; k < HPT; ++k) {
              const uint32_t i = tid + (uint32_t)k * BLOCK;
              if (i < nh) {
                R0[i] = r == 0 ? __builtin_bit_cast(uint32_t, v.x[k])
                        : r == 1 ? __builtin_bit_cast(uint32_t, v.z[k])
                        : r == 2 ? v.layer[k] : v.hid[k];
                R1[i] = r == 0 ? __builtin_bit_cast(uint32_t, v.y[k])
                        : r == 1 ? __builtin_bit_cast(uint32_t, v.phi[k])
                        : r == 2 ? v.det[k] : v.evt[k];
              }
            }
          } else {
            for (uint32_t i = tid; i < nh; i += BLOCK) {
              const uint32_t s = base + i;
              if (r == 0) {
                R0[i] = __builtin_bit_cast(uint32_t, a.ix[s]);
                R1[i] = __builtin_bit_cast(uint32_t, a.iy[s]);
              } else if (r == 1) {
                R0[i] = __builtin_bit_cast(uint32_t, a.iz[s]);
                R1[i] = __builtin_bit_cast(uint32_t, trax_atan2f(a.iy[s], a.ix[s]));
              } else if (r == 2) {
                R0[i] = a.ilayer[s];
                R1[i] = a.idet[s];
              } else {
                R0[i] = a.ihid[s];
                R1[i] = a.ievt[s];
              }
            }
          }
          __syncthreads();
          for (uint32_t p = tid; p < nh; p += BLOCK) {
            const uint32_t hit = R2[p];
            const uint32_t va = R0[hit], vb = R1[hit];
            const uint32_t dst = base + p;
            if (r == 0) {
              a.ox[dst] = __builtin_bit_cast(float, va);
              a.oy[dst] = __builtin_bit_cast(float, vb);
            } else if (r == 1) {
              a.oz[dst] = __builtin_bit_cast(float, va);
              a.ophi[dst] = __builtin_bit_cast(float, vb);
              a.ozphi[dst] = make_float2(__builtin_bit_cast(float, va), __builtin_bit_cast(float, vb));
              // x and y of this sorted position were stored by this thread in round 0
              a.oxyzphi[dst] = make_float4(a.ox[dst], a.oy[dst], __builtin_bit_cast(float, va),
                                           __builtin_bit_cast(float, vb));
            } else if (r == 2) {
              a.olayer[dst] = va;
              a.odet[dst] = vb;
            } else {
              a.ohid[dst] = va;
              a.oevt[dst] = vb;
              a.oorig[dst] = base + hit;
            }
          }
          __syncthreads();
        }
        continue;
      }
    }
    // Large segment (or a span too wide for packed keys): (cell, index)
    // pairs sorted in place in the output columns odet/ohid and oevt/olayer,
    // which are written last.
    uint32_t* cell_a = a.odet + base;
    uint32_t* idx_a = a.ohid + base;
    uint32_t* cell_b = a.oevt + base;
    uint32_t* idx_b = a.olayer + base;
    for (uint32_t i = tid; i < nh; i += BLOCK) {
      const uint32_t s = base + i;
      cell_a[i] = grid_cell_zphi(sp, a.iz[s], trax_atan2f(a.iy[s], a.ix[s]));
      idx_a[i] = i;
    }
    __syncthreads();
    const uint32_t passes = (sp.cell_bits + kDigitBits - 1u) / kDigitBits;
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
    // grid: lower bounds in the sorted cells, read before they are
    // overwritten below
    for (uint32_t c = tid; c < C; c += BLOCK) g[c] = base + lower_bound_u32(cell_a, nh, c);
    __syncthreads();
    for (uint32_t p0 = 0; p0 < nh; p0 += BLOCK) {
      // read this round's indexes before any thread overwrites the scratch
      // columns
      const uint32_t p = p0 + tid;
      const bool valid = p < nh;
      const uint32_t idx = valid ? idx_a[p] : 0u;
      __syncthreads();
      if (valid) {
        const uint32_t src = base + idx, dst = base + p;
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
        a.oxyzphi[dst] = make_float4(x, y, z, phi);
      }
      __syncthreads();
    }
  }
}

// Hits per thread kept in registers, per block size: the PU200 and uneven
// batches' largest segments (about 4,100 hits) fit 512 threads x 9.  Waves
// per SIMD the register allocation must allow: two 512-thread blocks per CU.
template <int NW>
constexpr int grid_hpt() {
  return NW == 8 ? 9 : 4;
}
template <int NW>
constexpr int grid_wpe() {
  return NW >= 8 ? 4 : 1;
}

template <int NW>
static void grid_sort_launch(uint32_t blocks, size_t lds, uint32_t cap, hipStream_t stream,
                             const SeedParams& sp, const GridArgs& a) {
  hipLaunchKernelGGL((k_grid_sort<NW, grid_hpt<NW>(), grid_wpe<NW>()>), dim3(blocks), dim3(64 * NW), lds, stream, sp, a, cap);
}

// Static LDS of a sort block (counters, wave partials) stays below this; a
// block uses at most half of a CU's 160 KiB.
constexpr size_t kSortStaticLds = 2048;
constexpr size_t kSortLdsPerBlock = 80u * 1024u;

hipError_t launch_grid_build(const SeedParams& sp, const GridLaunch& gl, hipStream_t stream) {
  GridArgs a;
  a.ix = gl.ix; a.iy = gl.iy; a.iz = gl.iz;
  a.ilayer = gl.ilayer; a.idet = gl.idet; a.ihid = gl.ihid; a.ievt = gl.ievt;
  a.ox = gl.ox; a.oy = gl.oy; a.oz = gl.oz;
  a.olayer = gl.olayer; a.odet = gl.odet; a.ohid = gl.ohid; a.oevt = gl.oevt;
  a.oorig = gl.oorig; a.ophi = gl.ophi; a.ozphi = gl.ozphi; a.oxyzphi = gl.oxyzphi;
  a.event_offset = gl.event_offset; a.layer_nhits = gl.layer_nhits; a.layer_offset = gl.layer_offset;
  a.n_hits = gl.n_hits;
  a.grid = gl.grid;
  a.flags = gl.flags;
  const uint32_t nseg = sp.n_events * sp.n_layers;
  if (nseg == 0) return hipSuccess;
  // 10 bytes of LDS per hit (two word regions, one 16-bit region), sized so
  // that two blocks share a CU
  const size_t budget = gl.lds_max < kSortLdsPerBlock ? gl.lds_max : kSortLdsPerBlock;
  const size_t room = budget > kSortStaticLds ? budget - kSortStaticLds : 0u;
  uint32_t cap = (uint32_t)(room / 10u);
  cap = cap > kLdsSortMax ? kLdsSortMax : cap & ~63u;
  // Batches of many small segments (thousands of events of a few hundred
  // hits) get less LDS per block, so that many blocks share a CU: the cap
  // covers the segment's cells plus four times the mean hit count (at least
  // 512 slots; 4,096 ttbar events with 459 cells and ~19 hits per segment:
  // 70 -> 55 us against a floor of 1,024).  A larger segment takes the radix
  // or in-place path, so the cap changes speed only.
  uint64_t fit = (((uint64_t)sp.cells + 4ull * (gl.n_hits / nseg)) + 63u) & ~63ull;
  fit = fit < 512u ? 512u : fit;
  if (fit < cap) cap = (uint32_t)fit;
  const size_t lds = (size_t)(cap > 0 ? cap : 64u) * 10u;
  const uint32_t blocks = nseg < 65535u ? nseg : 65535u;
  // automatic block size: at most two hits of the mean segment per thread,
  // so that segments of several times the mean keep their columns in
  // registers (1 wave for the few-hundred-hit segments of single
  // collisions, 8 from 1,024 hits on)
  int waves = gl.waves;
  if (waves == 0) {
    const uint32_t per = (gl.n_hits / nseg) / 256u;
    waves = per >= 4u ? 8 : per >= 2u ? 4 : per >= 1u ? 2 : 1;
  }
  switch (waves) {
    case 1: grid_sort_launch<1>(blocks, lds, cap, stream, sp, a); break;
    case 2: grid_sort_launch<2>(blocks, lds, cap, stream, sp, a); break;
    case 4: grid_sort_launch<4>(blocks, lds, cap, stream, sp, a); break;
    case 16: grid_sort_launch<16>(blocks, lds, cap, stream, sp, a); break;
    default: grid_sort_launch<8>(blocks, lds, cap, stream, sp, a); break;
  }
  return hipGetLastError();
}

hipError_t grid_build_set_lds_limit(size_t bytes) {
  hipError_t e = hipSuccess;
  const void* fns[] = {reinterpret_cast<const void*>(&k_grid_sort<1, grid_hpt<1>(), grid_wpe<1>()>),
                       reinterpret_cast<const void*>(&k_grid_sort<2, grid_hpt<2>(), grid_wpe<2>()>),
                       reinterpret_cast<const void*>(&k_grid_sort<4, grid_hpt<4>(), grid_wpe<4>()>),
                       reinterpret_cast<const void*>(&k_grid_sort<8, grid_hpt<8>(), grid_wpe<8>()>),
                       reinterpret_cast<const void*>(&k_grid_sort<16, grid_hpt<16>(), grid_wpe<16>()>)};
  const size_t dyn = bytes > kSortStaticLds ? bytes - kSortStaticLds : 0u;
  const size_t want = kSortLdsPerBlock - kSortStaticLds;
  for (const void* f : fns) {
    hipError_t r = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)(dyn < want ? dyn : want));
    if (r != hipSuccess) e = r;
  }
  return e;
}

}  // namespace trax
