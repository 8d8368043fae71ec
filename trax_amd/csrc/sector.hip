// Hit-pair generation by grid-cell spread (no geometric cut).
//
// Reference: src/algorithms/PairGeneratorSector.h:43-234 and
// PairGeneratorSector.cpp:5-151, the reference's second PairGenerator.  One
// work-group per (event, layer pair) strides over the inner-layer hits; each
// hit pairs with every outer-layer hit in the grid cells of its z row
// +- pairSpreadZ and phi cell +- pairSpreadPhi (wrapping around the phi
// ring), recorded in an n1*n2 bitmap plus per-work-item visit counts; after a
// scan, pairStore re-reads all n2 bits of every inner hit.
//
// Here the inner hits of all (event, layer pair) segments form one flat work
// list.  A hit's pairs are index ranges of the sorted outer layer (one or two
// per z row), so the count pass reads only the grid: no hit is loaded and no
// bitmap exists.  The store pass writes the same ranges.  Each outer index is
// emitted once per inner hit, as the reference's bitmap keeps it: when the
// phi window is wider than the ring (2 * spread + 1 > nPhi) the reference
// counts the repeated visits and leaves their slots unwritten; here they are
// not counted.
#include "common.h"
#include "kernels.h"

namespace trax {

struct SectorWindow {
  uint32_t zLowSector, zHighSector, phiLowSector, phiHighSector;
  bool wrap;
};

// pairCount window (PairGeneratorSector.h:100-115) with the OpenCL integer
// types: zLowSector - spreadZ is unsigned (it wraps below row 0 and the row
// loop then runs zero times); the phi sector is an int from float arithmetic.
__device__ __forceinline__ SectorWindow sector_window(const SeedParams& sp, uint32_t p, float z,
                                                      float phi) {
  const uint32_t spreadZ = sp.spread_z[p], spreadPhi = sp.spread_phi[p];
  const uint32_t nZ = sp.n_z, nPhi = sp.n_phi;
  SectorWindow w;
  uint32_t zLowSector = trax_f2u(__builtin_floorf((z - sp.min_z) / sp.sector_size_z));
  uint32_t zHighSector = zLowSector + spreadZ + 1u;
  w.zHighSector = zHighSector < nZ ? zHighSector : nZ;
  w.zLowSector = zLowSector - spreadZ;
  int32_t phiLowSector =
      trax_f2i(__builtin_floorf((phi - sp.min_phi) / sp.sector_size_phi) - (float)spreadPhi);
  uint32_t phiHighSector = (uint32_t)phiLowSector + 2u * spreadPhi + 1u;
  w.wrap = phiLowSector < 0 || phiHighSector > nPhi + 1u;
  w.phiLowSector = (uint32_t)phiLowSector + (uint32_t)(phiLowSector < 0) * nPhi;
  w.phiHighSector = phiHighSector - (uint32_t)(phiHighSector > nPhi + 1u) * nPhi;
  return w;
}

// Outer indexes paired with one inner hit, row by row: without wrap-around
// the cells [phiLow, phiHigh) of the row, i.e. [A, H); with it the row from
// phiLow to the row end (padding column excluded, [A, zEnd)) and from the row
// start to phiHigh ([zStart, H)), each index once.  f(lo, hi) per range.
template <typename F>
__device__ __forceinline__ void sector_ranges(const SeedParams& sp, const uint32_t* __restrict__ g2,
                                              const SectorWindow& w, F&& f) {
  const uint32_t nPhi = sp.n_phi;
  for (uint32_t zs = w.zLowSector; zs < w.zHighSector; ++zs) {
    const uint32_t row = zs * (nPhi + 1u);
    const uint32_t A = g2[row + w.phiLowSector];
    const uint32_t H = g2[row + w.phiHighSector];
    if (!w.wrap) {
      if (H > A) f(A, H);
      continue;
    }
    const uint32_t zStart = g2[row], zEnd = g2[row + nPhi];
    if (zEnd > A) f(A, zEnd);
    const uint32_t b = H < A ? H : A;  // [zStart, H) minus [A, zEnd)
    if (b > zStart) f(zStart, b);
    if (H > zEnd) f(zEnd > A ? zEnd : A, H);
  }
}

struct SectorItem {
  uint32_t i;
  const uint32_t* g2;
  SectorWindow win;
};

__device__ __forceinline__ SectorItem sector_item(const SeedParams& sp, const SortedHits& h,
                                                  const uint32_t* __restrict__ grid,
                                                  const uint32_t* offs, uint32_t w) {
  const uint32_t P = sp.n_triplets, K = sp.n_events * P, C = sp.cells, L = sp.n_layers;
  SectorItem it;
  const uint32_t k = segment_of(offs, K + 1, w);
  const uint32_t e = k / P, p = k - e * P;
  const uint32_t* g1 = grid + ((uint64_t)e * L + (sp.layer1[p] - 1)) * C;
  it.g2 = grid + ((uint64_t)e * L + (sp.layer2[p] - 1)) * C;
  it.i = g1[0] + (w - offs[k]);
  const float2 zp = load_zphi(h.zphi, it.i);
  it.win = sector_window(sp, p, zp.x, zp.y);
  return it;
}

// count: pairs of every inner hit; store: (inner, outer) records at the
// scanned positions.
template <bool STORE>
__global__ __launch_bounds__(256) void k_sector_pairs(SeedParams sp, SortedHits h,
                                                      const uint32_t* __restrict__ grid,
                                                      const uint32_t* __restrict__ work_off,
                                                      uint32_t* __restrict__ counts,
                                                      uint2* __restrict__ pairs, uint64_t capacity,
                                                      uint32_t items_max, uint32_t* __restrict__ flags) {
  extern __shared__ uint32_t s_offs[];  // K + 1 offsets when they fit (offsets_lds_bytes)
  const uint32_t K = sp.n_events * sp.n_triplets;
  const uint32_t* offs = stage_offsets(work_off, K + 1, s_offs);
  uint32_t W = work_off[K];
  if (W > items_max) {  // more inner hits than the counts array holds: hits beyond max_hits
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(flags, TRAX_FLAG_LAYOUT);
    W = items_max;
  }
  const XcdShare sh = xcd_share(W, blockDim.x, threadIdx.x);
  for (uint32_t w = sh.begin; w < sh.end; w += sh.stride) {
    const SectorItem it = sector_item(sp, h, grid, offs, w);
    if (!STORE) {
      uint32_t n = 0;
      sector_ranges(sp, it.g2, it.win, [&](uint32_t lo, uint32_t hi) { n += hi - lo; });
      counts[w] = n;
    } else {
      uint64_t pos = counts[w];
      bool overflow = false;
      sector_ranges(sp, it.g2, it.win, [&](uint32_t lo, uint32_t hi) {
        for (uint32_t j = lo; j < hi; ++j, ++pos) {
          if (pos < capacity)
            pairs[pos] = make_uint2(it.i, j);
          else
            overflow = true;
        }
      });
      if (overflow) atomicOr(flags, TRAX_FLAG_PAIR_OVERFLOW);
    }
  }
}

hipError_t launch_sector_pairs(const SeedParams& sp, const SortedHits& h, const uint32_t* grid,
                               const uint32_t* work_off, uint32_t* counts, uint2* pairs,
                               uint64_t capacity, uint32_t items_max, bool store, uint32_t* flags,
                               hipStream_t stream) {
  uint32_t blocks = (items_max + 255u) / 256u;
  if (blocks == 0) blocks = 1;
  if (blocks > 16384u) blocks = 16384u;
  blocks = xcd_grid(blocks);
  if (store)
    hipLaunchKernelGGL(k_sector_pairs<true>, dim3(blocks), dim3(256), offsets_lds_bytes(sp), stream, sp, h,
                       grid, work_off, counts, pairs, capacity, items_max, flags);
  else
    hipLaunchKernelGGL(k_sector_pairs<false>, dim3(blocks), dim3(256), offsets_lds_bytes(sp), stream, sp, h,
                       grid, work_off, counts, pairs, capacity, items_max, flags);
  return hipGetLastError();
}

}  // namespace trax
