// Hit-pair generation with a beam-spot constrained window.
//
// Reference: src/algorithms/PairGeneratorBeamspot.h:43-525 and
// PairGeneratorBeamspot.cpp:5-177.  There, one work-group per (event, layer
// pair) strides over the outer-layer hits, walks the inner-layer grid cells
// of each hit's (z, phi) window, and records valid pairs in an n1*n2 bitmap
// plus a per-work-item count; after a scan, pairStore re-reads all n1 bits of
// every outer hit to write the pairs.
//
// Here the outer hits of all (event, layer pair) segments form one flat work
// list (offsets from k_pair_plan), so a launch covers every event of a batch
// and uneven events cost nothing extra.  Count and store run the same window
// walk (no bitmap): count writes one uint32 per outer hit, the exclusive scan
// turns them into write positions, store re-walks and writes (inner, outer)
// pairs.  The window arithmetic is the reference's, expression for
// expression; the inner-hit phi comes from the phi column written by the grid
// build (the same atan2 of the same inputs).  The walk visits each inner index
// once: the bitmap of the reference also collapses the rare repeated visits
// of its wrap-around walk, so the emitted pair set is the same.
#include "common.h"
#include "kernels.h"

namespace trax {

struct PairWindow {
  float zLow, zHigh, phiLow, phiHigh;
  bool wrap;
  uint32_t zLowSector, zHighSector, phiLowSector, phiHighSector;
};

// PairGeneratorBeamspot.h:110-162 for outer hit (X, Y, Z) with phi = atan2(Y, X).
__device__ __forceinline__ PairWindow pair_window(const SeedParams& sp, uint32_t p, float X, float Y,
                                                  float Z, float phi, float minR1, float maxR1,
                                                  float acosMinR1, float acosMaxR1) {
  const float z0 = sp.z0[p], tip = sp.tip_cut[p], sigmaZ = sp.sigma_z[p];
  const float sigmaPhi = sp.pair_sigma_phi[p];
  const float R = sp.min_radius_curvature;
  const float PI = TRAX_PI_F;
  PairWindow w;
  float signRadius = trax_sign(Y);
  float r = signRadius * __builtin_sqrtf(X * X + Y * Y);
  float cotThetaHigh = (Z - z0) / r;
  float cotThetaLow = (Z + z0) / r;
  float tmp = signRadius * maxR1 * cotThetaHigh + z0;
  float zHigh = signRadius * minR1 * cotThetaHigh + z0;
  zHigh = trax_b2f(tmp < zHigh) * zHigh + trax_b2f(tmp > zHigh) * tmp;
  tmp = signRadius * maxR1 * cotThetaLow - z0;
  float zLow = signRadius * minR1 * cotThetaLow - z0;
  zLow = trax_b2f(tmp < zLow) * tmp + trax_b2f(tmp > zLow) * zLow;
  zLow -= sigmaZ;
  zHigh += sigmaZ;
  int32_t zl = trax_f2i(__builtin_floorf((zLow - sp.min_z) / sp.sector_size_z));
  w.zLowSector = (uint32_t)(zl > 0 ? zl : 0);
  uint32_t zh = trax_f2u(__builtin_floorf((zHigh - sp.min_z) / sp.sector_size_z)) + 1u;
  w.zHighSector = zh < sp.n_z ? zh : sp.n_z;

  // acos(maxR1 / (2 * R)) and acos(minR1 / (2 * R)) come precomputed per layer
  const float acosR = trax_acosf(r / (2 * R));
  tmp = __builtin_fabsf(acosR - acosMaxR1);
  float dPhi = __builtin_fabsf(acosR - acosMinR1);
  dPhi = trax_b2f(tmp < dPhi) * dPhi + trax_b2f(tmp > dPhi) * tmp;
  tmp = trax_atanf(tip * (r - minR1)) / (r * minR1);
  float phiHigh = trax_atanf(tip * (r - maxR1)) / (r * maxR1);
  dPhi += tmp > phiHigh ? tmp : phiHigh;
  float phiLow = phi - dPhi - sigmaPhi;
  phiHigh = phi + dPhi + sigmaPhi;
  w.wrap = phiLow < -PI || phiHigh > PI || phiLow > PI || phiHigh < -PI;
  phiLow -= (float)((phiLow > PI) * 2) * PI;
  phiLow += (float)((phiLow < -PI) * 2) * PI;
  phiHigh -= (float)((phiHigh > PI) * 2) * PI;
  phiHigh += (float)((phiHigh < -PI) * 2) * PI;
  uint32_t pl = trax_f2u(__builtin_floorf((phiLow - sp.min_phi) / sp.sector_size_phi));
  uint32_t ph = trax_f2u(__builtin_floorf((phiHigh - sp.min_phi) / sp.sector_size_phi)) + 1u;
  w.phiLowSector = pl > sp.n_phi ? sp.n_phi : pl;
  w.phiHighSector = ph > sp.n_phi + 1 ? sp.n_phi + 1 : ph;
  w.zLow = zLow;
  w.zHigh = zHigh;
  w.phiLow = phiLow;
  w.phiHigh = phiHigh;
  return w;
}

__device__ __forceinline__ bool window_phi_ok(bool wrap, float phiLow, float phiHigh, float actPhi) {
  const float PI = TRAX_PI_F;
  return (!wrap && (phiLow <= actPhi && actPhi <= phiHigh)) ||
         (wrap && ((phiLow <= actPhi && actPhi <= PI) || (-PI <= actPhi && actPhi <= phiHigh)));
}

// Per-(event, layer pair) work offsets: prefix sum of the outer-layer hit
// counts grid[C-1] - grid[0] (PairGeneratorBeamspot.h:70-73), or with `inner`
// of the inner layer's (PairGeneratorSector.h:67-70).
__global__ __launch_bounds__(256) void k_pair_plan(SeedParams sp, const uint32_t* __restrict__ grid,
                                                   uint32_t* __restrict__ work_off, bool inner) {
  __shared__ uint32_t s_wave[4];
  const uint32_t K = sp.n_events * sp.n_triplets, C = sp.cells, L = sp.n_layers;
  const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  uint32_t carry = 0;
  for (uint32_t base = 0; base < K; base += 256u) {
    const uint32_t k = base + tid;
    uint32_t v = 0;
    if (k < K) {
      const uint32_t e = k / sp.n_triplets, p = k % sp.n_triplets;
      const uint32_t layer = inner ? sp.layer1[p] : sp.layer2[p];
      const uint32_t* g = grid + ((uint64_t)e * L + (layer - 1)) * C;
      v = g[C - 1] - g[0];
    }
    uint32_t incl = wave_inclusive_scan(v);
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t wp = 0, tot = 0;
    for (int w = 0; w < 4; ++w) {
      wp += (w < (int)wave) ? s_wave[w] : 0u;
      tot += s_wave[w];
    }
    if (k < K) work_off[k] = carry + wp + incl - v;
    carry += tot;
    __syncthreads();
  }
  if (tid == 0) work_off[K] = carry;
}

// The same counts for many segments, one per thread (work_off[K] = 0); an
// exclusive scan of the K + 1 values then gives the offsets and the total.
__global__ __launch_bounds__(256) void k_pair_counts(SeedParams sp, const uint32_t* __restrict__ grid,
                                                     uint32_t* __restrict__ work_off, bool inner) {
  const uint32_t K = sp.n_events * sp.n_triplets, C = sp.cells, L = sp.n_layers;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k <= K; k += gridDim.x * blockDim.x) {
    uint32_t v = 0;
    if (k < K) {
      const uint32_t e = k / sp.n_triplets, p = k % sp.n_triplets;
      const uint32_t layer = inner ? sp.layer1[p] : sp.layer2[p];
      const uint32_t* g = grid + ((uint64_t)e * L + (layer - 1)) * C;
      v = g[C - 1] - g[0];
    }
    work_off[k] = v;
  }
}

// Inner hits of outer hit i that pass its window, in walk order: for each z
// row of the window, the phi cells [phiLowSector, phiHighSector), wrapping
// at the row end (PairGeneratorBeamspot.h:172-230).  Every index is visited
// once (the reference's bitmap collapses its repeated wrap-around visits).
// f(bits, ord) gets the visits' verdicts in groups (see scan_range below).
template <typename F>
__device__ __forceinline__ void pair_walk(const SeedParams& sp, const uint32_t* __restrict__ g1,
                                          const PairWindow& win, const SortedHits& h, F&& f) {
  const uint32_t nPhi = sp.n_phi;
  auto in_window = [&](float zi, float pi) {
    return win.zLow <= zi && zi <= win.zHigh && window_phi_ok(win.wrap, win.phiLow, win.phiHigh, pi);
  };
  uint32_t vis = 0;  // hits visited before the current range
  // f(bits, ord): the verdicts of up to four consecutive visits, bit u for
  // the visit with ordinal ord + u
  auto scan_range = [&](uint32_t a, uint32_t b, const float2* __restrict__ zphi, auto&& test, auto&& g) {
    uint32_t idx = a;
    for (; idx + 4 <= b; idx += 4) {
      float2 v[4];
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) v[u] = load_zphi(zphi, idx + u);
      uint32_t bits = 0;
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) bits |= test(v[u].x, v[u].y) ? 1u << u : 0u;
      g(bits, vis + (idx - a));
    }
    if (idx < b) {  // the last one to three hits, loaded together
      float2 v[3];
#pragma unroll
      for (uint32_t u = 0; u < 3; ++u) v[u] = load_zphi(zphi, idx + u < b ? idx + u : idx);
      uint32_t bits = 0;
#pragma unroll
      for (uint32_t u = 0; u < 3; ++u) bits |= idx + u < b && test(v[u].x, v[u].y) ? 1u << u : 0u;
      g(bits, vis + (idx - a));
    }
    vis += b > a ? b - a : 0u;
  };
  for (uint32_t zs = win.zLowSector; zs < win.zHighSector; ++zs) {
    const uint32_t row = zs * (nPhi + 1);
    const uint32_t zSectorStart = g1[row];
    const uint32_t zSectorEnd = g1[row + nPhi];  // (zs+1)(nPhi+1)-1: padding cell excluded
    const uint32_t len = zSectorEnd - zSectorStart;
    const uint32_t A = g1[row + win.phiLowSector];
    const uint32_t H = g1[row + win.phiHighSector];
    const uint32_t E = win.wrap ? zSectorEnd + H - zSectorStart : H;
    // indices j in [A, E) map to j - (j >= zSectorEnd) * len
    scan_range(A, E < zSectorEnd ? E : zSectorEnd, h.zphi, in_window, f);
    if (E > zSectorEnd) {
      const uint32_t m = E - len;
      scan_range(zSectorStart, m < A ? m : A, h.zphi, in_window, f);
      if (m > zSectorEnd) scan_range(zSectorEnd, m, h.zphi, in_window, f);
    }
  }
}

struct PairItem {
  uint32_t k, p, i;
  const uint32_t* g1;
  PairWindow win;
};

// Outer hit of work item w (segment k from the LDS-staged work offsets).
__device__ __forceinline__ PairItem pair_item(const SeedParams& sp, const SortedHits& h,
                                              const uint32_t* __restrict__ grid,
                                              const uint32_t* offs, uint32_t w) {
  const uint32_t P = sp.n_triplets, K = sp.n_events * P, C = sp.cells, L = sp.n_layers;
  PairItem it;
  it.k = segment_of(offs, K + 1, w);
  const uint32_t e = it.k / P;
  it.p = it.k - e * P;
  const uint32_t* g2 = grid + ((uint64_t)e * L + (sp.layer2[it.p] - 1)) * C;
  const uint32_t l1 = sp.layer1[it.p] - 1;
  it.g1 = grid + ((uint64_t)e * L + l1) * C;
  it.i = g2[0] + (w - offs[it.k]);
  const float4 o = load_xyzphi(h.xyzphi, it.i);
  it.win = pair_window(sp, it.p, o.x, o.y, o.z, o.w, sp.min_radius[l1],
                       sp.max_radius[l1], sp.acos_min_r[l1], sp.acos_max_r[l1]);
  return it;
}

// Work item w is part w % P of outer hit w / P, P = sp.pair_parts: the z
// rows of the hit's window split into P consecutive runs, so the items'
// outputs in item order are the hit's pairs in walk order, and a window of
// many rows is walked by two lanes (the passes end with their longest walks).
__device__ __forceinline__ uint32_t pair_items(const SeedParams& sp, const uint32_t* work_off, uint32_t K,
                                               uint32_t items_max, uint32_t* flags) {
  const uint32_t parts = sp.pair_parts;
  uint32_t W = work_off[K] * parts;
  if (W > items_max) {  // more outer hits than the counts array holds: hits beyond max_hits
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(flags, TRAX_FLAG_LAYOUT);
    W = items_max - items_max % parts;
  }
  return W;
}

__device__ __forceinline__ PairItem pair_work_item(const SeedParams& sp, const SortedHits& h,
                                                   const uint32_t* __restrict__ grid, const uint32_t* offs,
                                                   uint32_t w) {
  const uint32_t parts = sp.pair_parts;
  const uint32_t hit = parts == 2u ? w >> 1 : w, part = w - hit * parts;
  PairItem it = pair_item(sp, h, grid, offs, hit);
  if (parts == 2u) {
    const uint32_t lo = it.win.zLowSector, hi = it.win.zHighSector;
    const uint32_t mid = hi > lo ? lo + (hi - lo) / 2u : lo;
    it.win.zLowSector = part ? mid : lo;
    it.win.zHighSector = part ? hi : mid;
  }
  return it;
}

// Count pass: one lane per work item walks its window.  With `masks`, bit
// j of masks[w] says whether the item's j-th visited hit (j < 64) passed,
// so the store pass takes those hits without loading and testing them again.
__global__ __launch_bounds__(512) void k_pairs_count(SeedParams sp, SortedHits h,
                                                     const uint32_t* __restrict__ grid,
                                                     uint32_t* __restrict__ work_off,
                                                     uint32_t* __restrict__ counts,
                                                     unsigned long long* __restrict__ masks,
                                                     uint32_t items_max, uint32_t* __restrict__ flags) {
  extern __shared__ uint32_t s_offs[];  // K + 1 offsets when they fit (offsets_lds_bytes)
  const uint32_t K = sp.n_events * sp.n_triplets;
  const uint32_t* offs = stage_offsets(work_off, K + 1, s_offs);
  const uint32_t W = pair_items(sp, work_off, K, items_max, flags);
  if (blockIdx.x == 0 && threadIdx.x == 0) work_off[K + 1] = W;  // the scan's length - 1
  const XcdShare sh = xcd_share(W, blockDim.x, threadIdx.x);
  for (uint32_t w = sh.begin; w < sh.end; w += sh.stride) {
    const PairItem it = pair_work_item(sp, h, grid, offs, w);
    uint32_t n = 0;
    unsigned long long m = 0;
    pair_walk(sp, it.g1, it.win, h, [&](uint32_t bits, uint32_t ord) {
      m |= ord < 64u ? (unsigned long long)bits << ord : 0ull;
      n += (uint32_t)__popc(bits);
    });
    counts[w] = n;
    if (masks) masks[w] = m;
  }
}

// Store pass with lane groups.  A wave's 64 work items are consecutive, and
// so are their output runs; they are taken by sixteen groups of kStoreGroup
// lanes, each group claiming items from the wave's counter and walking them
// one after the other with its lanes over consecutive inner hits of a range
// (claimed rather than four fixed items per group: 15 us less on PU200, the
// groups that draw short walks take more items).  A ballot ranks the group's
// passing hits, so the records go out in walk order as contiguous runs
// (measured HBM write traffic 1.12x the records, against 2.4x when every
// lane wrote its own run: partially written lines were evicted and written
// again) and the (z, phi) loads cover consecutive hits.  Each lane first
// evaluates its own item's window and parks it in LDS for its group.
// PU200: 253 us against 286 us for one lane per item (eight lanes: 259 us,
// sixteen: 302 us).
constexpr uint32_t kStoreGroup = 4;
static_assert(64u % kStoreGroup == 0 && kStoreGroup <= 32u, "groups tile the wave");

struct StoreWindow {
  float zLow, zHigh, phiLow, phiHigh;
  uint32_t wrap, zLo, zHi, phiLoS, phiHiS, outer, pos;
  uint32_t g1;  // first cell of the inner layer's grid (offset into the grid)
  uint32_t mlo, mhi, known;  // the count pass's mask of the first `known` visits
};

__global__ __launch_bounds__(256) void k_pairs_store(SeedParams sp, SortedHits h,
                                                     const uint32_t* __restrict__ grid,
                                                     const uint32_t* __restrict__ work_off,
                                                     const uint32_t* __restrict__ counts,  // scanned
                                                     const unsigned long long* __restrict__ masks,
                                                     uint2* __restrict__ pairs, uint64_t capacity,
                                                     uint32_t items_max, uint32_t* __restrict__ flags) {
  extern __shared__ uint32_t s_offs[];  // K + 1 offsets when they fit (offsets_lds_bytes)
  __shared__ StoreWindow s_win[4][64];
  __shared__ uint32_t s_next[4];  // the wave's next unclaimed item
  const uint32_t K = sp.n_events * sp.n_triplets;
  const uint32_t* offs = stage_offsets(work_off, K + 1, s_offs);
  const uint32_t W = pair_items(sp, work_off, K, items_max, flags);
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  const uint32_t g = lane / kStoreGroup, sub = lane % kStoreGroup;
  const uint32_t gmask = (1u << kStoreGroup) - 1u, below = (1u << sub) - 1u;
  const uint32_t nPhi = sp.n_phi;
  const XcdShare sh = xcd_share(W, blockDim.x, threadIdx.x);
  bool overflow = false;
  StoreWindow* wins = s_win[wave];
  // the wave's rounds: items [wb, wb + 64), wb the same for all its lanes
  for (uint32_t wb = sh.begin - lane; wb < sh.end; wb += sh.stride) {
    const uint32_t w = wb + lane;
    StoreWindow sw{};
    if (w < sh.end) {
      const PairItem it = pair_work_item(sp, h, grid, offs, w);
      sw.zLow = it.win.zLow, sw.zHigh = it.win.zHigh, sw.phiLow = it.win.phiLow, sw.phiHigh = it.win.phiHigh;
      sw.wrap = it.win.wrap ? 1u : 0u;
      sw.zLo = it.win.zLowSector, sw.zHi = it.win.zHighSector;  // zLo == zHi: nothing to walk
      sw.phiLoS = it.win.phiLowSector, sw.phiHiS = it.win.phiHighSector;
      sw.outer = it.i;
      sw.pos = counts[w];
      sw.g1 = (uint32_t)(it.g1 - grid);
      if (masks) {
        const unsigned long long m = masks[w];
        sw.mlo = (uint32_t)m, sw.mhi = (uint32_t)(m >> 32), sw.known = 64u;
      }
    }
    wins[lane] = sw;
    if (lane == 0) s_next[wave] = 0u;
    wave_lds_sync();
    // groups claim the wave's items one at a time, so a group that drew
    // short walks takes more of them
    while (true) {
      uint32_t item = 0;
      if (sub == 0) item = atomicAdd(&s_next[wave], 1u);
      item = __shfl(item, (int)(g * kStoreGroup), 64);
      if (item >= 64u) break;
      const StoreWindow pw = wins[item];
      const uint32_t* g1 = grid + pw.g1;
      const bool wrap = pw.wrap != 0u;
      uint64_t pos = pw.pos;
      const unsigned long long known_mask = ((unsigned long long)pw.mhi << 32) | pw.mlo;
      uint32_t vis = 0;  // hits visited before the current range
      // inner hits [lo, hi) in index order, kStoreGroup at a time
      auto range = [&](uint32_t lo, uint32_t hi) {
        for (uint32_t i0 = lo; i0 < hi; i0 += kStoreGroup) {
          const uint32_t idx = i0 + sub;
          const uint32_t ord = vis + (idx - lo);
          bool ok = false;
          if (idx < hi) {
            if (ord < pw.known) {
              ok = (known_mask >> ord) & 1ull;
            } else {
              const float2 v = load_zphi(h.zphi, idx);
              ok = pw.zLow <= v.x && v.x <= pw.zHigh && window_phi_ok(wrap, pw.phiLow, pw.phiHigh, v.y);
            }
          }
          const uint32_t gb = (uint32_t)(__ballot(ok) >> (g * kStoreGroup)) & gmask;
          if (ok) {
            const uint64_t d = pos + (uint32_t)__popc(gb & below);
            if (d < capacity)
              pairs[d] = make_uint2(idx, pw.outer);
            else
              overflow = true;
          }
          pos += (uint32_t)__popc(gb);
        }
        vis += hi > lo ? hi - lo : 0u;
      };
      // the rows and ranges of pair_walk, in the same order
      for (uint32_t zs = pw.zLo; zs < pw.zHi; ++zs) {
        const uint32_t row = zs * (nPhi + 1);
        const uint32_t zSectorStart = g1[row];
        const uint32_t zSectorEnd = g1[row + nPhi];  // padding cell excluded
        const uint32_t len = zSectorEnd - zSectorStart;
        const uint32_t A = g1[row + pw.phiLoS];
        const uint32_t H = g1[row + pw.phiHiS];
        const uint32_t E = wrap ? zSectorEnd + H - zSectorStart : H;
        range(A, E < zSectorEnd ? E : zSectorEnd);
        if (E > zSectorEnd) {
          const uint32_t m = E - len;
          range(zSectorStart, m < A ? m : A);
          if (m > zSectorEnd) range(zSectorEnd, m);
        }
      }
    }
    wave_lds_sync();  // the windows are rewritten for the wave's next items
  }
  if (overflow) atomicOr(flags, TRAX_FLAG_PAIR_OVERFLOW);
}

// out[k] = scanned[min(idx[k] * mul, last)] for k in [0, n); total -> *d_total.
__global__ __launch_bounds__(256) void k_gather_offsets(const uint32_t* __restrict__ scanned,
                                                        uint32_t last,
                                                        const uint32_t* __restrict__ last_dev,
                                                        const uint32_t* __restrict__ idx, uint32_t n,
                                                        uint32_t mul, uint32_t* __restrict__ out,
                                                        uint32_t* __restrict__ d_total) {
  uint32_t lim = last;
  if (last_dev) lim = *last_dev < last ? *last_dev : last;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    uint32_t j = idx[k] * mul;
    j = j < lim ? j : lim;
    uint32_t v = scanned[j];
    out[k] = v;
    if (k == n - 1 && d_total) *d_total = v;
  }
}

// *out = arr[*d_index - 1]: the last element of a scan whose length is
// only known on the device.
__global__ void k_read_last(const uint32_t* __restrict__ arr, const uint32_t* __restrict__ d_index,
                            uint32_t* __restrict__ out) {
  if (threadIdx.x == 0) *out = arr[*d_index - 1u];
}

hipError_t launch_read_last(const uint32_t* arr, const uint32_t* d_index, uint32_t* out,
                            hipStream_t stream) {
  hipLaunchKernelGGL(k_read_last, dim3(1), dim3(64), 0, stream, arr, d_index, out);
  return hipGetLastError();
}

hipError_t launch_pair_plan(const SeedParams& sp, const uint32_t* grid, uint32_t* work_off,
                            ScanCtrl* ctrl, unsigned long long* status, int scan_sub, hipStream_t stream,
                            bool inner) {
  const uint32_t K = sp.n_events * sp.n_triplets;
  if (K <= kPlanOneBlock) {  // a few block-wide scan rounds
    hipLaunchKernelGGL(k_pair_plan, dim3(1), dim3(256), 0, stream, sp, grid, work_off, inner);
    return hipGetLastError();
  }
  const uint32_t blocks = (K + 1u + 255u) / 256u;
  hipLaunchKernelGGL(k_pair_counts, dim3(blocks < 4096u ? blocks : 4096u), dim3(256), 0, stream, sp, grid,
                     work_off, inner);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_scan_u32(work_off, work_off, (uint64_t)K + 1u, nullptr, 0, ctrl, status, stream, scan_sub);
}

hipError_t launch_pairs(const SeedParams& sp, const SortedHits& h, const uint32_t* grid,
                        uint32_t* work_off, uint32_t* counts, unsigned long long* masks, uint2* pairs,
                        uint64_t capacity, uint32_t items_max, bool store, uint32_t* flags, uint32_t block,
                        hipStream_t stream) {
  if (store) {  // 256-thread blocks (the LDS windows are laid out for four waves)
    uint32_t blocks = (uint32_t)(((uint64_t)items_max + 255u) / 256u);
    blocks = xcd_grid(blocks == 0 ? 1u : (blocks > 16384u ? 16384u : blocks));
    hipLaunchKernelGGL(k_pairs_store, dim3(blocks), dim3(256), offsets_lds_bytes(sp), stream, sp, h, grid, work_off,
                       counts, masks, pairs, capacity, items_max, flags);
  } else {
    uint32_t blocks = (uint32_t)(((uint64_t)items_max + block - 1) / block);
    blocks = xcd_grid(blocks == 0 ? 1u : (blocks > 16384u ? 16384u : blocks));
    hipLaunchKernelGGL(k_pairs_count, dim3(blocks), dim3(block), offsets_lds_bytes(sp), stream, sp, h, grid, work_off,
                       counts, masks, items_max, flags);
  }
  return hipGetLastError();
}


hipError_t launch_gather_offsets(const uint32_t* scanned, uint32_t last, const uint32_t* last_dev,
                                 const uint32_t* idx, uint32_t n, uint32_t* out, uint32_t* d_total,
                                 hipStream_t stream, uint32_t mul) {
  uint32_t blocks = (n + 255u) / 256u;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_gather_offsets, dim3(blocks), dim3(256), 0, stream, scanned, last, last_dev,
                     idx, n, mul, out, d_total);
  return hipGetLastError();
}

}  // namespace trax
