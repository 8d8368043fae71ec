// Triplet filter and compaction.
//
// Reference: src/algorithms/TripletThetaPhiFilter.h:50-313 and
// TripletThetaPhiFilter.cpp:10-164: filterCount sets one bit per accepted
// candidate, filterPopCount counts bits per 32-bit word, PrefixSum scans the
// counts, the total is read back to the host, filterStore compacts the
// accepted candidates in candidate order into three hit-index columns, and
// filterOffsetStore plus a racy monotonize pass derive per-(event, layer
// triplet) offsets.
//
// Here: a mask pass (the cuts, one bit per candidate, a popcount per 64-bit
// word), the exclusive scan of the word counts, and a store pass that
// writes the accepted candidates' (hit1, hit2, hit3) columns in candidate
// order and derives the per-segment offsets from the scanned counts and the
// mask (no monotonize pass needed).  The cuts are the reference's
// expressions in its order.
#include "common.h"
#include "cuts.h"
#include "kernels.h"

namespace trax {

// Pass 1 (filterCount + filterPopCount, TripletThetaPhiFilter.h:50-166):
// each wave takes 256 candidates = 4 mask words.  The phi cut runs on all of
// them; its survivors are queued densely in the wave's LDS slice for the
// theta cut, whose survivors are queued again for the circle fit, so the
// costlier stages run on full waves.  Each wave writes its 4 mask words
// (bit = accepted) and their popcounts.  No block barrier and no cross-block
// ordering: waves run at their own pace.
constexpr uint32_t kFilterWords = 4;  // mask words per wave task

// One filter stage over a wave's dense LDS queue of candidate numbers
// (0..255 within the task): keeps those that pass `cut`, compacted in order
// into `out`; returns how many.
template <typename Cut>
__device__ __forceinline__ uint32_t filter_stage(const uint8_t* in, uint32_t n_in, uint8_t* out,
                                                 Cut&& cut) {
  uint32_t n_out = 0;
  for (uint32_t j0 = 0; j0 < n_in; j0 += 64u) {
    const uint32_t j = j0 + lane_id();
    uint8_t e = 0;
    bool ok = false;
    if (j < n_in) {
      e = in[j];
      ok = cut(e);
    }
    const uint64_t m = __ballot(ok);
    if (ok) out[n_out + lanes_below(m)] = e;
    n_out += (uint32_t)__popcll(m);
  }
  wave_lds_sync();
  return n_out;
}

__global__ __launch_bounds__(256) void k_filter_mask(SeedParams sp, SortedHits h,
                                                     const uint2* __restrict__ pairs,
                                                     const uint2* __restrict__ cands,
                                                     const uint32_t* __restrict__ cand_off,
                                                     const uint32_t* __restrict__ d_ncands,
                                                     uint64_t cand_capacity,
                                                     unsigned long long* __restrict__ mask,
                                                     uint32_t* __restrict__ counts,
                                                     uint32_t* __restrict__ d_nwords) {
  __shared__ uint8_t s_q1[4][64 * kFilterWords];
  __shared__ uint8_t s_q2[4][64 * kFilterWords];
  __shared__ uint8_t s_pass[4][64 * kFilterWords];
  extern __shared__ uint32_t s_offs[];  // K + 1 offsets when they fit (offsets_lds_bytes)
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  const uint32_t P = sp.n_triplets, K = sp.n_events * P;
  const uint32_t* offs = stage_offsets(cand_off, K + 1, s_offs);
  const uint64_t nc64 = *d_ncands;
  const uint32_t nC = (uint32_t)(nc64 < cand_capacity ? nc64 : cand_capacity);
  const uint32_t nW = (nC + 63u) / 64u;
  // words [0, nW] are written (word nW is the zero that makes the exclusive
  // scan's last element the total)
  const uint32_t ntasks = (nW + 1u + kFilterWords - 1u) / kFilterWords;
  if (blockIdx.x == 0 && threadIdx.x == 0) *d_nwords = nW + 1u;
  const uint32_t wstride = gridDim.x * 4u;
  uint8_t* q1 = s_q1[wave];
  uint8_t* q2 = s_q2[wave];
  uint8_t* pass = s_pass[wave];
  for (uint32_t task = blockIdx.x * 4u + wave; task < ntasks; task += wstride) {
    const uint32_t c0 = task * 64u * kFilterWords;
    // stage 1: phi cut on every candidate of the task
    uint32_t nq = 0;
#pragma unroll
    for (uint32_t s = 0; s < kFilterWords; ++s) {
      const uint32_t c = c0 + s * 64u + lane;
      pass[s * 64u + lane] = 0;
      bool keep = false;
      if (c < nC) {
        const uint2 cd = cands[c];
        if (cd.x != TRAX_UNWRITTEN) {
          const uint2 pr = pairs[cd.x];
          const uint32_t lt = segment_of(offs, K + 1, c) % P;
          const float4 p1 = load_xyzphi(h.xyzphi, pr.x), p2 = load_xyzphi(h.xyzphi, pr.y),
                       p3 = load_xyzphi(h.xyzphi, cd.y);
          keep = filter_phi(sp, lt, p1.x, p1.y, p2.x, p2.y, p3.x, p3.y);
        }
      }
      const uint64_t m = __ballot(keep);
      if (keep) q1[nq + lanes_below(m)] = (uint8_t)(s * 64u + lane);
      nq += (uint32_t)__popcll(m);
    }
    wave_lds_sync();
    // stage 2: theta-ratio cut on the phi survivors
    nq = filter_stage(q1, nq, q2, [&](uint32_t e) {
      const uint32_t c = c0 + e;
      const uint2 cd = cands[c];
      const uint2 pr = pairs[cd.x];
      const uint32_t lt = segment_of(offs, K + 1, c) % P;
      const float4 p1 = load_xyzphi(h.xyzphi, pr.x), p2 = load_xyzphi(h.xyzphi, pr.y),
                   p3 = load_xyzphi(h.xyzphi, cd.y);
      return filter_theta(sp, lt, p1.x, p1.y, p1.z, p2.x, p2.y, p2.z, p3.x, p3.y, p3.z);
    });
    // stage 3: circle fit on the survivors of both angle cuts
    for (uint32_t j = lane; j < nq; j += 64u) {
      const uint32_t e = q2[j];
      const uint32_t c = c0 + e;
      const uint2 cd = cands[c];
      const uint2 pr = pairs[cd.x];
      const uint32_t lt = segment_of(offs, K + 1, c) % P;
      const float4 p1 = load_xyzphi(h.xyzphi, pr.x), p2 = load_xyzphi(h.xyzphi, pr.y),
                   p3 = load_xyzphi(h.xyzphi, cd.y);
      pass[e] = filter_circle(sp, lt, p1.x, p1.y, p2.x, p2.y, p3.x, p3.y) ? 1 : 0;
    }
    wave_lds_sync();
#pragma unroll
    for (uint32_t s = 0; s < kFilterWords; ++s) {
      const uint64_t m = __ballot(pass[s * 64u + lane] != 0);
      if (lane == s) {
        mask[task * kFilterWords + s] = m;
        counts[task * kFilterWords + s] = (uint32_t)__popcll(m);
      }
    }
    wave_lds_sync();  // the queues and pass flags are reused by the wave's next task
  }
}

// Pass 2 (filterStore + filterOffsetStore, TripletThetaPhiFilter.h:168-313):
// `counts` now holds the exclusive scan of the per-word popcounts.  One lane
// per mask word writes its accepted candidates' (hit1, hit2, hit3) in
// candidate order; the per-(event, layer triplet) output offsets are the
// scanned count at the segment's first candidate's word plus the accepted
// bits below it in that word.
__global__ __launch_bounds__(256) void k_filter_store(SeedParams sp, const uint2* __restrict__ pairs,
                                                      const uint2* __restrict__ cands,
                                                      const uint32_t* __restrict__ cand_off,
                                                      const uint32_t* __restrict__ d_ncands,
                                                      uint64_t cand_capacity,
                                                      const unsigned long long* __restrict__ mask,
                                                      const uint32_t* __restrict__ scanned,
                                                      uint32_t* __restrict__ t1,
                                                      uint32_t* __restrict__ t2,
                                                      uint32_t* __restrict__ t3,
                                                      uint64_t trk_capacity,
                                                      uint32_t* __restrict__ trk_off,
                                                      uint32_t* __restrict__ d_ntrk,
                                                      uint32_t* __restrict__ flags) {
  const uint32_t K = sp.n_events * sp.n_triplets;
  const uint64_t nc64 = *d_ncands;
  const uint32_t nC = (uint32_t)(nc64 < cand_capacity ? nc64 : cand_capacity);
  const uint32_t nW = (nC + 63u) / 64u;
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t first = blockIdx.x * blockDim.x + threadIdx.x;
  bool overflow = false;
  // one lane per 16 candidates: four lanes share a mask word, each starts at
  // the word's scanned count plus the bits below its quarter, so a dense word
  // is written by four short chains of dependent loads
  for (uint32_t u = first; u < 4u * nW; u += stride) {
    const uint32_t w = u >> 2, sh = (u & 3u) * 16u;
    const uint64_t word = mask[w];
    uint64_t m = (word >> sh) & 0xFFFFull;
    uint64_t pos = scanned[w] + (uint32_t)__popcll(word & ((1ull << sh) - 1ull));
    while (m) {
      const uint32_t b = (uint32_t)__builtin_ctzll(m) + sh;
      m &= m - 1;
      const uint2 cd = cands[w * 64u + b];
      const uint2 pr = pairs[cd.x];
      if (pos < trk_capacity) {
        t1[pos] = pr.x;
        t2[pos] = pr.y;
        t3[pos] = cd.y;
      } else {
        overflow = true;
      }
      ++pos;
    }
  }
  for (uint32_t k = first; k <= K; k += stride) {
    uint32_t c = cand_off[k];
    c = c < nC ? c : nC;
    const uint32_t w = c >> 6, b = c & 63u;
    const uint64_t below = mask[w] & ((1ull << b) - 1ull);
    trk_off[k] = scanned[w] + (uint32_t)__popcll(below);
  }
  if (first == 0) *d_ntrk = scanned[nW];
  if (overflow) atomicOr(flags, TRAX_FLAG_TRACKLET_OVERFLOW);
}

uint64_t filter_mask_words(uint64_t cand_capacity) { return (cand_capacity + 63u) / 64u + 8u; }

hipError_t launch_filter_mask(const SeedParams& sp, const SortedHits& h, const uint2* pairs,
                              const uint2* cands, const uint32_t* cand_off, const uint32_t* d_ncands,
                              uint64_t cand_capacity, unsigned long long* mask, uint32_t* counts,
                              uint32_t* d_nwords, hipStream_t stream) {
  const uint64_t tasks = ((cand_capacity + 63u) / 64u + kFilterWords) / kFilterWords;
  uint64_t b = (tasks + 3u) / 4u;
  if (b == 0) b = 1;
  const uint32_t blocks = (uint32_t)(b < 8192u ? b : 8192u);
  hipLaunchKernelGGL(k_filter_mask, dim3(blocks), dim3(256), offsets_lds_bytes(sp), stream, sp, h,
                     pairs, cands, cand_off, d_ncands, cand_capacity, mask, counts, d_nwords);
  return hipGetLastError();
}

hipError_t launch_filter_store(const SeedParams& sp, const uint2* pairs, const uint2* cands,
                               const uint32_t* cand_off, const uint32_t* d_ncands,
                               uint64_t cand_capacity, const unsigned long long* mask,
                               const uint32_t* scanned, uint32_t* t1, uint32_t* t2, uint32_t* t3,
                               uint64_t trk_capacity, uint32_t* trk_off, uint32_t* d_ntrk,
                               uint32_t* flags, hipStream_t stream) {
  const uint64_t quarters = 4u * ((cand_capacity + 63u) / 64u);
  const uint64_t K = (uint64_t)sp.n_events * sp.n_triplets + 1u;
  const uint64_t n = quarters > K ? quarters : K;
  uint64_t b = (n + 255u) / 256u;
  if (b == 0) b = 1;
  const uint32_t blocks = (uint32_t)(b < 4096u ? b : 4096u);
  hipLaunchKernelGGL(k_filter_store, dim3(blocks), dim3(256), 0, stream, sp, pairs, cands, cand_off,
                     d_ncands, cand_capacity, mask, scanned, t1, t2, t3, trk_capacity, trk_off, d_ntrk,
                     flags);
  return hipGetLastError();
}

}  // namespace trax
