// Triplet filter and compaction in one pass.
//
// Reference: src/algorithms/TripletThetaPhiFilter.h:50-313 and
// TripletThetaPhiFilter.cpp:10-164: filterCount sets one bit per accepted
// candidate, filterPopCount counts bits per 32-bit word, PrefixSum scans the
// counts, the total is read back to the host, filterStore compacts the
// accepted candidates in candidate order into three hit-index columns, and
// filterOffsetStore plus a racy monotonize pass derive per-(event, layer
// triplet) offsets.
//
// Here one kernel does it: each wave takes the next tile of 256 candidates
// (ticketed, so tiles resolve in order), evaluates the angle cuts, queues the
// survivors (about 7%) in LDS and runs the circle fit (TIP and radius cuts)
// on that dense queue, ranks the accepted candidates with wave64 ballots,
// resolves the tile's output base by decoupled look-back over its
// predecessors, and writes the (hit1, hit2, hit3) columns in candidate order.
// All expressions are the reference's, in its order.  The lane that owns the
// first candidate of an (event, layer triplet) segment writes that segment's
// output offset.
#include "common.h"
#include "kernels.h"

namespace trax {

// filterCount predicate, TripletThetaPhiFilter.h:68-147, split in two: the
// theta-ratio and phi cuts (cheap, most candidates stop here) and the
// circle-fit TIP and radius cuts.  The reference evaluates all of them and
// ANDs the results; the cuts are pure functions, so the product is the same.
__device__ __forceinline__ bool filter_angles(const SeedParams& sp, uint32_t lt, float x1, float y1,
                                              float z1, float x2, float y2, float z2, float x3,
                                              float y3, float z3) {
  const float dThetaCut = sp.dtheta_cut[lt], dPhiCut = sp.dphi_cut[lt];
  const float PI = TRAX_PI_F;
  bool valid = true;
  float angle1 = trax_atan2f(__builtin_sqrtf((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1)), (z2 - z1));
  float angle2 = trax_atan2f(__builtin_sqrtf((x3 - x2) * (x3 - x2) + (y3 - y2) * (y3 - y2)), (z3 - z2));
  float delta = __builtin_fabsf(angle2 / angle1);
  valid = valid && (1 - dThetaCut <= delta && delta <= 1 + dThetaCut);

  angle1 = trax_atan2f((y2 - y1), (x2 - x1));
  angle2 = trax_atan2f((y3 - y2), (x3 - x2));
  delta = angle2 - angle1;
  delta += (delta > PI) ? -2 * PI : (delta < -PI) ? 2 * PI : 0;
  valid = valid && (__builtin_fabsf(delta) <= dPhiCut);
  return valid;
}

// Circle through the three hits from the paraboloid (x, y, x^2 + y^2)
// plane: transverse impact parameter and radius cuts (:105-147).
__device__ __forceinline__ bool filter_circle(const SeedParams& sp, uint32_t lt, float x1, float y1,
                                              float x2, float y2, float x3, float y3) {
  const float tipCut = sp.tip_cut[lt];
  const float minRadius = sp.min_radius_curvature;
  const float p1x = x1, p1y = y1, p1z = x1 * x1 + y1 * y1;
  const float p2x = x2, p2y = y2, p2z = x2 * x2 + y2 * y2;
  const float p3x = x3, p3y = y3, p3z = x3 * x3 + y3 * y3;
  const float ax = p2x - p1x, ay = p2y - p1y, az = p2z - p1z;
  const float bx = p3x - p1x, by = p3y - p1y, bz = p3z - p1z;
  float nx = ay * bz - az * by;
  float ny = az * bx - ax * bz;
  float nz = ax * by - ay * bx;
  const float nlen = __builtin_sqrtf(nx * nx + ny * ny + nz * nz);
  if (nlen != 0.0f) {
    nx = nx / nlen;
    ny = ny / nlen;
    nz = nz / nlen;
  }
  const float cOx = (-nx) / (2 * nz);
  const float cOy = (-ny) / (2 * nz);
  const float c = -(nx * p1x + ny * p1y + nz * p1z);
  const float cR = __builtin_sqrtf((1 - nz * nz - 4 * c * nz) / (4 * nz * nz));
  float vx = -cOx, vy = -cOy;
  const float vlen = __builtin_sqrtf(vx * vx + vy * vy);
  if (vlen != 0.0f) {
    vx = vx / vlen;
    vy = vy / vlen;
  }
  const float pCAx = cOx + cR * vx;
  const float pCAy = cOy + cR * vy;
  const float tip = pCAx * pCAx + pCAy * pCAy;
  return (tip <= (tipCut * tipCut)) && (cR >= minRadius);
}

constexpr uint32_t kFilterSub = 4;
constexpr uint32_t kFilterTile = 64u * kFilterSub;  // candidates per wave tile

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__global__ __launch_bounds__(256) void k_filter(SeedParams sp, SortedHits h,
                                                const uint2* __restrict__ pairs,
                                                const uint2* __restrict__ cands,
                                                const uint32_t* __restrict__ cand_off,
                                                const uint32_t* __restrict__ d_ncands,
                                                uint64_t cand_capacity, uint32_t* __restrict__ t1,
                                                uint32_t* __restrict__ t2, uint32_t* __restrict__ t3,
                                                uint64_t trk_capacity,
                                                uint32_t* __restrict__ trk_off,
                                                uint32_t* __restrict__ d_ntrk, ScanCtrl* ctrl,
                                                unsigned long long* status,
                                                uint32_t* __restrict__ flags) {
  extern __shared__ uint32_t s_offs[];  // K + 1 offsets when they fit (offsets_lds_bytes)
  __shared__ uint16_t s_queue[4][kFilterTile];
  __shared__ uint8_t s_pass[4][kFilterTile];
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  const uint32_t P = sp.n_triplets, K = sp.n_events * P;
  const uint32_t epoch = launch_epoch(ctrl);
  const uint32_t* offs = stage_offsets(cand_off, K + 1, s_offs);
  const uint64_t nc64 = *d_ncands;
  const uint32_t nC = (uint32_t)(nc64 < cand_capacity ? nc64 : cand_capacity);
  uint32_t ntiles = (nC + kFilterTile - 1) / kFilterTile;
  if (ntiles == 0) ntiles = 1;  // tile 0 always runs so the offsets get written
  bool overflow = false;
  while (true) {  // wave tiles of 256 candidates (candidate = tile * 256 + s * 64 + lane)
    const uint32_t tile = wave_ticket(ctrl);
    if (tile >= ntiles) break;
    // stage 1: angle cuts; survivors queued in this wave's LDS slice
    bool keep[kFilterSub];
    uint32_t hh1[kFilterSub], hh2[kFilterSub], hh3[kFilterSub];
    uint32_t nq = 0;
#pragma unroll
    for (uint32_t s = 0; s < kFilterSub; ++s) {
      const uint32_t c = tile * kFilterTile + s * 64u + lane;
      keep[s] = false;
      hh1[s] = hh2[s] = hh3[s] = 0;
      if (c < nC) {
        const uint2 cd = cands[c];
        if (cd.x != TRAX_UNWRITTEN) {
          const uint2 pr = pairs[cd.x];
          hh1[s] = pr.x;
          hh2[s] = pr.y;
          hh3[s] = cd.y;
          const uint32_t lt = segment_of(offs, K + 1, c) % P;
          keep[s] = filter_angles(sp, lt, h.x[pr.x], h.y[pr.x], h.z[pr.x], h.x[pr.y], h.y[pr.y],
                                  h.z[pr.y], h.x[cd.y], h.y[cd.y], h.z[cd.y]);
        }
      }
      const uint64_t m = __ballot(keep[s]);
      if (keep[s]) s_queue[wave][nq + lanes_below(m)] = (uint16_t)(s * 64u + lane);
      nq += (uint32_t)__popcll(m);
    }
    wave_lds_sync();
    // stage 2: circle fit on the dense queue of survivors
    for (uint32_t j = lane; j < nq; j += 64u) {
      const uint32_t e = s_queue[wave][j];
      const uint32_t c = tile * kFilterTile + e;
      const uint2 cd = cands[c];
      const uint2 pr = pairs[cd.x];
      const uint32_t lt = segment_of(offs, K + 1, c) % P;
      s_pass[wave][e] = filter_circle(sp, lt, h.x[pr.x], h.y[pr.x], h.x[pr.y], h.y[pr.y], h.x[cd.y],
                                      h.y[cd.y]) ? 1 : 0;
    }
    wave_lds_sync();
    // stage 3: ordered ranks, wave look-back, store
    uint32_t local[kFilterSub], running = 0;
#pragma unroll
    for (uint32_t s = 0; s < kFilterSub; ++s) {
      keep[s] = keep[s] && s_pass[wave][s * 64u + lane];
      const uint64_t m = __ballot(keep[s]);
      local[s] = running + lanes_below(m);
      running += (uint32_t)__popcll(m);
    }
    const TileScan ts = wave_scan_lookback(lane == 0 ? running : 0u, tile, epoch, status);
    const uint32_t base = ts.base;
#pragma unroll
    for (uint32_t s = 0; s < kFilterSub; ++s) {
      const uint32_t c = tile * kFilterTile + s * 64u + lane;
      const uint64_t pos = (uint64_t)base + local[s];
      if (keep[s]) {
        if (pos < trk_capacity) {
          t1[pos] = hh1[s];
          t2[pos] = hh2[s];
          t3[pos] = hh3[s];
        } else {
          overflow = true;
        }
      }
      // segments starting at candidate c: their output offset is this position
      if (c < nC) write_segment_starts(offs, K, segment_of(offs, K + 1, c), c, (uint32_t)pos, trk_off);
    }
    if (tile == ntiles - 1 && lane == 0) {
      const uint32_t total = base + running;
      for (int64_t k = K; k >= 0 && offs[k] >= nC; --k) trk_off[k] = total;
      *d_ntrk = total;
    }
    wave_lds_sync();  // the queue is reused by the next tile
  }
  if (overflow) atomicOr(flags, TRAX_FLAG_TRACKLET_OVERFLOW);
  __syncthreads();
  launch_done(ctrl, epoch);
}

hipError_t launch_filter(const SeedParams& sp, const SortedHits& h, const uint2* pairs,
                         const uint2* cands, const uint32_t* cand_off, const uint32_t* d_ncands,
                         uint64_t cand_capacity, uint32_t* t1, uint32_t* t2, uint32_t* t3,
                         uint64_t trk_capacity, uint32_t* trk_off, uint32_t* d_ntrk, ScanCtrl* ctrl,
                         unsigned long long* status, uint32_t* flags, hipStream_t stream) {
  uint64_t b = (cand_capacity + kFilterTile - 1) / kFilterTile;
  if (b == 0) b = 1;
  const uint32_t blocks = (uint32_t)(b < kPersistentBlocks ? b : kPersistentBlocks);
  hipLaunchKernelGGL(k_filter, dim3(blocks), dim3(256), offsets_lds_bytes(sp), stream, sp, h, pairs, cands, cand_off,
                     d_ncands, cand_capacity, t1, t2, t3, trk_capacity, trk_off, d_ntrk, ctrl,
                     status, flags);
  return hipGetLastError();
}

}  // namespace trax
