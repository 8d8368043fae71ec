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
// Here one kernel does it: each 256-thread block takes the next tile of 1024
// candidates (ticketed, so tiles resolve in order), evaluates the cuts
// (theta ratio, phi difference, circle-fit transverse impact parameter and
// radius; the reference's expressions in its order), ranks the accepted ones
// with wave64 ballots, resolves the tile's output base by decoupled look-back
// over its predecessors, and writes the (hit1, hit2, hit3) columns in
// candidate order.  The thread that owns the first candidate of an
// (event, layer triplet) segment writes that segment's output offset.
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
constexpr uint32_t kFilterTile = 256u * kFilterSub;

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
  __shared__ uint32_t s_tile, s_epoch, s_excl;
  __shared__ uint32_t s_cnt[kFilterSub][4];
  __shared__ uint32_t s_nq;
  __shared__ uint16_t s_queue[kFilterTile];
  __shared__ uint8_t s_pass[kFilterTile];
  extern __shared__ uint32_t s_offs[];  // K + 1 offsets when they fit (offsets_lds_bytes)
  const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  const uint32_t P = sp.n_triplets, K = sp.n_events * P;
  const uint32_t* offs = stage_offsets(cand_off, K + 1, s_offs);
  if (tid == 0) s_epoch = __hip_atomic_load(&ctrl->epoch, __ATOMIC_RELAXED, TRAX_AGENT) & 0x3FFFFFFFu;
  const uint64_t nc64 = *d_ncands;
  const uint32_t nC = (uint32_t)(nc64 < cand_capacity ? nc64 : cand_capacity);
  uint32_t ntiles = (nC + kFilterTile - 1) / kFilterTile;
  if (ntiles == 0) ntiles = 1;  // tile 0 always runs so the offsets get written
  while (true) {
    if (tid == 0) s_tile = __hip_atomic_fetch_add(&ctrl->ticket, 1u, __ATOMIC_RELAXED, TRAX_AGENT);
    __syncthreads();
    const uint32_t tile = s_tile, epoch = s_epoch;
    if (tile >= ntiles) break;
    // stage 1: angle cuts for the tile's 1024 candidates; survivors queued in LDS
    bool keep[kFilterSub];
    uint32_t hh1[kFilterSub], hh2[kFilterSub], hh3[kFilterSub], lts[kFilterSub];
    if (threadIdx.x == 0) s_nq = 0;
    __syncthreads();
#pragma unroll
    for (uint32_t s = 0; s < kFilterSub; ++s) {
      const uint32_t c = tile * kFilterTile + s * 256u + tid;
      keep[s] = false;
      hh1[s] = hh2[s] = hh3[s] = 0;
      lts[s] = 0;
      if (c < nC) {
        const uint2 cd = cands[c];
        if (cd.x != TRAX_UNWRITTEN) {
          const uint2 pr = pairs[cd.x];
          hh1[s] = pr.x;
          hh2[s] = pr.y;
          hh3[s] = cd.y;
          lts[s] = segment_of(offs, K + 1, c) % P;
          keep[s] = filter_angles(sp, lts[s], h.x[pr.x], h.y[pr.x], h.z[pr.x], h.x[pr.y], h.y[pr.y],
                                  h.z[pr.y], h.x[cd.y], h.y[cd.y], h.z[cd.y]);
        }
      }
      const uint64_t m = __ballot(keep[s]);
      uint32_t base = 0;
      if (lane == 0 && m) base = atomicAdd(&s_nq, (uint32_t)__popcll(m));
      base = __shfl(base, 0, 64);
      if (keep[s]) s_queue[base + lanes_below(m)] = (s << 8) | tid;
    }
    __syncthreads();
    // stage 2: circle fit on the dense queue of survivors
    const uint32_t nq = s_nq;
    for (uint32_t j = tid; j < nq; j += 256u) {
      const uint32_t e = s_queue[j], s = e >> 8, t = e & 255u;
      const uint32_t c = tile * kFilterTile + s * 256u + t;
      const uint2 cd = cands[c];
      const uint2 pr = pairs[cd.x];
      const uint32_t lt = segment_of(offs, K + 1, c) % P;
      s_pass[s * 256u + t] = filter_circle(sp, lt, h.x[pr.x], h.y[pr.x], h.x[pr.y], h.y[pr.y],
                                           h.x[cd.y], h.y[cd.y]) ? 1 : 0;
    }
    __syncthreads();
    uint32_t rank[kFilterSub];
#pragma unroll
    for (uint32_t s = 0; s < kFilterSub; ++s) {
      keep[s] = keep[s] && s_pass[s * 256u + tid];
      const uint64_t m = __ballot(keep[s]);
      rank[s] = lanes_below(m);
      if (lane == 0) s_cnt[s][wave] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    uint32_t running = 0, local[kFilterSub];
#pragma unroll
    for (uint32_t s = 0; s < kFilterSub; ++s) {
      uint32_t wp = 0, tot = 0;
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        const uint32_t cw = s_cnt[s][w];
        wp += (w < (int)wave) ? cw : 0u;
        tot += cw;
      }
      local[s] = running + wp + rank[s];
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
    bool overflow = false;
#pragma unroll
    for (uint32_t s = 0; s < kFilterSub; ++s) {
      const uint32_t c = tile * kFilterTile + s * 256u + tid;
      const uint64_t pos = (uint64_t)excl_tile + local[s];
      if (keep[s]) {
        if (pos < trk_capacity) {
          t1[pos] = hh1[s];
          t2[pos] = hh2[s];
          t3[pos] = hh3[s];
        } else {
          overflow = true;
        }
      }
      if (c < nC) {
        // segment starts at candidate c: their output offset is this position
        uint32_t k = segment_of(offs, K + 1, c);
        while (offs[k] == c) {
          trk_off[k] = (uint32_t)pos;
          if (k == 0) break;
          --k;
        }
      }
    }
    if (overflow) atomicOr(flags, TRAX_FLAG_TRACKLET_OVERFLOW);
    if (tile == ntiles - 1 && tid == 0) {
      const uint32_t total = excl_tile + aggregate;
      // segments that start at or after the last candidate (empty tails)
      for (int64_t k = K; k >= 0 && offs[k] >= nC; --k) trk_off[k] = total;
      *d_ntrk = total;
    }
    __syncthreads();  // s_tile, s_cnt and s_excl are reused by the next tile
  }
  if (tid == 0) {
    const uint32_t epoch = s_epoch;
    uint32_t d = __hip_atomic_fetch_add(&ctrl->done, 1u, __ATOMIC_RELAXED, TRAX_AGENT);
    if (d == gridDim.x - 1) {
      __hip_atomic_store(&ctrl->done, 0u, __ATOMIC_RELAXED, TRAX_AGENT);
      __hip_atomic_store(&ctrl->ticket, 0u, __ATOMIC_RELAXED, TRAX_AGENT);
      __hip_atomic_store(&ctrl->epoch, (epoch + 1u) & 0x3FFFFFFFu, __ATOMIC_RELAXED, TRAX_AGENT);
    }
  }
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
