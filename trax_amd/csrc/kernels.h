// Host-side launchers of the gfx950 seeding kernels (internal to libtraxhip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"

namespace trax {

// Blocks of a persistent (ticketed) launch: 256 CUs x 8 resident blocks.
constexpr uint64_t kPersistentBlocks = 2048;
// Look-back status words reserved per context (tiles of a scan / filter launch).
constexpr uint64_t kStatusWords = 1ull << 22;
// Prediction tiles: pairs per tile (one wave) and stash slots per tile
// (== TRAX_PREDICT_TILE / TRAX_PREDICT_TILE_SLOTS, checked in the C ABI).
constexpr uint32_t kPredTile = 64;
constexpr uint32_t kTileSlots = 512;

uint64_t scan_tile_elems();
// Scans of at least this many elements use the two-pass form.
constexpr uint64_t kReduceScanMin = 1ull << 21;
// status: look-back words (single pass: one per tile; two-pass: one
// per chunk of 16384 elements), status_words of them.
hipError_t launch_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n_max, const uint32_t* d_n,
                           uint32_t n_add, ScanCtrl* ctrl, unsigned long long* status,
                           hipStream_t stream, int sub = 8, uint64_t status_words = kStatusWords);
// Elements per chunk of the two-pass form (one status word each).
constexpr uint64_t kScanChunk = 16384;
// Segmented fix-up after a global exclusive scan of data[0, n): bases[k] =
// data[off[k]] (gather), then data[i] -= bases[segment of i].
hipError_t launch_segment_bases(const uint32_t* data, uint64_t n, const uint32_t* off, uint32_t n_segments,
                                unsigned long long* bases, hipStream_t stream);
hipError_t launch_segment_fix(uint32_t* data, uint64_t n, const uint32_t* off, uint32_t n_segments,
                              const unsigned long long* bases, hipStream_t stream);
// Zeroes a scan control block (first use of a scratch block).
hipError_t launch_scan_ctrl_clear(ScanCtrl* ctrl, hipStream_t stream);

struct GridLaunch {
  const float *ix, *iy, *iz;
  const uint32_t *ilayer, *idet, *ihid, *ievt;
  float *ox, *oy, *oz;
  uint32_t *olayer, *odet, *ohid, *oevt, *oorig;
  float* ophi;
  float2* ozphi;
  float4* oxyzphi;
  const uint32_t *event_offset, *layer_nhits, *layer_offset;
  uint32_t n_hits;
  uint32_t* grid;
  uint32_t* flags;
  size_t lds_max;  // LDS a sort block may use
  int waves;       // waves per sort block: 1, 2, 4, 8 or 16; 0 = from the mean segment size
};
hipError_t launch_grid_build(const SeedParams& sp, const GridLaunch& gl, hipStream_t stream);
hipError_t grid_build_set_lds_limit(size_t bytes);

// Work offsets of the pair stage: per (event, layer pair) the hits of its outer
// (inner: innermost) layer, exclusive-scanned, total at [K].  One block for
// at most kPlanOneBlock segments, else a count kernel and the library scan.
constexpr uint32_t kPlanOneBlock = 1024;
hipError_t launch_pair_plan(const SeedParams& sp, const uint32_t* grid, uint32_t* work_off,
                            ScanCtrl* ctrl, unsigned long long* status, int scan_sub, hipStream_t stream,
                            bool inner = false);
// Sector pairs (count or store pass over the inner-hit work list).
hipError_t launch_sector_pairs(const SeedParams& sp, const SortedHits& h, const uint32_t* grid,
                               const uint32_t* work_off, uint32_t* counts, uint2* pairs,
                               uint64_t capacity, uint32_t items_max, bool store, uint32_t* flags,
                               hipStream_t stream);
// Beam-spot pairs: at most this many work items per outer hit
// (== TRAX_PAIR_PARTS, checked in the C ABI); the count pass also writes the
// item count to work_off[K + 1].  `masks` (one 64-bit word per work item,
// or null) carries the count pass's verdicts on each item's first 64
// visited hits to the store pass that follows it.
constexpr uint32_t kPairParts = 2;
hipError_t launch_pairs(const SeedParams& sp, const SortedHits& h, const uint32_t* grid,
                        uint32_t* work_off, uint32_t* counts, unsigned long long* masks, uint2* pairs,
                        uint64_t capacity, uint32_t items_max, bool store, uint32_t* flags, uint32_t block,
                        hipStream_t stream);
hipError_t launch_gather_offsets(const uint32_t* scanned, uint32_t last, const uint32_t* last_dev,
                                 const uint32_t* idx, uint32_t n, uint32_t* out, uint32_t* d_total,
                                 hipStream_t stream, uint32_t mul = 1);
hipError_t launch_read_last(const uint32_t* arr, const uint32_t* d_index, uint32_t* out,
                            hipStream_t stream);
hipError_t launch_predict(const SeedParams& sp, const SortedHits& h, const uint32_t* grid,
                          const uint2* pairs, const uint32_t* pair_off, const uint32_t* d_npairs,
                          uint64_t pair_capacity, uint32_t* counts, uint2* cands,
                          uint64_t cand_capacity, bool store, uint32_t* flags, uint32_t block,
                          hipStream_t stream);
// Tiled form of the predictor (walk + slab, scan, place).
hipError_t launch_predict_walk(const SeedParams& sp, const SortedHits& h, const uint32_t* grid,
                               const uint2* pairs, const uint32_t* pair_off, const uint32_t* d_npairs,
                               uint64_t pair_capacity, uint32_t* counts, uint32_t* slab,
                               uint32_t* tile_counts, hipStream_t stream);
// Candidate offsets per segment from the scanned tile counts; total -> *d_total.
hipError_t launch_pred_offsets(const uint32_t* tile_scanned, const uint32_t* counts, const uint32_t* pair_off,
                               const uint32_t* d_npairs, uint64_t pair_capacity, uint32_t n,
                               uint32_t* cand_off, uint32_t* d_total, hipStream_t stream);
hipError_t launch_predict_place(const SeedParams& sp, const SortedHits& h, const uint32_t* grid,
                                const uint2* pairs, const uint32_t* pair_off, const uint32_t* d_npairs,
                                uint64_t pair_capacity, const uint32_t* tile_scanned, const uint32_t* counts,
                                const uint32_t* slab, uint2* cands, uint64_t cand_capacity, uint32_t* flags,
                                hipStream_t stream);
// Prediction and filter in one pass: the walk (candidate budgets, tile
// totals, accepted records per tile in the slab), then after the scans the
// triplet offsets and the triplets.
// Compiled launch shapes of the fused walk: lane buffer depth, waves per
// block, waves per SIMD targeted by the register allocation.
struct FusedVariant {
  uint32_t buffer, block_waves, waves_per_eu;
};
constexpr uint32_t kNumFusedVariants = 7;
extern const FusedVariant kFusedVariants[kNumFusedVariants];
hipError_t launch_predict_filter_walk(const SeedParams& sp, const SortedHits& h, const uint32_t* grid,
                                      const uint2* pairs, const uint32_t* pair_off, const uint32_t* d_npairs,
                                      uint64_t pair_capacity, uint32_t* counts, uint32_t* tile_counts,
                                      uint4* acc_slab, uint32_t* acc_tiles, hipStream_t stream,
                                      uint32_t variant = 0);
hipError_t launch_acc_redo(const SeedParams& sp, const SortedHits& h, const uint32_t* grid, const uint2* pairs,
                           const uint32_t* pair_off, const uint32_t* d_npairs, uint64_t pair_capacity,
                           const uint32_t* counts, uint4* acc_slab, uint32_t* acc_tiles, hipStream_t stream);
hipError_t launch_acc_cut(const SeedParams& sp, const SortedHits& h, const uint32_t* d_npairs,
                          uint64_t pair_capacity, uint4* acc_slab, uint32_t* acc_tiles, hipStream_t stream);
hipError_t launch_fused_offsets(const uint32_t* tile_scanned, const uint32_t* counts, const uint32_t* acc_scanned,
                                const uint4* acc_slab, const uint32_t* pair_off, const uint32_t* d_npairs,
                                uint64_t pair_capacity, uint32_t n, uint32_t* cand_off, uint32_t* trk_off,
                                uint32_t* d_totals, hipStream_t stream);
hipError_t launch_acc_place(const SeedParams& sp, const SortedHits& h, const uint32_t* grid, const uint2* pairs,
                            const uint32_t* pair_off, const uint32_t* d_npairs, uint64_t pair_capacity,
                            const uint32_t* acc_scanned, const uint32_t* counts, const uint4* acc_slab,
                            uint32_t* t1, uint32_t* t2, uint32_t* t3, uint64_t trk_capacity, uint32_t* flags,
                            hipStream_t stream);
// Filter: mask pass, scan of the word counts, store pass.
uint64_t filter_mask_words(uint64_t cand_capacity);
hipError_t launch_filter_mask(const SeedParams& sp, const SortedHits& h, const uint2* pairs,
                              const uint2* cands, const uint32_t* cand_off, const uint32_t* d_ncands,
                              uint64_t cand_capacity, unsigned long long* mask, uint32_t* counts,
                              uint32_t* d_nwords, hipStream_t stream);
hipError_t launch_filter_store(const SeedParams& sp, const uint2* pairs, const uint2* cands,
                               const uint32_t* cand_off, const uint32_t* d_ncands,
                               uint64_t cand_capacity, const unsigned long long* mask,
                               const uint32_t* scanned, uint32_t* t1, uint32_t* t2, uint32_t* t3,
                               uint64_t trk_capacity, uint32_t* trk_off, uint32_t* d_ntrk,
                               uint32_t* flags, hipStream_t stream);
// Triplet connectivity: hash join of a base triplet set into a following one.
struct ConnTriplets {
  const uint32_t *hit1, *hit2, *hit3;
  uint32_t* con;  // base: read only; following: read-modify-write
  uint64_t n;
};
struct alignas(16) ConnSlot {
  unsigned long long key;
  uint32_t weight;
  uint32_t pad;
};
// Table slots for n_base base triplets (a power of two >= 2 * n_base); the
// table holds one more slot than that.
uint64_t connectivity_table_slots(uint64_t n_base);
// mode: 0 wide forward, 1 wide backward, 2 tight forward.
hipError_t launch_connectivity(const ConnTriplets& base, const ConnTriplets& following, int mode,
                               ConnSlot* table, uint64_t cap, uint32_t* flags, hipStream_t stream);
hipError_t launch_fp32_eval(int op, const float* a, const float* b, float* out, uint64_t n,
                            hipStream_t stream);

}  // namespace trax
