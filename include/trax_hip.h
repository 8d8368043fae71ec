/*
 * trax_hip.h -- C ABI of libtraxhip.so, the gfx950 (MI355X) track-seeding
 * path: grid build -> hit pairs -> triplet prediction -> triplet filter, with
 * the prefix sums between the stages.
 *
 * Every entry point takes plain device pointers and sizes, enqueues work on
 * the given HIP stream (`void*` = hipStream_t; NULL = default stream) and
 * returns an int status (TRAX_OK or a negative TRAX_ERR_*); the text of the
 * last error is kept per context.  Nothing here allocates device memory
 * except trax_ctx_create and trax_chain_graph_capture (the graph object);
 * callers own every buffer.  A context serialises its
 * launches on one stream at a time (as the reference does with its one
 * in-order OpenCL queue): use one context per concurrent stream.
 *
 * Each function replaces one piece of the reference's algorithm classes
 * (paths relative to the reference's src/), cited at the declaration.
 *
 * Hit layout (src/datastructures/HitCollection.h:23, HitCollection.cpp:106-146):
 * SoA columns x, y, z (float) and layer, detId, hitId, evt (uint32), ordered
 * event-major then layer-ascending; per event an offset, per (event, layer) a
 * hit count and an offset relative to the event.  evt[i] is the event's index
 * in the batch and layer[i] the 1-based detector layer.
 */
#ifndef TRAX_HIP_H
#define TRAX_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TRAX_ABI_VERSION 8

#define TRAX_OK 0
#define TRAX_ERR_INVALID (-1)  /* bad argument or configuration */
#define TRAX_ERR_HIP (-2)      /* HIP runtime error */
#define TRAX_ERR_NOMEM (-3)    /* device allocation failed (context creation) */
#define TRAX_ERR_CAPACITY (-4) /* an output buffer was too small (see trax_ctx_flags) */

/* Device status bits (trax_ctx_flags). */
#define TRAX_FLAG_PAIR_OVERFLOW 0x1u
#define TRAX_FLAG_CAND_OVERFLOW 0x2u
#define TRAX_FLAG_TRACKLET_OVERFLOW 0x4u
#define TRAX_FLAG_LAYOUT 0x8u
#define TRAX_FLAG_CONNECT_TABLE 0x10u /* connectivity table full (cannot happen at its sized capacity) */

#define TRAX_MAX_LAYERS 16
#define TRAX_MAX_TRIPLETS 8
#define TRAX_UNWRITTEN 0xFFFFFFFFu /* predictor slot the reference leaves unwritten */

typedef struct trax_ctx trax_ctx;

/* One row of the layer-triplet table; mirrors the columns of
 * src/datastructures/TripletConfiguration.h:9-33 (configs/layerTriplets.xml). */
typedef struct {
  uint32_t layer1, layer2, layer3; /* layer1 must equal row index + 1 */
  float dtheta_cut, sigma_z, dphi_cut, sigma_phi, tip_cut;
  uint32_t pair_spread_z;
  float z0; /* the reference parses it as uint (TripletConfiguration.cpp:91) */
  uint32_t pair_spread_phi;
  uint32_t n_candidates;
} trax_layer_triplet;

/* Grid binning (src/datastructures/Grid.h:5-39: MIN_Z=-300, MAX_Z=300,
 * MIN_PHI=-pi, MAX_PHI=pi), per-layer radius bands (GeometrySupplement,
 * index layer-1) and the layer triplets. */
typedef struct {
  uint32_t n_z, n_phi; /* n_z in 1..32767, n_phi in 1..16383 */
  uint32_t n_layers;   /* grid layers: detector layers 1..n_layers */
  float min_radius[TRAX_MAX_LAYERS];
  float max_radius[TRAX_MAX_LAYERS];
  uint32_t n_triplets;
  trax_layer_triplet triplets[TRAX_MAX_TRIPLETS];
  /* minPt * GEV_C / (Bz * Q) * 100 with Bz = 3.8112 (TripletConfiguration.h:56-70);
   * trax_min_radius_curvature() evaluates it the reference's way. */
  float min_radius_curvature;
  /* pair generator run by trax_seed_chain: TRAX_PAIRS_BEAMSPOT
   * (PairGeneratorBeamspot) or TRAX_PAIRS_SECTOR (PairGeneratorSector) */
  uint32_t pair_generator;
} trax_seed_config;

#define TRAX_PAIRS_BEAMSPOT 0u
#define TRAX_PAIRS_SECTOR 1u

/* Hit columns.  Input: x..evt.  Grid-sorted output: all of them, plus
 * orig_index (input index of each sorted hit), phi (atan2(y, x)), zphi
 * (2 floats per hit: z, phi interleaved, read by the window walks) and
 * xyzphi (4 floats per hit: x, y, z, phi, 16-byte aligned, read by the pair,
 * prediction and cut kernels as one load per hit). */
typedef struct {
  float *x, *y, *z;
  uint32_t *layer, *det_id, *hit_id, *evt;
  uint32_t *orig_index;
  float *phi;
  float *zphi;
  float *xyzphi;
} trax_hits;

/* Event layout, device arrays (EventSupplement.h / LayerSupplement.h). */
typedef struct {
  uint32_t n_events;
  uint32_t n_hits;
  const uint32_t *event_offset; /* [n_events] */
  const uint32_t *layer_nhits;  /* [n_events * n_layers] */
  const uint32_t *layer_offset; /* [n_events * n_layers], relative to the event */
} trax_layout;

/* The single-call predictor works on tiles of TRAX_PREDICT_TILE pairs; each
 * tile's candidate slots are staged in a stash slab of
 * TRAX_PREDICT_TILE_SLOTS uint32 (tiles with more slots walk again). */
#define TRAX_PREDICT_TILE 64
#define TRAX_PREDICT_TILE_SLOTS 512
/* The beam-spot pair passes split each outer hit's window into at most this
 * many runs of z rows (one count per run): two on grids of 128 or more z
 * rows, one otherwise. */
#define TRAX_PAIR_PARTS 2

/* Scratch of the count/store phases (device pointers, caller-owned; laid
 * out by trax_workspace_layout).  Every stage call checks the configuration
 * against the sizes recorded here (n_triplets, pair_items, max_events, and
 * for trax_seed_chain max_hits) and rejects one the layout cannot hold. */
typedef struct {
  uint32_t *work_offsets;  /* >= max_events * n_triplets + 2 */
  uint32_t *pair_counts;   /* >= pair_items + 1 */
  uint32_t *cand_counts;   /* >= pair_capacity + 1 */
  uint32_t max_events, max_hits;
  uint64_t pair_capacity;
  uint64_t cand_capacity;
  uint32_t *cand_stash;            /* >= ceil(pair_capacity / TRAX_PREDICT_TILE) * TRAX_PREDICT_TILE_SLOTS */
  unsigned long long *filter_mask; /* >= ceil(cand_capacity / 64) + 8 */
  uint32_t *filter_counts;         /* >= ceil(cand_capacity / 64) + 9 */
  uint32_t *cand_tile_counts;      /* >= ceil(pair_capacity / TRAX_PREDICT_TILE) + 2: candidate
                                      slots per tile of pairs, then their scan, then the tile count */
  uint32_t *trk_tile_counts;       /* >= ceil(pair_capacity / TRAX_PREDICT_TILE) + 2: accepted
                                      triplets per tile of pairs (trax_triplet_predict_filter) */
  uint32_t n_triplets; /* layer triplets the layout was sized for (work_offsets) */
  /* pair work items pair_counts holds: TRAX_PAIR_PARTS * max_hits times the
   * largest number of layer triplets sharing one middle layer (each such
   * triplet walks the middle layer's hits once) */
  uint32_t pair_items;
} trax_workspace;

/* Output buffers of the pair, prediction and filter stages.  Capacities are
 * in elements; an overflow raises a TRAX_FLAG_* bit and the stage's count
 * still reports the full size, so the caller can grow and rerun.
 * pairs/cands hold uint2 records: pairs (inner, outer sorted-hit index),
 * cands (pair index, third sorted-hit index).  Offsets have
 * n_events * n_triplets + 1 entries.  totals[0..2] = pairs, candidates,
 * triplets (device). */
typedef struct {
  uint32_t *pairs;
  uint64_t pair_capacity;
  uint32_t *pair_offsets;
  uint32_t *cands;
  uint64_t cand_capacity;
  uint32_t *cand_offsets;
  uint32_t *t1, *t2, *t3; /* TrackletHit1/2/3 (TrackletCollection.h:13-33) */
  uint64_t tracklet_capacity;
  uint32_t *tracklet_offsets;
  uint32_t *totals;
} trax_seed_outputs;

/* Launch shape of the kernels, per context (set by the autotuner). */
typedef struct {
  uint32_t grid_waves;    /* waves per grid-sort block: 1, 2, 4, 8 or 16; 0 (default) picks one
                             per batch from the mean (event, layer) segment size */
  uint32_t pairs_block;   /* threads per block, two-phase pair kernels: 64..512 (default 256) */
  uint32_t predict_block; /* threads per block, two-phase prediction kernels (default 256) */
  uint32_t scan_tile;     /* elements per single-pass scan tile: 4096, 8192 or 16384 (default 4096) */
  /* fused prediction + filter walk: lane buffer depth (store-window hits a
   * lane keeps before its wave's spill pool), waves per block, and waves per
   * SIMD the register allocation targets.  All three 0 = default (8, 4, 7);
   * otherwise one of the compiled combinations (8, 4, 6), (8, 4, 8),
   * (4, 4, 8), (4, 4, 7), (8, 2, 7), (12, 4, 6) */
  uint32_t fused_buffer, fused_block_waves, fused_waves_per_eu;
} trax_launch_config;

/* ---------------------------------------------------------------- context */
int trax_abi_version(void);
int trax_ctx_create(int device, trax_ctx **out);
int trax_ctx_destroy(trax_ctx *ctx);
const char *trax_ctx_last_error(const trax_ctx *ctx);
int trax_ctx_set_launch(trax_ctx *ctx, const trax_launch_config *cfg);
int trax_ctx_get_launch(const trax_ctx *ctx, trax_launch_config *cfg);
/* Synchronous read of the device status bits; clear != 0 resets them. */
int trax_ctx_flags(trax_ctx *ctx, void *stream, uint32_t *flags_out, int clear);

/* Kernel timer (no counterpart in the reference, which times whole stages
 * with host clocks, e.g. TripletThetaPhiPredictor.cpp:44-127): after
 * trax_ctx_kernel_timer(ctx, id), every eager launch of kernel `id` through
 * this context is bracketed by HIP events on its launch stream (launches
 * inside a stream capture are not timed); trax_ctx_kernel_time waits for
 * them and returns the summed device time and the launch count, then
 * resets.  id 0 turns the timer off. */
#define TRAX_KERNEL_GRID_BUILD 1      /* grid build: segment sort + grid fill */
#define TRAX_KERNEL_PAIRS 2           /* pair generation: the count and store kernels */
#define TRAX_KERNEL_PREDICT_WALK 3    /* prediction window walk: counts + stash slabs */
#define TRAX_KERNEL_PREDICT_PLACE 4   /* prediction slabs -> final candidate slots */
#define TRAX_KERNEL_FILTER_MASK 5     /* filter cuts -> mask words */
#define TRAX_KERNEL_FILTER_STORE 6    /* filter compaction */
#define TRAX_KERNEL_SCAN 7            /* every prefix sum the context launches */
#define TRAX_KERNEL_CONNECT 8         /* triplet connectivity: table clear, insert, lookup */
#define TRAX_KERNEL_PREDICT_FILTER 9  /* fused prediction + filter walk */
#define TRAX_KERNEL_TRIPLET_PLACE 10  /* fused path: accepted records -> triplets */
#define TRAX_KERNEL_ACC_CUT 11        /* fused path: theta and circle cuts over the phi survivors */
int trax_ctx_kernel_timer(trax_ctx *ctx, int kernel);
int trax_ctx_kernel_time(trax_ctx *ctx, double *total_ms, uint32_t *launches);

/* minRadiusCurvature(minPt) as TripletConfiguration.cpp:145-152 computes it. */
float trax_min_radius_curvature(float min_pt);

/* Carves one device allocation at `base` into a workspace for batches of at
 * most max_events events, max_hits hits, pair_capacity pairs and
 * cand_capacity candidates, and returns the bytes it needs (call with
 * base = NULL to size it). */
uint64_t trax_workspace_layout(const trax_seed_config *cfg, uint32_t max_events, uint32_t max_hits,
                               uint64_t pair_capacity, uint64_t cand_capacity, void *base,
                               trax_workspace *ws);

/* ------------------------------------------------------------- prefix sum */
/* In-place exclusive uint32 scan (mod 2^32) of n elements; replaces
 * PrefixSum::run(cl_mem, uint size, uint nThreads), PrefixSum.h:42.
 * d_data must be 16-byte aligned. */
int trax_prefix_sum(trax_ctx *ctx, uint32_t *d_data, uint64_t n, void *stream);

/* Out-of-place variant whose length is read on the device: n = min(*d_n +
 * n_add, n_max) (d_n may be NULL: n = n_max).  Graph-capture safe. */
int trax_prefix_sum_dev(trax_ctx *ctx, const uint32_t *d_in, uint32_t *d_out, uint64_t n_max,
                        const uint32_t *d_n, uint32_t n_add, void *stream);

/* trax_prefix_sum without a context and without allocating: no counterpart
 * in the reference; for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The scan's look-back control block and status words live in
 * a fixed per-device array of 512 KiB + 256 B, set up on the first call on a
 * device (the current device); n is limited to TRAX_PREFIX_SUM_NOALLOC_MAX
 * elements (larger n: TRAX_ERR_INVALID, nothing enqueued).  d_data must be
 * 16-byte aligned.  All calls on one device share that array: calls that
 * overlap in time on one device (different streams or host threads) race
 * and must be serialised by the caller. */
#define TRAX_PREFIX_SUM_NOALLOC_MAX (1ull << 30)
int trax_prefix_sum_noalloc(uint32_t *d_data, uint64_t n, void *stream);

/* Segmented in-place exclusive scan: the scan restarts at 0 at the start of
 * every segment.  d_seg_offsets (device, n_segments + 1 entries,
 * non-decreasing, first 0, last n) gives segment k = [off[k], off[k+1]),
 * e.g. one segment per event as the reference's per-event offset tables
 * (EventSupplement / LayerSupplement) lay them out; the reference scans
 * such tables segment by segment with PrefixSum::run(cl_mem, uint size,
 * uint nThreads, std::vector<cl_event>&), PrefixSum.h:42-43, one call per
 * range.  At most 2^22 segments (the context's look-back words hold the
 * segment bases).  No host read; graph-capture safe. */
int trax_prefix_sum_segmented(trax_ctx *ctx, uint32_t *d_data, uint64_t n, const uint32_t *d_seg_offsets,
                              uint32_t n_segments, void *stream);

/* ------------------------------------------------------------ grid build */
/* GridBuilder::run(HitCollection&, nThreads, EventSupplement, LayerSupplement,
 * Grid&), GridBuilder.h:42.  Counting sort of every (event, layer) segment
 * (at most 2^29 - 1 hits per batch) into (z, phi) cells; writes the sorted
 * hit columns (+ orig_index, phi, zphi) to
 * `sorted` and the cell starts (n_events * n_layers * (n_z+1)(n_phi+1)
 * uint32, global sorted indices) to d_grid.  Within a cell hits keep their
 * input order.  Needs no workspace: segments of more than 8192 hits are
 * sorted in place in the det_id/hit_id/evt/layer output columns before
 * those are written. */
int trax_grid_build(trax_ctx *ctx, const trax_seed_config *cfg, const trax_hits *in,
                    const trax_layout *layout, trax_hits *sorted, uint32_t *d_grid,
                    void *stream);

/* ------------------------------------------------------------------ pairs */
/* PairGeneratorBeamspot::run(...), PairGeneratorBeamspot.h:40-41, in two
 * phases.  count: per outer hit pair counts + prefix sum into the workspace;
 * the total lands in out->totals[0].  store: writes out->pairs and
 * out->pair_offsets (requires the count phase of the same batch). */
int trax_pairs_count(trax_ctx *ctx, const trax_seed_config *cfg, uint32_t n_events,
                     const trax_hits *sorted, const uint32_t *d_grid, const trax_workspace *ws,
                     trax_seed_outputs *out, void *stream);
int trax_pairs_store(trax_ctx *ctx, const trax_seed_config *cfg, uint32_t n_events,
                     const trax_hits *sorted, const uint32_t *d_grid, const trax_workspace *ws,
                     trax_seed_outputs *out, void *stream);

/* Single-call form: trax_pairs_count then trax_pairs_store with the total
 * kept on the device (no host read); writes out->pairs, out->pair_offsets and
 * out->totals[0].  Graph-capture safe.  When ws->cand_stash holds 8 bytes
 * per pair work item, the count pass leaves its per-hit verdicts there for
 * the store pass (overwriting the stash). */
int trax_pairs_beamspot(trax_ctx *ctx, const trax_seed_config *cfg, uint32_t n_events,
                        const trax_hits *sorted, const uint32_t *d_grid, const trax_workspace *ws,
                        trax_seed_outputs *out, void *stream);

/* Sector pairs: PairGeneratorSector::run(...), PairGeneratorSector.h:38-39
 * (the reference's second PairGenerator, same arguments as the beam-spot
 * one).  Every inner-layer hit pairs with every outer-layer hit in the grid
 * cells of its z row +- pair_spread_z and phi cell +- pair_spread_phi (with
 * the reference's wrap-around and unsigned row arithmetic); no geometric
 * cut.  Requires 2 * pair_spread_phi <= n_phi.  count: per inner hit counts
 * + prefix sum (total -> out->totals[0], offsets -> out->pair_offsets);
 * store: the (inner, outer) records.  trax_pairs_sector runs both (no host
 * read, graph-capture safe).  An outer hit is paired once per inner hit even
 * where the reference's window visits it twice (2 * spread + 1 > n_phi); the
 * reference counts those visits and leaves their slots unwritten. */
int trax_pairs_sector_count(trax_ctx *ctx, const trax_seed_config *cfg, uint32_t n_events,
                            const trax_hits *sorted, const uint32_t *d_grid, const trax_workspace *ws,
                            trax_seed_outputs *out, void *stream);
int trax_pairs_sector_store(trax_ctx *ctx, const trax_seed_config *cfg, uint32_t n_events,
                            const trax_hits *sorted, const uint32_t *d_grid, const trax_workspace *ws,
                            trax_seed_outputs *out, void *stream);
int trax_pairs_sector(trax_ctx *ctx, const trax_seed_config *cfg, uint32_t n_events,
                      const trax_hits *sorted, const uint32_t *d_grid, const trax_workspace *ws,
                      trax_seed_outputs *out, void *stream);

/* ------------------------------------------------------------- prediction */
/* TripletThetaPhiPredictor::run(...), TripletThetaPhiPredictor.h:43-44, in
 * two phases: count (+ prefix sum; total -> out->totals[1]) and store
 * (out->cands, out->cand_offsets).  The pair count is read on the device. */
int trax_predict_count(trax_ctx *ctx, const trax_seed_config *cfg, uint32_t n_events,
                       const trax_hits *sorted, const uint32_t *d_grid, const trax_workspace *ws,
                       trax_seed_outputs *out, void *stream);
int trax_predict_store(trax_ctx *ctx, const trax_seed_config *cfg, uint32_t n_events,
                       const trax_hits *sorted, const uint32_t *d_grid, const trax_workspace *ws,
                       trax_seed_outputs *out, void *stream);

/* Single-call form of the prediction (no host read; the pair total is read
 * on the device): one walk per pair counts its budget and stages its slots
 * per tile of pairs in ws->cand_stash, a prefix sum places the budgets, and
 * a placement pass copies each tile's slots to out->cands; writes
 * out->cand_offsets and out->totals[1].  Same output as trax_predict_count
 * + trax_predict_store. */
int trax_triplet_predict(trax_ctx *ctx, const trax_seed_config *cfg, uint32_t n_events,
                         const trax_hits *sorted, const uint32_t *d_grid, const trax_workspace *ws,
                         trax_seed_outputs *out, void *stream);

/* ----------------------------------------------------------------- filter */
/* TripletThetaPhiFilter::run(...), TripletThetaPhiFilter.h:46-48, in two
 * phases.  count: the cuts, one mask bit per candidate, per 64-bit word
 * popcounts and their prefix sum in the workspace (filterCount,
 * filterPopCount, PrefixSum); the triplet total lands in out->totals[2].
 * store: order-preserving compaction into t1/t2/t3 and the per-(event,
 * layer triplet) offsets (filterStore, filterOffsetStore). */
int trax_filter_count(trax_ctx *ctx, const trax_seed_config *cfg, uint32_t n_events,
                      const trax_hits *sorted, const trax_workspace *ws, trax_seed_outputs *out,
                      void *stream);
int trax_filter_store(trax_ctx *ctx, const trax_seed_config *cfg, uint32_t n_events,
                      const trax_workspace *ws, trax_seed_outputs *out, void *stream);

/* Both filter phases back to back (no host read). */
int trax_triplet_filter(trax_ctx *ctx, const trax_seed_config *cfg, uint32_t n_events,
                        const trax_hits *sorted, const trax_workspace *ws, trax_seed_outputs *out,
                        void *stream);

/* trax_triplet_predict followed by trax_triplet_filter, fused: no
 * counterpart call in the reference, whose TripletThetaPhiPredictor::run
 * (TripletThetaPhiPredictor.h:43-44) and TripletThetaPhiFilter::run
 * (TripletThetaPhiFilter.h:46-48) it replaces when the candidates
 * themselves are not wanted.  The filter's decision for a candidate depends
 * on its three hits and layer triplet only, so one walk per pair applies the
 * cuts to each candidate as the prediction finds it.  Writes t1/t2/t3,
 * tracklet_offsets, cand_offsets and totals[1..2] exactly as the two calls
 * would; out->cands is neither written nor read (cand_capacity may be 0)
 * and no candidate overflow can occur.  Needs ws->cand_stash,
 * ws->cand_tile_counts and ws->trk_tile_counts (16-byte aligned).  No host
 * read; graph-capture safe. */
int trax_triplet_predict_filter(trax_ctx *ctx, const trax_seed_config *cfg, uint32_t n_events,
                                const trax_hits *sorted, const uint32_t *d_grid,
                                const trax_workspace *ws, trax_seed_outputs *out, void *stream);

/* ------------------------------------------------------------ whole chain */
/* The four stages back to back with no host synchronisation (the stage
 * sequence of buildTriplets, src/triplet_builder/TripletBuilder.cpp:240-263):
 * trax_grid_build, trax_pairs_beamspot (or trax_pairs_sector, by
 * cfg->pair_generator), trax_triplet_predict, trax_triplet_filter; with
 * out->cands NULL, trax_triplet_predict_filter in place of the last two (the
 * same triplets and offsets, no candidate records).  Safe to capture in a
 * hipGraph: no allocation, no memset, no host read. */
int trax_seed_chain(trax_ctx *ctx, const trax_seed_config *cfg, const trax_hits *in,
                    const trax_layout *layout, trax_hits *sorted, uint32_t *d_grid,
                    const trax_workspace *ws, trax_seed_outputs *out, void *stream);

/* The whole chain as one HIP graph, for callers that run the same batch
 * shape again and again: trax_chain_graph_capture records trax_seed_chain
 * with these exact arguments (buffers, capacities, configuration) into a
 * graph and instantiates it; trax_chain_graph_launch replays it on `stream`
 * (one graph launch instead of ~13 kernel launches, no host work per
 * stage).  It replaces the reference's stage sequence with its three
 * blocking host reads (PairGeneratorBeamspot.cpp:120,
 * TripletThetaPhiPredictor.cpp:59, TripletThetaPhiFilter.cpp:82; called
 * from TripletBuilder.cpp:240-263).  The buffers named at capture are read
 * and written by every launch: keep them alive, and change their contents
 * (new hits of the same event layout sizes) between launches as needed; the
 * layout arrays are read on the device, so event sizes may change as long as
 * n_events and n_hits stay within what the buffers hold (n_events and n_hits
 * themselves are part of the recording).  Status bits land in the context
 * (trax_ctx_flags).  Capture runs on a private stream of the context's
 * device, so `stream` of the launch may be any stream of that device. */
typedef struct trax_graph trax_graph;
int trax_chain_graph_capture(trax_ctx *ctx, const trax_seed_config *cfg, const trax_hits *in,
                             const trax_layout *layout, trax_hits *sorted, uint32_t *d_grid,
                             const trax_workspace *ws, trax_seed_outputs *out, trax_graph **graph_out);
int trax_chain_graph_launch(trax_graph *graph, void *stream);
int trax_chain_graph_destroy(trax_graph *graph);

/* ----------------------------------------------------------- connectivity */
/* A triplet set on the device: the hit and connectivity columns of the
 * reference's TrackletCollection (src/datastructures/TrackletCollection.h:13-35:
 * TrackletHit1/2/3, TrackletConnectivity), n entries each. */
typedef struct {
  const uint32_t *hit1, *hit2, *hit3;
  uint32_t *connectivity; /* read only in the base set */
  uint64_t n;
} trax_tracklets;

#define TRAX_CONNECT_BACKWARD 0x1u /* iterateBackwards */
#define TRAX_CONNECT_TIGHT 0x2u    /* tightPacking (forward only) */

/* Device workspace bytes trax_triplet_connectivity needs for n_base base
 * triplets: a hash table of 16-byte slots, a power of two >= 2 * n_base
 * (at least 64) plus one. */
uint64_t trax_connectivity_workspace_bytes(uint64_t n_base);

/* TripletConnectivity::run(base, following, iterateBackwards, tightPacking),
 * src/algorithms/TripletConnectivity.h:54-56 / .cpp:21-71.  For every
 * following triplet g, following->connectivity[g] += the sum of
 * (base->connectivity[i] + 1) over the base triplets i joined to g (mod 2^32):
 *   wide, forward:   base.hit3[i] == following.hit1[g]
 *   wide, backward:  base.hit1[i] == following.hit3[g]
 *   tight (forward): base.hit2[i] == following.hit1[g] && base.hit3[i] == following.hit2[g]
 * as the reference's known answers state them (TripletConnectivityTest.h).
 * Only the columns the mode reads must be set.  Tight + backward is rejected,
 * as the reference asserts.  base and following may be the same set: every
 * base value is read before any following value is written.  d_workspace:
 * at least trax_connectivity_workspace_bytes(base->n) bytes, 16-byte aligned;
 * calls that share one workspace must not overlap. */
int trax_triplet_connectivity(trax_ctx *ctx, const trax_tracklets *base, trax_tracklets *following,
                              uint32_t mode, void *d_workspace, uint64_t workspace_bytes, void *stream);

/* ---------------------------------------------------------- test support */
/* Elementwise device evaluation of the FP32 built-in definitions used by the
 * kernels: op 0 atan2(a, b), 1 atan(a), 2 acos(a), 3 sign(a), 4 a / b,
 * 5 sqrt(a).  Lets tests pin the device arithmetic bit for bit. */
int trax_fp32_eval(trax_ctx *ctx, int op, const float *d_a, const float *d_b, float *d_out,
                   uint64_t n, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* TRAX_HIP_H */
