// CPU oracle for the trax track-seeding chain.
//
// TEST INFRASTRUCTURE ONLY.  This is a single-threaded C restatement of the
// reference's OpenCL kernels, statement by statement and in the same FP32
// operation order, used by tests/ (parity checks), __graft_entry__.smoke()
// and the CPU-baseline leg of bench.py.  Nothing in the product links or
// calls it.
//
// Stages restated (paths relative to the reference's src/):
//   grid build      algorithms/GridBuilder.h:48-257, GridBuilder.cpp:6-160
//   prefix sum      algorithms/PrefixSum.h:55-245, PrefixSum.cpp:8-126
//   pairs           algorithms/PairGeneratorBeamspot.h:43-246 (count),
//                   :448-525 (store), PairGeneratorBeamspot.cpp:5-177;
//                   or algorithms/PairGeneratorSector.h:43-234, .cpp:5-151
//   prediction      algorithms/TripletThetaPhiPredictor.h:46-248 (count),
//                   :250-481 (store), :483-497 (offsets), .cpp:10-131
//   filter          algorithms/TripletThetaPhiFilter.h:50-313, .cpp:10-164
//   connectivity    algorithms/TripletConnectivity.h:56-108, .cpp:21-71
//                   (semantics of test/TripletConnectivityTest.h)
//
// Where the reference is undefined or non-deterministic the oracle fixes one
// behaviour, and the HIP path implements the same one:
//   * within-cell hit order is the input order (stable counting sort); the
//     reference uses atomic_inc order (GridBuilder.h:243-244);
//   * out-of-range grid coordinates are clamped to the padding row/column and
//     float->int conversions saturate (NaN -> 0);
//   * result slots that the reference allocates but never writes (the
//     predictor store finding fewer hits than its count, TripletThetaPhi-
//     Predictor.h:394; duplicate pair visits, PairGeneratorBeamspot.h:183-216)
//     hold 0xFFFFFFFF and are counted in *_unwritten; later stages skip them;
//   * the racy offset monotonize kernels are evaluated as the sequential
//     forward fill they intend (TripletThetaPhiPredictor.h:493-496).
#pragma once

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_MAX_LAYERS 16
#define ORACLE_MAX_TRIPLETS 8
#define ORACLE_UNWRITTEN 0xFFFFFFFFu
#define ORACLE_PAIRS_BEAMSPOT 0u  // PairGeneratorBeamspot
#define ORACLE_PAIRS_SECTOR 1u    // PairGeneratorSector

typedef struct {
  uint32_t n_events;
  uint32_t n_layers;  // grid layers: detector layers 1..n_layers
  uint32_t n_z, n_phi;
  float min_z, sector_size_z, min_phi, sector_size_phi;
  float min_radius[ORACLE_MAX_LAYERS];  // index: layer - 1
  float max_radius[ORACLE_MAX_LAYERS];
  uint32_t n_triplets;
  uint32_t layer1[ORACLE_MAX_TRIPLETS], layer2[ORACLE_MAX_TRIPLETS], layer3[ORACLE_MAX_TRIPLETS];
  float dtheta_cut[ORACLE_MAX_TRIPLETS], sigma_z[ORACLE_MAX_TRIPLETS];
  float dphi_cut[ORACLE_MAX_TRIPLETS], sigma_phi[ORACLE_MAX_TRIPLETS];
  float tip_cut[ORACLE_MAX_TRIPLETS], z0[ORACLE_MAX_TRIPLETS];
  uint32_t pair_spread_z[ORACLE_MAX_TRIPLETS], pair_spread_phi[ORACLE_MAX_TRIPLETS];
  float min_radius_curvature;
  uint32_t threads;  // reference work-group size (orders pairs; no effect on sets)
  uint32_t pair_generator;  // ORACLE_PAIRS_BEAMSPOT or ORACLE_PAIRS_SECTOR
} oracle_cfg;

typedef struct {
  uint32_t n_hits;
  const float *x, *y, *z;
  const uint32_t *layer, *det_id, *hit_id, *evt;
  const uint32_t *event_offset;  // [n_events]
  const uint32_t *layer_nhits;   // [n_events * n_layers]
  const uint32_t *layer_offset;  // [n_events * n_layers], relative to the event
} oracle_input;

typedef struct {
  // grid stage: sorted hits and cell starts
  uint32_t n_hits;
  float *x, *y, *z;
  uint32_t *layer, *det_id, *hit_id, *evt, *orig_index;
  uint64_t grid_size;
  uint32_t *grid;
  // pairs: pairs[2k] = inner sorted-hit index, pairs[2k+1] = outer
  uint32_t n_pairs, n_pairs_unwritten, n_pairs_raw_overcount;
  uint32_t *pairs, *pair_offsets;  // offsets: [n_events * n_triplets + 1]
  // triplet candidates: cands[2k] = pair index, cands[2k+1] = third hit
  uint32_t n_cands, n_cands_unwritten;
  uint32_t *cands, *cand_offsets;
  // filtered triplets
  uint32_t n_tracklets;
  uint32_t *t1, *t2, *t3, *tracklet_offsets;
} oracle_result;

// Exclusive uint32 prefix sum in place (mod 2^32).
void oracle_scan_u32(uint32_t *data, uint64_t n);

// Whole chain; stop_after: 0 grid, 1 pairs, 2 prediction, 3 filter.
int oracle_run_chain(const oracle_cfg *cfg, const oracle_input *in, int stop_after,
                     oracle_result *out);
void oracle_free_result(oracle_result *r);

// Elementwise evaluation of the FP32 built-in definitions, for the math tests.
// op: 0 atan2(a, b), 1 atan(a), 2 acos(a), 3 sign(a), 4 a / b, 5 sqrt(a).
void oracle_eval_f32(int op, const float *a, const float *b, float *out, uint64_t n);

// Grid cell of one hit (GridBuilder.h:80-84); phi_out receives atan2(y, x).
uint32_t oracle_cell_of(const oracle_cfg *cfg, float x, float y, float z, float *phi_out);

// Window of one outer hit (pair stage) or one pair (prediction; store != 0
// selects the store-side variant).  out: zLow, zHigh, phiLow, phiHigh, wrap,
// zLowSector, zHighSector, phiLowSector, phiHighSector.
void oracle_pair_window(const oracle_cfg *cfg, uint32_t p, float x, float y, float z, float *out);
// Sector window of one inner hit: out = zLowSector, zHighSector, phiLowSector,
// phiHighSector, wrap (after the reference's corrections).
void oracle_sector_window(const oracle_cfg *cfg, uint32_t p, float x, float y, float z, uint32_t *out);
void oracle_predict_window(const oracle_cfg *cfg, uint32_t lt, int store, const float *p1,
                           const float *p2, float minR, float maxR, float *out);
// Filter quantities of one triplet: out = theta ratio, |dphi|, tip^2, circle radius, pass.
void oracle_filter_eval(const oracle_cfg *cfg, uint32_t lt, const float *h1, const float *h2,
                        const float *h3, float *out);

// TripletConnectivity::run (TripletConnectivity.cpp:21-71): every following
// triplet g against every base triplet i, fcon[g] += joined * (bcon[i] + 1),
// joined = bh3[i] == fh1[g] (wide, forward), bh1[i] == fh3[g] (wide,
// backward), or bh2[i] == fh1[g] && bh3[i] == fh2[g] (tight, forward only).
// The base connectivity is read from a copy taken before the first update,
// so base and following may alias.  Returns -1 for tight + backward.
int oracle_connectivity(const uint32_t *bh1, const uint32_t *bh2, const uint32_t *bh3,
                        const uint32_t *bcon, uint64_t nb, const uint32_t *fh1, const uint32_t *fh2,
                        const uint32_t *fh3, uint32_t *fcon, uint64_t nf, int backward, int tight);

#ifdef __cplusplus
}
#endif
