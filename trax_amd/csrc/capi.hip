// C ABI of libtraxhip.so (declarations and reference citations in
// include/trax_hip.h).  Validates arguments, turns the public configuration
// into the by-value kernel parameter block, and enqueues the kernels.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/trax_hip.h"
#include "common.h"
#include "kernels.h"

static_assert(TRAX_MAX_LAYERS == 16 && TRAX_MAX_TRIPLETS == 8, "table sizes");
static_assert(TRAX_PAIR_PARTS == trax::kPairParts, "pair parts");
static_assert(TRAX_PREDICT_TILE == trax::kPredTile && TRAX_PREDICT_TILE_SLOTS == trax::kTileSlots,
              "predictor tile shape");

struct trax_ctx {
  int device = 0;
  std::string err;
  void* block = nullptr;  // one device allocation: ctrl | flags | status words
  trax::ScanCtrl* ctrl = nullptr;
  uint32_t* flags = nullptr;
  unsigned long long* status = nullptr;
  size_t lds_max = 65536;
  trax_launch_config launch = {0, 256, 256, 4096, 0, 0, 0};
  uint32_t fused_variant = 0;  // index into trax::kFusedVariants
  // kernel timer: HIP events recorded on the launch stream around every
  // eager launch of one selected kernel
  int timed_kernel = 0;
  std::vector<hipEvent_t> ev_start, ev_stop;
  size_t ev_used = 0;
  double timed_ms = 0.0;
  uint32_t timed_launches = 0;
};

namespace {

constexpr uint64_t kAlign = 256;

uint64_t align_up(uint64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

int fail(trax_ctx* ctx, int code, const char* msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int hip_fail(trax_ctx* ctx, hipError_t e, const char* where) {
  if (ctx) {
    ctx->err = std::string(where) + ": " + hipGetErrorString(e);
  }
  return TRAX_ERR_HIP;
}

#define TRAX_HIP_CHECK(ctx, expr)                          \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) return hip_fail(ctx, e_, #expr); \
  } while (0)

// Folds the recorded event pairs into the timer's running total.
int timer_drain(trax_ctx* ctx) {
  for (size_t i = 0; i < ctx->ev_used; ++i) {
    TRAX_HIP_CHECK(ctx, hipEventSynchronize(ctx->ev_stop[i]));
    float ms = 0.0f;
    TRAX_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev_start[i], ctx->ev_stop[i]));
    ctx->timed_ms += ms;
    ++ctx->timed_launches;
  }
  ctx->ev_used = 0;
  return TRAX_OK;
}

// Returns the event pair to bracket this launch with, or -1 when the kernel
// is not the timed one or the stream is being captured.
int timer_slot(trax_ctx* ctx, int kernel, hipStream_t s) {
  if (ctx->timed_kernel != kernel) return -1;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return -1;
  if (ctx->ev_used == ctx->ev_start.size() && timer_drain(ctx) != TRAX_OK) return -1;
  return (int)ctx->ev_used++;
}

#define TRAX_TIMED(ctx, kernel, stream, expr)                                       \
  do {                                                                            \
    const int slot_ = timer_slot(ctx, kernel, stream);                            \
    if (slot_ >= 0) TRAX_HIP_CHECK(ctx, hipEventRecord(ctx->ev_start[slot_], stream)); \
    TRAX_HIP_CHECK(ctx, expr);                                                    \
    if (slot_ >= 0) TRAX_HIP_CHECK(ctx, hipEventRecord(ctx->ev_stop[slot_], stream));  \
  } while (0)

uint32_t bits_for(uint32_t n) {  // bits needed to hold values 0..n-1
  uint32_t b = 0;
  while (b < 32 && (1ull << b) < n) ++b;
  return b;
}

// Validate the configuration and build the kernel parameter block.
int make_params(trax_ctx* ctx, const trax_seed_config* cfg, uint32_t n_events, trax::SeedParams* sp) {
  if (!cfg) return fail(ctx, TRAX_ERR_INVALID, "null configuration");
  // window sector bounds are packed in 16-bit fields (phi up to n_phi + 1)
  if (cfg->n_z == 0 || cfg->n_phi == 0 || cfg->n_z > 32767 || cfg->n_phi > 16383)
    return fail(ctx, TRAX_ERR_INVALID, "n_z must be in 1..32767 and n_phi in 1..16383");
  if (cfg->n_layers == 0 || cfg->n_layers > TRAX_MAX_LAYERS)
    return fail(ctx, TRAX_ERR_INVALID, "n_layers must be in 1..16");
  if (cfg->n_triplets == 0 || cfg->n_triplets > TRAX_MAX_TRIPLETS)
    return fail(ctx, TRAX_ERR_INVALID, "n_triplets must be in 1..8");
  const uint64_t cells = (uint64_t)(cfg->n_z + 1) * (cfg->n_phi + 1);
  if (cells > (1u << 24)) return fail(ctx, TRAX_ERR_INVALID, "too many grid cells");
  if ((uint64_t)n_events * cfg->n_layers * cells > 0xFFFFFFFFull)
    return fail(ctx, TRAX_ERR_INVALID, "grid larger than 2^32 cells");
  memset(sp, 0, sizeof(*sp));
  sp->n_events = n_events;
  sp->n_layers = cfg->n_layers;
  sp->n_z = cfg->n_z;
  sp->n_phi = cfg->n_phi;
  sp->n_triplets = cfg->n_triplets;
  sp->cells = (uint32_t)cells;
  sp->cell_bits = bits_for((uint32_t)cells);
  // GridConfig (Grid.h:9-15, 33-37): float arithmetic on the float constants
  const float min_z = -300.0f, max_z = 300.0f;
  const float max_phi = (float)M_PI, min_phi = (float)(-M_PI);
  sp->min_z = min_z;
  sp->sector_size_z = (max_z - min_z) / (float)cfg->n_z;
  sp->min_phi = min_phi;
  sp->sector_size_phi = (max_phi - min_phi) / (float)cfg->n_phi;
  sp->min_radius_curvature = cfg->min_radius_curvature;
  for (uint32_t l = 0; l < TRAX_MAX_LAYERS; ++l) {
    sp->min_radius[l] = cfg->min_radius[l];
    sp->max_radius[l] = cfg->max_radius[l];
    const float twoR = 2 * sp->min_radius_curvature;
    sp->acos_min_r[l] = trax_acosf(sp->min_radius[l] / twoR);
    sp->acos_max_r[l] = trax_acosf(sp->max_radius[l] / twoR);
  }
  for (uint32_t p = 0; p < cfg->n_triplets; ++p) {
    const trax_layer_triplet& t = cfg->triplets[p];
    // The predictor and filter identify the layer triplet by its innermost
    // layer (TripletThetaPhiPredictor.h:79, TripletThetaPhiFilter.h:73).
    if (t.layer1 != p + 1) return fail(ctx, TRAX_ERR_INVALID, "layer triplet p must have layer1 == p + 1");
    if (t.layer2 < 1 || t.layer2 > cfg->n_layers || t.layer3 < 1 || t.layer3 > cfg->n_layers ||
        t.layer1 > cfg->n_layers)
      return fail(ctx, TRAX_ERR_INVALID, "layer triplet names a layer outside 1..n_layers");
    sp->layer1[p] = t.layer1;
    sp->layer2[p] = t.layer2;
    sp->layer3[p] = t.layer3;
    sp->dtheta_cut[p] = t.dtheta_cut;
    sp->sigma_z[p] = t.sigma_z;
    sp->dphi_cut[p] = t.dphi_cut;
    sp->sigma_phi[p] = t.sigma_phi;
    sp->tip_cut[p] = t.tip_cut;
    sp->z0[p] = t.z0;
    // the pair kernel's sigmaPhi argument is bound to the uint pairSpreadPhi
    // column (PairGeneratorBeamspot.cpp:55): its bits are read as a float
    memcpy(&sp->pair_sigma_phi[p], &t.pair_spread_phi, 4);
    sp->spread_z[p] = t.pair_spread_z;
    sp->spread_phi[p] = t.pair_spread_phi;
    // the sector window stays inside its grid row only for 2 * spread <= nPhi
    // (the reference indexes neighbouring rows beyond that)
    if (cfg->pair_generator == TRAX_PAIRS_SECTOR && 2ull * t.pair_spread_phi > cfg->n_phi)
      return fail(ctx, TRAX_ERR_INVALID, "sector pairs need 2 * pair_spread_phi <= n_phi");
  }
  if (cfg->pair_generator != TRAX_PAIRS_BEAMSPOT && cfg->pair_generator != TRAX_PAIRS_SECTOR)
    return fail(ctx, TRAX_ERR_INVALID, "unknown pair generator");
  // Beam-spot windows span many z rows only on fine z binnings; there two
  // lanes share each outer hit's window (PU200, 256 z rows: pair count
  // 148 -> 142 us, store 314 -> 309 us), while on coarse ones (the ttbar
  // configuration's 50 z rows) the second window evaluation costs more than
  // the shorter walks save (84 -> 102 us).
  sp->pair_parts = cfg->n_z >= 128 ? 2u : 1u;
  return TRAX_OK;
}

trax::SortedHits sorted_view(const trax_hits* h) {
  trax::SortedHits s;
  s.x = h->x;
  s.y = h->y;
  s.z = h->z;
  s.phi = h->phi;
  s.layer = h->layer;
  s.evt = h->evt;
  s.zphi = reinterpret_cast<const float2*>(h->zphi);
  s.xyzphi = reinterpret_cast<const float4*>(h->xyzphi);
  return s;
}

bool sorted_ok(const trax_hits* h) {
  return h && h->x && h->y && h->z && h->phi && h->zphi && h->xyzphi &&
         !(reinterpret_cast<uintptr_t>(h->xyzphi) & 15u);
}

int check_ctx(trax_ctx* ctx) {
  if (!ctx) return TRAX_ERR_INVALID;
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
  return TRAX_OK;
}

// Largest number of layer triplets that share one middle layer: the
// beam-spot pair passes walk a middle layer's hits once per such triplet.
uint32_t layer2_multiplicity(const trax_seed_config* cfg) {
  uint32_t m = 1;
  for (uint32_t p = 0; p < cfg->n_triplets && p < TRAX_MAX_TRIPLETS; ++p) {
    uint32_t c = 0;
    for (uint32_t q = 0; q < cfg->n_triplets && q < TRAX_MAX_TRIPLETS; ++q)
      c += cfg->triplets[q].layer2 == cfg->triplets[p].layer2 ? 1u : 0u;
    m = c > m ? c : m;
  }
  return m;
}

// Pair work items a batch of max_hits hits can need under cfg.
uint64_t pair_items_needed(const trax_seed_config* cfg, uint32_t max_hits) {
  return (uint64_t)TRAX_PAIR_PARTS * max_hits * layer2_multiplicity(cfg);
}

// The workspace must have been laid out for (at least) this configuration:
// its work offsets for n_triplets per event, its pair counts for the pair
// work items, its event count.
int check_ws(trax_ctx* ctx, const trax_workspace* ws, uint32_t n_events, const trax_seed_config* cfg) {
  if (!ws || !ws->work_offsets || !ws->pair_counts || !ws->cand_counts)
    return fail(ctx, TRAX_ERR_INVALID, "incomplete workspace");
  if (n_events > ws->max_events) return fail(ctx, TRAX_ERR_INVALID, "more events than the workspace holds");
  if (cfg->n_triplets > ws->n_triplets)
    return fail(ctx, TRAX_ERR_INVALID, "workspace laid out for fewer layer triplets than the configuration has");
  if (pair_items_needed(cfg, ws->max_hits) > ws->pair_items)
    return fail(ctx, TRAX_ERR_INVALID,
                "workspace pair counts too small for this configuration (layer triplets sharing a middle layer)");
  if ((((uintptr_t)ws->pair_counts) | ((uintptr_t)ws->cand_counts)) & 15u)
    return fail(ctx, TRAX_ERR_INVALID, "workspace counts must be 16-byte aligned");
  return TRAX_OK;
}

// Pair work items of the beam-spot passes: one or two per middle-layer hit
// and layer triplet walking that layer (validated against the workspace by
// check_ws).  The count kernel clamps larger work lists to this and raises
// TRAX_FLAG_LAYOUT (hits beyond max_hits).
uint32_t pair_items_max(const trax::SeedParams& sp, const trax_seed_config* cfg, const trax_workspace* ws) {
  const uint64_t need = (uint64_t)sp.pair_parts * ws->max_hits * layer2_multiplicity(cfg);
  return need < ws->pair_items ? (uint32_t)need : ws->pair_items;
}

// Sector pairs: one work item per inner-layer hit (inner layers are distinct).
uint32_t sector_items_max(const trax_workspace* ws) {
  return ws->max_hits < ws->pair_items ? ws->max_hits : ws->pair_items;
}

int scan_sub(const trax_ctx* ctx) { return (int)(ctx->launch.scan_tile / 1024u); }

uint64_t pair_cap(const trax_workspace* ws, const trax_seed_outputs* out) {
  return out->pair_capacity < ws->pair_capacity ? out->pair_capacity : ws->pair_capacity;
}

}  // namespace

// ------------------------------------------------------------------ context
extern "C" int trax_abi_version(void) { return TRAX_ABI_VERSION; }

extern "C" int trax_ctx_create(int device, trax_ctx** out) {
  if (!out) return TRAX_ERR_INVALID;
  *out = nullptr;
  trax_ctx* ctx = new trax_ctx();
  ctx->device = device;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) {
    delete ctx;
    return TRAX_ERR_HIP;
  }
  const size_t bytes = 256 + trax::kStatusWords * sizeof(unsigned long long);
  e = hipMalloc(&ctx->block, bytes);
  if (e != hipSuccess) {
    delete ctx;
    return TRAX_ERR_NOMEM;
  }
  e = hipMemset(ctx->block, 0, bytes);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    (void)hipFree(ctx->block);
    delete ctx;
    return TRAX_ERR_HIP;
  }
  char* b = static_cast<char*>(ctx->block);
  ctx->ctrl = reinterpret_cast<trax::ScanCtrl*>(b);
  ctx->flags = reinterpret_cast<uint32_t*>(b + 64);
  ctx->status = reinterpret_cast<unsigned long long*>(b + 256);
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess &&
      lds > 0)
    ctx->lds_max = (size_t)lds;
  // leave room for the block's static LDS (scan scratch)
  ctx->lds_max = ctx->lds_max > 1024 ? ctx->lds_max - 1024 : 0;
  if (trax::grid_build_set_lds_limit(ctx->lds_max) != hipSuccess) {
    (void)hipGetLastError();
    ctx->lds_max = 64512;
  }
  *out = ctx;
  return TRAX_OK;
}

extern "C" int trax_ctx_destroy(trax_ctx* ctx) {
  if (!ctx) return TRAX_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->block) (void)hipFree(ctx->block);
  for (hipEvent_t e : ctx->ev_start) (void)hipEventDestroy(e);
  for (hipEvent_t e : ctx->ev_stop) (void)hipEventDestroy(e);
  delete ctx;
  return TRAX_OK;
}

extern "C" int trax_ctx_kernel_timer(trax_ctx* ctx, int kernel) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (kernel < 0 || kernel > TRAX_KERNEL_ACC_CUT) return fail(ctx, TRAX_ERR_INVALID, "unknown kernel id");
  if (ctx->ev_start.empty()) {
    const size_t n = 256;
    for (size_t i = 0; i < n; ++i) {
      hipEvent_t a, b;
      TRAX_HIP_CHECK(ctx, hipEventCreate(&a));
      TRAX_HIP_CHECK(ctx, hipEventCreate(&b));
      ctx->ev_start.push_back(a);
      ctx->ev_stop.push_back(b);
    }
  }
  if ((rc = timer_drain(ctx))) return rc;
  ctx->timed_kernel = kernel;
  ctx->timed_ms = 0.0;
  ctx->timed_launches = 0;
  return TRAX_OK;
}

extern "C" int trax_ctx_kernel_time(trax_ctx* ctx, double* total_ms, uint32_t* launches) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if ((rc = timer_drain(ctx))) return rc;
  if (total_ms) *total_ms = ctx->timed_ms;
  if (launches) *launches = ctx->timed_launches;
  ctx->timed_ms = 0.0;
  ctx->timed_launches = 0;
  return TRAX_OK;
}

extern "C" int trax_ctx_set_launch(trax_ctx* ctx, const trax_launch_config* cfg) {
  if (!ctx || !cfg) return TRAX_ERR_INVALID;
  const uint32_t w = cfg->grid_waves;
  if (!(w == 0 || w == 1 || w == 2 || w == 4 || w == 8 || w == 16))
    return fail(ctx, TRAX_ERR_INVALID, "grid_waves must be 0 (automatic), 1, 2, 4, 8 or 16");
  for (uint32_t b : {cfg->pairs_block, cfg->predict_block})
    if (!(b == 64 || b == 128 || b == 256 || b == 512))
      return fail(ctx, TRAX_ERR_INVALID, "block sizes must be 64, 128, 256 or 512");
  if (!(cfg->scan_tile == 4096 || cfg->scan_tile == 8192 || cfg->scan_tile == 16384))
    return fail(ctx, TRAX_ERR_INVALID, "scan_tile must be 4096, 8192 or 16384");
  uint32_t variant = 0;
  if (cfg->fused_buffer || cfg->fused_block_waves || cfg->fused_waves_per_eu) {
    variant = trax::kNumFusedVariants;
    for (uint32_t v = 0; v < trax::kNumFusedVariants; ++v) {
      const trax::FusedVariant& f = trax::kFusedVariants[v];
      if (f.buffer == cfg->fused_buffer && f.block_waves == cfg->fused_block_waves &&
          f.waves_per_eu == cfg->fused_waves_per_eu)
        variant = v;
    }
    if (variant == trax::kNumFusedVariants)
      return fail(ctx, TRAX_ERR_INVALID, "fused walk shape is not one of the compiled combinations");
  }
  ctx->launch = *cfg;
  ctx->fused_variant = variant;
  return TRAX_OK;
}

extern "C" int trax_ctx_get_launch(const trax_ctx* ctx, trax_launch_config* cfg) {
  if (!ctx || !cfg) return TRAX_ERR_INVALID;
  *cfg = ctx->launch;
  return TRAX_OK;
}

extern "C" const char* trax_ctx_last_error(const trax_ctx* ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

extern "C" int trax_ctx_flags(trax_ctx* ctx, void* stream, uint32_t* flags_out, int clear) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  uint32_t v = 0;
  TRAX_HIP_CHECK(ctx, hipMemcpyAsync(&v, ctx->flags, 4, hipMemcpyDeviceToHost, s));
  TRAX_HIP_CHECK(ctx, hipStreamSynchronize(s));
  if (clear && v) {
    const uint32_t zero = 0;
    TRAX_HIP_CHECK(ctx, hipMemcpyAsync(ctx->flags, &zero, 4, hipMemcpyHostToDevice, s));
    TRAX_HIP_CHECK(ctx, hipStreamSynchronize(s));
  }
  if (flags_out) *flags_out = v;
  return TRAX_OK;
}

extern "C" float trax_min_radius_curvature(float min_pt) {
  // TripletConfiguration.cpp:145-152 with Bz = IDEAL_BZ passed through a float
  const double Q = 1.602177E-19;
  const double GEV_C = 5.344286E-19;
  const float bz = (float)3.8112;
  return (float)(min_pt * GEV_C / (bz * Q) * 1E2);
}

extern "C" uint64_t trax_workspace_layout(const trax_seed_config* cfg, uint32_t max_events,
                                          uint32_t max_hits, uint64_t pair_capacity,
                                          uint64_t cand_capacity, void* base, trax_workspace* ws) {
  if (!cfg || cfg->n_triplets > TRAX_MAX_TRIPLETS) return 0;
  const uint64_t items = pair_items_needed(cfg, max_hits);
  if (items >= 0xFFFFFFFFull) return 0;
  const uint64_t words = trax::filter_mask_words(cand_capacity);
  uint64_t off = 0;
  const uint64_t o_work = off;
  off = align_up(off + ((uint64_t)max_events * cfg->n_triplets + 2) * 4);
  const uint64_t o_pc = off;
  off = align_up(off + (items + 1) * 4);
  const uint64_t o_cc = off;
  off = align_up(off + (pair_capacity + 1) * 4);
  const uint64_t o_stash = off;
  off = align_up(off + (pair_capacity + TRAX_PREDICT_TILE - 1) / TRAX_PREDICT_TILE *
                           TRAX_PREDICT_TILE_SLOTS * 4);
  const uint64_t o_mask = off;
  off = align_up(off + words * 8);
  const uint64_t o_fc = off;
  off = align_up(off + (words + 1) * 4);
  const uint64_t o_tc = off;
  off = align_up(off + ((pair_capacity + TRAX_PREDICT_TILE - 1) / TRAX_PREDICT_TILE + 2) * 4);
  const uint64_t o_ac = off;
  off = align_up(off + ((pair_capacity + TRAX_PREDICT_TILE - 1) / TRAX_PREDICT_TILE + 2) * 4);
  if (ws) {
    char* b = static_cast<char*>(base);
    ws->work_offsets = b ? reinterpret_cast<uint32_t*>(b + o_work) : nullptr;
    ws->pair_counts = b ? reinterpret_cast<uint32_t*>(b + o_pc) : nullptr;
    ws->cand_counts = b ? reinterpret_cast<uint32_t*>(b + o_cc) : nullptr;
    ws->max_events = max_events;
    ws->max_hits = max_hits;
    ws->pair_capacity = pair_capacity;
    ws->cand_capacity = cand_capacity;
    ws->cand_stash = b ? reinterpret_cast<uint32_t*>(b + o_stash) : nullptr;
    ws->filter_mask = b ? reinterpret_cast<unsigned long long*>(b + o_mask) : nullptr;
    ws->filter_counts = b ? reinterpret_cast<uint32_t*>(b + o_fc) : nullptr;
    ws->cand_tile_counts = b ? reinterpret_cast<uint32_t*>(b + o_tc) : nullptr;
    ws->trk_tile_counts = b ? reinterpret_cast<uint32_t*>(b + o_ac) : nullptr;
    ws->n_triplets = cfg->n_triplets;
    ws->pair_items = (uint32_t)items;
  }
  return off;
}

// --------------------------------------------------------------- prefix sum
extern "C" int trax_prefix_sum_dev(trax_ctx* ctx, const uint32_t* d_in, uint32_t* d_out,
                                   uint64_t n_max, const uint32_t* d_n, uint32_t n_add,
                                   void* stream) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (n_max == 0 && !d_n) return TRAX_OK;
  if (!d_in || !d_out) return fail(ctx, TRAX_ERR_INVALID, "null scan buffer");
  if ((((uintptr_t)d_in) | ((uintptr_t)d_out)) & 15u)
    return fail(ctx, TRAX_ERR_INVALID, "scan buffers must be 16-byte aligned");
  if (n_max > 0xFFFFFFFFull) return fail(ctx, TRAX_ERR_INVALID, "scan longer than 2^32 elements");
  TRAX_TIMED(ctx, TRAX_KERNEL_SCAN, static_cast<hipStream_t>(stream), trax::launch_scan_u32(d_in, d_out, n_max, d_n, n_add, ctx->ctrl, ctx->status,
                                            static_cast<hipStream_t>(stream), scan_sub(ctx)));
  return TRAX_OK;
}

extern "C" int trax_prefix_sum(trax_ctx* ctx, uint32_t* d_data, uint64_t n, void* stream) {
  if (n == 0) return check_ctx(ctx);
  return trax_prefix_sum_dev(ctx, d_data, d_data, n, nullptr, 0, stream);
}

extern "C" int trax_prefix_sum_segmented(trax_ctx* ctx, uint32_t* d_data, uint64_t n,
                                         const uint32_t* d_seg_offsets, uint32_t n_segments, void* stream) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (n == 0) return TRAX_OK;
  if (!d_data || !d_seg_offsets) return fail(ctx, TRAX_ERR_INVALID, "null scan buffer or segment offsets");
  if (n_segments == 0) return fail(ctx, TRAX_ERR_INVALID, "a non-empty scan needs at least one segment");
  if ((uint64_t)n_segments > trax::kStatusWords)
    return fail(ctx, TRAX_ERR_INVALID, "more than 2^22 segments");
  // the global scan, then each element minus the scan's value at its
  // segment's start (gathered into the look-back words once the scan is done)
  if ((rc = trax_prefix_sum(ctx, d_data, n, stream))) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  TRAX_HIP_CHECK(ctx, trax::launch_segment_bases(d_data, n, d_seg_offsets, n_segments, ctx->status, s));
  TRAX_HIP_CHECK(ctx, trax::launch_segment_fix(d_data, n, d_seg_offsets, n_segments, ctx->status, s));
  return TRAX_OK;
}

// --------------------------------------------------------------- grid build
extern "C" int trax_grid_build(trax_ctx* ctx, const trax_seed_config* cfg, const trax_hits* in,
                               const trax_layout* layout, trax_hits* sorted, uint32_t* d_grid,
                               void* stream) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!layout) return fail(ctx, TRAX_ERR_INVALID, "null layout");
  trax::SeedParams sp;
  rc = make_params(ctx, cfg, layout->n_events, &sp);
  if (rc) return rc;
  if (layout->n_events == 0) return TRAX_OK;
  if (!in || !in->x || !in->y || !in->z || !in->layer || !in->det_id || !in->hit_id || !in->evt)
    return fail(ctx, TRAX_ERR_INVALID, "incomplete input hits");
  if (!sorted || !sorted->x || !sorted->y || !sorted->z || !sorted->layer || !sorted->det_id ||
      !sorted->hit_id || !sorted->evt || !sorted->orig_index || !sorted->phi || !sorted->zphi ||
      !sorted->xyzphi || (reinterpret_cast<uintptr_t>(sorted->xyzphi) & 15u))
    return fail(ctx, TRAX_ERR_INVALID, "incomplete sorted-hit buffers");
  if (!d_grid || !layout->event_offset || !layout->layer_nhits || !layout->layer_offset)
    return fail(ctx, TRAX_ERR_INVALID, "null grid or layout array");
  if (layout->n_hits >= (1u << 29)) return fail(ctx, TRAX_ERR_INVALID, "at most 2^29 - 1 hits per batch");
  trax::GridLaunch gl;
  gl.ix = in->x; gl.iy = in->y; gl.iz = in->z;
  gl.ilayer = in->layer; gl.idet = in->det_id; gl.ihid = in->hit_id; gl.ievt = in->evt;
  gl.ox = sorted->x; gl.oy = sorted->y; gl.oz = sorted->z;
  gl.olayer = sorted->layer; gl.odet = sorted->det_id; gl.ohid = sorted->hit_id;
  gl.oevt = sorted->evt; gl.oorig = sorted->orig_index; gl.ophi = sorted->phi;
  gl.ozphi = reinterpret_cast<float2*>(sorted->zphi);
  gl.oxyzphi = reinterpret_cast<float4*>(sorted->xyzphi);
  gl.event_offset = layout->event_offset;
  gl.layer_nhits = layout->layer_nhits;
  gl.layer_offset = layout->layer_offset;
  gl.n_hits = layout->n_hits;
  gl.grid = d_grid;
  gl.flags = ctx->flags;
  gl.waves = (int)ctx->launch.grid_waves;
  gl.lds_max = ctx->lds_max;
  TRAX_TIMED(ctx, TRAX_KERNEL_GRID_BUILD, static_cast<hipStream_t>(stream), trax::launch_grid_build(sp, gl, static_cast<hipStream_t>(stream)));
  return TRAX_OK;
}

// -------------------------------------------------------------------- pairs
// The predictor's stash is idle during the pair stage: when it holds one
// 64-bit word per pair work item, count and store run back to back share
// the count pass's verdicts through it (trax_pairs_beamspot).
static unsigned long long* pair_masks(const trax_workspace* ws, uint32_t items_max) {
  const uint64_t stash_bytes = (ws->pair_capacity + TRAX_PREDICT_TILE - 1) / TRAX_PREDICT_TILE *
                               TRAX_PREDICT_TILE_SLOTS * 4ull;
  if (!ws->cand_stash || (reinterpret_cast<uintptr_t>(ws->cand_stash) & 7u) ||
      stash_bytes < (uint64_t)items_max * 8u)
    return nullptr;
  return reinterpret_cast<unsigned long long*>(ws->cand_stash);
}

static int pairs_count_impl(trax_ctx* ctx, const trax_seed_config* cfg, uint32_t n_events,
                            const trax_hits* sorted, const uint32_t* d_grid, const trax_workspace* ws,
                            trax_seed_outputs* out, void* stream, bool masks) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  trax::SeedParams sp;
  if ((rc = make_params(ctx, cfg, n_events, &sp))) return rc;
  if ((rc = check_ws(ctx, ws, n_events, cfg))) return rc;
  if (!sorted_ok(sorted) || !d_grid || !out || !out->pair_offsets || !out->totals)
    return fail(ctx, TRAX_ERR_INVALID, "null pair-stage buffer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint32_t K = n_events * cfg->n_triplets;
  const trax::SortedHits h = sorted_view(sorted);
  TRAX_HIP_CHECK(ctx, trax::launch_pair_plan(sp, d_grid, ws->work_offsets, ctx->ctrl, ctx->status, scan_sub(ctx),
                                             s));
  const uint32_t items_max = pair_items_max(sp, cfg, ws);
  unsigned long long* m = masks ? pair_masks(ws, items_max) : nullptr;
  TRAX_TIMED(ctx, TRAX_KERNEL_PAIRS, s, trax::launch_pairs(sp, h, d_grid, ws->work_offsets, ws->pair_counts, m,
                                         nullptr, 0, items_max, false, ctx->flags, ctx->launch.pairs_block,
                                         s));
  TRAX_TIMED(ctx, TRAX_KERNEL_SCAN, s, trax::launch_scan_u32(ws->pair_counts, ws->pair_counts, (uint64_t)items_max + 1,
                                            ws->work_offsets + K + 1, 1, ctx->ctrl, ctx->status, s,
                                            scan_sub(ctx)));
  // pair_offsets[k] = scanned count at the first item of segment k
  TRAX_HIP_CHECK(ctx, trax::launch_gather_offsets(ws->pair_counts, items_max, ws->work_offsets + K + 1,
                                                  ws->work_offsets, K + 1, out->pair_offsets,
                                                  out->totals + 0, s, sp.pair_parts));
  return TRAX_OK;
}

static int pairs_store_impl(trax_ctx* ctx, const trax_seed_config* cfg, uint32_t n_events,
                            const trax_hits* sorted, const uint32_t* d_grid, const trax_workspace* ws,
                            trax_seed_outputs* out, void* stream, bool masks) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  trax::SeedParams sp;
  if ((rc = make_params(ctx, cfg, n_events, &sp))) return rc;
  if ((rc = check_ws(ctx, ws, n_events, cfg))) return rc;
  if (!sorted_ok(sorted) || !d_grid || !out || (!out->pairs && out->pair_capacity))
    return fail(ctx, TRAX_ERR_INVALID, "null pair-stage buffer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint32_t items_max = pair_items_max(sp, cfg, ws);
  TRAX_TIMED(ctx, TRAX_KERNEL_PAIRS, s, trax::launch_pairs(sp, sorted_view(sorted), d_grid, ws->work_offsets,
                                         ws->pair_counts, masks ? pair_masks(ws, items_max) : nullptr,
                                         reinterpret_cast<uint2*>(out->pairs), out->pair_capacity, items_max,
                                         true, ctx->flags, ctx->launch.pairs_block, s));
  return TRAX_OK;
}

extern "C" int trax_pairs_count(trax_ctx* ctx, const trax_seed_config* cfg, uint32_t n_events,
                                const trax_hits* sorted, const uint32_t* d_grid,
                                const trax_workspace* ws, trax_seed_outputs* out, void* stream) {
  return pairs_count_impl(ctx, cfg, n_events, sorted, d_grid, ws, out, stream, false);
}

extern "C" int trax_pairs_store(trax_ctx* ctx, const trax_seed_config* cfg, uint32_t n_events,
                                const trax_hits* sorted, const uint32_t* d_grid,
                                const trax_workspace* ws, trax_seed_outputs* out, void* stream) {
  return pairs_store_impl(ctx, cfg, n_events, sorted, d_grid, ws, out, stream, false);
}

extern "C" int trax_pairs_beamspot(trax_ctx* ctx, const trax_seed_config* cfg, uint32_t n_events,
                                   const trax_hits* sorted, const uint32_t* d_grid,
                                   const trax_workspace* ws, trax_seed_outputs* out, void* stream) {
  // count -> scan -> store with the total on the device.  Measured against a
  // single pass with tile look-back, this is faster on large batches (PU200
  // 64 events: 0.445 vs 0.554 ms; 4096 ttbar events: 0.146 vs 0.245 ms): in
  // the single pass each tile's stores wait for its predecessors' counts.
  int rc = pairs_count_impl(ctx, cfg, n_events, sorted, d_grid, ws, out, stream, true);
  if (rc) return rc;
  return pairs_store_impl(ctx, cfg, n_events, sorted, d_grid, ws, out, stream, true);
}

// ------------------------------------------------------------- sector pairs
extern "C" int trax_pairs_sector_count(trax_ctx* ctx, const trax_seed_config* cfg, uint32_t n_events,
                                       const trax_hits* sorted, const uint32_t* d_grid,
                                       const trax_workspace* ws, trax_seed_outputs* out, void* stream) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  trax::SeedParams sp;
  if ((rc = make_params(ctx, cfg, n_events, &sp))) return rc;
  if ((rc = check_ws(ctx, ws, n_events, cfg))) return rc;
  if (!sorted_ok(sorted) || !d_grid || !out || !out->pair_offsets || !out->totals)
    return fail(ctx, TRAX_ERR_INVALID, "null pair-stage buffer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint32_t K = n_events * cfg->n_triplets;
  TRAX_HIP_CHECK(ctx, trax::launch_pair_plan(sp, d_grid, ws->work_offsets, ctx->ctrl, ctx->status, scan_sub(ctx),
                                             s, true));
  TRAX_TIMED(ctx, TRAX_KERNEL_PAIRS, s, trax::launch_sector_pairs(sp, sorted_view(sorted), d_grid, ws->work_offsets,
                                                ws->pair_counts, nullptr, 0, sector_items_max(ws), false,
                                                ctx->flags, s));
  TRAX_TIMED(ctx, TRAX_KERNEL_SCAN, s, trax::launch_scan_u32(ws->pair_counts, ws->pair_counts, (uint64_t)sector_items_max(ws) + 1,
                                            ws->work_offsets + K, 1, ctx->ctrl, ctx->status, s,
                                            scan_sub(ctx)));
  // pair_offsets[k] = scanned count at the first inner hit of segment k
  TRAX_HIP_CHECK(ctx, trax::launch_gather_offsets(ws->pair_counts, sector_items_max(ws), ws->work_offsets + K,
                                                  ws->work_offsets, K + 1, out->pair_offsets,
                                                  out->totals + 0, s));
  return TRAX_OK;
}

extern "C" int trax_pairs_sector_store(trax_ctx* ctx, const trax_seed_config* cfg, uint32_t n_events,
                                       const trax_hits* sorted, const uint32_t* d_grid,
                                       const trax_workspace* ws, trax_seed_outputs* out, void* stream) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  trax::SeedParams sp;
  if ((rc = make_params(ctx, cfg, n_events, &sp))) return rc;
  if ((rc = check_ws(ctx, ws, n_events, cfg))) return rc;
  if (!sorted_ok(sorted) || !d_grid || !out || (!out->pairs && out->pair_capacity))
    return fail(ctx, TRAX_ERR_INVALID, "null pair-stage buffer");
  TRAX_TIMED(ctx, TRAX_KERNEL_PAIRS, static_cast<hipStream_t>(stream),
             trax::launch_sector_pairs(sp, sorted_view(sorted), d_grid, ws->work_offsets, ws->pair_counts,
                                       reinterpret_cast<uint2*>(out->pairs), out->pair_capacity,
                                       sector_items_max(ws), true, ctx->flags, static_cast<hipStream_t>(stream)));
  return TRAX_OK;
}

extern "C" int trax_pairs_sector(trax_ctx* ctx, const trax_seed_config* cfg, uint32_t n_events,
                                 const trax_hits* sorted, const uint32_t* d_grid,
                                 const trax_workspace* ws, trax_seed_outputs* out, void* stream) {
  int rc = trax_pairs_sector_count(ctx, cfg, n_events, sorted, d_grid, ws, out, stream);
  if (rc) return rc;
  return trax_pairs_sector_store(ctx, cfg, n_events, sorted, d_grid, ws, out, stream);
}

// ---------------------------------------------------------------- predictor
extern "C" int trax_triplet_predict(trax_ctx* ctx, const trax_seed_config* cfg, uint32_t n_events,
                                    const trax_hits* sorted, const uint32_t* d_grid,
                                    const trax_workspace* ws, trax_seed_outputs* out, void* stream) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  trax::SeedParams sp;
  if ((rc = make_params(ctx, cfg, n_events, &sp))) return rc;
  if ((rc = check_ws(ctx, ws, n_events, cfg))) return rc;
  if (!ws->cand_stash || !ws->cand_tile_counts || (reinterpret_cast<uintptr_t>(ws->cand_tile_counts) & 15u))
    return fail(ctx, TRAX_ERR_INVALID, "workspace has no predictor stash or tile counts");
  if (!sorted_ok(sorted) || !d_grid || !out || !out->pair_offsets || !out->cand_offsets ||
      !out->totals || (!out->pairs && out->pair_capacity) || (!out->cands && out->cand_capacity))
    return fail(ctx, TRAX_ERR_INVALID, "null prediction-stage buffer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint32_t K = n_events * cfg->n_triplets;
  const uint64_t pc = pair_cap(ws, out);
  const trax::SortedHits h = sorted_view(sorted);
  const uint2* pairs = reinterpret_cast<const uint2*>(out->pairs);
  // the stash holds ceil(ws->pair_capacity / 64) slabs; pc <= ws->pair_capacity
  // the walk leaves per-pair counts, per-tile totals and the tile count; only
  // the tile totals are scanned (one entry per 64 pairs)
  const uint64_t tiles_max = (pc + trax::kPredTile - 1) / trax::kPredTile;
  uint32_t* tiles = ws->cand_tile_counts;
  TRAX_TIMED(ctx, TRAX_KERNEL_PREDICT_WALK, s,
             trax::launch_predict_walk(sp, h, d_grid, pairs, out->pair_offsets, out->totals + 0, pc,
                                       ws->cand_counts, ws->cand_stash, tiles, s));
  TRAX_TIMED(ctx, TRAX_KERNEL_SCAN, s,
             trax::launch_scan_u32(tiles, tiles, tiles_max + 1, tiles + tiles_max + 1, 1, ctx->ctrl,
                                   ctx->status, s, scan_sub(ctx)));
  TRAX_HIP_CHECK(ctx, trax::launch_pred_offsets(tiles, ws->cand_counts, out->pair_offsets, out->totals + 0, pc,
                                                K + 1, out->cand_offsets, out->totals + 1, s));
  TRAX_TIMED(ctx, TRAX_KERNEL_PREDICT_PLACE, s,
             trax::launch_predict_place(sp, h, d_grid, pairs, out->pair_offsets, out->totals + 0, pc, tiles,
                                        ws->cand_counts, ws->cand_stash,
                                        reinterpret_cast<uint2*>(out->cands), out->cand_capacity,
                                        ctx->flags, s));
  return TRAX_OK;
}

extern "C" int trax_predict_count(trax_ctx* ctx, const trax_seed_config* cfg, uint32_t n_events,
                                  const trax_hits* sorted, const uint32_t* d_grid,
                                  const trax_workspace* ws, trax_seed_outputs* out, void* stream) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  trax::SeedParams sp;
  if ((rc = make_params(ctx, cfg, n_events, &sp))) return rc;
  if ((rc = check_ws(ctx, ws, n_events, cfg))) return rc;
  if (!sorted_ok(sorted) || !d_grid || !out || !out->pair_offsets || !out->cand_offsets ||
      !out->totals || (!out->pairs && out->pair_capacity))
    return fail(ctx, TRAX_ERR_INVALID, "null prediction-stage buffer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint32_t K = n_events * cfg->n_triplets;
  const uint64_t pc = pair_cap(ws, out);
  TRAX_HIP_CHECK(ctx, trax::launch_predict(sp, sorted_view(sorted), d_grid,
                                           reinterpret_cast<const uint2*>(out->pairs), out->pair_offsets,
                                           out->totals + 0, pc, ws->cand_counts, nullptr, 0, false,
                                           ctx->flags, ctx->launch.predict_block, s));
  TRAX_TIMED(ctx, TRAX_KERNEL_SCAN, s, trax::launch_scan_u32(ws->cand_counts, ws->cand_counts, pc + 1, out->totals + 0, 1,
                                            ctx->ctrl, ctx->status, s, scan_sub(ctx)));
  TRAX_HIP_CHECK(ctx, trax::launch_gather_offsets(ws->cand_counts, (uint32_t)pc, out->totals + 0,
                                                  out->pair_offsets, K + 1, out->cand_offsets,
                                                  out->totals + 1, s));
  return TRAX_OK;
}

extern "C" int trax_predict_store(trax_ctx* ctx, const trax_seed_config* cfg, uint32_t n_events,
                                  const trax_hits* sorted, const uint32_t* d_grid,
                                  const trax_workspace* ws, trax_seed_outputs* out, void* stream) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  trax::SeedParams sp;
  if ((rc = make_params(ctx, cfg, n_events, &sp))) return rc;
  if ((rc = check_ws(ctx, ws, n_events, cfg))) return rc;
  if (!sorted_ok(sorted) || !d_grid || !out || !out->pair_offsets || !out->totals ||
      (!out->cands && out->cand_capacity) || (!out->pairs && out->pair_capacity))
    return fail(ctx, TRAX_ERR_INVALID, "null prediction-stage buffer");
  TRAX_HIP_CHECK(ctx, trax::launch_predict(sp, sorted_view(sorted), d_grid,
                                           reinterpret_cast<const uint2*>(out->pairs), out->pair_offsets,
                                           out->totals + 0, pair_cap(ws, out), ws->cand_counts,
                                           reinterpret_cast<uint2*>(out->cands), out->cand_capacity, true,
                                           ctx->flags, ctx->launch.predict_block,
                                           static_cast<hipStream_t>(stream)));
  return TRAX_OK;
}

// ------------------------------------------------------------------- filter
namespace {
int filter_args(trax_ctx* ctx, const trax_workspace* ws, const trax_seed_outputs* out, bool store) {
  if (!ws || !ws->filter_mask || !ws->filter_counts)
    return fail(ctx, TRAX_ERR_INVALID, "workspace has no filter mask");
  if (!out || !out->cand_offsets || !out->totals || (!out->cands && out->cand_capacity) ||
      (!out->pairs && out->pair_capacity))
    return fail(ctx, TRAX_ERR_INVALID, "null filter-stage buffer");
  if (store && (!out->tracklet_offsets || ((!out->t1 || !out->t2 || !out->t3) && out->tracklet_capacity)))
    return fail(ctx, TRAX_ERR_INVALID, "null filter output buffer");
  if (out->cand_capacity > ws->cand_capacity)
    return fail(ctx, TRAX_ERR_INVALID, "candidate capacity larger than the workspace's");
  if (out->cand_capacity >= 0xFFFFFFC0ull)
    return fail(ctx, TRAX_ERR_INVALID, "candidate capacity must stay below 2^32 - 64");
  return TRAX_OK;
}
}  // namespace

extern "C" int trax_filter_count(trax_ctx* ctx, const trax_seed_config* cfg, uint32_t n_events,
                                 const trax_hits* sorted, const trax_workspace* ws,
                                 trax_seed_outputs* out, void* stream) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  trax::SeedParams sp;
  if ((rc = make_params(ctx, cfg, n_events, &sp))) return rc;
  if ((rc = filter_args(ctx, ws, out, false))) return rc;
  if (!sorted || !sorted->x || !sorted->y || !sorted->z)
    return fail(ctx, TRAX_ERR_INVALID, "null filter-stage buffer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint64_t words = trax::filter_mask_words(ws->cand_capacity);
  uint32_t* d_nwords = ws->filter_counts + words;
  TRAX_TIMED(ctx, TRAX_KERNEL_FILTER_MASK, s, trax::launch_filter_mask(sp, sorted_view(sorted), reinterpret_cast<const uint2*>(out->pairs),
                                               reinterpret_cast<const uint2*>(out->cands), out->cand_offsets,
                                               out->totals + 1, out->cand_capacity, ws->filter_mask,
                                               ws->filter_counts, d_nwords, s));
  TRAX_TIMED(ctx, TRAX_KERNEL_SCAN, s, trax::launch_scan_u32(ws->filter_counts, ws->filter_counts, words, d_nwords, 0,
                                            ctx->ctrl, ctx->status, s, scan_sub(ctx)));
  // total = scanned count at word ceil(n_cands / 64) (the zero word)
  TRAX_HIP_CHECK(ctx, trax::launch_read_last(ws->filter_counts, d_nwords, out->totals + 2, s));
  return TRAX_OK;
}

extern "C" int trax_filter_store(trax_ctx* ctx, const trax_seed_config* cfg, uint32_t n_events,
                                 const trax_workspace* ws, trax_seed_outputs* out, void* stream) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  trax::SeedParams sp;
  if ((rc = make_params(ctx, cfg, n_events, &sp))) return rc;
  if ((rc = filter_args(ctx, ws, out, true))) return rc;
  TRAX_TIMED(ctx, TRAX_KERNEL_FILTER_STORE, static_cast<hipStream_t>(stream), trax::launch_filter_store(sp, reinterpret_cast<const uint2*>(out->pairs),
                                                reinterpret_cast<const uint2*>(out->cands), out->cand_offsets,
                                                out->totals + 1, out->cand_capacity, ws->filter_mask,
                                                ws->filter_counts, out->t1, out->t2, out->t3,
                                                out->tracklet_capacity, out->tracklet_offsets,
                                                out->totals + 2, ctx->flags,
                                                static_cast<hipStream_t>(stream)));
  return TRAX_OK;
}

extern "C" int trax_triplet_filter(trax_ctx* ctx, const trax_seed_config* cfg, uint32_t n_events,
                                   const trax_hits* sorted, const trax_workspace* ws,
                                   trax_seed_outputs* out, void* stream) {
  int rc = trax_filter_count(ctx, cfg, n_events, sorted, ws, out, stream);
  if (rc) return rc;
  return trax_filter_store(ctx, cfg, n_events, ws, out, stream);
}

// ------------------------------------------------------ fused predict+filter
extern "C" int trax_triplet_predict_filter(trax_ctx* ctx, const trax_seed_config* cfg, uint32_t n_events,
                                           const trax_hits* sorted, const uint32_t* d_grid,
                                           const trax_workspace* ws, trax_seed_outputs* out, void* stream) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  trax::SeedParams sp;
  if ((rc = make_params(ctx, cfg, n_events, &sp))) return rc;
  if ((rc = check_ws(ctx, ws, n_events, cfg))) return rc;
  if (!ws->cand_stash || !ws->cand_tile_counts || !ws->trk_tile_counts ||
      (reinterpret_cast<uintptr_t>(ws->cand_tile_counts) & 15u) ||
      (reinterpret_cast<uintptr_t>(ws->trk_tile_counts) & 15u) || (reinterpret_cast<uintptr_t>(ws->cand_stash) & 15u))
    return fail(ctx, TRAX_ERR_INVALID, "workspace has no predictor stash or tile counts");
  if (!sorted_ok(sorted) || !d_grid || !out || !out->pair_offsets || !out->cand_offsets || !out->totals ||
      !out->tracklet_offsets || (!out->pairs && out->pair_capacity) ||
      ((!out->t1 || !out->t2 || !out->t3) && out->tracklet_capacity))
    return fail(ctx, TRAX_ERR_INVALID, "null prediction or filter buffer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint32_t K = n_events * cfg->n_triplets;
  const uint64_t pc = pair_cap(ws, out);
  const trax::SortedHits h = sorted_view(sorted);
  const uint2* pairs = reinterpret_cast<const uint2*>(out->pairs);
  const uint64_t tiles_max = (pc + trax::kPredTile - 1) / trax::kPredTile;
  uint32_t* tiles = ws->cand_tile_counts;
  uint32_t* acc = ws->trk_tile_counts;
  uint4* slab = reinterpret_cast<uint4*>(ws->cand_stash);
  TRAX_TIMED(ctx, TRAX_KERNEL_PREDICT_FILTER, s,
             trax::launch_predict_filter_walk(sp, h, d_grid, pairs, out->pair_offsets, out->totals + 0, pc,
                                              ws->cand_counts, tiles, slab, acc, s, ctx->fused_variant));
  TRAX_TIMED(ctx, TRAX_KERNEL_ACC_CUT, s,
             trax::launch_acc_cut(sp, h, out->totals + 0, pc, slab, acc, s));
  TRAX_HIP_CHECK(ctx, trax::launch_acc_redo(sp, h, d_grid, pairs, out->pair_offsets, out->totals + 0, pc,
                                            ws->cand_counts, slab, acc, s));
  TRAX_TIMED(ctx, TRAX_KERNEL_SCAN, s,
             trax::launch_scan_u32(tiles, tiles, tiles_max + 1, tiles + tiles_max + 1, 1, ctx->ctrl,
                                   ctx->status, s, scan_sub(ctx)));
  TRAX_TIMED(ctx, TRAX_KERNEL_SCAN, s,
             trax::launch_scan_u32(acc, acc, tiles_max + 1, acc + tiles_max + 1, 1, ctx->ctrl, ctx->status, s,
                                   scan_sub(ctx)));
  TRAX_HIP_CHECK(ctx, trax::launch_fused_offsets(tiles, ws->cand_counts, acc, slab, out->pair_offsets,
                                                 out->totals + 0, pc, K + 1, out->cand_offsets,
                                                 out->tracklet_offsets, out->totals + 1, s));
  TRAX_TIMED(ctx, TRAX_KERNEL_TRIPLET_PLACE, s,
             trax::launch_acc_place(sp, h, d_grid, pairs, out->pair_offsets, out->totals + 0, pc, acc,
                                    ws->cand_counts, slab, out->t1, out->t2, out->t3, out->tracklet_capacity,
                                    ctx->flags, s));
  return TRAX_OK;
}

// -------------------------------------------------------------- whole chain
extern "C" int trax_seed_chain(trax_ctx* ctx, const trax_seed_config* cfg, const trax_hits* in,
                               const trax_layout* layout, trax_hits* sorted, uint32_t* d_grid,
                               const trax_workspace* ws, trax_seed_outputs* out, void* stream) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!layout || !ws || !out) return fail(ctx, TRAX_ERR_INVALID, "null layout, workspace or outputs");
  if (layout->n_hits > ws->max_hits) return fail(ctx, TRAX_ERR_INVALID, "more hits than the workspace holds");
  {  // refuse a workspace that cannot hold the batch before anything is enqueued
    trax::SeedParams sp;
    if ((rc = make_params(ctx, cfg, layout->n_events, &sp))) return rc;
    if ((rc = check_ws(ctx, ws, layout->n_events, cfg))) return rc;
  }
  rc = trax_grid_build(ctx, cfg, in, layout, sorted, d_grid, stream);
  if (rc) return rc;
  const uint32_t E = layout->n_events;
  rc = cfg->pair_generator == TRAX_PAIRS_SECTOR
           ? trax_pairs_sector(ctx, cfg, E, sorted, d_grid, ws, out, stream)
           : trax_pairs_beamspot(ctx, cfg, E, sorted, d_grid, ws, out, stream);
  if (rc) return rc;
  if (!out->cands) return trax_triplet_predict_filter(ctx, cfg, E, sorted, d_grid, ws, out, stream);
  if ((rc = trax_triplet_predict(ctx, cfg, E, sorted, d_grid, ws, out, stream))) return rc;
  return trax_triplet_filter(ctx, cfg, E, sorted, ws, out, stream);
}

// ------------------------------------------------------------ chain graphs
struct trax_graph {
  trax_ctx* ctx = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" int trax_chain_graph_capture(trax_ctx* ctx, const trax_seed_config* cfg, const trax_hits* in,
                                        const trax_layout* layout, trax_hits* sorted, uint32_t* d_grid,
                                        const trax_workspace* ws, trax_seed_outputs* out, trax_graph** graph_out) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!graph_out) return fail(ctx, TRAX_ERR_INVALID, "null graph handle");
  *graph_out = nullptr;
  hipStream_t cs = nullptr;
  TRAX_HIP_CHECK(ctx, hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  // the timer brackets eager launches only; a capture records none of its events
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(cs);
    return hip_fail(ctx, e, "hipStreamBeginCapture");
  }
  rc = trax_seed_chain(ctx, cfg, in, layout, sorted, d_grid, ws, out, cs);
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(cs, &g);
  (void)hipStreamDestroy(cs);
  if (rc) {  // the chain refused its arguments: its error text stays
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return hip_fail(ctx, e, "hipStreamEndCapture");
  }
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return hip_fail(ctx, e, "hipGraphInstantiate");
  }
  trax_graph* h = new trax_graph();
  h->ctx = ctx;
  h->graph = g;
  h->exec = x;
  *graph_out = h;
  return TRAX_OK;
}

extern "C" int trax_chain_graph_launch(trax_graph* graph, void* stream) {
  if (!graph || !graph->exec) return TRAX_ERR_INVALID;
  int rc = check_ctx(graph->ctx);
  if (rc) return rc;
  TRAX_HIP_CHECK(graph->ctx, hipGraphLaunch(graph->exec, static_cast<hipStream_t>(stream)));
  return TRAX_OK;
}

extern "C" int trax_chain_graph_destroy(trax_graph* graph) {
  if (!graph) return TRAX_OK;
  if (graph->ctx) (void)hipSetDevice(graph->ctx->device);
  if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
  if (graph->graph) (void)hipGraphDestroy(graph->graph);
  delete graph;
  return TRAX_OK;
}

// ------------------------------------------------------------- test support
namespace trax {

__global__ void k_fp32_eval(int op, const float* __restrict__ a, const float* __restrict__ b,
                            float* __restrict__ out, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    float r;
    switch (op) {
      case 0: r = trax_atan2f(a[i], b[i]); break;
      case 1: r = trax_atanf(a[i]); break;
      case 2: r = trax_acosf(a[i]); break;
      case 3: r = trax_sign(a[i]); break;
      case 4: r = a[i] / b[i]; break;
      case 5: r = __builtin_sqrtf(a[i]); break;
      default: r = 0.0f;
    }
    out[i] = r;
  }
}

hipError_t launch_fp32_eval(int op, const float* a, const float* b, float* out, uint64_t n,
                            hipStream_t stream) {
  uint64_t blocks = (n + 255) / 256;
  if (blocks == 0) return hipSuccess;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_fp32_eval, dim3((uint32_t)blocks), dim3(256), 0, stream, op, a, b, out, n);
  return hipGetLastError();
}

}  // namespace trax

// ------------------------------------------------------------ connectivity
extern "C" uint64_t trax_connectivity_workspace_bytes(uint64_t n_base) {
  if (n_base > (1ull << 40)) return 0;
  return (trax::connectivity_table_slots(n_base) + 1) * sizeof(trax::ConnSlot);
}

extern "C" int trax_triplet_connectivity(trax_ctx* ctx, const trax_tracklets* base, trax_tracklets* following,
                                         uint32_t mode, void* d_workspace, uint64_t workspace_bytes,
                                         void* stream) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!base || !following) return fail(ctx, TRAX_ERR_INVALID, "null triplet set");
  if (mode & ~(TRAX_CONNECT_BACKWARD | TRAX_CONNECT_TIGHT))
    return fail(ctx, TRAX_ERR_INVALID, "unknown connectivity mode bits");
  const bool backward = mode & TRAX_CONNECT_BACKWARD, tight = mode & TRAX_CONNECT_TIGHT;
  // TripletConnectivity.cpp:56: assert(!iterateBackwards) for tight packing
  if (tight && backward) return fail(ctx, TRAX_ERR_INVALID, "tight packing is forward only");
  if (base->n > (1ull << 40)) return fail(ctx, TRAX_ERR_INVALID, "base set larger than 2^40 triplets");
  if (base->n == 0 || following->n == 0) return TRAX_OK;
  const bool base_cols = tight ? (base->hit2 && base->hit3) : backward ? base->hit1 != nullptr : base->hit3 != nullptr;
  const bool follow_cols = tight ? (following->hit1 && following->hit2)
                                 : backward ? following->hit3 != nullptr : following->hit1 != nullptr;
  if (!base_cols || !base->connectivity) return fail(ctx, TRAX_ERR_INVALID, "base set lacks a column the mode reads");
  if (!follow_cols || !following->connectivity)
    return fail(ctx, TRAX_ERR_INVALID, "following set lacks a column the mode reads");
  const uint64_t cap = trax::connectivity_table_slots(base->n);
  if (!d_workspace || workspace_bytes < (cap + 1) * sizeof(trax::ConnSlot))
    return fail(ctx, TRAX_ERR_INVALID, "connectivity workspace smaller than trax_connectivity_workspace_bytes");
  if (reinterpret_cast<uintptr_t>(d_workspace) & 15u)
    return fail(ctx, TRAX_ERR_INVALID, "connectivity workspace must be 16-byte aligned");
  trax::ConnTriplets b{base->hit1, base->hit2, base->hit3, base->connectivity, base->n};
  trax::ConnTriplets f{following->hit1, following->hit2, following->hit3, following->connectivity, following->n};
  hipStream_t s = static_cast<hipStream_t>(stream);
  TRAX_TIMED(ctx, TRAX_KERNEL_CONNECT, s,
             trax::launch_connectivity(b, f, tight ? 2 : (backward ? 1 : 0),
                                       static_cast<trax::ConnSlot*>(d_workspace), cap, ctx->flags, s));
  return TRAX_OK;
}

extern "C" int trax_fp32_eval(trax_ctx* ctx, int op, const float* d_a, const float* d_b, float* d_out,
                              uint64_t n, void* stream) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (op < 0 || op > 5) return fail(ctx, TRAX_ERR_INVALID, "unknown op");
  if (n && (!d_a || !d_b || !d_out)) return fail(ctx, TRAX_ERR_INVALID, "null buffer");
  TRAX_HIP_CHECK(ctx, trax::launch_fp32_eval(op, d_a, d_b, d_out, n, static_cast<hipStream_t>(stream)));
  return TRAX_OK;
}
