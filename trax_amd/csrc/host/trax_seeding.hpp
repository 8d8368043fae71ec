// C++ host interface of the seeding path, over the C ABI of libtraxhip.so.
//
// Same class names and run() shapes as the reference's algorithm classes
// (paths relative to the reference's src/):
//   GridBuilder::run                  algorithms/GridBuilder.h:42
//   PairGeneratorBeamspot::run        algorithms/PairGeneratorBeamspot.h:40-41
//   PairGeneratorSector::run          algorithms/PairGeneratorSector.h:38-39
//   TripletThetaPhiPredictor::run     algorithms/TripletThetaPhiPredictor.h:43-44
//   TripletThetaPhiFilter::run        algorithms/TripletThetaPhiFilter.h:46-48
//   PrefixSum::run                    algorithms/PrefixSum.h:42
//   TripletConnectivity::run          algorithms/TripletConnectivity.h:54-56
// and the data holders they exchange (HitCollection, Grid, Pairing,
// TrackletCollection, TripletConfigurations, GeometrySupplement,
// EventSupplement, LayerSupplement).
//
// Differences that a caller of the reference notices:
//  * errors throw trax::Error (the reference has no status codes);
//  * the nThreads argument is accepted and ignored: the gfx950 kernels pick
//    their own launch shape, and pair order no longer depends on it;
//  * GridBuilder::run keeps the unsorted input and adds the sorted columns
//    (plus orig_index and phi) next to it, instead of swapping buffers;
//  * run() results are returned as std::unique_ptr instead of raw new.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/trax_hip.h"

namespace trax {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(e));
}

// Owning device buffer of T.
template <typename T>
class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(size_t n) { resize(n); }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  DeviceArray(DeviceArray&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr, o.n_ = 0; }
  DeviceArray& operator=(DeviceArray&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_, n_ = o.n_;
      o.p_ = nullptr, o.n_ = 0;
    }
    return *this;
  }
  ~DeviceArray() { release(); }
  void resize(size_t n) {
    if (n == n_ && p_) return;
    release();
    n_ = n;
    hip_check(hipMalloc(&p_, (n ? n : 1) * sizeof(T)), "hipMalloc");
  }
  void upload(const std::vector<T>& h, hipStream_t s) {
    resize(h.size());
    if (!h.empty()) hip_check(hipMemcpyAsync(p_, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s), "upload");
  }
  std::vector<T> download(size_t n, hipStream_t s) const {
    std::vector<T> h(n);
    if (n) hip_check(hipMemcpyAsync(h.data(), p_, n * sizeof(T), hipMemcpyDeviceToHost, s), "download");
    hip_check(hipStreamSynchronize(s), "sync");
    return h;
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }

 private:
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* p_ = nullptr;
  size_t n_ = 0;
};

// The device context all algorithm classes share (the reference's
// clever::context): one HIP device, one stream, the trax context.
class Context {
 public:
  explicit Context(int device = 0) {
    hip_check(hipSetDevice(device), "hipSetDevice");
    hip_check(hipStreamCreate(&stream_), "hipStreamCreate");
    if (trax_ctx_create(device, &ctx_) != TRAX_OK) throw Error("trax_ctx_create failed");
  }
  ~Context() {
    trax_ctx_destroy(ctx_);
    (void)hipStreamDestroy(stream_);
  }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  trax_ctx* ctx() const { return ctx_; }
  hipStream_t stream() const { return stream_; }
  void check(int rc, const char* what) const {
    if (rc != TRAX_OK) throw Error(std::string(what) + ": " + trax_ctx_last_error(ctx_));
  }
  void finish() const { hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize"); }
  uint32_t flags(bool clear = true) const {
    uint32_t f = 0;
    check(trax_ctx_flags(ctx_, stream_, &f, clear ? 1 : 0), "trax_ctx_flags");
    return f;
  }

 private:
  trax_ctx* ctx_ = nullptr;
  hipStream_t stream_ = nullptr;
};

// GridConfig (Grid.h:5-39).
struct GridConfig {
  uint32_t nEvents = 0, nLayers = 0, nSectorsZ = 0, nSectorsPhi = 0;
  static constexpr float MIN_Z = -300.0f, MAX_Z = 300.0f;
  float sectorSizeZ() const { return (MAX_Z - MIN_Z) / (float)nSectorsZ; }
  float sectorSizePhi() const { return ((float)M_PI - (float)(-M_PI)) / (float)nSectorsPhi; }
  uint32_t cells() const { return (nSectorsZ + 1) * (nSectorsPhi + 1); }
};

// GeometrySupplement (GeometrySupplement.h:12-81): per-layer radius band.
struct GeometrySupplement {
  std::vector<float> minRadius, maxRadius;  // index: layer - 1
};

// TripletConfigurations (TripletConfiguration.h:39-72).
class TripletConfigurations {
 public:
  explicit TripletConfigurations(float minPt) : minPt_(minPt) {}
  void addWithValue(uint32_t l1, uint32_t l2, uint32_t l3, float dThetaCut, float sigmaZ, float dPhiCut,
                    float sigmaPhi, float tipCut, uint32_t pairSpreadZ, float z0, uint32_t pairSpreadPhi,
                    uint32_t nCandidates) {
    rows_.push_back({l1, l2, l3, dThetaCut, sigmaZ, dPhiCut, sigmaPhi, tipCut, pairSpreadZ, z0, pairSpreadPhi,
                     nCandidates});
  }
  size_t size() const { return rows_.size(); }
  float minRadiusCurvature() const { return trax_min_radius_curvature(minPt_); }
  const std::vector<trax_layer_triplet>& rows() const { return rows_; }
  uint32_t maxLayer() const {
    uint32_t m = 0;
    for (auto& r : rows_) m = r.layer3 > m ? r.layer3 : m;
    return m;
  }

 private:
  float minPt_;
  std::vector<trax_layer_triplet> rows_;
};

inline trax_seed_config make_config(const GridConfig& g, const GeometrySupplement& geom,
                                    const TripletConfigurations& t) {
  trax_seed_config c;
  std::memset(&c, 0, sizeof(c));
  c.n_z = g.nSectorsZ;
  c.n_phi = g.nSectorsPhi;
  c.n_layers = g.nLayers;
  for (size_t l = 0; l < geom.minRadius.size() && l < TRAX_MAX_LAYERS; ++l) {
    c.min_radius[l] = geom.minRadius[l];
    c.max_radius[l] = geom.maxRadius[l];
  }
  if (t.size() > TRAX_MAX_TRIPLETS) throw Error("too many layer triplets");
  c.n_triplets = (uint32_t)t.size();
  for (size_t p = 0; p < t.size(); ++p) c.triplets[p] = t.rows()[p];
  c.min_radius_curvature = t.minRadiusCurvature();
  return c;
}

// EventSupplement / LayerSupplement (EventSupplement.h, LayerSupplement.h).
struct EventSupplement {
  std::vector<uint32_t> offset, nHits;
  DeviceArray<uint32_t> dOffset;
};
struct LayerSupplement {
  uint32_t nLayers = 0;
  std::vector<uint32_t> offset, nHits;  // [event * nLayers + layer - 1]
  DeviceArray<uint32_t> dOffset, dNHits;
};

// HitCollection (HitCollection.h:23): the seven SoA columns, on the host and
// the device; after GridBuilder::run also the grid-sorted columns.
class HitCollection {
 public:
  std::vector<float> x, y, z;
  std::vector<uint32_t> layer, detId, hitId, evt;

  size_t size() const { return x.size(); }
  void addWithValue(float gx, float gy, float gz, uint32_t l, uint32_t det, uint32_t hid, uint32_t e) {
    x.push_back(gx), y.push_back(gy), z.push_back(gz);
    layer.push_back(l), detId.push_back(det), hitId.push_back(hid), evt.push_back(e);
  }
  void toDevice(const Context& c) {
    hipStream_t s = c.stream();
    dx.upload(x, s), dy.upload(y, s), dz.upload(z, s);
    dlayer.upload(layer, s), ddet.upload(detId, s), dhid.upload(hitId, s), devt.upload(evt, s);
    const size_t n = size();
    sx.resize(n), sy.resize(n), sz.resize(n), sphi.resize(n), szphi.resize(2 * n), sxyzphi.resize(4 * n);
    slayer.resize(n), sdet.resize(n), shid.resize(n), sevt.resize(n), sorig.resize(n);
  }
  trax_hits input() const {
    return {dx.get(), dy.get(), dz.get(), dlayer.get(), ddet.get(), dhid.get(), devt.get(), nullptr, nullptr,
            nullptr, nullptr};
  }
  trax_hits sorted() const {
    return {sx.get(), sy.get(), sz.get(), slayer.get(), sdet.get(), shid.get(), sevt.get(), sorig.get(), sphi.get(),
            szphi.get(), sxyzphi.get()};
  }
  // device columns: unsorted input and grid-sorted output
  DeviceArray<float> dx, dy, dz, sx, sy, sz, sphi, szphi, sxyzphi;
  DeviceArray<uint32_t> dlayer, ddet, dhid, devt, slayer, sdet, shid, sevt, sorig;
};

// Grid (Grid.h:49-68): cell starts, [event][layer][z][phi].
class Grid {
 public:
  explicit Grid(const GridConfig& cfg) : config(cfg), boundary((size_t)cfg.nEvents * cfg.nLayers * cfg.cells()) {}
  GridConfig config;
  DeviceArray<uint32_t> boundary;
  size_t size() const { return boundary.size(); }
};

// Pairing (Pairings.h:9-59): uint2 records plus per-(event, triplet) offsets.
class Pairing {
 public:
  Pairing(uint64_t n, uint32_t nEvents, uint32_t nLayerTriplets)
      : pairing(2 * (n ? n : 1)), pairingOffsets((size_t)nEvents * nLayerTriplets + 1), count(n) {}
  DeviceArray<uint32_t> pairing;  // x = first index, y = second index
  DeviceArray<uint32_t> pairingOffsets;
  uint64_t count;
};

// TrackletCollection (TrackletCollection.h:13-77): hit1/2/3, the
// connectivity column (zero at construction) and the offsets.
class TrackletCollection {
 public:
  TrackletCollection(uint64_t n, uint32_t nEvents, uint32_t nLayerTriplets)
      : hit1(n ? n : 1), hit2(n ? n : 1), hit3(n ? n : 1), connectivity(n ? n : 1),
        trackletOffsets((size_t)nEvents * nLayerTriplets + 1), count(n) {
    hip_check(hipMemset(connectivity.get(), 0, connectivity.size() * sizeof(uint32_t)), "hipMemset");
  }
  trax_tracklets view() const {
    return trax_tracklets{hit1.get(), hit2.get(), hit3.get(), connectivity.get(), count};
  }
  DeviceArray<uint32_t> hit1, hit2, hit3, connectivity, trackletOffsets;
  uint64_t count;
};

// Shared per-context scratch for the count phases (kept between runs).
class Workspace {
 public:
  void ensure(const trax_seed_config& cfg, uint32_t events, uint32_t hits, uint64_t pairs,
              uint64_t cands = 1) {
    // a layout of the same sizes for this configuration: its layer-triplet
    // count and pair work items must fit too, not only the capacities
    trax_workspace need{};
    if (trax_workspace_layout(&cfg, std::max(events, ws_.max_events), std::max(hits, ws_.max_hits), 1, 1, nullptr,
                              &need) == 0)
      throw Error("workspace layout rejected the configuration");
    if (events <= ws_.max_events && hits <= ws_.max_hits && pairs <= ws_.pair_capacity &&
        cands <= ws_.cand_capacity && cfg.n_triplets <= ws_.n_triplets && need.pair_items <= ws_.pair_items &&
        mem_.get())
      return;
    events = std::max(events, ws_.max_events);
    hits = std::max(hits, ws_.max_hits);
    pairs = std::max(pairs, ws_.pair_capacity);
    cands = std::max(cands, ws_.cand_capacity);
    const uint64_t n = trax_workspace_layout(&cfg, events, hits, pairs, cands, nullptr, &ws_);
    mem_.resize(n);
    trax_workspace_layout(&cfg, events, hits, pairs, cands, mem_.get(), &ws_);
  }
  const trax_workspace* get() const { return &ws_; }

 private:
  DeviceArray<unsigned char> mem_;
  trax_workspace ws_{};
};

// Totals the stages leave on the device (pairs, candidates, triplets).
struct Totals {
  DeviceArray<uint32_t> d{4};
  uint32_t read(const Context& c, int i) const {
    uint32_t v = 0;
    hip_check(hipMemcpyAsync(&v, d.get() + i, 4, hipMemcpyDeviceToHost, c.stream()), "read total");
    hip_check(hipStreamSynchronize(c.stream()), "sync");
    return v;
  }
};

// Common base of the algorithm classes (KernelWrapper.hpp:12-40).
class KernelWrapper {
 public:
  explicit KernelWrapper(Context& c) : ctx(c) {}

 protected:
  Context& ctx;
};

class PrefixSum : public KernelWrapper {
 public:
  using KernelWrapper::KernelWrapper;
  // In-place exclusive scan of n uint32 (PrefixSum.h:42).
  void run(uint32_t* d_input, uint64_t size, uint32_t /*nThreads*/) {
    ctx.check(trax_prefix_sum(ctx.ctx(), d_input, size, ctx.stream()), "PrefixSum::run");
  }
};

class GridBuilder : public KernelWrapper {
 public:
  GridBuilder(Context& c, Workspace& w) : KernelWrapper(c), ws(w) {}
  // GridBuilder.h:42
  void run(HitCollection& hits, uint32_t /*nThreads*/, const EventSupplement& es, const LayerSupplement& ls,
           Grid& grid, const GeometrySupplement& geom, const TripletConfigurations& t) {
    const trax_seed_config cfg = make_config(grid.config, geom, t);
    ws.ensure(cfg, grid.config.nEvents, (uint32_t)hits.size(), 1);
    trax_hits in = hits.input(), out = hits.sorted();
    trax_layout lay{grid.config.nEvents, (uint32_t)hits.size(), es.dOffset.get(), ls.dNHits.get(), ls.dOffset.get()};
    ctx.check(trax_grid_build(ctx.ctx(), &cfg, &in, &lay, &out, grid.boundary.get(), ctx.stream()),
              "GridBuilder::run");
  }

 private:
  Workspace& ws;
};

class PairGeneratorBeamspot : public KernelWrapper {
 public:
  PairGeneratorBeamspot(Context& c, Workspace& w, Totals& t) : KernelWrapper(c), ws(w), tot(t) {}
  // PairGeneratorBeamspot.h:40-41: count, read the total, allocate, store.
  std::unique_ptr<Pairing> run(HitCollection& hits, const GeometrySupplement& geom, uint32_t /*nThreads*/,
                               const TripletConfigurations& t, const Grid& grid) {
    const trax_seed_config cfg = make_config(grid.config, geom, t);
    const uint32_t E = grid.config.nEvents, P = (uint32_t)t.size();
    ws.ensure(cfg, E, (uint32_t)hits.size(), 1);
    trax_hits s = hits.sorted();
    auto offs = std::make_unique<Pairing>(0, E, P);
    trax_seed_outputs o{};
    o.pair_offsets = offs->pairingOffsets.get();
    o.totals = tot.d.get();
    ctx.check(trax_pairs_count(ctx.ctx(), &cfg, E, &s, grid.boundary.get(), ws.get(), &o, ctx.stream()),
              "pairs count");
    const uint32_t n = tot.read(ctx, 0);
    auto pairs = std::make_unique<Pairing>(n, E, P);
    std::swap(pairs->pairingOffsets, offs->pairingOffsets);
    o.pairs = pairs->pairing.get();
    o.pair_capacity = n;
    o.pair_offsets = pairs->pairingOffsets.get();
    ctx.check(trax_pairs_store(ctx.ctx(), &cfg, E, &s, grid.boundary.get(), ws.get(), &o, ctx.stream()),
              "pairs store");
    return pairs;
  }

 private:
  Workspace& ws;
  Totals& tot;
};

class PairGeneratorSector : public KernelWrapper {
 public:
  PairGeneratorSector(Context& c, Workspace& w, Totals& t) : KernelWrapper(c), ws(w), tot(t) {}
  // PairGeneratorSector.h:38-39 (same arguments as the beam-spot generator):
  // count, read the total, allocate, store.
  std::unique_ptr<Pairing> run(HitCollection& hits, const GeometrySupplement& geom, uint32_t /*nThreads*/,
                               const TripletConfigurations& t, const Grid& grid) {
    trax_seed_config cfg = make_config(grid.config, geom, t);
    cfg.pair_generator = TRAX_PAIRS_SECTOR;
    const uint32_t E = grid.config.nEvents, P = (uint32_t)t.size();
    ws.ensure(cfg, E, (uint32_t)hits.size(), 1);
    trax_hits s = hits.sorted();
    auto offs = std::make_unique<Pairing>(0, E, P);
    trax_seed_outputs o{};
    o.pair_offsets = offs->pairingOffsets.get();
    o.totals = tot.d.get();
    ctx.check(trax_pairs_sector_count(ctx.ctx(), &cfg, E, &s, grid.boundary.get(), ws.get(), &o, ctx.stream()),
              "sector pairs count");
    const uint32_t n = tot.read(ctx, 0);
    auto pairs = std::make_unique<Pairing>(n, E, P);
    std::swap(pairs->pairingOffsets, offs->pairingOffsets);
    o.pairs = pairs->pairing.get();
    o.pair_capacity = n;
    o.pair_offsets = pairs->pairingOffsets.get();
    ctx.check(trax_pairs_sector_store(ctx.ctx(), &cfg, E, &s, grid.boundary.get(), ws.get(), &o, ctx.stream()),
              "sector pairs store");
    return pairs;
  }

 private:
  Workspace& ws;
  Totals& tot;
};

class TripletThetaPhiPredictor : public KernelWrapper {
 public:
  TripletThetaPhiPredictor(Context& c, Workspace& w, Totals& t) : KernelWrapper(c), ws(w), tot(t) {}
  // TripletThetaPhiPredictor.h:43-44 (detector geometry / dictionary are
  // not read by the reference kernels and are not taken here).
  std::unique_ptr<Pairing> run(HitCollection& hits, const GeometrySupplement& geom, int /*nThreads*/,
                               const TripletConfigurations& t, const Grid& grid, const Pairing& pairs) {
    const trax_seed_config cfg = make_config(grid.config, geom, t);
    const uint32_t E = grid.config.nEvents, P = (uint32_t)t.size();
    ws.ensure(cfg, E, (uint32_t)hits.size(), pairs.count);
    trax_hits s = hits.sorted();
    auto cands = std::make_unique<Pairing>(0, E, P);
    trax_seed_outputs o{};
    o.pairs = pairs.pairing.get();
    o.pair_capacity = pairs.count;
    o.pair_offsets = pairs.pairingOffsets.get();
    o.cand_offsets = cands->pairingOffsets.get();
    o.totals = tot.d.get();
    ctx.check(trax_predict_count(ctx.ctx(), &cfg, E, &s, grid.boundary.get(), ws.get(), &o, ctx.stream()),
              "predict count");
    const uint32_t n = tot.read(ctx, 1);
    auto out = std::make_unique<Pairing>(n, E, P);
    std::swap(out->pairingOffsets, cands->pairingOffsets);
    o.cands = out->pairing.get();
    o.cand_capacity = n;
    o.cand_offsets = out->pairingOffsets.get();
    ctx.check(trax_predict_store(ctx.ctx(), &cfg, E, &s, grid.boundary.get(), ws.get(), &o, ctx.stream()),
              "predict store");
    return out;
  }

 private:
  Workspace& ws;
  Totals& tot;
};

class TripletThetaPhiFilter : public KernelWrapper {
 public:
  TripletThetaPhiFilter(Context& c, Workspace& w, Totals& t) : KernelWrapper(c), ws(w), tot(t) {}
  // TripletThetaPhiFilter.h:46-48 / .cpp:10-164: count (cuts, mask
  // popcounts, prefix sum), read the triplet total, allocate, store.
  std::unique_ptr<TrackletCollection> run(HitCollection& hits, const Grid& grid, const Pairing& pairs,
                                          const Pairing& cands, int /*nThreads*/, const TripletConfigurations& t,
                                          const GeometrySupplement& geom) {
    const trax_seed_config cfg = make_config(grid.config, geom, t);
    const uint32_t E = grid.config.nEvents, P = (uint32_t)t.size();
    ws.ensure(cfg, E, (uint32_t)hits.size(), 1, cands.count);
    trax_hits s = hits.sorted();
    trax_seed_outputs o{};
    o.pairs = pairs.pairing.get();
    o.pair_capacity = pairs.count;
    o.cands = cands.pairing.get();
    o.cand_capacity = cands.count;
    o.cand_offsets = cands.pairingOffsets.get();
    o.totals = tot.d.get();
    ctx.check(trax_filter_count(ctx.ctx(), &cfg, E, &s, ws.get(), &o, ctx.stream()), "filter count");
    const uint32_t n = tot.read(ctx, 2);
    auto trk = std::make_unique<TrackletCollection>(n, E, P);
    o.t1 = trk->hit1.get(), o.t2 = trk->hit2.get(), o.t3 = trk->hit3.get();
    o.tracklet_capacity = n;
    o.tracklet_offsets = trk->trackletOffsets.get();
    ctx.check(trax_filter_store(ctx.ctx(), &cfg, E, ws.get(), &o, ctx.stream()), "filter store");
    trk->count = n;
    return trk;
  }

 private:
  Workspace& ws;
  Totals& tot;
};

// The predictor and the filter as one pass (trax_triplet_predict_filter; no
// class of the reference corresponds): for callers that use the triplets
// only.  The triplet count is known only after the pass, so the pass runs
// once with a first guess of the capacity and again when that overflowed.
// `candOffsets` receives the candidate offsets the predictor would have
// written.
class TripletThetaPhiPredictFilter : public KernelWrapper {
 public:
  TripletThetaPhiPredictFilter(Context& c, Workspace& w, Totals& t) : KernelWrapper(c), ws(w), tot(t) {}
  std::unique_ptr<TrackletCollection> run(HitCollection& hits, const GeometrySupplement& geom,
                                          int /*nThreads*/, const TripletConfigurations& t, const Grid& grid,
                                          const Pairing& pairs, DeviceArray<uint32_t>* candOffsets = nullptr) {
    const trax_seed_config cfg = make_config(grid.config, geom, t);
    const uint32_t E = grid.config.nEvents, P = (uint32_t)t.size();
    ws.ensure(cfg, E, (uint32_t)hits.size(), pairs.count);
    trax_hits s = hits.sorted();
    DeviceArray<uint32_t> own((size_t)E * P + 1);
    DeviceArray<uint32_t>& co = candOffsets ? *candOffsets : own;
    if (co.size() < (size_t)E * P + 1) co.resize((size_t)E * P + 1);
    uint64_t cap = pairs.count / 16 + 1024;
    // an earlier stage's status bits would be lost to the overflow check below
    if (const uint32_t f = ctx.flags(true)) throw Error("device status flags " + std::to_string(f));
    for (;;) {
      auto trk = std::make_unique<TrackletCollection>(cap, E, P);
      trax_seed_outputs o{};
      o.pairs = pairs.pairing.get();
      o.pair_capacity = pairs.count;
      o.pair_offsets = pairs.pairingOffsets.get();
      o.cand_offsets = co.get();
      o.t1 = trk->hit1.get(), o.t2 = trk->hit2.get(), o.t3 = trk->hit3.get();
      o.tracklet_capacity = cap;
      o.tracklet_offsets = trk->trackletOffsets.get();
      o.totals = tot.d.get();
      ctx.check(trax_triplet_predict_filter(ctx.ctx(), &cfg, E, &s, grid.boundary.get(), ws.get(), &o,
                                            ctx.stream()),
                "predict + filter");
      const uint32_t n = tot.read(ctx, 2);
      const uint32_t f = ctx.flags(true);
      if (!(f & TRAX_FLAG_TRACKLET_OVERFLOW) && n <= cap) {
        trk->count = n;
        return trk;
      }
      cap = n;
    }
  }

 private:
  Workspace& ws;
  Totals& tot;
};

// The whole chain over one resident event group recorded as a HIP graph
// (trax_chain_graph_capture) and replayed: for callers that seed batch
// after batch of the same shape without the three blocking host reads of
// the stage-by-stage classes (TripletBuilder.cpp:240-263).  Capacities are
// fixed at construction; after a launch, status() reports an overflow and
// the caller rebuilds with larger capacities.
class SeedChainGraph {
 public:
  SeedChainGraph(Context& c, HitCollection& hits, const EventSupplement& es, const LayerSupplement& ls, Grid& grid,
                 const GeometrySupplement& geom, const TripletConfigurations& t, uint64_t pairCapacity,
                 uint64_t tripletCapacity, uint32_t pairGenerator = TRAX_PAIRS_BEAMSPOT)
      : ctx(c),
        cfg_(with_generator(make_config(grid.config, geom, t), pairGenerator)),
        pairs_(pairCapacity, grid.config.nEvents, (uint32_t)t.size()),
        tracklets_(tripletCapacity, grid.config.nEvents, (uint32_t)t.size()),
        candOffsets_((size_t)grid.config.nEvents * t.size() + 1) {
    const uint32_t E = grid.config.nEvents;
    ws_.ensure(cfg_, E, (uint32_t)hits.size(), pairCapacity);
    in_ = hits.input();
    sorted_ = hits.sorted();
    layout_ = trax_layout{E, (uint32_t)hits.size(), es.dOffset.get(), ls.dNHits.get(), ls.dOffset.get()};
    out_ = trax_seed_outputs{};
    out_.pairs = pairs_.pairing.get();
    out_.pair_capacity = pairCapacity;
    out_.pair_offsets = pairs_.pairingOffsets.get();
    out_.cand_offsets = candOffsets_.get();  // no candidate records: the fused prediction + filter
    out_.t1 = tracklets_.hit1.get(), out_.t2 = tracklets_.hit2.get(), out_.t3 = tracklets_.hit3.get();
    out_.tracklet_capacity = tripletCapacity;
    out_.tracklet_offsets = tracklets_.trackletOffsets.get();
    out_.totals = tot_.d.get();
    ctx.check(trax_chain_graph_capture(ctx.ctx(), &cfg_, &in_, &layout_, &sorted_, grid.boundary.get(), ws_.get(),
                                       &out_, &graph_),
              "trax_chain_graph_capture");
  }
  ~SeedChainGraph() { trax_chain_graph_destroy(graph_); }
  SeedChainGraph(const SeedChainGraph&) = delete;
  SeedChainGraph& operator=(const SeedChainGraph&) = delete;

  void launch() { ctx.check(trax_chain_graph_launch(graph_, ctx.stream()), "trax_chain_graph_launch"); }
  // Waits for the stream; the device status bits (0 = no overflow), cleared.
  uint32_t status() {
    ctx.finish();
    return ctx.flags(true);
  }
  uint32_t pairCount() const { return tot_.read(ctx, 0); }
  uint32_t tripletCount() const { return tot_.read(ctx, 2); }
  const Pairing& pairs() const { return pairs_; }
  const TrackletCollection& tracklets() const { return tracklets_; }
  const DeviceArray<uint32_t>& candOffsets() const { return candOffsets_; }

 private:
  static trax_seed_config with_generator(trax_seed_config c, uint32_t g) {
    c.pair_generator = g;
    return c;
  }
  Context& ctx;
  trax_seed_config cfg_;
  Workspace ws_;
  Totals tot_;
  Pairing pairs_;
  TrackletCollection tracklets_;
  DeviceArray<uint32_t> candOffsets_;
  trax_hits in_{}, sorted_{};
  trax_layout layout_{};
  trax_seed_outputs out_{};
  trax_graph* graph_ = nullptr;
};

// TripletConnectivity (TripletConnectivity.h:27-56, .cpp:21-71): adds to
// following.connectivity, per triplet, the connectivity + 1 of every base
// triplet whose outer hit is its inner hit (backward: whose inner hit is its
// outer hit; tightPacking: whose two outer hits are its two inner hits).
class TripletConnectivity : public KernelWrapper {
 public:
  using KernelWrapper::KernelWrapper;
  void run(const TrackletCollection& base, TrackletCollection& following, bool iterateBackwards = false,
           bool tightPacking = false) {
    const uint64_t need = trax_connectivity_workspace_bytes(base.count);
    if (need > ws_.size()) ws_.resize(need);
    const trax_tracklets b = base.view();
    trax_tracklets f = following.view();
    const uint32_t mode = (iterateBackwards ? TRAX_CONNECT_BACKWARD : 0u) | (tightPacking ? TRAX_CONNECT_TIGHT : 0u);
    ctx.check(trax_triplet_connectivity(ctx.ctx(), &b, &f, mode, ws_.get(), ws_.size(), ctx.stream()),
              "TripletConnectivity::run");
  }

 private:
  DeviceArray<unsigned char> ws_;
};

}  // namespace trax
