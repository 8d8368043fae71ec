// trax_seed: runs the seeding stages through the C++ algorithm classes, in
// the order buildTriplets uses them (src/triplet_builder/TripletBuilder.cpp:
// 240-263): GridBuilder, PairGeneratorBeamspot, TripletThetaPhiPredictor,
// TripletThetaPhiFilter, one event group per input file.
//
//   trax_seed <events.bin> <results.bin> [beamspot|sector] [fused] [graph]
//
// "sector" runs PairGeneratorSector in place of PairGeneratorBeamspot;
// "fused" runs the predictor and the filter as one pass
// (TripletThetaPhiPredictFilter): the results then hold no candidate
// records (count 0), only the candidate offsets.  "graph" then records the
// whole chain once as a HIP graph (SeedChainGraph, capacities from the
// stage-by-stage run) and replays it three times; every replay must equal
// the stage-by-stage pairs, triplets and offsets, and the results written
// are the last replay's (no candidate records, as with "fused").
//
// events.bin (little endian): "TRXE", u32 version=1, u32 nEvents, nLayers,
// nHits, nSectorsZ, nSectorsPhi, nTriplets, f32 minPt, f32 minRadius[nLayers],
// f32 maxRadius[nLayers], nTriplets rows of (u32 l1, l2, l3, f32 dThetaCut,
// sigmaZ, dPhiCut, sigmaPhi, tipCut, u32 pairSpreadZ, f32 z0, u32
// pairSpreadPhi, u32 nCandidates), u32 eventOffset[E], layerNHits[E*L],
// layerOffset[E*L], f32 x[n], y[n], z[n], u32 layer[n], detId[n], hitId[n],
// evt[n].
// results.bin: "TRXR", u32 nHits, origIndex[nHits], u32 gridSize,
// grid[gridSize], then for pairs, candidates and triplets the count and the
// records (pairs/cands: 2 u32 each, triplets: 3 u32 each) and the
// nEvents*nTriplets+1 offsets.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "trax_seeding.hpp"

namespace {

struct Reader {
  FILE* f;
  template <typename T>
  T get() {
    T v;
    if (fread(&v, sizeof(T), 1, f) != 1) throw trax::Error("truncated input");
    return v;
  }
  template <typename T>
  std::vector<T> vec(size_t n) {
    std::vector<T> v(n);
    if (n && fread(v.data(), sizeof(T), n, f) != n) throw trax::Error("truncated input");
    return v;
  }
};

template <typename T>
void put(FILE* f, const std::vector<T>& v) {
  if (!v.empty() && fwrite(v.data(), sizeof(T), v.size(), f) != v.size()) throw trax::Error("write failed");
}
void put_u32(FILE* f, uint32_t v) {
  if (fwrite(&v, 4, 1, f) != 1) throw trax::Error("write failed");
}

}  // namespace

int main(int argc, char** argv) {
  bool sector = false, fused = false, graph = false, bad = false;
  for (int i = 3; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "sector") sector = true;
    else if (a == "fused") fused = true;
    else if (a == "graph") graph = true;
    else if (a != "beamspot") bad = true;
  }
  if (argc < 3 || argc > 6 || bad) {
    fprintf(stderr, "usage: %s <events.bin> <results.bin> [beamspot|sector] [fused] [graph]\n", argv[0]);
    return 2;
  }
  try {
    FILE* in = fopen(argv[1], "rb");
    if (!in) throw trax::Error(std::string("cannot open ") + argv[1]);
    Reader r{in};
    if (r.get<uint32_t>() != 0x45585254u /* "TRXE" */ || r.get<uint32_t>() != 1u)
      throw trax::Error("not a trax event file");
    const uint32_t E = r.get<uint32_t>(), L = r.get<uint32_t>(), N = r.get<uint32_t>();
    trax::GridConfig gc;
    gc.nEvents = E;
    gc.nLayers = L;
    gc.nSectorsZ = r.get<uint32_t>();
    gc.nSectorsPhi = r.get<uint32_t>();
    const uint32_t P = r.get<uint32_t>();
    trax::TripletConfigurations trip(r.get<float>());
    trax::GeometrySupplement geom;
    geom.minRadius = r.vec<float>(L);
    geom.maxRadius = r.vec<float>(L);
    for (uint32_t p = 0; p < P; ++p) {
      const uint32_t l1 = r.get<uint32_t>(), l2 = r.get<uint32_t>(), l3 = r.get<uint32_t>();
      const float dth = r.get<float>(), sz = r.get<float>(), dph = r.get<float>(), sph = r.get<float>(),
                  tip = r.get<float>();
      const uint32_t psz = r.get<uint32_t>();
      const float z0 = r.get<float>();
      const uint32_t psp = r.get<uint32_t>(), nc = r.get<uint32_t>();
      trip.addWithValue(l1, l2, l3, dth, sz, dph, sph, tip, psz, z0, psp, nc);
    }
    trax::EventSupplement es;
    trax::LayerSupplement ls;
    ls.nLayers = L;
    es.offset = r.vec<uint32_t>(E);
    ls.nHits = r.vec<uint32_t>((size_t)E * L);
    ls.offset = r.vec<uint32_t>((size_t)E * L);
    trax::HitCollection hits;
    hits.x = r.vec<float>(N);
    hits.y = r.vec<float>(N);
    hits.z = r.vec<float>(N);
    hits.layer = r.vec<uint32_t>(N);
    hits.detId = r.vec<uint32_t>(N);
    hits.hitId = r.vec<uint32_t>(N);
    hits.evt = r.vec<uint32_t>(N);
    fclose(in);

    trax::Context ctx(0);
    trax::Workspace ws;
    trax::Totals tot;
    hits.toDevice(ctx);
    es.dOffset.upload(es.offset, ctx.stream());
    ls.dNHits.upload(ls.nHits, ctx.stream());
    ls.dOffset.upload(ls.offset, ctx.stream());
    trax::Grid grid(gc);

    trax::GridBuilder gridBuilder(ctx, ws);
    gridBuilder.run(hits, 256, es, ls, grid, geom, trip);
    std::unique_ptr<trax::Pairing> pairs;
    if (sector) {
      trax::PairGeneratorSector pairGen(ctx, ws, tot);
      pairs = pairGen.run(hits, geom, 256, trip, grid);
    } else {
      trax::PairGeneratorBeamspot pairGen(ctx, ws, tot);
      pairs = pairGen.run(hits, geom, 256, trip, grid);
    }
    std::unique_ptr<trax::Pairing> cands;
    std::unique_ptr<trax::TrackletCollection> trk;
    if (fused) {  // no candidate records: an empty candidate section with its offsets
      cands = std::make_unique<trax::Pairing>(0, E, P);
      trax::TripletThetaPhiPredictFilter predictFilter(ctx, ws, tot);
      trk = predictFilter.run(hits, geom, 256, trip, grid, *pairs, &cands->pairingOffsets);
    } else {
      trax::TripletThetaPhiPredictor predictor(ctx, ws, tot);
      cands = predictor.run(hits, geom, 256, trip, grid, *pairs);
      trax::TripletThetaPhiFilter filter(ctx, ws, tot);
      trk = filter.run(hits, grid, *pairs, *cands, 256, trip, geom);
    }
    ctx.finish();
    if (uint32_t f = ctx.flags()) throw trax::Error("device status flags " + std::to_string(f));

    const size_t K1 = (size_t)E * P + 1;
    hipStream_t s = ctx.stream();
    if (graph) {
      // the whole chain as one graph, replayed three times against the
      // stage-by-stage results
      const std::vector<uint32_t> wantPairs = pairs->pairing.download(2 * pairs->count, s);
      const std::vector<uint32_t> wantPairOff = pairs->pairingOffsets.download(K1, s);
      const std::vector<uint32_t> want1 = trk->hit1.download(trk->count, s), want2 = trk->hit2.download(trk->count, s),
                                  want3 = trk->hit3.download(trk->count, s);
      const std::vector<uint32_t> wantOff = trk->trackletOffsets.download(K1, s);
      const std::vector<uint32_t> wantCandOff = cands->pairingOffsets.download(K1, s);
      auto g = std::make_unique<trax::SeedChainGraph>(ctx, hits, es, ls, grid, geom, trip, pairs->count ? pairs->count : 1,
                                                      trk->count ? trk->count : 1,
                                                      sector ? TRAX_PAIRS_SECTOR : TRAX_PAIRS_BEAMSPOT);
      for (int rep = 0; rep < 3; ++rep) {
        g->launch();
        if (uint32_t f = g->status()) throw trax::Error("graph replay: device status flags " + std::to_string(f));
        const uint32_t np = g->pairCount(), nt = g->tripletCount();
        if (np != pairs->count || nt != trk->count) throw trax::Error("graph replay: different totals");
        if (g->pairs().pairing.download(2 * (size_t)np, s) != wantPairs ||
            g->pairs().pairingOffsets.download(K1, s) != wantPairOff ||
            g->tracklets().hit1.download(nt, s) != want1 || g->tracklets().hit2.download(nt, s) != want2 ||
            g->tracklets().hit3.download(nt, s) != want3 ||
            g->tracklets().trackletOffsets.download(K1, s) != wantOff ||
            g->candOffsets().download(K1, s) != wantCandOff)
          throw trax::Error("graph replay " + std::to_string(rep) + " differs from the stage-by-stage run");
      }
      // results of the last replay (no candidate records)
      auto gp = std::make_unique<trax::Pairing>(pairs->count, E, P);
      auto gt = std::make_unique<trax::TrackletCollection>(trk->count, E, P);
      auto gc = std::make_unique<trax::Pairing>(0, E, P);
      auto copy = [&](const trax::DeviceArray<uint32_t>& from, trax::DeviceArray<uint32_t>& to, size_t n) {
        if (n) trax::hip_check(hipMemcpyAsync(to.get(), from.get(), n * 4, hipMemcpyDeviceToDevice, s), "copy");
      };
      copy(g->pairs().pairing, gp->pairing, 2 * pairs->count);
      copy(g->pairs().pairingOffsets, gp->pairingOffsets, K1);
      copy(g->tracklets().hit1, gt->hit1, trk->count);
      copy(g->tracklets().hit2, gt->hit2, trk->count);
      copy(g->tracklets().hit3, gt->hit3, trk->count);
      copy(g->tracklets().trackletOffsets, gt->trackletOffsets, K1);
      copy(g->candOffsets(), gc->pairingOffsets, K1);
      ctx.finish();
      pairs = std::move(gp);
      trk = std::move(gt);
      cands = std::move(gc);
    }
    FILE* out = fopen(argv[2], "wb");
    if (!out) throw trax::Error(std::string("cannot open ") + argv[2]);
    put_u32(out, 0x52585254u /* "TRXR" */);
    put_u32(out, N);
    put(out, hits.sorig.download(N, s));
    put_u32(out, (uint32_t)grid.size());
    put(out, grid.boundary.download(grid.size(), s));
    put_u32(out, (uint32_t)pairs->count);
    put(out, pairs->pairing.download(2 * pairs->count, s));
    put(out, pairs->pairingOffsets.download(K1, s));
    put_u32(out, (uint32_t)cands->count);
    put(out, cands->pairing.download(2 * cands->count, s));
    put(out, cands->pairingOffsets.download(K1, s));
    put_u32(out, (uint32_t)trk->count);
    put(out, trk->hit1.download(trk->count, s));
    put(out, trk->hit2.download(trk->count, s));
    put(out, trk->hit3.download(trk->count, s));
    put(out, trk->trackletOffsets.download(K1, s));
    fclose(out);
    printf("trax_seed: %u events, %u hits, %llu pairs, %llu candidates, %llu triplets\n", E, N,
           (unsigned long long)pairs->count, (unsigned long long)cands->count, (unsigned long long)trk->count);
    return 0;
  } catch (const std::exception& e) {
    fprintf(stderr, "trax_seed: %s\n", e.what());
    return 1;
  }
}
