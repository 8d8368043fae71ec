// trax_connect: chains triplet sets through the C++ TripletConnectivity class
// the way the reference's tests do (src/algorithms/test/
// TripletConnectivityTest.h: run(A, B), run(B, C), run(C, D), ...).
//
//   trax_connect <sets.bin> <connectivity.bin>
//
// sets.bin (little endian): "TRXC", u32 version=1, u32 nSets, u32 mode
// (bit 0 iterateBackwards, bit 1 tightPacking), then per set u32 n and the
// columns hit1[n], hit2[n], hit3[n], connectivity[n].
// connectivity.bin: per set, the connectivity[n] column after the chain.
#include <cstdio>
#include <string>
#include <vector>

#include "trax_seeding.hpp"

namespace {

std::vector<uint32_t> read_u32(FILE* f, size_t n) {
  std::vector<uint32_t> v(n);
  if (n && fread(v.data(), 4, n, f) != n) throw trax::Error("truncated input");
  return v;
}

uint32_t read_one(FILE* f) { return read_u32(f, 1)[0]; }

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s <sets.bin> <connectivity.bin>\n", argv[0]);
    return 2;
  }
  try {
    FILE* in = fopen(argv[1], "rb");
    if (!in) throw trax::Error(std::string("cannot open ") + argv[1]);
    if (read_one(in) != 0x43585254u /* "TRXC" */ || read_one(in) != 1u) throw trax::Error("not a triplet-set file");
    const uint32_t n_sets = read_one(in), mode = read_one(in);
    trax::Context ctx(0);
    std::vector<std::unique_ptr<trax::TrackletCollection>> sets;
    for (uint32_t k = 0; k < n_sets; ++k) {
      const uint32_t n = read_one(in);
      auto t = std::make_unique<trax::TrackletCollection>(n, 1, 1);
      t->hit1.upload(read_u32(in, n), ctx.stream());
      t->hit2.upload(read_u32(in, n), ctx.stream());
      t->hit3.upload(read_u32(in, n), ctx.stream());
      t->connectivity.upload(read_u32(in, n), ctx.stream());
      t->count = n;
      sets.push_back(std::move(t));
    }
    fclose(in);
    trax::TripletConnectivity conn(ctx);
    for (uint32_t k = 0; k + 1 < n_sets; ++k) conn.run(*sets[k], *sets[k + 1], mode & 1u, mode & 2u);
    ctx.finish();
    FILE* out = fopen(argv[2], "wb");
    if (!out) throw trax::Error(std::string("cannot open ") + argv[2]);
    for (auto& t : sets) {
      auto v = t->connectivity.download(t->count, ctx.stream());
      if (!v.empty() && fwrite(v.data(), 4, v.size(), out) != v.size()) throw trax::Error("write failed");
    }
    fclose(out);
    printf("trax_connect: %u sets chained\n", n_sets);
    return 0;
  } catch (const std::exception& e) {
    fprintf(stderr, "trax_connect: %s\n", e.what());
    return 1;
  }
}
