// Triplet connectivity: how many triplet chains lead into each triplet.
//
// Reference: src/algorithms/TripletConnectivity.h:56-108 (kernels
// tripletConnectivityWide / tripletConnectivityTight), TripletConnectivity.cpp:21-71
// (run) and the known answers of src/algorithms/test/TripletConnectivityTest.h.
// For every following triplet g
//
//   con_f[g] += sum of (con_b[i] + 1) over the base triplets i joined to g   (mod 2^32)
//
// where i joins g when
//   wide, forward    base[i].hit3 == following[g].hit1
//   wide, backward   base[i].hit1 == following[g].hit3
//   tight (forward)  base[i].hit2 == following[g].hit1 && base[i].hit3 == following[g].hit2
//
// The reference loops every following triplet over every base triplet (N x M
// compares per call).  Here the base is folded into an open-addressing hash
// table keyed on the joining hit (or hit pair), one summed weight con_b + 1 per
// key, and each following triplet looks its key up once: O(N + M) work over a
// table of 2N..4N 16-byte slots, which stays in L2 / Infinity Cache at the
// sizes seeding produces.  The insert kernel reads every base value before the
// lookup kernel writes any following value, so base and following may be the
// same arrays (in the reference that case races).
#include "common.h"
#include "kernels.h"

namespace trax {

namespace {

constexpr unsigned long long kConnEmpty = ~0ull;

// fmix64, the finaliser of MurmurHash3 (Austin Appleby, public domain).
__device__ __forceinline__ unsigned long long conn_hash(unsigned long long k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

// MODE: 0 wide forward, 1 wide backward, 2 tight forward.
template <int MODE>
__device__ __forceinline__ unsigned long long base_key(const ConnTriplets& b, uint64_t i) {
  if (MODE == 2) return ((unsigned long long)b.hit2[i] << 32) | b.hit3[i];
  if (MODE == 1) return b.hit1[i];
  return b.hit3[i];
}

template <int MODE>
__device__ __forceinline__ unsigned long long follow_key(const ConnTriplets& f, uint64_t g) {
  if (MODE == 2) return ((unsigned long long)f.hit1[g] << 32) | f.hit2[g];
  if (MODE == 1) return f.hit3[g];
  return f.hit1[g];
}

// Every slot empty with weight 0; slot `cap` holds the weight of the one key
// equal to the empty marker (a tight key whose two hits are both 0xFFFFFFFF).
__global__ __launch_bounds__(256) void k_conn_clear(ConnSlot* __restrict__ table, uint64_t cap) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s <= cap; s += stride) {
    table[s].key = kConnEmpty;
    table[s].weight = 0u;
    table[s].pad = 0u;
  }
}

template <int MODE>
__global__ __launch_bounds__(256) void k_conn_insert(ConnTriplets b, ConnSlot* __restrict__ table, uint64_t mask,
                                                     uint32_t* __restrict__ flags) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < b.n; i += stride) {
    const unsigned long long key = base_key<MODE>(b, i);
    const uint32_t w = b.con[i] + 1u;
    if (key == kConnEmpty) {
      atomicAdd(&table[mask + 1].weight, w);
      continue;
    }
    uint64_t s = conn_hash(key) & mask;
    bool placed = false;
    // at most half the slots are taken, so an empty or matching slot comes
    // within the table length
    for (uint64_t probe = 0; probe <= mask; ++probe) {
      const unsigned long long prev = atomicCAS(&table[s].key, kConnEmpty, key);
      if (prev == kConnEmpty || prev == key) {
        atomicAdd(&table[s].weight, w);
        placed = true;
        break;
      }
      s = (s + 1) & mask;
    }
    if (!placed) atomicOr(flags, TRAX_FLAG_CONNECT_TABLE);
  }
}

template <int MODE>
__global__ __launch_bounds__(256) void k_conn_lookup(ConnTriplets f, const ConnSlot* __restrict__ table,
                                                     uint64_t mask) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < f.n; g += stride) {
    const unsigned long long key = follow_key<MODE>(f, g);
    uint32_t add = 0u;
    if (key == kConnEmpty) {
      add = table[mask + 1].weight;
    } else {
      uint64_t s = conn_hash(key) & mask;
      for (uint64_t probe = 0; probe <= mask; ++probe) {
        const unsigned long long k = table[s].key;
        if (k == key) {
          add = table[s].weight;
          break;
        }
        if (k == kConnEmpty) break;
        s = (s + 1) & mask;
      }
    }
    if (add) f.con[g] += add;
  }
}

uint32_t conn_blocks(uint64_t n) {
  const uint64_t b = (n + 255) / 256;
  return (uint32_t)(b == 0 ? 1 : (b > 8192 ? 8192 : b));
}

template <int MODE>
hipError_t launch_mode(const ConnTriplets& base, const ConnTriplets& following, ConnSlot* table, uint64_t cap,
                       uint32_t* flags, hipStream_t stream) {
  hipLaunchKernelGGL(k_conn_clear, dim3(conn_blocks(cap + 1)), dim3(256), 0, stream, table, cap);
  hipLaunchKernelGGL(k_conn_insert<MODE>, dim3(conn_blocks(base.n)), dim3(256), 0, stream, base, table, cap - 1,
                     flags);
  hipLaunchKernelGGL(k_conn_lookup<MODE>, dim3(conn_blocks(following.n)), dim3(256), 0, stream, following,
                     (const ConnSlot*)table, cap - 1);
  return hipGetLastError();
}

}  // namespace

uint64_t connectivity_table_slots(uint64_t n_base) {
  uint64_t cap = 64;
  while (cap < 2 * n_base) cap <<= 1;
  return cap;
}

hipError_t launch_connectivity(const ConnTriplets& base, const ConnTriplets& following, int mode,
                               ConnSlot* table, uint64_t cap, uint32_t* flags, hipStream_t stream) {
  switch (mode) {
    case 0: return launch_mode<0>(base, following, table, cap, flags, stream);
    case 1: return launch_mode<1>(base, following, table, cap, flags, stream);
    default: return launch_mode<2>(base, following, table, cap, flags, stream);
  }
}

}  // namespace trax
