// Prefix sum whose scratch lives in a fixed device array: for callers that
// must not allocate (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  Same kernels as trax_prefix_sum; the
// look-back control block and status words come from `scratch` instead of a
// context's allocation.
#include <mutex>

#include "../../include/trax_hip.h"
#include "common.h"
#include "kernels.h"

namespace {

// Look-back words in the fixed scratch: the single-pass form needs one per
// 8192-element tile (below 2^21 elements: at most 256), the large-array form
// one per 16384-element chunk, so 2^16 words cover 2^30 elements.
constexpr uint64_t kNoAllocStatusWords = 1ull << 16;
constexpr uint64_t kNoAllocCtrlBytes = 256;
constexpr uint64_t kNoAllocBytes = kNoAllocCtrlBytes + kNoAllocStatusWords * 8;  // 512 KiB + 256 B
static_assert(TRAX_PREFIX_SUM_NOALLOC_MAX == kNoAllocStatusWords * trax::kScanChunk, "maximum size");
static_assert(sizeof(trax::ScanCtrl) <= kNoAllocCtrlBytes, "control block");

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices];  // device address of `scratch` per device, once set up

}  // namespace

// One copy per device, held until the process ends.  In host code this name
// is a host-side stand-in: only the address hipGetSymbolAddress returns may
// be handed to kernels or copies.
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

extern "C" int trax_prefix_sum_noalloc(uint32_t* d_data, uint64_t n, void* stream) {
  if (n == 0) return TRAX_OK;
  if (!d_data || (reinterpret_cast<uintptr_t>(d_data) & 15u)) return TRAX_ERR_INVALID;
  if (n > TRAX_PREFIX_SUM_NOALLOC_MAX) return TRAX_ERR_INVALID;  // beyond the fixed scratch
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return TRAX_ERR_HIP;
  if (dev < 0 || dev >= kMaxDevices) return TRAX_ERR_INVALID;
  hipStream_t s = static_cast<hipStream_t>(stream);
  void* ptr = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
      // first call on this device (current): its copy's address, and a zeroed
      // control block (tile tickets, finished-block count, epoch) ahead of
      // the first scan on the caller's stream
      void* p = nullptr;
      if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return TRAX_ERR_HIP;
      if (trax::launch_scan_ctrl_clear(static_cast<trax::ScanCtrl*>(p), s) != hipSuccess) return TRAX_ERR_HIP;
      g_scratch[dev] = p;
    }
    ptr = g_scratch[dev];
  }
  trax::ScanCtrl* ctrl = static_cast<trax::ScanCtrl*>(ptr);
  unsigned long long* status =
      reinterpret_cast<unsigned long long*>(static_cast<unsigned char*>(ptr) + kNoAllocCtrlBytes);
  if (trax::launch_scan_u32(d_data, d_data, n, nullptr, 0, ctrl, status, s, 8, kNoAllocStatusWords) != hipSuccess)
    return TRAX_ERR_HIP;
  return TRAX_OK;
}
