// Sort + dedup that allocates nothing: srd_prim_sort_dedup_noalloc.
//
// Callers that cannot call hipMalloc -- code under HIP graph capture, tight
// serving loops, harnesses that forbid allocation -- pass device columns in
// and device buffers out; every intermediate (packed keys in two buffers,
// digit histograms, the radix passes' look-back words, dedup tile counts,
// the device-side plan) lives in `scratch`, a file-scope device array of
// kNoAllocScratchBytes that each GPU the library runs on holds until the
// process ends.  The sort is the device-planned radix sort + dedup
// (build_index_dev): row count, key layout and pass schedule are decided by
// kernels, so the call enqueues a fixed launch sequence and never reads
// anything back.
//
// The scratch is always carved for kNoAllocMaxRows rows whatever n is, so
// its look-back words sit at fixed addresses and are only ever written as
// look-back words; a new look-back epoch per pass (a counter kept in the
// scratch and advanced by a kernel) makes the words of earlier calls read as
// "not yet published", and the words are cleared by a kernel at set-up and
// whenever the 16-bit epoch wraps.  Every call on a GPU
// shares that GPU's array: calls that overlap on one GPU (two streams, two
// threads) race on it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/srdatalog_prims.h"
#include "srd_runtime.h"

namespace srd {
void capi_set_error(const std::string& msg);
}

namespace {
constexpr uint64_t kNoAllocMaxRows = uint64_t(1) << 20;
// devsort_scratch_bytes(2^20) = 34,643,712 bytes (checked at set-up)
constexpr uint64_t kNoAllocScratchBytes = uint64_t(34) << 20;
constexpr int kNoAllocMaxDevices = 64;
}  // namespace

// Host code never uses this name as an address: it is resolved per device
// with hipGetSymbolAddress at set-up.
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {

std::mutex g_na_mu;
void* g_na_ptr[kNoAllocMaxDevices] = {};

// Device address of `scratch` on the current device `dev`, resolved once per
// device; the first call also clears the look-back words on stream s.
void* noalloc_scratch(int dev, hipStream_t s) {
  if (g_na_ptr[dev]) return g_na_ptr[dev];
  if (srd::devsort_scratch_bytes(kNoAllocMaxRows) > kNoAllocScratchBytes)
    throw std::runtime_error("srd_prim_sort_dedup_noalloc: scratch array smaller than the sort needs");
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    throw std::runtime_error(std::string("srd_prim_sort_dedup_noalloc: hipGetSymbolAddress failed: ") +
                             hipGetErrorString(e));
  srd::devsort_clear_lookback(p, kNoAllocMaxRows, s);
  g_na_ptr[dev] = p;
  return p;
}

}  // namespace

extern "C" int srd_prim_sort_dedup_noalloc(const uint32_t* cols, uint64_t stride, uint64_t n,
                                           uint32_t k, uint32_t* out_cols, uint64_t out_stride,
                                           uint32_t* root_vals, uint32_t* root_offs,
                                           uint64_t* counts, void* stream) {
  try {
    if (k == 0 || k > 4) throw std::runtime_error("srd_prim_sort_dedup_noalloc: arity must be in [1, 4]");
    if (n > kNoAllocMaxRows)
      throw std::runtime_error("srd_prim_sort_dedup_noalloc: " + std::to_string(n) +
                               " rows exceed the maximum of " + std::to_string(kNoAllocMaxRows));
    if (!cols || !out_cols || !root_vals || !root_offs || !counts)
      throw std::runtime_error("srd_prim_sort_dedup_noalloc: null buffer");
    if (k > 1 && (stride < n || out_stride < n))
      throw std::runtime_error("srd_prim_sort_dedup_noalloc: column stride below the row count");
    int dev = 0;
    SRD_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices)
      throw std::runtime_error("srd_prim_sort_dedup_noalloc: device id out of range");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint32_t* in[4] = {};
    for (uint32_t c = 0; c < k; ++c) in[c] = cols + uint64_t(c) * stride;
    std::lock_guard<std::mutex> g(g_na_mu);
    void* sp = noalloc_scratch(dev, s);
    // w = 1 holds any two 32-bit columns; three or four need two key words
    srd::build_index_dev(in, k, kNoAllocMaxRows, nullptr, n, k <= 2 ? 1u : 2u, sp, out_cols,
                         out_stride, root_vals, root_offs, counts, /*ticket=*/true, s,
                         /*timed=*/false);
    srd::capi_set_error("");
    return 0;
  } catch (const std::exception& e) {
    srd::capi_set_error(e.what());
    return 1;
  } catch (...) {
    srd::capi_set_error("unknown error");
    return 2;
  }
}
