// Host-only build of the runtime's device memory pool (srd_pool.cpp) over
// the simulated hipMalloc of hip/hip_runtime.h, with a C ABI for the tests.
// Streams are small integers cast to hipStream_t.
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "srd_runtime.h"

namespace srd {
void fail(const std::string& msg) { throw std::runtime_error(msg); }
void check(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace srd

SimDevice& sim_device() {
  static SimDevice d;
  return d;
}

static hipStream_t stream_of(uint64_t id) { return reinterpret_cast<hipStream_t>(id); }

static std::vector<uint64_t>& event_log() {
  static std::vector<uint64_t> v;
  return v;
}

extern "C" {
// Records and waits, as op * 2^32 + stream id, in call order.
void sim_stream_event(int op, hipStream_t s) {
  event_log().push_back((uint64_t(op) << 32) | reinterpret_cast<uint64_t>(s));
}
// Copies up to cap entries of the log; returns its length.
uint64_t sim_event_log(uint64_t* out, uint64_t cap) {
  const std::vector<uint64_t>& v = event_log();
  for (uint64_t i = 0; i < cap && i < v.size(); ++i) out[i] = v[i];
  return v.size();
}
void sim_pool_register(uint64_t stream) { srd::pool_register_stream(stream_of(stream)); }
// 0 on failure (the pool raised an error).
uint64_t sim_pool_alloc(uint64_t bytes, uint64_t stream) {
  try {
    return reinterpret_cast<uint64_t>(srd::pool_alloc(bytes, stream_of(stream)));
  } catch (const std::exception&) {
    return 0;
  }
}
void sim_pool_free(uint64_t p, uint64_t bytes, uint64_t stream) {
  srd::pool_free(reinterpret_cast<void*>(p), bytes, stream_of(stream));
}
void sim_pool_retire(uint64_t stream) { srd::pool_retire_stream(stream_of(stream)); }
void sim_pool_budget(uint64_t bytes) { srd::pool_set_budget(bytes); }
// in_use, peak, budget, allocs, driver_allocs, reserved, cached, trims,
// cross_stream
void sim_pool_stats(uint64_t* out) {
  const srd::PoolStats s = srd::pool_stats();
  const uint64_t v[9] = {s.in_use, s.peak,   s.budget, s.allocs,      s.driver_allocs,
                         s.reserved, s.cached, s.trims, s.cross_stream};
  for (int i = 0; i < 9; ++i) out[i] = v[i];
}
void sim_device_capacity(uint64_t bytes) { sim_device().capacity = bytes; }
// used, hipMalloc calls, hipFree calls, device synchronisations, blocks
void sim_device_stats(uint64_t* out) {
  const SimDevice& d = sim_device();
  out[0] = d.used;
  out[1] = uint64_t(d.mallocs);
  out[2] = uint64_t(d.frees);
  out[3] = uint64_t(d.syncs);
  out[4] = d.blocks.size();
}
// 1 if [p, p + bytes) lies inside one block handed out by hipMalloc
int sim_device_contains(uint64_t p, uint64_t bytes) {
  const SimDevice& d = sim_device();
  auto it = d.blocks.upper_bound(reinterpret_cast<void*>(p));
  if (it == d.blocks.begin()) return 0;
  --it;
  const uint64_t b = reinterpret_cast<uint64_t>(it->first);
  return p >= b && p + bytes <= b + it->second ? 1 : 0;
}
}
