// Device memory pool of the runtime (host code only: compiled by hipcc into
// the runtime library, and by g++ against a simulated allocator for the
// host-only tests).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <iterator>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "srd_runtime.h"

namespace srd {

// ================================================================ pool ====
// A caching allocator over hipMalloc, in the manner of the reference's RMM
// pool (gpu/device_array_rmm_impl.h:102-160: memory obtained from the driver
// is kept while the process runs, and free ranges are split and coalesced).
//
// Requests above 1 MiB are carved out of arenas (hipMalloc'd regions of at
// least 1 GiB): best fit over the free ranges, the rest of the range stays
// free, and a freed range merges with free neighbours of the same arena.  A
// free range is tagged with the stream it was freed on: a later allocation
// on that stream reuses it without synchronising (stream order makes the
// reuse safe); once the stream is retired (synchronised) the range serves
// any stream and merges with other such ranges.  Before going to the driver,
// a request also takes a range freed on another stream the pool knows to
// live until the process ends (the library's own, registered when created),
// after making its stream wait for an event recorded on that stream: runs
// on concurrent streams then share one cache instead of each growing its
// own (unless a stream capture is open, which such a wait would join).
// So the memory one run
// obtained serves any mix of sizes in the next, and the reservation stays
// near the peak in use.  Smaller requests are cached per power-of-two class.
//
// The driver is avoided in steady state: the stream-ordered hipMallocAsync
// pool this replaced blocked the host for 0.1-5 s per run once many blocks
// had passed through it, and memory handed back with hipFree costs a clear
// when the driver hands it out again (3.4 s for 247 GB measured).  Only when
// hipMalloc fails are cached blocks and wholly free arenas handed back, least
// recently freed first, until the request succeeds.
namespace {
std::mutex g_pool_mu;
PoolStats g_stats;
bool g_pool_ready = false;
int g_device = 0;
uint64_t g_free_seq = 0;
uint64_t g_cached = 0;                          // bytes on the free lists

constexpr uint64_t kSmallMax = uint64_t(1) << 20;
constexpr uint64_t kArenaMin = uint64_t(1) << 30;
constexpr uint64_t kGranule = uint64_t(2) << 20;  // large sizes are multiples

struct FreeBlock {
  void* p;
  hipStream_t s;
  bool any_stream;  // its stream was synchronised when retired
  uint64_t seq;     // order of the frees (oldest first when handing back)
};
std::multimap<uint64_t, FreeBlock> g_small;     // class bytes -> block

struct Arena {
  uint64_t bytes;
};
std::map<char*, Arena> g_arenas;                // base -> arena

struct Range {
  uint64_t bytes;
  char* arena;      // base of the arena holding it
  hipStream_t s;
  bool any_stream;
  uint64_t seq;
};
std::map<char*, Range> g_ranges;                // start -> free range (address order)
std::set<hipStream_t> g_known;                  // streams that are never destroyed
std::multimap<uint64_t, char*> g_by_size;       // bytes -> start of a free range

struct Live {
  uint64_t bytes;   // block bytes
  bool small;
};
std::map<void*, Live> g_live;

uint64_t env_bytes(const char* a, const char* b) {
  const char* v = std::getenv(a);
  if (!v) v = std::getenv(b);
  if (!v || !*v) return 0;
  return std::strtoull(v, nullptr, 10);
}

// Size classes: powers of two up to 1 MiB, 2 MiB multiples beyond.
uint64_t block_bytes(uint64_t bytes) {
  if (bytes <= kSmallMax) {
    uint64_t b = 512;
    while (b < bytes) b <<= 1;
    return b;
  }
  return (bytes + kGranule - 1) / kGranule * kGranule;
}

void by_size_erase(uint64_t bytes, char* start) {
  for (auto it = g_by_size.lower_bound(bytes); it != g_by_size.end() && it->first == bytes; ++it)
    if (it->second == start) {
      g_by_size.erase(it);
      return;
    }
}

void range_erase(std::map<char*, Range>::iterator it) {
  by_size_erase(it->second.bytes, it->first);
  g_ranges.erase(it);
}

std::map<char*, Range>::iterator range_insert(char* start, const Range& r) {
  g_by_size.emplace(r.bytes, start);
  return g_ranges.emplace(start, r).first;
}

bool mergeable(const Range& a, const Range& b) {
  return a.arena == b.arena && (a.any_stream ? b.any_stream : (!b.any_stream && a.s == b.s));
}

// Merge the free range at `it` with free neighbours it may join.
void coalesce(std::map<char*, Range>::iterator it) {
  if (it != g_ranges.begin()) {
    auto prev = std::prev(it);
    if (prev->first + prev->second.bytes == it->first && mergeable(prev->second, it->second)) {
      Range r = prev->second;
      r.bytes += it->second.bytes;
      r.seq = std::max(r.seq, it->second.seq);
      char* start = prev->first;
      range_erase(it);
      range_erase(prev);
      it = range_insert(start, r);
    }
  }
  auto next = std::next(it);
  if (next != g_ranges.end() && it->first + it->second.bytes == next->first &&
      mergeable(it->second, next->second)) {
    Range r = it->second;
    r.bytes += next->second.bytes;
    r.seq = std::max(r.seq, next->second.seq);
    char* start = it->first;
    range_erase(next);
    range_erase(it);
    range_insert(start, r);
  }
}

// Best-fit free range for `b` bytes usable on stream s (lock held).
void* take_range(uint64_t b, hipStream_t s) {
  for (auto it = g_by_size.lower_bound(b); it != g_by_size.end(); ++it) {
    auto rit = g_ranges.find(it->second);
    Range r = rit->second;
    if (!(r.any_stream || r.s == s)) continue;
    char* start = rit->first;
    range_erase(rit);
    if (r.bytes > b) {
      Range rest = r;
      rest.bytes = r.bytes - b;
      range_insert(start + b, rest);
    }
    g_cached -= b;
    return start;
  }
  return nullptr;
}

bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) {
    (void)hipGetLastError();
    return true;
  }
  return st != hipStreamCaptureStatusNone;
}

// Best-fit free range for `b` bytes left by another registered stream; s
// is made to wait for that stream's work enqueued so far, which includes
// every use of the range (lock held; no other thread's capture open).
void* take_range_cross(uint64_t b, hipStream_t s) {
  for (auto it = g_by_size.lower_bound(b); it != g_by_size.end(); ++it) {
    auto rit = g_ranges.find(it->second);
    const Range r = rit->second;
    if (r.any_stream || r.s == s || !g_known.count(r.s) || capturing(r.s)) continue;
    hipEvent_t ev = nullptr;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    const bool ok = hipEventRecord(ev, r.s) == hipSuccess && hipStreamWaitEvent(s, ev, 0) == hipSuccess;
    (void)hipEventDestroy(ev);  // released once the recorded work completes
    if (!ok) {
      (void)hipGetLastError();
      return nullptr;
    }
    char* start = rit->first;
    range_erase(rit);
    if (r.bytes > b) {
      Range rest = r;  // still tagged with r.s
      rest.bytes = r.bytes - b;
      range_insert(start + b, rest);
    }
    g_cached -= b;
    g_stats.cross_stream++;
    return start;
  }
  return nullptr;
}

// Whether no other thread may have a stream capture open and s is not
// capturing: a capture holds capture_mutex() throughout, so a lock attempt
// that succeeds rules out other threads (this thread's own capture is
// checked per stream).  The caller releases `cg`.
bool no_capture(hipStream_t s, std::unique_lock<std::recursive_mutex>& cg) {
  cg = std::unique_lock<std::recursive_mutex>(capture_mutex(), std::try_to_lock);
  return cg.owns_lock() && !capturing(s);
}

void* take_small(uint64_t b, hipStream_t s) {
  for (auto it = g_small.lower_bound(b); it != g_small.end() && it->first == b; ++it)
    if (it->second.s == s || it->second.any_stream) {
      void* p = it->second.p;
      g_cached -= b;
      g_small.erase(it);
      return p;
    }
  return nullptr;
}

// Every cached block and range serves any stream once the device is idle;
// merge the ranges that then touch (lock held).
void settle_all() {
  for (auto& kv : g_small) kv.second.any_stream = true;
  for (auto& kv : g_ranges) kv.second.any_stream = true;
  auto it = g_ranges.begin();
  while (it != g_ranges.end()) {
    auto nx = std::next(it);
    if (nx != g_ranges.end() && it->first + it->second.bytes == nx->first &&
        mergeable(it->second, nx->second)) {
      Range r = it->second;
      r.bytes += nx->second.bytes;
      r.seq = std::max(r.seq, nx->second.seq);
      char* start = it->first;
      range_erase(nx);
      range_erase(it);
      it = range_insert(start, r);
    } else {
      it = nx;
    }
  }
}

// Hand back cached small blocks and wholly free arenas, least recently freed
// first, until at least `need` bytes went back (lock held, device idle).
uint64_t hand_back(uint64_t need) {
  struct Cand {
    uint64_t seq;
    void* p;
    uint64_t bytes;
    bool small;
  };
  std::vector<Cand> c;
  for (auto& kv : g_small) c.push_back({kv.second.seq, kv.second.p, kv.first, true});
  for (auto& kv : g_ranges) {
    const Range& r = kv.second;
    if (r.any_stream && kv.first == r.arena && r.bytes == g_arenas[r.arena].bytes)
      c.push_back({r.seq, kv.first, r.bytes, false});
  }
  std::sort(c.begin(), c.end(), [](const Cand& x, const Cand& y) { return x.seq < y.seq; });
  uint64_t freed = 0;
  for (const Cand& x : c) {
    if (freed >= need) break;
    if (x.small) {
      for (auto it = g_small.lower_bound(x.bytes); it != g_small.end() && it->first == x.bytes; ++it)
        if (it->second.p == x.p) {
          g_small.erase(it);
          break;
        }
    } else {
      range_erase(g_ranges.find(static_cast<char*>(x.p)));
      g_arenas.erase(static_cast<char*>(x.p));
    }
    (void)hipFree(x.p);
    freed += x.bytes;
    g_cached -= x.bytes;
  }
  return freed;
}

// New driver memory for a request of b block bytes (lock held; may release
// and retake it to hand cached memory back).
void* driver_alloc(uint64_t b, bool small, hipStream_t s, std::unique_lock<std::mutex>& g) {
  // large requests open an arena whose rest (fresh memory, no pending
  // work) is free for any stream
  uint64_t got = small ? b : std::max(b, kArenaMin);
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, got);
  if (e != hipSuccess && got > b) {
    (void)hipGetLastError();
    got = b;
    e = hipMalloc(&p, got);
  }
  if (e != hipSuccess) {
    // Hand cached memory back once the device is idle, then retry.  No
    // other thread's stream capture may be open meanwhile (lock order:
    // capture, then pool).
    (void)hipGetLastError();
    g.unlock();
    std::lock_guard<std::recursive_mutex> cg(capture_mutex());
    g.lock();
    (void)hipDeviceSynchronize();
    settle_all();
    if (void* q = small ? take_small(b, s) : take_range(b, s)) return q;
    g_stats.trims++;
    got = b;
    // hand back a little at a time (an eighth of the request per retry), so
    // that no more cached memory goes back than the request needs
    while (e != hipSuccess && hand_back(b / 8 + 1) > 0) {
      p = nullptr;
      e = hipMalloc(&p, got);
      if (e != hipSuccess) (void)hipGetLastError();
    }
    if (e != hipSuccess)
      fail("srdatalog pool: hipMalloc(" + std::to_string(b) + ") failed: " + hipGetErrorString(e));
  }
  g_stats.reserved += got;
  g_stats.driver_allocs++;
  if (!small) {
    char* base = static_cast<char*>(p);
    g_arenas[base] = Arena{got};
    if (got > b) {
      g_cached += got - b;
      range_insert(base + b, Range{got - b, base, nullptr, true, ++g_free_seq});
    }
  }
  return p;
}
}  // namespace

void pool_init(int device) {
  std::lock_guard<std::mutex> g(g_pool_mu);
  if (g_pool_ready) return;
  SRD_CHECK(hipSetDevice(device));
  g_device = device;
  // the environment's budget, unless one was set before the first allocation
  if (const uint64_t b = env_bytes("SRDATALOG_POOL_MAX_SIZE", "SRDATALOG_RMM_POOL_MAX_SIZE"))
    g_stats.budget = b;
  g_pool_ready = true;
}

std::recursive_mutex& capture_mutex() {
  static std::recursive_mutex m;
  return m;
}

void* pool_alloc(uint64_t bytes, hipStream_t s) {
  if (!g_pool_ready) pool_init(0);
  const uint64_t b = block_bytes(std::max<uint64_t>(bytes, 1));
  const bool small = b <= kSmallMax;
  std::unique_lock<std::mutex> g(g_pool_mu);
  if (g_stats.budget && g_stats.in_use + bytes > g_stats.budget)
    fail("srdatalog pool: allocation of " + std::to_string(bytes) + " bytes exceeds the budget of " +
         std::to_string(g_stats.budget) + " bytes (" + std::to_string(g_stats.in_use) + " in use)");
  void* p = small ? take_small(b, s) : take_range(b, s);
  if (!p && !small && !g_known.empty()) {
    std::unique_lock<std::recursive_mutex> cg;
    if (no_capture(s, cg)) p = take_range_cross(b, s);
  }
  if (!p) p = driver_alloc(b, small, s, g);
  g_live[p] = Live{b, small};
  g_stats.in_use += bytes;
  g_stats.peak = std::max(g_stats.peak, g_stats.in_use);
  g_stats.allocs++;
  return p;
}

void pool_free(void* p, uint64_t bytes, hipStream_t s) {
  if (!p) return;
  std::lock_guard<std::mutex> g(g_pool_mu);
  auto it = g_live.find(p);
  if (it == g_live.end()) return;  // not from this pool
  const Live l = it->second;
  g_live.erase(it);
  g_stats.in_use -= bytes;
  g_cached += l.bytes;
  if (l.small) {
    g_small.emplace(l.bytes, FreeBlock{p, s, false, ++g_free_seq});
    return;
  }
  char* c = static_cast<char*>(p);
  char* arena = std::prev(g_arenas.upper_bound(c))->first;
  coalesce(range_insert(c, Range{l.bytes, arena, s, false, ++g_free_seq}));
}

void pool_retire_stream(hipStream_t s) {
  (void)hipStreamSynchronize(s);
  std::lock_guard<std::mutex> g(g_pool_mu);
  for (auto& kv : g_small)
    if (kv.second.s == s) kv.second.any_stream = true;
  std::vector<char*> mine;
  for (auto& kv : g_ranges)
    if (!kv.second.any_stream && kv.second.s == s) {
      kv.second.any_stream = true;
      mine.push_back(kv.first);
    }
  for (char* c : mine) {
    auto it = g_ranges.find(c);
    if (it != g_ranges.end()) coalesce(it);
  }
}

void pool_register_stream(hipStream_t s) {
  std::lock_guard<std::mutex> g(g_pool_mu);
  g_known.insert(s);
}

PoolStats pool_stats() {
  std::lock_guard<std::mutex> g(g_pool_mu);
  PoolStats st = g_stats;
  st.cached = g_cached;
  return st;
}

void pool_set_budget(uint64_t bytes) {
  std::lock_guard<std::mutex> g(g_pool_mu);
  g_stats.budget = bytes;
}

void pool_reset_peak() {
  std::lock_guard<std::mutex> g(g_pool_mu);
  g_stats.peak = g_stats.in_use;
}

}  // namespace srd
