// Stream-ordered set primitives of the MI355X runtime (see srd_runtime.h).
//
// Design (wave64, gfx950):
//  * Every streaming kernel works on 4096-element tiles: 256 threads = 4
//    waves, each wave owns 1024 consecutive elements and reads them as 16
//    coalesced 64-lane rows ("wave-striped"), so element order inside a
//    wave is (item, lane) and stable ranks come from one ballot per row.
//  * sort = LSD radix over the tightly packed key (columns concatenated at
//    their significant bit widths, column 0 most significant) with 8-bit
//    digits.  The pack kernel also builds every pass's digit histogram, and
//    each pass is ONE kernel: each tile ranks its keys with an 8-ballot
//    digit match, publishes its per-digit count and resolves its global
//    offset with a decoupled look-back over earlier tiles, then writes its
//    keys from an LDS-sorted copy so that each digit run is a coalesced
//    store.  Tile order is the block id when the run is alone on the GPU and
//    a ticket from an atomic counter when other runs may share it.
//  * dedup unpacks the sorted keys straight into the index columns and
//    emits the distinct column-0 values with their row offsets in the same
//    pass.
//  * merge and the dense set difference are merge-path tiles (diagonal
//    split points, both pieces staged in LDS); the sparse set difference
//    stages, per 4096-row tile of NEW, the window of FULL it covers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "srd_devloop.h"
#include "srd_runtime.h"

namespace srd {

void fail(const std::string& msg) { throw std::runtime_error(msg); }

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(std::string(what) + ": " + hipGetErrorString(e));
}

// ------------------------------------------------- streams, pinned memory --
// Engines come and go (one per session, several per serving process); their
// streams and pinned host buffers are kept for the next one instead of being
// destroyed and freed, which synchronises and unmaps.
namespace {
std::mutex g_res_mu;
std::vector<hipStream_t> g_streams;
std::multimap<uint64_t, void*> g_pinned;  // bytes -> pinned block
std::atomic<int> g_running{0};
}  // namespace

hipStream_t stream_acquire() {
  {
    std::lock_guard<std::mutex> g(g_res_mu);
    if (!g_streams.empty()) {
      hipStream_t s = g_streams.back();
      g_streams.pop_back();
      return s;
    }
  }
  hipStream_t s;
  SRD_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  pool_register_stream(s);  // never destroyed: released streams return to g_streams
  return s;
}

void stream_release(hipStream_t s) {
  if (!s) return;
  pool_retire_stream(s);
  std::lock_guard<std::mutex> g(g_res_mu);
  g_streams.push_back(s);
}

void* pinned_acquire(uint64_t bytes, uint64_t* got) {
  {
    std::lock_guard<std::mutex> g(g_res_mu);
    auto it = g_pinned.lower_bound(bytes);
    if (it != g_pinned.end() && it->first <= 2 * bytes + 4096) {
      void* p = it->second;
      *got = it->first;
      g_pinned.erase(it);
      return p;
    }
  }
  void* p = nullptr;
  {
    std::lock_guard<std::recursive_mutex> cg(capture_mutex());
    SRD_CHECK(hipHostMalloc(&p, bytes, hipHostMallocDefault));
  }
  *got = bytes;
  return p;
}

void pinned_release(void* p, uint64_t bytes) {
  if (!p) return;
  std::lock_guard<std::mutex> g(g_res_mu);
  g_pinned.emplace(bytes, p);
}

bool run_enter() { return g_running.fetch_add(1) == 0; }
void run_exit() { g_running.fetch_sub(1); }

Workspace::~Workspace() {
  pinned_release(pinned, pinned_words * 4);
  pinned_release(stage_host, stage_host_bytes);
}

uint32_t* pinned_words(Workspace& ws, uint64_t words) {
  if (words > ws.pinned_words) {
    pinned_release(ws.pinned, ws.pinned_words * 4);
    ws.pinned = nullptr;
    ws.pinned_words = 0;
    uint64_t got = 0;
    ws.pinned = static_cast<uint32_t*>(pinned_acquire(std::max<uint64_t>(words, 1024) * 4, &got));
    ws.pinned_words = got / 4;
  }
  return ws.pinned;
}

// ============================================================== timing ====
// Several engines may run on their own threads (serving); the scope nesting
// is per thread, the collected times are shared under a lock.
namespace {
struct PendingTime {
  std::string name;
  hipEvent_t a, b;
  double bytes;
};
std::atomic<bool> g_timing{false};
std::mutex g_time_mu;
std::vector<PendingTime> g_pending;
std::vector<hipEvent_t> g_free_events;
std::map<std::string, KernelTime> g_times;
thread_local std::string t_cur_name;
thread_local hipEvent_t t_cur_start;
thread_local int t_depth = 0;  // nested scopes: only the outermost one is recorded

hipEvent_t get_event() {
  {
    std::lock_guard<std::mutex> g(g_time_mu);
    if (!g_free_events.empty()) {
      hipEvent_t e = g_free_events.back();
      g_free_events.pop_back();
      return e;
    }
  }
  hipEvent_t e;
  SRD_CHECK(hipEventCreate(&e));
  return e;
}
}  // namespace

void timing_enable(bool on) {
  g_timing = on;
  t_depth = 0;
}
bool timing_enabled() { return g_timing; }

void timing_begin(const char* name, hipStream_t s) {
  if (!g_timing) return;
  if (t_depth++ > 0) return;  // inside another timed scope: that one holds the time
  t_cur_name = name;
  t_cur_start = get_event();
  SRD_CHECK(hipEventRecord(t_cur_start, s));
}

void timing_end(hipStream_t s, double alg_bytes) {
  if (!g_timing || t_depth == 0) return;
  if (--t_depth > 0) return;
  hipEvent_t e = get_event();
  SRD_CHECK(hipEventRecord(e, s));
  std::lock_guard<std::mutex> g(g_time_mu);
  g_pending.push_back({t_cur_name, t_cur_start, e, alg_bytes});
}

std::vector<KernelTime> timing_collect() {
  std::lock_guard<std::mutex> g(g_time_mu);
  for (auto& p : g_pending) {
    SRD_CHECK(hipEventSynchronize(p.b));
    float ms = 0;
    SRD_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
    auto& t = g_times[p.name];
    t.name = p.name;
    t.launches++;
    t.ms += ms;
    t.bytes += p.bytes;
    g_free_events.push_back(p.a);
    g_free_events.push_back(p.b);
  }
  g_pending.clear();
  std::vector<KernelTime> out;
  for (auto& kv : g_times) out.push_back(kv.second);
  return out;
}

void timing_reset() {
  timing_collect();
  std::lock_guard<std::mutex> g(g_time_mu);
  g_times.clear();
}

// A null name makes the scope a no-op (callers that must not create events).
struct TimeScope {
  hipStream_t s;
  double bytes;
  bool on;
  TimeScope(const char* n, hipStream_t st, double b = 0) : s(st), bytes(b), on(n != nullptr) {
    if (on) timing_begin(n, st);
  }
  ~TimeScope() {
    if (on) timing_end(s, bytes);
  }
};

// ================================================================ fill ====
template <class T>
__global__ void fill_kernel(T* p, T v, uint64_t n) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

static uint32_t grid_for(uint64_t n, uint32_t per_block, uint32_t cap = 65535u * 16u) {
  uint64_t g = (n + per_block - 1) / per_block;
  if (g == 0) g = 1;
  return uint32_t(std::min<uint64_t>(g, cap));
}

void fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s) {
  if (!n) return;
  fill_kernel<uint32_t><<<grid_for(n, 1024, 8192), 256, 0, s>>>(p, v, n);
  SRD_CHECK(hipGetLastError());
}

void fill_u64(uint64_t* p, uint64_t v, uint64_t n, hipStream_t s) {
  if (!n) return;
  fill_kernel<uint64_t><<<grid_for(n, 1024, 8192), 256, 0, s>>>(p, v, n);
  SRD_CHECK(hipGetLastError());
}

// ============================================================ gather ====
struct GatherArgs {
  const uint64_t* recs;
  const uint32_t* rec_ctr;
  uint64_t rec_cap;
  const uint32_t* stage;
  uint64_t stage_stride;
  const uint64_t* offsets;
  GatherDest d[2];
  uint32_t nd;
};

// One wave per segment record; lanes copy consecutive rows (coalesced).
__global__ __launch_bounds__(kBlock) void gather_segments_kernel(GatherArgs g) {
  const uint64_t nrec = min<uint64_t>(*g.rec_ctr, g.rec_cap);
  const uint32_t l = lane_id();
  const uint64_t waves = uint64_t(gridDim.x) * kWavesPerBlock;
  for (uint64_t r = uint64_t(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave; r < nrec;
       r += waves) {
    const uint64_t slice = g.recs[3 * r], src = g.recs[3 * r + 1], lc = g.recs[3 * r + 2];
    const uint64_t cnt = lc & 0xFFFFFFFFu, local = lc >> 32;
    const uint64_t dst0 = g.offsets[slice] + local;
    for (uint32_t d = 0; d < g.nd; ++d) {
      const GatherDest& D = g.d[d];
      for (uint32_t c = 0; c < D.k; ++c) {
        const uint32_t* sc = g.stage + (uint64_t(d) * 8 + c) * g.stage_stride + src;
        uint32_t* dc = D.cols + uint64_t(c) * D.stride + D.base + dst0;
        // eight loads in flight per lane before the stores that use them
        // (measured on the triangle join's 424 M rows: 4 loads 2.7 ms, 8 loads
        // 2.6 ms, 16 loads 3.2 ms; all columns' loads at once, four per
        // column, 4.1 ms; non-temporal loads and stores, 3.5 ms)
        constexpr int U = 8;
        for (uint64_t j0 = 0; j0 < cnt; j0 += uint64_t(kWave) * U) {
          uint32_t v[U];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const uint64_t j = j0 + uint64_t(u) * kWave + l;
            v[u] = j < cnt ? sc[j] : 0u;
          }
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const uint64_t j = j0 + uint64_t(u) * kWave + l;
            if (j < cnt) dc[j] = v[u];
          }
        }
      }
    }
  }
}

void gather_segments(const uint64_t* recs, const uint32_t* rec_ctr, uint64_t rec_cap,
                     const uint32_t* stage, uint64_t stage_stride, const uint64_t* offsets,
                     const GatherDest* dests, uint32_t n_dests, uint64_t total, hipStream_t s) {
  if (n_dests > 2) fail("gather_segments: at most two destinations");
  GatherArgs g{};
  g.recs = recs;
  g.rec_ctr = rec_ctr;
  g.rec_cap = rec_cap;
  g.stage = stage;
  g.stage_stride = stage_stride;
  g.offsets = offsets;
  g.nd = n_dests;
  double bytes = 0;
  for (uint32_t d = 0; d < n_dests; ++d) {
    g.d[d] = dests[d];
    bytes += 8.0 * double(total) * dests[d].k;
  }
  TimeScope t("join.gather", s, bytes);
  gather_segments_kernel<<<2048, kBlock, 0, s>>>(g);
  SRD_CHECK(hipGetLastError());
}

// ===================================================== dense bitmaps ====
// Block b covers words [b * kDenseBlockWords, (b + 1) * kDenseBlockWords),
// thread t of it the 32 consecutive words at t * 32, so a thread's bits, and
// the block's, are consecutive keys.
constexpr uint32_t kDenseThreadWords = uint32_t(kDenseBlockWords / kBlock);
static_assert(kDenseThreadWords == 32, "eight 16-byte loads per thread");

__device__ __forceinline__ void dense_load(const uint32_t* bits, uint64_t w0, uint32_t (&v)[32]) {
  const uint4* p = reinterpret_cast<const uint4*>(bits + w0);
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const uint4 q = p[i];
    v[4 * i] = q.x;
    v[4 * i + 1] = q.y;
    v[4 * i + 2] = q.z;
    v[4 * i + 3] = q.w;
  }
}

__device__ __forceinline__ bool dense_marked(const uint32_t* sum, uint32_t b) {
  return (sum[b >> 5] >> (b & 31u)) & 1u;
}

__global__ __launch_bounds__(kBlock) void dense_count_kernel(const uint32_t* __restrict__ bits,
                                                             const uint32_t* __restrict__ sum,
                                                             uint32_t* __restrict__ blk) {
  __shared__ uint32_t wsum[kWavesPerBlock];
  uint32_t v[32] = {};
  if (dense_marked(sum, blockIdx.x))
    dense_load(bits, blockIdx.x * kDenseBlockWords + uint64_t(threadIdx.x) * kDenseThreadWords, v);
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 32; ++i) c += __popc(v[i]);
  uint32_t tot;
  block_excl_scan<uint32_t>(c, wsum, &tot);
  if (threadIdx.x == 0) blk[blockIdx.x] = tot;
}

// Writes the tuples of the block's set bits at its scanned offset (thread
// by thread, in key order) and clears the words it read.
__global__ __launch_bounds__(kBlock) void dense_emit_kernel(uint32_t* __restrict__ bits,
                                                            const uint32_t* __restrict__ sum,
                                                            const uint64_t* __restrict__ blk_off,
                                                            uint32_t k, uint32_t shift,
                                                            uint32_t* __restrict__ out,
                                                            uint64_t stride) {
  __shared__ uint32_t wsum[kWavesPerBlock];
  const uint64_t w0 = blockIdx.x * kDenseBlockWords + uint64_t(threadIdx.x) * kDenseThreadWords;
  uint32_t v[32] = {};
  if (dense_marked(sum, blockIdx.x)) dense_load(bits, w0, v);
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 32; ++i) c += __popc(v[i]);
  uint32_t tot;
  uint64_t pos = blk_off[blockIdx.x] + block_excl_scan<uint32_t>(c, wsum, &tot);
  if (c == 0) return;
  const uint64_t mask = (uint64_t(1) << shift) - 1;
  for (int i = 0; i < 32; ++i) {
    uint32_t x = v[i];
    if (!x) continue;
    bits[w0 + i] = 0;
    while (x) {
      const uint64_t key = ((w0 + uint64_t(i)) << 5) | uint64_t(__builtin_ctz(x));
      x &= x - 1;
      for (uint32_t col = 0; col < k; ++col)
        out[uint64_t(col) * stride + pos] = uint32_t((key >> (shift * (k - 1 - col))) & mask);
      ++pos;
    }
  }
}

uint64_t dense_extract(uint32_t* bits, uint32_t* sum, uint64_t words, uint32_t k, uint32_t shift,
                       DevBuf<uint32_t>& out, uint64_t& stride, Workspace& ws, hipStream_t s) {
  if (words == 0 || words % kDenseBlockWords) fail("dense_extract: bitmap size not whole blocks");
  if (k == 0 || k > 4 || shift == 0 || uint64_t(shift) * k > 63)
    fail("dense_extract: key layout out of range");
  const uint64_t nb = words / kDenseBlockWords;
  if (nb >= (uint64_t(1) << 31)) fail("dense_extract: bitmap too large");
  DevBuf<uint32_t> blk(nb, s);
  DevBuf<uint64_t> off(nb + 1, s);
  uint64_t n = 0;
  {
    TimeScope t("new.dense_count", s, double(words) * 4);
    dense_count_kernel<<<uint32_t(nb), kBlock, 0, s>>>(bits, sum, blk.p);
    SRD_CHECK(hipGetLastError());
    exclusive_scan_u32(blk.p, off.p, nb, ws, s);
    uint32_t* hw = pinned_words(ws, 2);
    SRD_CHECK(hipMemcpyAsync(hw, off.p + nb, 8, hipMemcpyDeviceToHost, s));
    SRD_CHECK(hipStreamSynchronize(s));
    n = uint64_t(hw[0]) | (uint64_t(hw[1]) << 32);
  }
  stride = round_stride(std::max<uint64_t>(n, 1));
  out.alloc(stride * k, s);
  TimeScope t("new.dense_emit", s, double(words) * 4 + double(n) * 4 * k);
  dense_emit_kernel<<<uint32_t(nb), kBlock, 0, s>>>(bits, sum, off.p, k, shift, out.p, stride);
  SRD_CHECK(hipGetLastError());
  fill_u32(sum, 0, dense_sum_words(words), s);
  return n;
}

// ========================================================= calibration ====
// Plain streaming copy: reads and writes exactly n words, so profiler HBM
// counters read on it give the bytes-per-counter-unit scale.
__global__ __launch_bounds__(kBlock) void srd_calib_copy_kernel(const uint4* __restrict__ src,
                                                                uint4* __restrict__ dst,
                                                                uint64_t n16) {
  for (uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x; i < n16;
       i += uint64_t(gridDim.x) * kBlock)
    dst[i] = src[i];
}

// The same copy with 8-byte words, the width of the sort's keys and of the
// merge-path and join loads: its counters check that the scale of the
// 16-byte copy holds for narrower accesses.
__global__ __launch_bounds__(kBlock) void srd_calib_copy8_kernel(const uint64_t* __restrict__ src,
                                                                 uint64_t* __restrict__ dst,
                                                                 uint64_t n8) {
  for (uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x; i < n8;
       i += uint64_t(gridDim.x) * kBlock)
    dst[i] = src[i];
}

void calib_copy(uint64_t bytes, int reps, hipStream_t s, double* ms_per_rep) {
  const uint64_t n16 = (bytes + 15) / 16;
  DevBuf<uint4> a(n16, s), b(n16, s);
  hipEvent_t e0, e1;
  SRD_CHECK(hipEventCreate(&e0));
  SRD_CHECK(hipEventCreate(&e1));
  const uint32_t grid = 256u * 8u;
  srd_calib_copy_kernel<<<grid, kBlock, 0, s>>>(a.p, b.p, n16);  // warm-up (first touch)
  SRD_CHECK(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) srd_calib_copy_kernel<<<grid, kBlock, 0, s>>>(a.p, b.p, n16);
  SRD_CHECK(hipEventRecord(e1, s));
  SRD_CHECK(hipEventSynchronize(e1));
  float ms = 0;
  SRD_CHECK(hipEventElapsedTime(&ms, e0, e1));
  *ms_per_rep = reps ? ms / reps : 0;
  for (int r = 0; r < reps; ++r)
    srd_calib_copy8_kernel<<<grid, kBlock, 0, s>>>(reinterpret_cast<const uint64_t*>(a.p),
                                                   reinterpret_cast<uint64_t*>(b.p), 2 * n16);
  SRD_CHECK(hipGetLastError());
  SRD_CHECK(hipStreamSynchronize(s));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
}

// Synthetic relation columns for primitive benchmarks, generated in place.
// Row i has the 64-bit key  mul*i + off  (range == 0: strictly ascending)
// or  mix(i + off) % range  (range > 0: unordered, with repeats).  The key is
// split into k fields: the low k-1 fields take fbits bits each, column 0 the
// rest, so the lexicographic row order equals the key order.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;  // splitmix64 finaliser
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(kBlock) void pattern_kernel(uint32_t* __restrict__ cols,
                                                         uint64_t stride, uint64_t n, uint32_t k,
                                                         uint32_t fbits, uint64_t mul,
                                                         uint64_t off, uint64_t range) {
  const uint64_t fmask = (uint64_t(1) << fbits) - 1;
  for (uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * kBlock) {
    uint64_t key = range ? mix64(i + off) % range : mul * i + off;
    for (int c = int(k) - 1; c > 0; --c) {
      cols[uint64_t(c) * stride + i] = uint32_t(key & fmask);
      key >>= fbits;
    }
    cols[i] = uint32_t(key);
  }
}

void fill_pattern(uint32_t* cols, uint64_t stride, uint64_t n, uint32_t k, uint32_t fbits,
                  uint64_t mul, uint64_t off, uint64_t range, hipStream_t s) {
  if (n == 0) return;
  if (k == 0 || k > 8 || fbits == 0 || fbits > 32)
    fail("fill_pattern: bad arity or field width");
  pattern_kernel<<<256u * 8u, kBlock, 0, s>>>(cols, stride, n, k, fbits, mul, off, range);
  SRD_CHECK(hipGetLastError());
}

// Rows of a column-0-sorted index falling in each instance's id range
// [bounds[g], bounds[g+1]): two searches per instance.
__global__ __launch_bounds__(kBlock) void instance_count_kernel(const uint32_t* __restrict__ col0,
                                                                uint32_t n,
                                                                const uint32_t* __restrict__ bounds,
                                                                uint32_t n_inst,
                                                                uint64_t* __restrict__ out) {
  const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
  if (g >= n_inst) return;
  const uint32_t lo = n ? lower_bound(col0, 0, n, bounds[g]) : 0;
  const uint32_t hi = n ? lower_bound(col0, lo, n, bounds[g + 1]) : 0;
  out[g] = hi - lo;
}

__global__ __launch_bounds__(kBlock) void slice_bounds_kernel(const uint64_t* __restrict__ pref,
                                                              uint64_t n, uint64_t q, uint64_t rem,
                                                              uint64_t S, uint64_t* __restrict__ out) {
  const uint64_t s = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (s > S) return;
  const uint64_t target = s * q + (s < rem ? s : rem);  // s*w/S split without overflow
  uint64_t lo = 0, len = n + 1;                          // first i in [0, n] with pref[i] >= target
  while (len > 0) {
    const uint64_t half = len >> 1;
    if (pref[lo + half] < target) {
      lo += half + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  out[s] = lo;
}

void slice_bounds(const uint64_t* pref, uint64_t n, uint64_t w, uint64_t S, uint64_t* out,
                  hipStream_t s) {
  if (S == 0) fail("slice_bounds: no slices");
  slice_bounds_kernel<<<uint32_t((S + 1 + kBlock - 1) / kBlock), kBlock, 0, s>>>(pref, n, w / S,
                                                                               w % S, S, out);
  SRD_CHECK(hipGetLastError());
}

// ------------------------------------------------- heavy-key bitmaps ----
__global__ __launch_bounds__(kBlock) void heavy_flag_kernel(const uint32_t* __restrict__ offs,
                                                            uint32_t r, uint32_t min_len,
                                                            uint32_t* __restrict__ flag) {
  const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
  if (g < r) flag[g] = offs[g + 1] - offs[g] >= min_len ? 1u : 0u;
}

__global__ __launch_bounds__(kBlock) void heavy_scatter_kernel(
    const uint32_t* __restrict__ offs, uint32_t r, const uint32_t* __restrict__ flag,
    const uint64_t* __restrict__ pos, uint32_t* __restrict__ starts, uint32_t* __restrict__ ends) {
  const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
  if (g < r && flag[g]) {
    starts[pos[g]] = offs[g];
    ends[pos[g]] = offs[g + 1];
  }
}

// Blocks kHeavyParts*i .. kHeavyParts*i+kHeavyParts-1 set the bits of heavy
// key i's column-1 values, each over one part of its rows (the longest keys
// hold tens of thousands of rows).
constexpr uint32_t kHeavyParts = 8;
__global__ __launch_bounds__(kBlock) void heavy_set_kernel(const uint32_t* __restrict__ col1,
                                                           const uint32_t* __restrict__ starts,
                                                           const uint32_t* __restrict__ ends,
                                                           uint32_t words,
                                                           uint32_t* __restrict__ bits) {
  const uint32_t i = blockIdx.x / kHeavyParts, part = blockIdx.x % kHeavyParts;
  uint32_t* b = bits + uint64_t(i) * words;
  const uint32_t s = starts[i], len = ends[i] - s;
  const uint32_t lo = s + uint32_t(uint64_t(len) * part / kHeavyParts);
  const uint32_t hi = s + uint32_t(uint64_t(len) * (part + 1) / kHeavyParts);
  for (uint32_t row = lo + threadIdx.x; row < hi; row += kBlock) {
    const uint32_t v = col1[row];
    atomicOr(&b[v >> 5], 1u << (v & 31u));
  }
}

// Heavy key i's bitmap built by block i in LDS (zeroed there, one LDS atomic
// OR per row, written out whole once): no fill of the bitmaps beforehand and
// no global atomics.  Taken when a bitmap fits kHeavyLdsWords words (column-1
// values below 1,179,648); wider domains take heavy_set_kernel.
constexpr uint32_t kHeavyLdsWords = 36864;  // 144 KiB of the CU's 160 KiB
constexpr int kHeavyLdsBlock = 1024;
__global__ __launch_bounds__(kHeavyLdsBlock) void heavy_set_lds_kernel(
    const uint32_t* __restrict__ col1, const uint32_t* __restrict__ starts,
    const uint32_t* __restrict__ ends, uint32_t words, uint32_t* __restrict__ bits) {
  __shared__ uint32_t bm[kHeavyLdsWords];
  for (uint32_t w = threadIdx.x; w < words; w += kHeavyLdsBlock) bm[w] = 0;
  __syncthreads();
  const uint32_t e = ends[blockIdx.x];
  for (uint32_t row = starts[blockIdx.x] + threadIdx.x; row < e; row += kHeavyLdsBlock) {
    const uint32_t v = col1[row];
    atomicOr(&bm[v >> 5], 1u << (v & 31u));
  }
  __syncthreads();
  uint32_t* out = bits + uint64_t(blockIdx.x) * words;
  for (uint32_t w = threadIdx.x; w < words; w += kHeavyLdsBlock) out[w] = bm[w];
}

void build_heavy_bitmaps(const Index& x, uint32_t min_len, uint64_t max_bytes, Workspace& ws,
                         hipStream_t s) {
  if (!x.heavy) x.heavy = std::make_shared<HeavyBits>();
  HeavyBits& h = *x.heavy;
  h.n = 0;
  h.built_for = x.cols.p;
  h.built_n = x.n;
  if (x.arity != 2 || x.r == 0 || x.col_max.size() < 2 || max_bytes == 0) return;
  const uint32_t words = (x.col_max[1] >> 5) + 1;
  DevBuf<uint32_t> flag(x.r, s), ends;
  DevBuf<uint64_t> pos(uint64_t(x.r) + 1, s);
  for (;;) {
    heavy_flag_kernel<<<(x.r + kBlock - 1) / kBlock, kBlock, 0, s>>>(x.root_offs.p, x.r, min_len,
                                                                     flag.p);
    SRD_CHECK(hipGetLastError());
    exclusive_scan_u32(flag.p, pos.p, x.r, ws, s);
    uint32_t* hw = pinned_words(ws, 2);
    SRD_CHECK(hipMemcpyAsync(hw, pos.p + x.r, 8, hipMemcpyDeviceToHost, s));
    SRD_CHECK(hipStreamSynchronize(s));
    const uint32_t n = hw[0];  // at most r
    if (n == 0) return;
    if (uint64_t(n) * words * 4 <= max_bytes) {
      h.starts.alloc(n, s);
      ends.alloc(n, s);
      heavy_scatter_kernel<<<(x.r + kBlock - 1) / kBlock, kBlock, 0, s>>>(
          x.root_offs.p, x.r, flag.p, pos.p, h.starts.p, ends.p);
      SRD_CHECK(hipGetLastError());
      h.bits.alloc(uint64_t(n) * words, s);
      if (words <= kHeavyLdsWords) {
        heavy_set_lds_kernel<<<n, kHeavyLdsBlock, 0, s>>>(x.col(1), h.starts.p, ends.p, words,
                                                          h.bits.p);
      } else {
        fill_u32(h.bits.p, 0, uint64_t(n) * words, s);
        heavy_set_kernel<<<n * kHeavyParts, kBlock, 0, s>>>(x.col(1), h.starts.p, ends.p, words,
                                                             h.bits.p);
      }
      SRD_CHECK(hipGetLastError());
      h.n = n;
      h.words = words;
      h.min_len = min_len;
      return;
    }
    if (min_len >= (1u << 30)) return;
    min_len *= 2;  // too many heavy keys for the budget: keep the longer ones
  }
}

// pos[vals[g]] = g; values beyond the map (a column-0 bound that does not
// hold) are skipped rather than written out of bounds.
__global__ __launch_bounds__(kBlock) void root_pos_kernel(const uint32_t* __restrict__ vals,
                                                          uint32_t r, uint64_t entries,
                                                          uint32_t* __restrict__ pos) {
  const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
  if (g >= r) return;
  const uint32_t v = vals[g];
  if (v < entries) pos[v] = g;
}

void build_root_pos(const Index& x, uint64_t max_entries, hipStream_t s) {
  if (!x.rootpos) x.rootpos = std::make_shared<RootPos>();
  RootPos& m = *x.rootpos;
  m.ok = false;
  m.built_for = x.root_vals.p;
  m.built_r = x.r;
  if (x.r == 0 || x.col_max.empty()) return;
  const uint64_t entries = uint64_t(x.col_max[0]) + 1;  // bound on every root value
  if (entries > max_entries) return;
  m.pos.reserve(entries, s);
  fill_u32(m.pos.p, ~0u, entries, s);
  root_pos_kernel<<<(x.r + kBlock - 1) / kBlock, kBlock, 0, s>>>(x.root_vals.p, x.r, entries,
                                                                 m.pos.p);
  SRD_CHECK(hipGetLastError());
  m.max = uint32_t(entries - 1);
  m.ok = true;
}

// out[0] = rows below bounds[0], out[1] = rows at or above bounds[n_inst].
__global__ void outside_count_kernel(const uint32_t* __restrict__ col0, uint32_t n,
                                     const uint32_t* __restrict__ bounds, uint32_t n_inst,
                                     uint64_t* __restrict__ out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  out[0] = n ? lower_bound(col0, 0, n, bounds[0]) : 0;
  out[1] = n ? n - lower_bound(col0, 0, n, bounds[n_inst]) : 0;
}

void outside_counts(const uint32_t* col0, uint64_t n, const uint32_t* bounds, uint32_t n_inst,
                    uint64_t* out, hipStream_t s) {
  if (n > 0xffffffffull) fail("outside_counts: more than 2^32 rows");
  outside_count_kernel<<<1, 64, 0, s>>>(col0, uint32_t(n), bounds, n_inst, out);
  SRD_CHECK(hipGetLastError());
}

void instance_counts(const uint32_t* col0, uint64_t n, const uint32_t* bounds, uint32_t n_inst,
                     uint64_t* out, hipStream_t s) {
  if (n_inst == 0) return;
  if (n > 0xffffffffull) fail("instance_counts: more than 2^32 rows");
  instance_count_kernel<<<(n_inst + kBlock - 1) / kBlock, kBlock, 0, s>>>(col0, uint32_t(n),
                                                                           bounds, n_inst, out);
  SRD_CHECK(hipGetLastError());
}

// Row-major copy of an index for the host: out[i*k + c] = column c of row i,
// minus the base of the row's instance (the largest bounds[g] <= column 0)
// when an instance table is given.
__global__ __launch_bounds__(kBlock) void interleave_kernel(const uint32_t* __restrict__ cols,
                                                            uint64_t stride, uint64_t n,
                                                            uint32_t k,
                                                            const uint32_t* __restrict__ bounds,
                                                            uint32_t n_bounds,
                                                            uint32_t* __restrict__ out) {
  for (uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * kBlock) {
    uint32_t base = 0;
    if (n_bounds) {
      const uint32_t g = upper_bound(bounds, 0, n_bounds, cols[i]);
      base = g ? bounds[g - 1] : 0u;
    }
    for (uint32_t c = 0; c < k; ++c) out[i * k + c] = cols[uint64_t(c) * stride + i] - base;
  }
}

// Device-to-host copy into caller memory that is usually pageable and not yet
// touched: the bytes go through two pinned chunks, the DMA of chunk c+1
// running while host threads copy chunk c out (the first touch of the
// destination pages is then spread over the threads too).  One pageable
// hipMemcpy of a fresh 1 GB array runs at about 6 GB/s here.
static void copy_out(uint8_t* dst, const uint8_t* src, uint64_t bytes, Workspace& ws,
                     hipStream_t s) {
  constexpr uint64_t kChunk = uint64_t(64) << 20;
  if (bytes <= (uint64_t(4) << 20)) {
    SRD_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    SRD_CHECK(hipStreamSynchronize(s));
    return;
  }
  if (!ws.stage_host)
    ws.stage_host = static_cast<uint8_t*>(pinned_acquire(2 * kChunk, &ws.stage_host_bytes));
  const unsigned hw = std::thread::hardware_concurrency();
  const unsigned nt = std::max(1u, std::min(8u, hw ? hw / 2 : 4u));
  hipEvent_t ev[2];
  SRD_CHECK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  SRD_CHECK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  const uint64_t chunks = (bytes + kChunk - 1) / kChunk;
  auto issue = [&](uint64_t c) {
    const uint64_t off = c * kChunk, len = std::min(kChunk, bytes - off);
    SRD_CHECK(hipMemcpyAsync(ws.stage_host + (c & 1) * kChunk, src + off, len,
                             hipMemcpyDeviceToHost, s));
    SRD_CHECK(hipEventRecord(ev[c & 1], s));
  };
  issue(0);
  for (uint64_t c = 0; c < chunks; ++c) {
    SRD_CHECK(hipEventSynchronize(ev[c & 1]));
    if (c + 1 < chunks) issue(c + 1);  // the other chunk was copied out last round
    const uint64_t off = c * kChunk, len = std::min(kChunk, bytes - off);
    const uint8_t* from = ws.stage_host + (c & 1) * kChunk;
    const uint64_t part = (len + nt - 1) / nt;
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; ++t) {
      const uint64_t a = std::min<uint64_t>(len, t * part), b = std::min<uint64_t>(len, a + part);
      if (a < b) th.emplace_back([=] { std::memcpy(dst + off + a, from + a, b - a); });
    }
    std::memcpy(dst + off, from, std::min<uint64_t>(len, part));
    for (auto& x : th) x.join();
  }
  SRD_CHECK(hipEventDestroy(ev[0]));
  SRD_CHECK(hipEventDestroy(ev[1]));
}

void export_rows(const Index& x, uint32_t* host_out, const uint32_t* dev_bounds,
                 uint32_t n_bounds, Workspace& ws, hipStream_t s) {
  if (x.n == 0) return;
  DevBuf<uint32_t> buf(x.n * x.arity, s);
  interleave_kernel<<<grid_for(x.n, kBlock * 4, 8192), kBlock, 0, s>>>(
      x.cols.p, x.stride, x.n, x.arity, dev_bounds, n_bounds, buf.p);
  SRD_CHECK(hipGetLastError());
  copy_out(reinterpret_cast<uint8_t*>(host_out), reinterpret_cast<const uint8_t*>(buf.p),
           x.n * x.arity * 4, ws, s);
}

// ================================================================ scan ====
constexpr int kItems = 16;
constexpr uint32_t kTile = kBlock * kItems;  // 4096

// Position of (item i, lane l) of wave w inside a tile.
__device__ __forceinline__ uint32_t tile_pos(uint32_t w, uint32_t i, uint32_t l) {
  return w * (kWave * kItems) + i * kWave + l;
}

template <class T>
__global__ __launch_bounds__(kBlock) void scan_reduce_kernel(const T* __restrict__ in, uint64_t n,
                                                             uint64_t* __restrict__ partial,
                                                             const uint64_t* __restrict__ n_dev = nullptr) {
  __shared__ uint64_t wsum[kWavesPerBlock];
  if (n_dev) n = min(n, *n_dev);  // device-sized: entries past the count are 0
  const uint64_t base = uint64_t(blockIdx.x) * kTile;
  uint64_t s = 0;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    uint64_t idx = base + uint64_t(i) * kBlock + threadIdx.x;
    if (idx < n) s += uint64_t(in[idx]);
  }
  s = wave_sum(s);
  if (lane_id() == 0) wsum[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (int w = 0; w < kWavesPerBlock; ++w) t += wsum[w];
    partial[blockIdx.x] = t;
  }
}

// Exclusive scan of one tile per block; carry[blockIdx] (if non-null) is
// added.  The block owning the last element writes out[n] = total.
template <class T>
__global__ __launch_bounds__(kBlock) void scan_tile_kernel(const T* __restrict__ in, uint64_t n,
                                                           const uint64_t* __restrict__ carry,
                                                           uint64_t* __restrict__ out,
                                                           const uint64_t* __restrict__ n_dev = nullptr) {
  __shared__ uint64_t wtot[kWavesPerBlock];
  if (n_dev) n = min(n, *n_dev);
  const uint32_t w = threadIdx.x / kWave, l = lane_id();
  const uint64_t base = uint64_t(blockIdx.x) * kTile;
  uint64_t v[kItems];
  uint64_t run = 0;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    uint64_t idx = base + tile_pos(w, i, l);
    v[i] = idx < n ? uint64_t(in[idx]) : 0;
    run += v[i];
  }
  run = wave_sum(run);
  if (l == 0) wtot[w] = run;
  __syncthreads();
  uint64_t off = carry ? carry[blockIdx.x] : 0;
  uint64_t total = off;
  for (uint32_t j = 0; j < kWavesPerBlock; ++j) {
    if (j < w) off += wtot[j];
    total += wtot[j];
  }
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    uint64_t inc = wave_incl_scan(v[i]);
    uint64_t idx = base + tile_pos(w, i, l);
    if (idx < n) out[idx] = off + inc - v[i];
    off += read_lane_u64(inc, kWave - 1);
  }
  uint64_t last_block = (n == 0) ? 0 : (n - 1) / kTile;
  if (blockIdx.x == last_block && threadIdx.x == 0) out[n] = total;
}

template <class T>
static void scan_impl(const T* in, uint64_t* out, uint64_t n, hipStream_t s) {
  if (n <= kTile) {
    scan_tile_kernel<T><<<1, kBlock, 0, s>>>(in, n, nullptr, out);
    SRD_CHECK(hipGetLastError());
    return;
  }
  uint64_t tiles = (n + kTile - 1) / kTile;
  DevBuf<uint64_t> partial(tiles, s), carry(tiles + 1, s);
  scan_reduce_kernel<T><<<uint32_t(tiles), kBlock, 0, s>>>(in, n, partial.p);
  SRD_CHECK(hipGetLastError());
  scan_impl<uint64_t>(partial.p, carry.p, tiles, s);
  scan_tile_kernel<T><<<uint32_t(tiles), kBlock, 0, s>>>(in, n, carry.p, out);
  SRD_CHECK(hipGetLastError());
}

void exclusive_scan_u32(const uint32_t* in, uint64_t* out, uint64_t n, Workspace&, hipStream_t s) {
  TimeScope t("scan", s, double(n) * 12);
  scan_impl<uint32_t>(in, out, n, s);
}

void exclusive_scan_u64(const uint64_t* in, uint64_t* out, uint64_t n, Workspace&, hipStream_t s) {
  TimeScope t("scan", s, double(n) * 16);
  scan_impl<uint64_t>(in, out, n, s);
}

// ======================================================== sort: packing ====
constexpr int kMaxCols = 8;
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kMaxPasses = 16;

struct PackSpec {
  const uint32_t* in[kMaxCols];
  uint32_t shift[kMaxCols];  // bit offset of column c in the 128-bit key
  uint32_t k;
  uint32_t passes;
};

// State of a device-planned sort (build_index_dev): the row count, key
// layout and pass schedule are computed on the device, so the sort needs
// no host read-back and its launches can be enqueued or captured blind.
struct DevPlan {
  uint64_t n;                  // rows to sort
  uint64_t u, r;               // outputs: distinct rows, distinct column-0 values
  uint32_t shift[kMaxCols];    // bit offset of column c in the packed key
  uint32_t bits[kMaxCols];     // significant bits of column c
  uint32_t k, passes, flags, err;
  uint32_t pass[kMaxPasses];   // bit 0: pass runs, bit 1: it reads the second buffer
  uint32_t final_buf;          // buffer holding the sorted keys
  uint32_t epoch0;             // look-back epoch of pass p: epoch0 + p
  uint32_t wrap;               // the epochs wrapped: the look-back words are cleared first
  uint32_t pad_;
};

// Block-level reduction of per-thread column maxima and order flags, then one
// atomic per column per block (every wave hitting the same words with
// atomics would serialise them in L2).
__device__ __forceinline__ void colmax_flush(const uint32_t* m, uint32_t k, bool unsorted,
                                             bool dup, uint32_t* __restrict__ out_max) {
  __shared__ uint32_t s_max[kWavesPerBlock][kMaxCols + 1];
  const uint32_t w = threadIdx.x / kWave, lane = lane_id();
  for (uint32_t c = 0; c < k; ++c) {
    const uint32_t v = wave_max(m[c]);
    if (lane == 0) s_max[w][c] = v;
  }
  const uint32_t fl = (ballot(unsorted) ? 1u : 0u) | (ballot(dup) ? 2u : 0u);
  if (lane == 0) s_max[w][kMaxCols] = fl;
  __syncthreads();
  if (threadIdx.x <= k) {
    const uint32_t c = threadIdx.x == k ? uint32_t(kMaxCols) : threadIdx.x;
    uint32_t v = 0;
    for (int j = 0; j < kWavesPerBlock; ++j) v = c == kMaxCols ? (v | s_max[j][c]) : max(v, s_max[j][c]);
    if (v) {
      if (c == kMaxCols) atomicOr(&out_max[c], v);
      else atomicMax(&out_max[c], v);
    }
  }
}

// Per-column maxima plus order flags in out_max[kMaxCols]: bit 0 = some row
// is lexicographically smaller than its predecessor, bit 1 = some row equals
// its predecessor.  0 means strictly increasing: already a valid index.
__global__ __launch_bounds__(kBlock) void col_max_kernel(PackSpec spec, uint64_t n,
                                                         uint32_t* __restrict__ out_max) {
  uint32_t m[kMaxCols];
  for (uint32_t c = 0; c < spec.k; ++c) m[c] = 0;
  bool unsorted = false, dup = false;
  for (uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * kBlock) {
    int order = 0;  // compare row i with row i-1
    for (uint32_t c = 0; c < spec.k; ++c) {
      uint32_t v = spec.in[c][i];
      m[c] = max(m[c], v);
      if (i > 0 && order == 0) {
        uint32_t p = spec.in[c][i - 1];
        order = v < p ? -1 : (v > p ? 1 : 0);
      }
    }
    unsorted |= order < 0;
    dup |= (i > 0 && order == 0);
  }
  colmax_flush(m, spec.k, unsorted, dup, out_max);
}

// col_max_kernel for 16-byte aligned columns: each lane reads rows
// 4g..4g+3 of every column with one 16-byte load; the row before 4g is the
// previous lane's last row (a shuffle), or, for the first lane of a wave, a
// direct load.  Rows past the last full group of 4 are done one per thread.
__global__ __launch_bounds__(kBlock) void col_max_vec_kernel(PackSpec spec, uint64_t n,
                                                             uint32_t* __restrict__ out_max) {
  uint32_t m[kMaxCols];
  for (uint32_t c = 0; c < spec.k; ++c) m[c] = 0;
  bool unsorted = false, dup = false;
  const uint32_t lane = lane_id();
  const uint64_t ngroups = n >> 2;
  const uint64_t stride = uint64_t(gridDim.x) * kBlock;
  for (uint64_t g = blockIdx.x * uint64_t(kBlock) + threadIdx.x; g < ngroups; g += stride) {
    int ord[4] = {0, 0, 0, 0};
    for (uint32_t c = 0; c < spec.k; ++c) {
      const uint4 v = reinterpret_cast<const uint4*>(spec.in[c])[g];
      m[c] = max(m[c], max(max(v.x, v.y), max(v.z, v.w)));
      uint32_t p0 = __shfl_up(v.w, 1, kWave);
      if (lane == 0) p0 = g > 0 ? spec.in[c][4 * g - 1] : 0u;
      const uint32_t cur[4] = {v.x, v.y, v.z, v.w};
      const uint32_t prv[4] = {p0, v.x, v.y, v.z};
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (ord[j] == 0) ord[j] = cur[j] < prv[j] ? -1 : (cur[j] > prv[j] ? 1 : 0);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const bool has_prev = g > 0 || j > 0;
      unsorted |= has_prev && ord[j] < 0;
      dup |= has_prev && ord[j] == 0;
    }
  }
  const uint64_t t = blockIdx.x * uint64_t(kBlock) + threadIdx.x;
  const uint64_t i = 4 * ngroups + t;
  if (i < n) {
    int order = 0;
    for (uint32_t c = 0; c < spec.k; ++c) {
      const uint32_t v = spec.in[c][i];
      m[c] = max(m[c], v);
      if (i > 0 && order == 0) {
        const uint32_t p = spec.in[c][i - 1];
        order = v < p ? -1 : (v > p ? 1 : 0);
      }
    }
    unsorted |= order < 0;
    dup |= i > 0 && order == 0;
  }
  colmax_flush(m, spec.k, unsorted, dup, out_max);
}

template <int W>
__device__ __forceinline__ uint32_t digit_of(uint64_t lo, uint64_t hi, uint32_t shift) {
  if (W == 1) return uint32_t(lo >> shift) & (kRadix - 1);
  if (shift >= 64) return uint32_t(hi >> (shift - 64)) & (kRadix - 1);
  uint64_t v = lo >> shift;
  if (shift > 56) v |= hi << (64 - shift);
  return uint32_t(v) & (kRadix - 1);
}

// Pack columns into W-word keys and accumulate every pass's digit histogram.
template <int W>
__global__ __launch_bounds__(kBlock) void pack_hist_kernel(PackSpec spec, uint64_t n,
                                                           uint64_t* __restrict__ lo_out,
                                                           uint64_t* __restrict__ hi_out,
                                                           uint32_t* __restrict__ hist,
                                                           uint32_t* __restrict__ tile_ctr) {
  __shared__ uint32_t h[kMaxPasses * kRadix];
  for (uint32_t i = threadIdx.x; i < spec.passes * kRadix; i += kBlock) h[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x < kMaxPasses) tile_ctr[threadIdx.x] = 0;
  __syncthreads();
  // wave-uniform steps (hist_add_runs needs the whole wave)
  const uint64_t stride = uint64_t(gridDim.x) * kBlock;
  for (uint64_t b = blockIdx.x * uint64_t(kBlock) + (threadIdx.x & ~uint32_t(kWave - 1)); b < n;
       b += stride) {
    const uint64_t i = b + lane_id();
    const bool valid = i < n;
    uint64_t lo = 0, hi = 0;
    if (valid) {
      for (uint32_t c = 0; c < spec.k; ++c) {
        uint64_t v = spec.in[c][i];
        uint32_t s = spec.shift[c];
        if (s >= 64) {
          hi |= v << (s - 64);
        } else {
          lo |= v << s;
          if (W == 2 && s > 32) hi |= v >> (64 - s);
        }
      }
      lo_out[i] = lo;
      if (W == 2) hi_out[i] = hi;
    }
    for (uint32_t p = 0; p < spec.passes; ++p)
      hist_add_runs(&h[p * kRadix], digit_of<W>(lo, hi, p * kRadixBits), valid);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < spec.passes * kRadix; i += kBlock)
    if (h[i]) atomicAdd(&hist[i], h[i]);
}

// ====================================================== sort: onesweep ====
// Look-back word: [63:48] epoch, [47:46] flag (1 = tile aggregate,
// 2 = inclusive prefix), [45:0] count.
constexpr uint64_t kLbAgg = 1, kLbPrefix = 2;
constexpr uint64_t kLbMask = (uint64_t(1) << 46) - 1;
// Polls before a look-back gives up and raises the error flag.  Tiles wait
// only on tiles of smaller order -- smaller block ids, which every XCD
// dispatches first, when the run is alone on the GPU, else smaller tickets,
// which are held only by running blocks -- so a wait is short; the limit
// bounds a broken invariant to seconds.
constexpr uint32_t kSpinLimit = 1u << 22;
constexpr int kLbWin = 4;  // onesweep look-back words loaded per round trip (4: 449 us per
                           // 64 M-key pass, 8: 455, 16: 468)

__device__ __forceinline__ uint64_t lb_pack(uint32_t epoch, uint64_t flag, uint64_t v) {
  return (uint64_t(epoch) << 48) | (flag << 46) | v;
}

__device__ __forceinline__ void lb_store(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t lb_load(uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Lanes (among `valid`) whose digit equals this lane's digit.
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid) {
  uint64_t m = ballot(valid);
#pragma unroll
  for (int b = 0; b < kRadixBits; ++b) {
    uint64_t bb = ballot((d >> b) & 1u);
    m &= ((d >> b) & 1u) ? bb : ~bb;
  }
  return m;
}

// Decoupled look-back by one whole wave: the 64 lanes read the status words
// of the 64 tiles before `tile` at once (lane 0 the nearest), wait until every
// word up to the nearest inclusive prefix is published, and add the
// aggregates up to and including that prefix; with no prefix in the window
// they add all 64 aggregates and move 64 tiles further back.  Status of tile
// t is at base[t * stride].  Returns the exclusive prefix of `tile`; raises
// *err after kSpinLimit unsuccessful polls.  All 64 lanes must call it.
__device__ __forceinline__ uint64_t wave_lookback(uint64_t* base, uint64_t stride, uint32_t tile,
                                                  uint32_t epoch, uint32_t* err) {
  const uint32_t l = lane_id();
  uint64_t excl = 0;
  int64_t hi = int64_t(tile) - 1;
  uint32_t spins = 0;
  while (hi >= 0) {
    const int64_t idx = hi - int64_t(l);
    const uint64_t v = idx >= 0 ? lb_load(base + uint64_t(idx) * stride)
                                : lb_pack(epoch, kLbPrefix, 0);
    const uint64_t flag = (v >> 46) & 3;
    const bool ready = uint32_t(v >> 48) == epoch && flag != 0;
    const uint64_t pm = ballot(ready && flag == kLbPrefix);
    const uint64_t need = pm ? (((pm & (~pm + 1)) << 1) - 1) : ~0ull;
    const uint64_t rm = ballot(ready);
    if ((rm & need) != need) {
      if (++spins > kSpinLimit) {
        if (l == 0) atomicOr(err, 1u);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    const uint64_t mine = ((need >> l) & 1ull) ? (v & kLbMask) : 0;
    excl += wave_sum<uint64_t>(mine);
    if (pm) break;
    hi -= 64;
  }
  return excl;
}

// Keys per thread of a onesweep tile (W = 1: 64-bit keys, W = 2: 128-bit).
#ifndef SRD_OS_ITEMS1
#define SRD_OS_ITEMS1 16
#endif
constexpr int os_items(int W) { return W == 1 ? SRD_OS_ITEMS1 : 8; }
// Threads of a onesweep block: kBlock or 2 * kBlock.  The first kRadix
// threads own one digit each for the tile's histogram and look-back.
constexpr uint64_t os_tile(int W, int B) { return uint64_t(B) * os_items(W); }
// Fewest keys of any onesweep tile: sizes look-back tables for every launch.
constexpr uint64_t kOsMinTile = os_tile(2, kBlock);

template <int W, int B>
__global__ __launch_bounds__(B) void onesweep_kernel(
    const uint64_t* __restrict__ in_lo, const uint64_t* __restrict__ in_hi,
    uint64_t* __restrict__ out_lo, uint64_t* __restrict__ out_hi, uint32_t n, uint32_t shift,
    const uint32_t* __restrict__ bin_start, uint64_t* __restrict__ lookback,
    uint32_t* __restrict__ tile_ctr, uint32_t epoch, uint32_t* __restrict__ err, bool ticket,
    const DevPlan* __restrict__ dp, uint32_t pass_no) {
  constexpr int ITEMS = os_items(W);
  constexpr uint32_t TILE = B * ITEMS;
  constexpr int WAVES = B / kWave;
  static_assert(B % kRadix == 0, "every digit needs an owner thread");
  if (dp) {
    // device-planned sort: a skipped pass leaves the keys where they are;
    // an active one reads the buffer the schedule names
    const uint32_t ps = dp->pass[pass_no];
    if (!(ps & 1u)) return;
    n = uint32_t(dp->n);
    epoch = dp->epoch0 + pass_no;
    if (ps & 2u) {
      const uint64_t* t0 = in_lo;
      const uint64_t* t1 = in_hi;
      in_lo = out_lo;
      in_hi = out_hi;
      out_lo = const_cast<uint64_t*>(t0);
      out_hi = const_cast<uint64_t*>(t1);
    }
  }
  __shared__ uint64_t s_keys[W][TILE];
  __shared__ uint32_t s_wcnt[WAVES][kRadix];
  __shared__ uint32_t s_local[kRadix];
  __shared__ uint64_t s_global[kRadix];
  __shared__ uint32_t s_wsum[WAVES];

  __shared__ uint32_t s_tile;
  const uint32_t tid = threadIdx.x, w = tid / kWave, l = lane_id();
  // Tile order.  Alone on the GPU, tile = block id: each XCD dispatches its
  // blocks in id order, so every predecessor of a running tile is running,
  // done, or next in its XCD's queue behind blocks that finish.  When other
  // runs' kernels may hold CUs meanwhile, a block takes a ticket once it
  // runs instead (tile_ctr, zeroed by pack_hist_kernel), so a tile only ever
  // waits on tiles that are already running; the ticket's same-address
  // atomic costs ~10 ns per tile, which is why it is not always taken.
  for (uint32_t i = tid; i < WAVES * kRadix; i += B) (&s_wcnt[0][0])[i] = 0;
  if (ticket && tid == 0) s_tile = atomicAdd(tile_ctr, 1u);
  __syncthreads();
  const uint32_t tile = ticket ? s_tile : blockIdx.x;
  const uint64_t base = uint64_t(tile) * TILE;
  if (base >= n) return;  // launched for a capacity above the rows (device-planned sorts)
  const uint32_t count = uint32_t(min<uint64_t>(TILE, n - base));

  uint64_t klo[ITEMS], khi[ITEMS];
  uint32_t rank[ITEMS], dig[ITEMS];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    uint32_t p = w * (kWave * ITEMS) + i * kWave + l;
    bool valid = p < count;
    klo[i] = valid ? in_lo[base + p] : 0;
    khi[i] = (W == 2 && valid) ? in_hi[base + p] : 0;
    uint32_t d = valid ? digit_of<W>(klo[i], khi[i], shift) : 0;
    dig[i] = d;
    uint64_t peers = match_digit(d, valid);
    uint32_t before = valid ? s_wcnt[w][d] : 0;
    rank[i] = before + popc_below(peers);
    if (valid && popc_below(peers) == 0) s_wcnt[w][d] = before + __popcll(peers);
  }
  __syncthreads();

  // Thread d < kRadix owns digit d: per-wave exclusive offsets and the tile
  // total.
  const bool owner = B == kRadix || tid < kRadix;
  const uint32_t d = owner ? tid : 0;
  uint32_t tot = 0;
  if (owner) {
#pragma unroll
    for (int j = 0; j < WAVES; ++j) {
      uint32_t c = s_wcnt[j][d];
      s_wcnt[j][d] = tot;
      tot += c;
    }
  }
  uint64_t* my_lb = lookback + uint64_t(tile) * kRadix + d;
  uint64_t excl = 0;
  if (!owner) {
  } else if (tile == 0) {
    lb_store(my_lb, lb_pack(epoch, kLbPrefix, tot));
  } else {
    lb_store(my_lb, lb_pack(epoch, kLbAgg, tot));
    // Walk back kLbWin tiles per round trip: the words of the kLbWin nearest
    // unresolved predecessors are loaded together, then summed nearest first
    // up to the first inclusive prefix; a window with a word not yet
    // published is reloaded from that word on.
    int64_t t = int64_t(tile) - 1;
    uint32_t spins = 0;
    while (t >= 0) {
      uint64_t v[kLbWin];
#pragma unroll
      for (int k = 0; k < kLbWin; ++k)
        v[k] = t - k >= 0 ? lb_load(lookback + uint64_t(t - k) * kRadix + d)
                          : lb_pack(epoch, kLbPrefix, 0);
      bool done = false;
      int used = 0;
#pragma unroll
      for (int k = 0; k < kLbWin; ++k) {
        if (done || used != k) continue;
        const uint64_t flag = (v[k] >> 46) & 3;
        if (uint32_t(v[k] >> 48) != epoch || flag == 0) continue;  // not yet published
        excl += v[k] & kLbMask;
        used = k + 1;
        if (flag == kLbPrefix) done = true;
      }
      if (done) break;
      t -= used;
      if (used == 0) {
        if (++spins > kSpinLimit) {
          atomicOr(err, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    lb_store(my_lb, lb_pack(epoch, kLbPrefix, excl + tot));
  }
  // Local start of digit d inside the LDS-sorted tile.
  uint32_t wtotal;
  uint32_t lstart = block_excl_scan<uint32_t, WAVES>(tot, s_wsum, &wtotal);
  if (owner) {
    s_local[d] = lstart;
    s_global[d] = uint64_t(bin_start[d]) + excl - lstart;
  }
  __syncthreads();

#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    uint32_t p = w * (kWave * ITEMS) + i * kWave + l;
    if (p < count) {
      uint32_t lp = s_local[dig[i]] + s_wcnt[w][dig[i]] + rank[i];
      s_keys[0][lp] = klo[i];
      if (W == 2) s_keys[W - 1][lp] = khi[i];
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    uint32_t p = i * B + tid;
    if (p < count) {
      uint64_t lo = s_keys[0][p];
      uint64_t hi = (W == 2) ? s_keys[W - 1][p] : 0;
      uint64_t g = s_global[digit_of<W>(lo, hi, shift)] + p;
      if (g < n) {  // always true when the look-back resolved; err reports otherwise
        out_lo[g] = lo;
        if (W == 2) out_hi[g] = hi;
      }
    }
  }
}

// Threads per onesweep block for a sort of n keys (SRDATALOG_OS_BLOCK =
// 256, 512 or 1024 forces one).  Larger blocks sort larger tiles: fewer
// tiles to chain through the look-back and longer runs of one digit in each
// tile's scattered write, at the same 16 waves per CU (four 256-thread
// blocks, two of 512 with 77 KB of LDS each, or one of 1024 with 147 KB).
// Measured per pass (k = 2): 576 / 494 / 455 us at 64 M keys and 40 / 35 /
// 34 us at 4 M for 256 / 512 / 1024; below 4 M keys the 1024-thread
// tiles are too few to fill 256 CUs (30 us against 26 us for 512 at 1 M).
static int os_block_for(uint64_t n) {
  static const int forced = [] {
    const char* e = getenv("SRDATALOG_OS_BLOCK");
    const int v = e ? atoi(e) : 0;
    return (v == kBlock || v == 2 * kBlock || v == 4 * kBlock) ? v : 0;
  }();
  if (forced) return forced;
  return n >= (uint64_t(1) << 22) ? 4 * kBlock : 2 * kBlock;
}

static uint64_t os_tiles(int W, int B, uint64_t n) {
  const uint64_t t = os_tile(W, B);
  return (n + t - 1) / t;
}

// One onesweep pass over `tiles` tiles of B threads (os_tiles(W, B, n)).
static void onesweep_launch(int W, int B, uint64_t tiles, const uint64_t* ilo, const uint64_t* ihi,
                            uint64_t* olo, uint64_t* ohi, uint32_t n, uint32_t shift,
                            const uint32_t* starts, uint64_t* lookback, uint32_t* ctr,
                            uint32_t epoch, uint32_t* err, bool ticket, const DevPlan* dp,
                            uint32_t pass_no, hipStream_t s) {
#define SRD_OS_LAUNCH(WW, BB)                                                                 \
  onesweep_kernel<WW, BB><<<uint32_t(tiles), BB, 0, s>>>(ilo, ihi, olo, ohi, n, shift, starts, \
                                                          lookback, ctr, epoch, err, ticket, dp, \
                                                          pass_no)
  const int b = B;
  if (W == 1) {
    if (b == kBlock) SRD_OS_LAUNCH(1, kBlock);
    else if (b == 2 * kBlock) SRD_OS_LAUNCH(1, 2 * kBlock);
    else SRD_OS_LAUNCH(1, 4 * kBlock);
  } else {
    if (b == kBlock) SRD_OS_LAUNCH(2, kBlock);
    else if (b == 2 * kBlock) SRD_OS_LAUNCH(2, 2 * kBlock);
    else SRD_OS_LAUNCH(2, 4 * kBlock);
  }
#undef SRD_OS_LAUNCH
  SRD_CHECK(hipGetLastError());
}

__global__ void bin_scan_kernel(const uint32_t* __restrict__ hist, uint32_t passes,
                                uint32_t* __restrict__ starts) {
  // one block of kRadix threads per pass
  __shared__ uint32_t wsum[kWavesPerBlock];
  uint32_t p = blockIdx.x;
  uint32_t v = hist[p * kRadix + threadIdx.x];
  uint32_t tot;
  uint32_t ex = block_excl_scan<uint32_t>(v, wsum, &tot);
  starts[p * kRadix + threadIdx.x] = ex;
}

// ========================================================= sort: dedup ====
template <int W>
__device__ __forceinline__ void load_key(const uint64_t* lo, const uint64_t* hi, uint64_t i,
                                         uint64_t& a, uint64_t& b) {
  a = lo[i];
  b = (W == 2) ? hi[i] : 0;
}

// Column 0 of the packed key: bits [s0, s0+32).
template <int W>
__device__ __forceinline__ uint32_t key_col(uint64_t lo, uint64_t hi, uint32_t s, uint32_t bits) {
  if (bits == 0) return 0;
  uint64_t v;
  if (W == 1 || s + bits <= 64) v = lo >> s;
  else if (s >= 64) v = hi >> (s - 64);
  else v = (lo >> s) | (hi << (64 - s));
  return bits >= 32 ? uint32_t(v) : uint32_t(v) & ((1u << bits) - 1u);
}

struct UnpackSpec {
  uint32_t shift[kMaxCols];
  uint32_t bits[kMaxCols];
  uint32_t k;
};

template <int W>
__global__ __launch_bounds__(kBlock) void dedup_count_kernel(const uint64_t* __restrict__ lo,
                                                             const uint64_t* __restrict__ hi,
                                                             uint64_t n, UnpackSpec spec,
                                                             uint32_t* __restrict__ tile_u,
                                                             uint32_t* __restrict__ tile_r) {
  __shared__ uint32_t su[kWavesPerBlock], sr[kWavesPerBlock];
  const uint32_t w = threadIdx.x / kWave, l = lane_id();
  const uint64_t base = uint64_t(blockIdx.x) * kTile;
  uint32_t cu = 0, cr = 0;
  for (int i = 0; i < kItems; ++i) {
    uint64_t idx = base + tile_pos(w, i, l);
    bool fu = false, fr = false;
    if (idx < n) {
      uint64_t a, b, pa = 0, pb = 0;
      load_key<W>(lo, hi, idx, a, b);
      if (idx > 0) load_key<W>(lo, hi, idx - 1, pa, pb);
      fu = idx == 0 || a != pa || b != pb;
      fr = idx == 0 || key_col<W>(a, b, spec.shift[0], spec.bits[0]) !=
                           key_col<W>(pa, pb, spec.shift[0], spec.bits[0]);
    }
    cu += __popcll(ballot(fu));
    cr += __popcll(ballot(fr));
  }
  if (l == 0) su[w] = cu, sr[w] = cr;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t a = 0, b = 0;
    for (int j = 0; j < kWavesPerBlock; ++j) a += su[j], b += sr[j];
    tile_u[blockIdx.x] = a;
    tile_r[blockIdx.x] = b;
  }
}

template <int W>
__global__ __launch_bounds__(kBlock) void dedup_write_kernel(
    const uint64_t* __restrict__ lo, const uint64_t* __restrict__ hi, uint64_t n, UnpackSpec spec,
    const uint64_t* __restrict__ uoff, const uint64_t* __restrict__ roff,
    uint32_t* __restrict__ cols, uint64_t stride, uint32_t* __restrict__ root_vals,
    uint32_t* __restrict__ root_offs) {
  __shared__ uint32_t su[kWavesPerBlock], sr[kWavesPerBlock];
  const uint32_t w = threadIdx.x / kWave, l = lane_id();
  const uint64_t base = uint64_t(blockIdx.x) * kTile;
  uint64_t a[kItems], b[kItems];
  uint64_t mu[kItems], mr[kItems];
  uint32_t cu = 0, cr = 0;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    uint64_t idx = base + tile_pos(w, i, l);
    bool fu = false, fr = false;
    a[i] = b[i] = 0;
    if (idx < n) {
      uint64_t pa = 0, pb = 0;
      load_key<W>(lo, hi, idx, a[i], b[i]);
      if (idx > 0) load_key<W>(lo, hi, idx - 1, pa, pb);
      fu = idx == 0 || a[i] != pa || b[i] != pb;
      fr = idx == 0 || key_col<W>(a[i], b[i], spec.shift[0], spec.bits[0]) !=
                           key_col<W>(pa, pb, spec.shift[0], spec.bits[0]);
    }
    mu[i] = ballot(fu);
    mr[i] = ballot(fr);
    cu += __popcll(mu[i]);
    cr += __popcll(mr[i]);
  }
  if (l == 0) su[w] = cu, sr[w] = cr;
  __syncthreads();
  uint64_t pu = uoff[blockIdx.x], pr = roff[blockIdx.x];
  for (uint32_t j = 0; j < w; ++j) pu += su[j], pr += sr[j];
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    if ((mu[i] >> l) & 1) {
      uint64_t row = pu + popc_below(mu[i]);
      for (uint32_t c = 0; c < spec.k; ++c)
        cols[uint64_t(c) * stride + row] = key_col<W>(a[i], b[i], spec.shift[c], spec.bits[c]);
      if ((mr[i] >> l) & 1) {
        uint64_t g = pr + popc_below(mr[i]);
        root_vals[g] = key_col<W>(a[i], b[i], spec.shift[0], spec.bits[0]);
        root_offs[g] = uint32_t(row);
      }
    }
    pu += __popcll(mu[i]);
    pr += __popcll(mr[i]);
  }
}

// ==================================================== sort: host driver ====
static uint32_t bit_width(uint32_t v) { return v ? 32 - __builtin_clz(v) : 0; }

void build_index(const uint32_t* const* in, uint32_t k, uint64_t n, Index& out, Workspace& ws,
                 hipStream_t s, const std::vector<uint32_t>* col_max, const DevBuf<uint32_t>* adopt,
                 uint64_t adopt_stride, bool presorted) {
  if (k == 0 || k > uint32_t(kMaxCols)) fail("build_index: arity must be in [1, 8]");
  if (n >= (uint64_t(1) << 32)) fail("build_index: more than 2^32-1 rows");
  out = Index();
  out.arity = k;
  out.col_max.assign(k, 0);
  if (n == 0) return;

  PackSpec ps{};
  ps.k = k;
  for (uint32_t c = 0; c < k; ++c) ps.in[c] = in[c];

  // Scratch layout in ws.small: [hist P*R][starts P*R][tile ctr P][err 1]
  // [col max 8][dedup tile counts 2*tiles].
  const uint64_t dtiles = (n + kTile - 1) / kTile;
  ws.small.reserve(uint64_t(2 * kMaxPasses) * kRadix + kMaxPasses + 1 + kMaxCols + 1 + 2 * dtiles, s);
  uint32_t* hist = ws.small.p;
  uint32_t* starts = hist + kMaxPasses * kRadix;
  uint32_t* ctr = starts + kMaxPasses * kRadix;
  uint32_t* err = ctr + kMaxPasses;
  uint32_t* dmax = err + 1;
  uint32_t* tu = dmax + kMaxCols + 1;
  uint32_t* tr = tu + dtiles;

  // 1. column maxima -> packed bit layout; input already in order -> no passes
  std::vector<uint32_t> mx(k, 0);
  bool already_sorted = false;
  uint32_t* hp = nullptr;
  if (col_max && col_max->size() == k) {
    mx = *col_max;
    if (presorted) {
      out.n = n;
      out.col_max = mx;
      if (adopt && adopt->p) {
        out.cols = *adopt;
        out.stride = adopt_stride;
      } else {
        out.stride = round_stride(n);
        out.cols.alloc(out.stride * k, s);
        for (uint32_t c = 0; c < k; ++c)
          SRD_CHECK(hipMemcpyAsync(out.cols.p + uint64_t(c) * out.stride, in[c], n * 4,
                                   hipMemcpyDeviceToDevice, s));
      }
      build_roots(out, ws, s);
      return;
    }
  } else {
    fill_u32(dmax, 0, kMaxCols + 1, s);
    {
      TimeScope t("sort.colmax", s, double(n) * 4 * k);
      bool aligned = true;
      for (uint32_t c = 0; c < k; ++c) aligned &= (reinterpret_cast<uintptr_t>(in[c]) & 15u) == 0;
      if (aligned)
        col_max_vec_kernel<<<grid_for(std::max<uint64_t>(n / 4, 1), kBlock * 4, 1024), kBlock, 0,
                             s>>>(ps, n, dmax);
      else
        col_max_kernel<<<grid_for(n, kBlock * 16, 1024), kBlock, 0, s>>>(ps, n, dmax);
      SRD_CHECK(hipGetLastError());
    }
    hp = pinned_words(ws, kMaxCols + 1);
    SRD_CHECK(hipMemcpyAsync(hp, dmax, (kMaxCols + 1) * 4, hipMemcpyDeviceToHost, s));
    SRD_CHECK(hipStreamSynchronize(s));
    for (uint32_t c = 0; c < k; ++c) mx[c] = hp[c];
    already_sorted = (hp[kMaxCols] & 1u) == 0;
    if (hp[kMaxCols] == 0) {
      // strictly increasing: the rows are already a sorted, duplicate-free index
      out.n = n;
      out.col_max = mx;
      if (adopt && adopt->p) {
        out.cols = *adopt;
        out.stride = adopt_stride;
      } else {
        out.stride = round_stride(n);
        out.cols.alloc(out.stride * k, s);
        for (uint32_t c = 0; c < k; ++c)
          SRD_CHECK(hipMemcpyAsync(out.cols.p + uint64_t(c) * out.stride, in[c], n * 4,
                                   hipMemcpyDeviceToDevice, s));
      }
      build_roots(out, ws, s);
      return;
    }
  }
  UnpackSpec us{};
  us.k = k;
  uint32_t total_bits = 0;
  for (int c = int(k) - 1; c >= 0; --c) {
    uint32_t b = bit_width(mx[c]);
    us.shift[c] = total_bits;
    us.bits[c] = b;
    ps.shift[c] = total_bits;
    total_bits += b;
  }
  if (total_bits > 128) fail("build_index: packed key wider than 128 bits");
  const int W = total_bits > 64 ? 2 : 1;
  const uint32_t passes = (total_bits + kRadixBits - 1) / kRadixBits;
  ps.passes = passes;

  for (int b = 0; b < 2; ++b)
    for (int wd = 0; wd < W; ++wd) ws.keys[b][wd].reserve(n, s);
  fill_u32(hist, 0, uint64_t(kMaxPasses) * kRadix, s);
  fill_u32(err, 0, 1, s);

  // 2. pack + histograms
  {
    TimeScope t("sort.pack", s, double(n) * (4.0 * k + 8.0 * W));
    uint32_t g = grid_for(n, kBlock * 8, 4096);
    if (W == 1)
      pack_hist_kernel<1><<<g, kBlock, 0, s>>>(ps, n, ws.keys[0][0].p, nullptr, hist, ctr);
    else
      pack_hist_kernel<2><<<g, kBlock, 0, s>>>(ps, n, ws.keys[0][0].p, ws.keys[0][1].p, hist, ctr);
    SRD_CHECK(hipGetLastError());
  }
  int cur = 0;
  if (passes > 0) {
    bin_scan_kernel<<<passes, kRadix, 0, s>>>(hist, passes, starts);
    SRD_CHECK(hipGetLastError());
    // Skip passes whose digit is constant over all keys.
    hp = pinned_words(ws, uint64_t(passes) * kRadix);
    SRD_CHECK(hipMemcpyAsync(hp, hist, uint64_t(passes) * kRadix * 4, hipMemcpyDeviceToHost, s));
    SRD_CHECK(hipStreamSynchronize(s));
    std::vector<bool> trivial(passes, false);
    for (uint32_t p = 0; p < passes; ++p)
      for (int d = 0; d < kRadix; ++d)
        if (hp[p * kRadix + d] == n || already_sorted) trivial[p] = true;

    const int os_b = os_block_for(n);
    const uint64_t tiles = os_tiles(W, os_b, n);
    uint64_t lb_words = tiles * kRadix;
    if (lb_words > ws.lookback.n) {
      ws.lookback.alloc(lb_words + lb_words / 4, s);
      fill_u64(ws.lookback.p, 0, ws.lookback.n, s);
      ws.epoch = 0;
    }
    for (uint32_t p = 0; p < passes; ++p) {
      if (trivial[p]) continue;
      if (++ws.epoch >= 0xFFFF) {
        fill_u64(ws.lookback.p, 0, ws.lookback.n, s);
        ws.epoch = 1;
      }
      TimeScope t("sort.onesweep", s, double(n) * 16.0 * W);
      const uint64_t* ilo = ws.keys[cur][0].p;
      const uint64_t* ihi = W == 2 ? ws.keys[cur][1].p : nullptr;
      uint64_t* olo = ws.keys[cur ^ 1][0].p;
      uint64_t* ohi = W == 2 ? ws.keys[cur ^ 1][1].p : nullptr;
      onesweep_launch(W, os_b, tiles, ilo, ihi, olo, ohi, uint32_t(n), p * kRadixBits,
                      starts + p * kRadix, ws.lookback.p, ctr + p, ws.epoch, err,
                      ws.ticket_order, nullptr, 0, s);
      cur ^= 1;
    }
  }

  // 3. dedup + unpack + roots
  const uint64_t tiles = dtiles;
  ws.tile_scan.reserve(2 * (tiles + 1), s);
  uint64_t* su = ws.tile_scan.p;
  uint64_t* sr = su + tiles + 1;
  const uint64_t* klo = ws.keys[cur][0].p;
  const uint64_t* khi = W == 2 ? ws.keys[cur][1].p : nullptr;
  {
    TimeScope t("sort.dedup_count", s, double(n) * 8 * W);
    if (W == 1)
      dedup_count_kernel<1><<<uint32_t(tiles), kBlock, 0, s>>>(klo, khi, n, us, tu, tr);
    else
      dedup_count_kernel<2><<<uint32_t(tiles), kBlock, 0, s>>>(klo, khi, n, us, tu, tr);
    SRD_CHECK(hipGetLastError());
  }
  scan_impl<uint32_t>(tu, su, tiles, s);
  scan_impl<uint32_t>(tr, sr, tiles, s);
  uint64_t* h64 = reinterpret_cast<uint64_t*>(pinned_words(ws, 8));
  SRD_CHECK(hipMemcpyAsync(&h64[0], su + tiles, 8, hipMemcpyDeviceToHost, s));
  SRD_CHECK(hipMemcpyAsync(&h64[1], sr + tiles, 8, hipMemcpyDeviceToHost, s));
  SRD_CHECK(hipMemcpyAsync(&h64[2], err, 4, hipMemcpyDeviceToHost, s));
  SRD_CHECK(hipStreamSynchronize(s));
  if (uint32_t(h64[2]) != 0) fail("build_index: radix sort look-back did not resolve");
  const uint64_t u = h64[0], r = h64[1];
  out.n = u;
  out.r = uint32_t(r);
  out.stride = round_stride(u);
  out.cols.alloc(out.stride * k, s);
  out.root_vals.alloc(r, s);
  out.root_offs.alloc(r + 1, s);
  {
    TimeScope t("sort.dedup_write", s, double(n) * 8 * W + double(u) * 4 * k + double(r) * 8);
    if (W == 1)
      dedup_write_kernel<1><<<uint32_t(tiles), kBlock, 0, s>>>(klo, khi, n, us, su, sr, out.cols.p,
                                                               out.stride, out.root_vals.p,
                                                               out.root_offs.p);
    else
      dedup_write_kernel<2><<<uint32_t(tiles), kBlock, 0, s>>>(klo, khi, n, us, su, sr, out.cols.p,
                                                               out.stride, out.root_vals.p,
                                                               out.root_offs.p);
    SRD_CHECK(hipGetLastError());
  }
  fill_u32(out.root_offs.p + r, uint32_t(u), 1, s);
  out.col_max = mx;
}

void build_index_from(const Index& src, const std::vector<uint32_t>& perm, Index& out,
                      Workspace& ws, hipStream_t s) {
  if (perm.size() != src.arity) fail("build_index_from: permutation arity mismatch");
  std::vector<const uint32_t*> in(src.arity);
  std::vector<uint32_t> mx(src.arity);
  for (uint32_t c = 0; c < src.arity; ++c) {
    in[c] = src.col(perm[c]);
    mx[c] = src.col_max.size() == src.arity ? src.col_max[perm[c]] : 0;
  }
  if (src.n == 0) {
    out = Index();
    out.arity = src.arity;
    out.col_max.assign(src.arity, 0);
    return;
  }
  bool known = src.col_max.size() == src.arity;
  build_index(in.data(), src.arity, src.n, out, ws, s, known ? &mx : nullptr);
}

// ======================================== device-planned sort + dedup ====
// build_index_dev: the same radix sort + dedup as build_index, with no host
// read-back and no allocation.  The row count may live in device memory;
// column maxima, the packed key layout, which passes run and which buffer
// each reads are decided by device kernels (DevPlan), and every launch is
// shaped by a host-known row capacity, so the whole sequence can be
// enqueued (or captured) without knowing the data.  Passes whose digit is
// constant and every pass over already-ordered input are skipped on the
// device; a skipped pass leaves the keys in place, so the buffer the next
// pass reads is the parity of the passes run before it.

// Scratch carved from one caller-provided buffer (see devsort_scratch_bytes).
struct DevSortScratch {
  DevPlan* plan;
  uint64_t* keys[2][2];     // [buffer][word], cap entries each
  uint32_t* hist;           // kMaxPasses * kRadix
  uint32_t* starts;         // kMaxPasses * kRadix
  uint32_t* colmax;         // kMaxCols + 1 (order flags last)
  uint32_t* tickets;        // kMaxPasses onesweep tile tickets
  uint32_t* epoch_ctr;      // last look-back epoch handed out (device-resident, so
                            // replays of a captured sort never reuse one)
  uint64_t* lookback;       // lb_tiles * kRadix
  uint64_t lookback_words;
  uint32_t* tile_u;         // dedup tiles
  uint32_t* tile_r;
  uint64_t* scan_u;         // dedup tiles + 1
  uint64_t* scan_r;
  uint64_t* scan_tmp;       // multi-level scan partials (scan_noalloc_words)
};

static uint64_t align256(uint64_t b) { return (b + 255) & ~uint64_t(255); }

// u64 words of partial sums a scan of n entries needs beyond its output.
static uint64_t scan_noalloc_words(uint64_t n) {
  if (n <= kTile) return 0;
  const uint64_t tiles = (n + kTile - 1) / kTile;
  return 2 * tiles + 1 + scan_noalloc_words(tiles);
}

// Exclusive scan (out[n] = total) using `tmp` for the partial sums: the
// multi-level scan_impl without its allocations.
template <class T>
static void scan_noalloc(const T* in, uint64_t* out, uint64_t n, uint64_t* tmp, hipStream_t s) {
  if (n <= kTile) {
    scan_tile_kernel<T><<<1, kBlock, 0, s>>>(in, n, nullptr, out);
    SRD_CHECK(hipGetLastError());
    return;
  }
  const uint64_t tiles = (n + kTile - 1) / kTile;
  uint64_t* partial = tmp;
  uint64_t* carry = tmp + tiles;
  scan_reduce_kernel<T><<<uint32_t(tiles), kBlock, 0, s>>>(in, n, partial);
  SRD_CHECK(hipGetLastError());
  scan_noalloc<uint64_t>(partial, carry, tiles, tmp + 2 * tiles + 1, s);
  scan_tile_kernel<T><<<uint32_t(tiles), kBlock, 0, s>>>(in, n, carry, out);
  SRD_CHECK(hipGetLastError());
}

// The layout depends on the capacity only (both key words are always laid
// out), so one scratch serves every key width without moving its look-back
// words: a look-back word is only ever written as one.
uint64_t devsort_scratch_bytes(uint64_t cap) {
  const uint64_t lb_tiles = (cap + kOsMinTile - 1) / kOsMinTile;
  const uint64_t dt = (cap + kTile - 1) / kTile;
  uint64_t b = align256(sizeof(DevPlan));
  b += 4 * align256(cap * 8);
  b += 2 * align256(uint64_t(kMaxPasses) * kRadix * 4);
  b += align256((kMaxCols + 1) * 4) + align256(kMaxPasses * 4) + align256(4);
  b += align256(lb_tiles * kRadix * 8);
  b += 2 * align256(dt * 4) + 2 * align256((dt + 1) * 8);
  b += align256(std::max<uint64_t>(scan_noalloc_words(dt), 1) * 8);
  return b;
}

static DevSortScratch carve_devsort(void* base, uint64_t cap) {
  uint8_t* p = static_cast<uint8_t*>(base);
  auto take = [&](uint64_t bytes) {
    uint8_t* q = p;
    p += align256(bytes);
    return q;
  };
  const uint64_t lb_tiles = (cap + kOsMinTile - 1) / kOsMinTile;
  const uint64_t dt = (cap + kTile - 1) / kTile;
  DevSortScratch x{};
  x.plan = reinterpret_cast<DevPlan*>(take(sizeof(DevPlan)));
  for (int b = 0; b < 2; ++b)
    for (uint32_t wd = 0; wd < 2; ++wd)
      x.keys[b][wd] = reinterpret_cast<uint64_t*>(take(cap * 8));
  x.hist = reinterpret_cast<uint32_t*>(take(uint64_t(kMaxPasses) * kRadix * 4));
  x.starts = reinterpret_cast<uint32_t*>(take(uint64_t(kMaxPasses) * kRadix * 4));
  x.colmax = reinterpret_cast<uint32_t*>(take((kMaxCols + 1) * 4));
  x.tickets = reinterpret_cast<uint32_t*>(take(kMaxPasses * 4));
  x.epoch_ctr = reinterpret_cast<uint32_t*>(take(4));
  x.lookback_words = lb_tiles * kRadix;
  x.lookback = reinterpret_cast<uint64_t*>(take(x.lookback_words * 8));
  x.tile_u = reinterpret_cast<uint32_t*>(take(dt * 4));
  x.tile_r = reinterpret_cast<uint32_t*>(take(dt * 4));
  x.scan_u = reinterpret_cast<uint64_t*>(take((dt + 1) * 8));
  x.scan_r = reinterpret_cast<uint64_t*>(take((dt + 1) * 8));
  x.scan_tmp = reinterpret_cast<uint64_t*>(take(std::max<uint64_t>(scan_noalloc_words(dt), 1) * 8));
  return x;
}

// Row count, cleared histograms, maxima and tickets.  n_dev (may be null)
// overrides n_host and is clamped to the capacity (err bit 1 when above).
// Error bits of a plan: 0 look-back did not resolve, 1 rows above the
// capacity, 2 key wider than the launch's key words.
__global__ __launch_bounds__(kBlock) void devsort_init_kernel(DevPlan* __restrict__ plan,
                                                             const uint64_t* __restrict__ n_dev,
                                                             uint64_t n_host, uint64_t cap, uint32_t k,
                                                             uint32_t* __restrict__ hist,
                                                             uint32_t* __restrict__ colmax,
                                                             uint32_t* __restrict__ tickets,
                                                             uint32_t* __restrict__ epoch_ctr) {
  for (uint32_t i = threadIdx.x; i < uint32_t(kMaxPasses) * kRadix; i += kBlock) hist[i] = 0;
  if (threadIdx.x <= uint32_t(kMaxCols)) colmax[threadIdx.x] = 0;
  if (threadIdx.x < uint32_t(kMaxPasses)) tickets[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    uint64_t n = n_dev ? *n_dev : n_host;
    uint32_t err = 0;
    if (n > cap) {
      n = cap;
      err = 2;
    }
    plan->n = n;
    plan->u = plan->r = 0;
    plan->k = k;
    plan->passes = 0;
    plan->flags = 0;
    plan->err = err;
    plan->final_buf = 0;
    for (int p = 0; p < kMaxPasses; ++p) plan->pass[p] = 0;
    // kMaxPasses epochs per sort; epochs live in 16 bits of a look-back word
    uint32_t e = *epoch_ctr;
    plan->wrap = 0;
    if (e + 2u * kMaxPasses >= 0xFFFFu) {
      e = 0;
      plan->wrap = 1;
    }
    plan->epoch0 = e + 1u;
    *epoch_ctr = e + uint32_t(kMaxPasses);
  }
}

// Clear the look-back words when the plan's epochs wrapped (else nothing).
__global__ __launch_bounds__(kBlock) void devsort_wrap_clear_kernel(const DevPlan* __restrict__ plan,
                                                                   uint64_t* __restrict__ lookback,
                                                                   uint64_t words) {
  if (!plan->wrap) return;
  for (uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x; i < words;
       i += uint64_t(gridDim.x) * kBlock)
    lookback[i] = 0;
}

// Column maxima and order flags (as col_max_kernel) over the planned rows.
__global__ __launch_bounds__(kBlock) void devsort_colmax_kernel(PackSpec spec,
                                                               const DevPlan* __restrict__ plan,
                                                               uint32_t* __restrict__ out_max) {
  const uint64_t n = plan->n;
  uint32_t m[kMaxCols];
  for (uint32_t c = 0; c < spec.k; ++c) m[c] = 0;
  bool unsorted = false, dup = false;
  for (uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * kBlock) {
    int order = 0;
    for (uint32_t c = 0; c < spec.k; ++c) {
      const uint32_t v = spec.in[c][i];
      m[c] = max(m[c], v);
      if (i > 0 && order == 0) {
        const uint32_t pv = spec.in[c][i - 1];
        order = v < pv ? -1 : (v > pv ? 1 : 0);
      }
    }
    unsorted |= order < 0;
    dup |= (i > 0 && order == 0);
  }
  colmax_flush(m, spec.k, unsorted, dup, out_max);
}

// Key layout from the maxima (column k-1 in the lowest bits); err bit 2 when
// the key is wider than the W words the launch was built for.
__global__ void devsort_layout_kernel(DevPlan* __restrict__ plan,
                                      const uint32_t* __restrict__ colmax, uint32_t w) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint32_t k = plan->k;
  uint32_t total = 0;
  for (int c = int(k) - 1; c >= 0; --c) {
    const uint32_t v = colmax[c];
    const uint32_t b = v ? 32u - uint32_t(__builtin_clz(v)) : 0u;
    plan->shift[c] = total;
    plan->bits[c] = b;
    total += b;
  }
  if (total > 64u * w) plan->err |= 4u;
  plan->passes = total > 64u * w ? 0u : (total + kRadixBits - 1) / kRadixBits;
  plan->flags = colmax[kMaxCols];
}

// Pack the planned rows into keys and build every pass's histogram.
template <int W>
__global__ __launch_bounds__(kBlock) void devsort_pack_kernel(PackSpec spec,
                                                             const DevPlan* __restrict__ plan,
                                                             uint64_t* __restrict__ lo_out,
                                                             uint64_t* __restrict__ hi_out,
                                                             uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[kMaxPasses * kRadix];
  const uint64_t n = plan->n;
  const uint32_t passes = plan->passes;
  uint32_t shift[kMaxCols];
  for (uint32_t c = 0; c < spec.k; ++c) shift[c] = plan->shift[c];
  for (uint32_t i = threadIdx.x; i < passes * kRadix; i += kBlock) h[i] = 0;
  __syncthreads();
  // wave-uniform steps (hist_add_runs needs the whole wave)
  const uint64_t stride = uint64_t(gridDim.x) * kBlock;
  for (uint64_t b = blockIdx.x * uint64_t(kBlock) + (threadIdx.x & ~uint32_t(kWave - 1)); b < n;
       b += stride) {
    const uint64_t i = b + lane_id();
    const bool valid = i < n;
    uint64_t lo = 0, hi = 0;
    if (valid) {
      for (uint32_t c = 0; c < spec.k; ++c) {
        const uint64_t v = spec.in[c][i];
        const uint32_t sh = shift[c];
        if (sh >= 64) {
          hi |= v << (sh - 64);
        } else {
          lo |= v << sh;
          if (W == 2 && sh > 32) hi |= v >> (64 - sh);
        }
      }
      lo_out[i] = lo;
      if (W == 2) hi_out[i] = hi;
    }
    for (uint32_t p = 0; p < passes; ++p)
      hist_add_runs(&h[p * kRadix], digit_of<W>(lo, hi, p * kRadixBits), valid);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < passes * kRadix; i += kBlock)
    if (h[i]) atomicAdd(&hist[i], h[i]);
}

// Digit starts of every pass, and the schedule: a pass runs unless its
// digit is constant over the keys or the rows were already in order; the
// buffer it reads alternates with the passes that run.
__global__ __launch_bounds__(kRadix) void devsort_schedule_kernel(DevPlan* __restrict__ plan,
                                                                 const uint32_t* __restrict__ hist,
                                                                 uint32_t* __restrict__ starts) {
  __shared__ uint32_t wsum[kWavesPerBlock];
  __shared__ uint32_t s_trivial[kMaxPasses];
  const uint32_t passes = plan->passes;
  const uint64_t n = plan->n;
  const bool ordered = (plan->flags & 1u) == 0;
  if (threadIdx.x < uint32_t(kMaxPasses)) s_trivial[threadIdx.x] = 0;
  __syncthreads();
  for (uint32_t p = 0; p < passes; ++p) {
    const uint32_t v = hist[p * kRadix + threadIdx.x];
    uint32_t tot;
    starts[p * kRadix + threadIdx.x] = block_excl_scan<uint32_t>(v, wsum, &tot);
    if (uint64_t(v) == n) s_trivial[p] = 1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t buf = 0;
    for (uint32_t p = 0; p < uint32_t(kMaxPasses); ++p) {
      const bool run = p < passes && n > 0 && !ordered && !s_trivial[p];
      plan->pass[p] = (run ? 1u : 0u) | (buf << 1);
      if (run) buf ^= 1u;
    }
    plan->final_buf = buf;
  }
}

// dedup_count_kernel / dedup_write_kernel over the planned rows and layout.
template <int W>
__global__ __launch_bounds__(kBlock) void devsort_dedup_count_kernel(
    const uint64_t* __restrict__ lo0, const uint64_t* __restrict__ hi0,
    const uint64_t* __restrict__ lo1, const uint64_t* __restrict__ hi1,
    const DevPlan* __restrict__ plan, uint32_t* __restrict__ tile_u,
    uint32_t* __restrict__ tile_r) {
  __shared__ uint32_t su[kWavesPerBlock], sr[kWavesPerBlock];
  const uint64_t n = plan->n;
  const bool b1 = plan->final_buf != 0;
  const uint64_t* lo = b1 ? lo1 : lo0;
  const uint64_t* hi = b1 ? hi1 : hi0;
  const uint32_t s0 = plan->shift[0], bits0 = plan->bits[0];
  const uint32_t w = threadIdx.x / kWave, l = lane_id();
  const uint64_t base = uint64_t(blockIdx.x) * kTile;
  uint32_t cu = 0, cr = 0;
  for (int i = 0; i < kItems; ++i) {
    const uint64_t idx = base + tile_pos(w, i, l);
    bool fu = false, fr = false;
    if (idx < n) {
      uint64_t a, b, pa = 0, pb = 0;
      load_key<W>(lo, hi, idx, a, b);
      if (idx > 0) load_key<W>(lo, hi, idx - 1, pa, pb);
      fu = idx == 0 || a != pa || b != pb;
      fr = idx == 0 || key_col<W>(a, b, s0, bits0) != key_col<W>(pa, pb, s0, bits0);
    }
    cu += __popcll(ballot(fu));
    cr += __popcll(ballot(fr));
  }
  if (l == 0) su[w] = cu, sr[w] = cr;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t a = 0, b = 0;
    for (int j = 0; j < kWavesPerBlock; ++j) a += su[j], b += sr[j];
    tile_u[blockIdx.x] = a;
    tile_r[blockIdx.x] = b;
  }
}

template <int W>
__global__ __launch_bounds__(kBlock) void devsort_dedup_write_kernel(
    const uint64_t* __restrict__ lo0, const uint64_t* __restrict__ hi0,
    const uint64_t* __restrict__ lo1, const uint64_t* __restrict__ hi1,
    const DevPlan* __restrict__ plan, const uint64_t* __restrict__ uoff,
    const uint64_t* __restrict__ roff, uint32_t* __restrict__ cols, uint64_t stride,
    uint32_t* __restrict__ root_vals, uint32_t* __restrict__ root_offs) {
  __shared__ uint32_t su[kWavesPerBlock], sr[kWavesPerBlock];
  const uint64_t n = plan->n;
  const bool b1 = plan->final_buf != 0;
  const uint64_t* lo = b1 ? lo1 : lo0;
  const uint64_t* hi = b1 ? hi1 : hi0;
  const uint32_t k = plan->k;
  uint32_t sh[kMaxCols], bt[kMaxCols];
  for (uint32_t c = 0; c < k; ++c) sh[c] = plan->shift[c], bt[c] = plan->bits[c];
  const uint32_t w = threadIdx.x / kWave, l = lane_id();
  const uint64_t base = uint64_t(blockIdx.x) * kTile;
  uint64_t a[kItems], b[kItems], mu[kItems], mr[kItems];
  uint32_t cu = 0, cr = 0;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const uint64_t idx = base + tile_pos(w, i, l);
    bool fu = false, fr = false;
    a[i] = b[i] = 0;
    if (idx < n) {
      uint64_t pa = 0, pb = 0;
      load_key<W>(lo, hi, idx, a[i], b[i]);
      if (idx > 0) load_key<W>(lo, hi, idx - 1, pa, pb);
      fu = idx == 0 || a[i] != pa || b[i] != pb;
      fr = idx == 0 || key_col<W>(a[i], b[i], sh[0], bt[0]) != key_col<W>(pa, pb, sh[0], bt[0]);
    }
    mu[i] = ballot(fu);
    mr[i] = ballot(fr);
    cu += __popcll(mu[i]);
    cr += __popcll(mr[i]);
  }
  if (l == 0) su[w] = cu, sr[w] = cr;
  __syncthreads();
  uint64_t pu = uoff[blockIdx.x], pr = roff[blockIdx.x];
  for (uint32_t j = 0; j < w; ++j) pu += su[j], pr += sr[j];
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    if ((mu[i] >> l) & 1) {
      const uint64_t row = pu + popc_below(mu[i]);
      for (uint32_t c = 0; c < k; ++c)
        cols[uint64_t(c) * stride + row] = key_col<W>(a[i], b[i], sh[c], bt[c]);
      if ((mr[i] >> l) & 1) {
        const uint64_t g = pr + popc_below(mr[i]);
        root_vals[g] = key_col<W>(a[i], b[i], sh[0], bt[0]);
        root_offs[g] = uint32_t(row);
      }
    }
    pu += __popcll(mu[i]);
    pr += __popcll(mr[i]);
  }
}

// Totals into the plan, root_offs[r] = u, and the caller's count words.
__global__ void devsort_finish_kernel(DevPlan* __restrict__ plan,
                                      const uint64_t* __restrict__ su,
                                      const uint64_t* __restrict__ sr, uint64_t tiles,
                                      uint32_t* __restrict__ root_offs,
                                      uint64_t* __restrict__ counts_out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint64_t u = su[tiles], r = sr[tiles];
  plan->u = u;
  plan->r = r;
  root_offs[r] = uint32_t(u);
  if (counts_out) {
    counts_out[0] = u;
    counts_out[1] = r;
    counts_out[2] = plan->err;
  }
}

void build_index_dev(const uint32_t* const* in, uint32_t k, uint64_t cap, const uint64_t* n_dev,
                     uint64_t n_host, uint32_t w, void* scratch, uint32_t* out_cols,
                     uint64_t out_stride, uint32_t* root_vals, uint32_t* root_offs,
                     uint64_t* counts_out, bool ticket, hipStream_t s, bool timed) {
  if (k == 0 || k > 4) fail("build_index_dev: arity must be in [1, 4]");
  if (w != 1 && w != 2) fail("build_index_dev: key words must be 1 or 2");
  if (cap == 0) fail("build_index_dev: zero capacity");
  if (cap >= (uint64_t(1) << 32)) fail("build_index_dev: capacity of 2^32 rows or more");
  DevSortScratch x = carve_devsort(scratch, cap);
  auto tname = [&](const char* nm) { return timed ? nm : nullptr; };
  PackSpec ps{};
  ps.k = k;
  for (uint32_t c = 0; c < k; ++c) ps.in[c] = in[c];
  devsort_init_kernel<<<1, kBlock, 0, s>>>(x.plan, n_dev, n_host, cap, k, x.hist, x.colmax,
                                           x.tickets, x.epoch_ctr);
  SRD_CHECK(hipGetLastError());
  devsort_wrap_clear_kernel<<<grid_for(x.lookback_words, kBlock * 16, 1024), kBlock, 0, s>>>(
      x.plan, x.lookback, x.lookback_words);
  SRD_CHECK(hipGetLastError());
  {
    TimeScope t(tname("sort.colmax"), s);
    devsort_colmax_kernel<<<grid_for(cap, kBlock * 16, 1024), kBlock, 0, s>>>(ps, x.plan, x.colmax);
    SRD_CHECK(hipGetLastError());
  }
  devsort_layout_kernel<<<1, 64, 0, s>>>(x.plan, x.colmax, w);
  SRD_CHECK(hipGetLastError());
  {
    TimeScope t(tname("sort.pack"), s);
    const uint32_t g = grid_for(cap, kBlock * 8, 4096);
    if (w == 1)
      devsort_pack_kernel<1><<<g, kBlock, 0, s>>>(ps, x.plan, x.keys[0][0], nullptr, x.hist);
    else
      devsort_pack_kernel<2><<<g, kBlock, 0, s>>>(ps, x.plan, x.keys[0][0], x.keys[0][1], x.hist);
    SRD_CHECK(hipGetLastError());
  }
  devsort_schedule_kernel<<<1, kRadix, 0, s>>>(x.plan, x.hist, x.starts);
  SRD_CHECK(hipGetLastError());
  // every pass the widest layout could need is launched; the plan skips
  // the ones the data does not
  const uint32_t max_passes = (32u * k + kRadixBits - 1) / kRadixBits;
  const int os_b = os_block_for(cap);
  const uint64_t tiles = os_tiles(int(w), os_b, cap);
  for (uint32_t p = 0; p < std::min<uint32_t>(max_passes, uint32_t(w) * 64 / kRadixBits); ++p) {
    TimeScope t(tname("sort.onesweep"), s);
    onesweep_launch(int(w), os_b, tiles, x.keys[0][0], w == 2 ? x.keys[0][1] : nullptr, x.keys[1][0],
                    w == 2 ? x.keys[1][1] : nullptr, uint32_t(cap), p * kRadixBits,
                    x.starts + p * kRadix, x.lookback, x.tickets + p, 0u, &x.plan->err, ticket,
                    x.plan, p, s);
  }
  const uint64_t dt = (cap + kTile - 1) / kTile;
  {
    TimeScope t(tname("sort.dedup_count"), s);
    if (w == 1)
      devsort_dedup_count_kernel<1><<<uint32_t(dt), kBlock, 0, s>>>(
          x.keys[0][0], nullptr, x.keys[1][0], nullptr, x.plan, x.tile_u, x.tile_r);
    else
      devsort_dedup_count_kernel<2><<<uint32_t(dt), kBlock, 0, s>>>(
          x.keys[0][0], x.keys[0][1], x.keys[1][0], x.keys[1][1], x.plan, x.tile_u, x.tile_r);
    SRD_CHECK(hipGetLastError());
  }
  scan_noalloc<uint32_t>(x.tile_u, x.scan_u, dt, x.scan_tmp, s);
  scan_noalloc<uint32_t>(x.tile_r, x.scan_r, dt, x.scan_tmp, s);
  {
    TimeScope t(tname("sort.dedup_write"), s);
    if (w == 1)
      devsort_dedup_write_kernel<1><<<uint32_t(dt), kBlock, 0, s>>>(
          x.keys[0][0], nullptr, x.keys[1][0], nullptr, x.plan, x.scan_u, x.scan_r, out_cols,
          out_stride, root_vals, root_offs);
    else
      devsort_dedup_write_kernel<2><<<uint32_t(dt), kBlock, 0, s>>>(
          x.keys[0][0], x.keys[0][1], x.keys[1][0], x.keys[1][1], x.plan, x.scan_u, x.scan_r,
          out_cols, out_stride, root_vals, root_offs);
    SRD_CHECK(hipGetLastError());
  }
  devsort_finish_kernel<<<1, 64, 0, s>>>(x.plan, x.scan_u, x.scan_r, dt, root_offs, counts_out);
  SRD_CHECK(hipGetLastError());
}

void devsort_clear_lookback(void* scratch, uint64_t cap, hipStream_t s) {
  DevSortScratch x = carve_devsort(scratch, cap);
  fill_u64(x.lookback, 0, x.lookback_words, s);
  fill_u32(x.epoch_ctr, 0, 1, s);
}

// ============================================================== roots ====
// One read of column 0: each tile writes the positions of its key starts
// at its own first row of `pos` (a tile has at most kTile of them) and their
// count; root_gather_kernel packs them after a scan of the counts.
__global__ __launch_bounds__(kBlock) void root_mark_kernel(const uint32_t* __restrict__ c0,
                                                           uint64_t n,
                                                           uint32_t* __restrict__ pos,
                                                           uint32_t* __restrict__ tile_r,
                                                           const uint64_t* __restrict__ n_dev = nullptr) {
  __shared__ uint32_t sr[kWavesPerBlock];
  if (n_dev) n = min(n, *n_dev);  // tiles past the count record 0 starts
  const uint32_t w = threadIdx.x / kWave, l = lane_id();
  const uint64_t base = uint64_t(blockIdx.x) * kTile;
  uint64_t m[kItems];
  uint32_t cr = 0;
  // each value's predecessor comes from the lane below; lane 0 loads it
  uint32_t v[kItems], pv[kItems];
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const uint64_t idx = base + tile_pos(w, i, l);
    v[i] = idx < n ? c0[idx] : 0u;
    pv[i] = (l == 0 && idx > 0 && idx - 1 < n) ? c0[idx - 1] : 0u;
  }
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const uint64_t idx = base + tile_pos(w, i, l);
    const uint32_t up = __shfl_up(v[i], 1, kWave);
    const uint32_t prev = l == 0 ? pv[i] : up;
    const bool f = idx < n && (idx == 0 || v[i] != prev);
    m[i] = ballot(f);
    cr += __popcll(m[i]);
  }
  if (l == 0) sr[w] = cr;
  __syncthreads();
  uint32_t pr = 0, tot = 0;
  for (uint32_t j = 0; j < kWavesPerBlock; ++j) {
    if (j < w) pr += sr[j];
    tot += sr[j];
  }
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    if ((m[i] >> l) & 1) pos[base + pr + popc_below(m[i])] = uint32_t(base + tile_pos(w, i, l));
    pr += __popcll(m[i]);
  }
  if (threadIdx.x == 0) tile_r[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kBlock) void root_gather_kernel(const uint32_t* __restrict__ c0,
                                                             const uint32_t* __restrict__ pos,
                                                             const uint32_t* __restrict__ tile_r,
                                                             const uint64_t* __restrict__ roff,
                                                             uint32_t* __restrict__ root_vals,
                                                             uint32_t* __restrict__ root_offs) {
  const uint64_t base = uint64_t(blockIdx.x) * kTile;
  const uint64_t dst = roff[blockIdx.x];
  const uint32_t cnt = tile_r[blockIdx.x];
  for (uint32_t j = threadIdx.x; j < cnt; j += kBlock) {
    const uint32_t p = pos[base + j];
    root_offs[dst + j] = p;
    root_vals[dst + j] = c0[p];
  }
}

// Build root arrays of an index whose columns are filled.
void build_roots(Index& x, Workspace& ws, hipStream_t s) {
  x.root_vals.release();
  x.root_offs.release();
  x.r = 0;
  if (x.n == 0) return;
  const uint64_t tiles = (x.n + kTile - 1) / kTile;
  DevBuf<uint32_t> tr(tiles, s);
  DevBuf<uint64_t> sr(tiles + 1, s);
  DevBuf<uint32_t> pos(x.n, s);
  {
    TimeScope t("roots", s, double(x.n) * 4);
    root_mark_kernel<<<uint32_t(tiles), kBlock, 0, s>>>(x.col(0), x.n, pos.p, tr.p);
    SRD_CHECK(hipGetLastError());
  }
  scan_impl<uint32_t>(tr.p, sr.p, tiles, s);
  uint64_t* h64 = reinterpret_cast<uint64_t*>(pinned_words(ws, 2));
  SRD_CHECK(hipMemcpyAsync(h64, sr.p + tiles, 8, hipMemcpyDeviceToHost, s));
  SRD_CHECK(hipStreamSynchronize(s));
  x.r = uint32_t(h64[0]);
  x.root_vals.alloc(x.r, s);
  x.root_offs.alloc(x.r + 1, s);
  {
    TimeScope t("roots", s, double(x.r) * 12);
    root_gather_kernel<<<uint32_t(tiles), kBlock, 0, s>>>(x.col(0), pos.p, tr.p, sr.p,
                                                          x.root_vals.p, x.root_offs.p);
    SRD_CHECK(hipGetLastError());
  }
  fill_u32(x.root_offs.p + x.r, uint32_t(x.n), 1, s);
}

// ============================================ windowed lexicographic search ====
// Stage capacity (rows) of the other input's window in LDS.
constexpr uint32_t kStageRows = 2048;

template <int K>
struct RowK {
  uint32_t v[K];
};

template <int K>
__device__ __forceinline__ RowK<K> load_row(const uint32_t* cols, uint64_t stride, uint64_t i) {
  RowK<K> r;
#pragma unroll
  for (int c = 0; c < K; ++c) r.v[c] = cols[uint64_t(c) * stride + i];
  return r;
}

template <int K>
__device__ __forceinline__ bool row_less(const RowK<K>& a, const RowK<K>& b) {
#pragma unroll
  for (int c = 0; c < K; ++c)
    if (a.v[c] != b.v[c]) return a.v[c] < b.v[c];
  return false;
}

template <int K>
__device__ __forceinline__ bool row_eq(const RowK<K>& a, const RowK<K>& b) {
#pragma unroll
  for (int c = 0; c < K; ++c)
    if (a.v[c] != b.v[c]) return false;
  return true;
}

// First row in [lo,hi) of (cols,stride) not less than key.
template <int K>
__device__ __forceinline__ uint64_t lex_lower_bound(const uint32_t* cols, uint64_t stride,
                                                    uint64_t lo, uint64_t hi, const RowK<K>& key) {
  uint64_t n = hi - lo;
  while (n > 0) {
    uint64_t half = n >> 1, mid = lo + half;
    if (row_less<K>(load_row<K>(cols, stride, mid), key)) {
      lo = mid + 1;
      n -= half + 1;
    } else {
      n = half;
    }
  }
  return lo;
}

// First row in [lo,hi) greater than key.
template <int K>
__device__ __forceinline__ uint64_t lex_upper_bound(const uint32_t* cols, uint64_t stride,
                                                    uint64_t lo, uint64_t hi, const RowK<K>& key) {
  uint64_t n = hi - lo;
  while (n > 0) {
    uint64_t half = n >> 1, mid = lo + half;
    if (!row_less<K>(key, load_row<K>(cols, stride, mid))) {
      lo = mid + 1;
      n -= half + 1;
    } else {
      n = half;
    }
  }
  return lo;
}

// Per tile of `a` (kTile rows): window of `b` covering the tile's key range,
// staged in LDS when it fits.  mode 0: flags[i] = a[i] not in b (set
// difference); mode 1: scatter a[i] to out at i + #(b < a[i]) (merge).
template <int K, int MODE>
__global__ __launch_bounds__(kBlock) void window_kernel(
    const uint32_t* __restrict__ a, uint64_t sa, uint64_t na, const uint32_t* __restrict__ b,
    uint64_t sb, uint64_t nb, uint8_t* __restrict__ flags, uint32_t* __restrict__ tile_cnt,
    uint32_t* __restrict__ out, uint64_t so) {
  __shared__ uint32_t stage[K][kStageRows];
  __shared__ uint64_t s_lo, s_hi;
  __shared__ uint32_t s_cnt[kWavesPerBlock];
  const uint32_t w = threadIdx.x / kWave, l = lane_id();
  const uint64_t base = uint64_t(blockIdx.x) * kTile;
  const uint64_t last = min<uint64_t>(base + kTile, na) - 1;
  if (threadIdx.x == 0) {
    s_lo = lex_lower_bound<K>(b, sb, 0, nb, load_row<K>(a, sa, base));
    s_hi = lex_upper_bound<K>(b, sb, s_lo, nb, load_row<K>(a, sa, last));
  }
  __syncthreads();
  const uint64_t wlo = s_lo, whi = s_hi;
  const bool staged = whi - wlo <= kStageRows;
  if (staged) {
    for (uint64_t j = threadIdx.x; j < whi - wlo; j += kBlock)
#pragma unroll
      for (int c = 0; c < K; ++c) stage[c][j] = b[uint64_t(c) * sb + wlo + j];
  }
  __syncthreads();
  uint32_t cnt = 0;
#pragma unroll 4
  for (int i = 0; i < kItems; ++i) {
    uint64_t idx = base + tile_pos(w, i, l);
    bool keep = false;
    if (idx < na) {
      RowK<K> key = load_row<K>(a, sa, idx);
      uint64_t pos;
      if (staged) {
        uint64_t lo = 0, n = whi - wlo;
        while (n > 0) {
          uint64_t half = n >> 1, mid = lo + half;
          RowK<K> m;
#pragma unroll
          for (int c = 0; c < K; ++c) m.v[c] = stage[c][mid];
          if (row_less<K>(m, key)) {
            lo = mid + 1;
            n -= half + 1;
          } else {
            n = half;
          }
        }
        pos = wlo + lo;
      } else {
        pos = lex_lower_bound<K>(b, sb, wlo, whi, key);
      }
      if (MODE == 0) {
        bool found = false;
        if (pos < whi) {
          RowK<K> m;
          if (staged) {
#pragma unroll
            for (int c = 0; c < K; ++c) m.v[c] = stage[c][pos - wlo];
          } else {
            m = load_row<K>(b, sb, pos);
          }
          found = row_eq<K>(m, key);
        }
        keep = !found;
        flags[idx] = keep ? 1 : 0;
      } else {
        uint64_t o = idx + pos;
#pragma unroll
        for (int c = 0; c < K; ++c) out[uint64_t(c) * so + o] = key.v[c];
      }
    }
    if (MODE == 0) cnt += __popcll(ballot(keep));
  }
  if (MODE == 0) {
    if (l == 0) s_cnt[w] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) tile_cnt[blockIdx.x] = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
  }
}

// ====================================================== merge path ====
// Merge and (dense) set difference of two sorted, duplicate-free k-column
// row sets A and B by merge-path partitioning (Odeh, Green et al., "Merge
// Path", 2012): the merged order is cut into tiles of kMpTile positions at
// diagonal split points, each tile stages its piece of A and of B in LDS
// with coalesced loads and merges them there, and the output leaves through
// LDS with coalesced stores.  Ties are broken B-first (a row of A comes
// before a row of B only when strictly less), so a row of A that is also in
// B directly follows its equal B row; for the first rows of a tile that B row
// may be the last one before the tile, B[j0-1], which the tile stages too.
#ifndef SRD_MP_ITEMS
#define SRD_MP_ITEMS 8
#endif
constexpr int kMpItems = SRD_MP_ITEMS;
constexpr uint32_t kMpTile = kBlock * kMpItems;
// Staged rows are stored one padding word per 32: thread t works on rows
// near t*kMpItems, a stride that would put a wave's accesses in a few LDS
// banks; the padding spreads them over all banks.
constexpr uint32_t kMpLds = kMpTile + 2 + (kMpTile + 2) / 32 + 1;
__device__ __forceinline__ uint32_t mp_pad(uint32_t r) { return r + (r >> 5); }

// Rows of A among the first d rows of the B-first merged order.
template <int K>
__device__ __forceinline__ uint64_t mp_split(const uint32_t* a, uint64_t sa, uint64_t na,
                                             const uint32_t* b, uint64_t sb, uint64_t nb,
                                             uint64_t d) {
  uint64_t lo = d > nb ? d - nb : 0, hi = min<uint64_t>(d, na);
  while (lo < hi) {
    const uint64_t i = (lo + hi + 1) >> 1;  // A[i-1] and B[d-i] both exist
    if (row_less<K>(load_row<K>(a, sa, i - 1), load_row<K>(b, sb, d - i))) lo = i;
    else hi = i - 1;
  }
  return lo;
}

template <int K>
__global__ __launch_bounds__(kBlock) void mp_partition_kernel(
    const uint32_t* __restrict__ a, uint64_t sa, uint64_t na, const uint32_t* __restrict__ b,
    uint64_t sb, uint64_t nb, uint64_t tiles, uint64_t* __restrict__ split,
    const uint64_t* __restrict__ nab = nullptr) {
  if (nab) na = nab[0], nb = nab[1];  // device-sized: split points past the rows are na
  const uint64_t t = blockIdx.x * uint64_t(kBlock) + threadIdx.x;
  if (t > tiles) return;
  const uint64_t d = min<uint64_t>(t * kMpTile, na + nb);
  split[t] = mp_split<K>(a, sa, na, b, sb, nb, d);
}

// LDS row r of the staged tile: A rows at [0, an), then B[j0-1] at an (when
// j0 > 0), then B rows.
template <int K>
__device__ __forceinline__ RowK<K> lds_row(const uint32_t (*st)[kMpLds], uint32_t r) {
  RowK<K> x;
  const uint32_t q = mp_pad(r);
#pragma unroll
  for (int c = 0; c < K; ++c) x.v[c] = st[c][q];
  return x;
}

// MODE 1: out gets the merged rows at d0 + position.  MODE 0: out gets the
// rows of A not in B, compacted; tile offsets come from a decoupled
// look-back in ticket order, and the last tile stores the total.  MODE 2:
// only count the kept rows of each tile (tile_cnt[tile]); MODE 3: write
// them at the scanned offsets tile_off[tile] (the two-pass variant).
// MODE 4: write the kept rows at the tile's first A row i0 (tiles own
// disjoint A ranges, and a tile keeps at most its A rows, so no offset is
// needed) and their count to tile_cnt[tile]; mp_gather_kernel then packs
// them (the gathered variant: no look-back, no second merge).
template <int K, int MODE>
__global__ __launch_bounds__(kBlock) void mp_tile_kernel(
    const uint32_t* __restrict__ a, uint64_t sa, uint64_t na, const uint32_t* __restrict__ b,
    uint64_t sb, uint64_t nb, const uint64_t* __restrict__ split, uint32_t* __restrict__ out,
    uint64_t so, uint64_t* __restrict__ lookback, uint32_t* __restrict__ ctr, uint32_t epoch,
    uint64_t* __restrict__ total_out, uint32_t* __restrict__ tile_cnt,
    const uint64_t* __restrict__ tile_off, bool ticket, const uint64_t* __restrict__ nab = nullptr) {
  constexpr bool DIFF = MODE != 1;
  __shared__ uint32_t st[K][kMpLds];
  __shared__ uint32_t s_wsum[kWavesPerBlock];
  __shared__ uint64_t s_prefix;
  __shared__ uint32_t s_tile;
  const uint32_t tid = threadIdx.x;
  // look-back order: block id, or a ticket when other runs share the GPU
  // (the reasoning of onesweep_kernel); ctr[0] is zeroed before the launch
  uint32_t tile = blockIdx.x;
  if (MODE == 0 && ticket) {
    if (tid == 0) s_tile = atomicAdd(&ctr[0], 1u);
    __syncthreads();
    tile = s_tile;
  }
  if (nab) {
    // device-sized launch (MODE 1 and 4): tiles past the rows do nothing
    na = nab[0];
    nb = nab[1];
    if (uint64_t(tile) * kMpTile >= na + nb) {
      if ((MODE == 2 || MODE == 4) && tid == 0) tile_cnt[tile] = 0;
      return;
    }
  }
  const uint64_t d0 = uint64_t(tile) * kMpTile;
  const uint64_t d1 = min<uint64_t>(d0 + kMpTile, na + nb);
  const uint64_t i0 = split[tile], i1 = split[tile + 1];
  const uint64_t j0 = d0 - i0, j1 = d1 - i1;
  const uint32_t an = uint32_t(i1 - i0), bn = uint32_t(j1 - j0);
  const uint32_t pre = (DIFF && j0 > 0) ? 1u : 0u;  // B[j0-1] staged at an
  const uint32_t bb = an + pre;                          // first staged B row of the tile
  for (uint32_t r = tid; r < an + pre + bn; r += kBlock) {
#pragma unroll
    for (int c = 0; c < K; ++c) {
      uint32_t v;
      if (r < an) v = a[uint64_t(c) * sa + i0 + r];
      else v = b[uint64_t(c) * sb + j0 - pre + (r - an)];
      st[c][mp_pad(r)] = v;
    }
  }
  __syncthreads();

  // this thread's merged positions [p0, p0 + kMpItems) of the tile
  const uint32_t tn = an + bn;
  const uint32_t p0 = min<uint32_t>(tid * kMpItems, tn);
  uint32_t lo = p0 > bn ? p0 - bn : 0, hi = min<uint32_t>(p0, an);
  while (lo < hi) {
    const uint32_t i = (lo + hi + 1) >> 1;
    if (row_less<K>(lds_row<K>(st, i - 1), lds_row<K>(st, bb + (p0 - i)))) lo = i;
    else hi = i - 1;
  }
  uint32_t ai = lo, bi = p0 - lo;
  RowK<K> rows[kMpItems];
  bool keep[kMpItems];
  uint32_t kept = 0;
#pragma unroll
  for (int it = 0; it < kMpItems; ++it) {
    keep[it] = false;
    if (p0 + uint32_t(it) >= tn) continue;
    const bool take_a =
        bi >= bn || (ai < an && row_less<K>(lds_row<K>(st, ai), lds_row<K>(st, bb + bi)));
    if (take_a) {
      rows[it] = lds_row<K>(st, ai);
      if (DIFF) {
        // in B iff the last B row at or before this position equals it
        const bool has_prev = bi > 0 || pre;
        keep[it] = !(has_prev && row_eq<K>(lds_row<K>(st, bb + bi - 1), rows[it]));
        kept += keep[it] ? 1u : 0u;
      }
      ++ai;
    } else {
      if (MODE == 1) rows[it] = lds_row<K>(st, bb + bi);
      ++bi;
    }
  }
  __syncthreads();  // every thread is done reading the staged input

  if (MODE == 1) {
#pragma unroll
    for (int it = 0; it < kMpItems; ++it)
      if (p0 + uint32_t(it) < tn)
#pragma unroll
        for (int c = 0; c < K; ++c) st[c][mp_pad(p0 + it)] = rows[it].v[c];
    __syncthreads();
    for (uint32_t r = tid; r < tn; r += kBlock)
#pragma unroll
      for (int c = 0; c < K; ++c) out[uint64_t(c) * so + d0 + r] = st[c][mp_pad(r)];
    return;
  }

  uint32_t tile_tot;
  uint32_t off = block_excl_scan<uint32_t>(kept, s_wsum, &tile_tot);
  if (MODE == 2) {
    if (tid == 0) tile_cnt[tile] = tile_tot;
    return;
  }
#pragma unroll
  for (int it = 0; it < kMpItems; ++it)
    if (keep[it]) {
#pragma unroll
      for (int c = 0; c < K; ++c) st[c][mp_pad(off)] = rows[it].v[c];
      ++off;
    }
  if (MODE == 3 || MODE == 4) {
    if (tid == 0) {
      s_prefix = MODE == 3 ? tile_off[tile] : i0;
      if (MODE == 4) tile_cnt[tile] = tile_tot;
    }
  } else if (tid < kWave) {  // wave 0 resolves the tile's output offset
    uint64_t* mine = lookback + tile;
    if (tid == 0) lb_store(mine, lb_pack(epoch, tile == 0 ? kLbPrefix : kLbAgg, tile_tot));
    const uint64_t excl = tile == 0 ? 0 : wave_lookback(lookback, 1, tile, epoch, &ctr[1]);
    if (tid == 0) {
      if (tile != 0) lb_store(mine, lb_pack(epoch, kLbPrefix, excl + tile_tot));
      s_prefix = excl;
      if (d1 == na + nb) *total_out = excl + tile_tot;
    }
  }
  __syncthreads();
  const uint64_t base = s_prefix;
  for (uint32_t r = tid; r < tile_tot; r += kBlock)
#pragma unroll
    for (int c = 0; c < K; ++c) out[uint64_t(c) * so + base + r] = st[c][mp_pad(r)];
}

// Launch the partition and the tiles; returns the number of tiles.
template <int K, int MODE>
static uint64_t merge_path(const Index& a, const Index& b, uint32_t* out, uint64_t so,
                           Workspace& ws, hipStream_t s) {
  const uint64_t tiles = (a.n + b.n + kMpTile - 1) / kMpTile;
  ws.mp_split.reserve(tiles + 1, s);
  mp_partition_kernel<K><<<uint32_t((tiles + 1 + kBlock - 1) / kBlock), kBlock, 0, s>>>(
      a.cols.p, a.stride, a.n, b.cols.p, b.stride, b.n, tiles, ws.mp_split.p);
  SRD_CHECK(hipGetLastError());
  if (MODE == 0) {
    if (tiles > ws.mp_lb.n) {
      ws.mp_lb.alloc(tiles + tiles / 4 + 64, s);
      fill_u64(ws.mp_lb.p, 0, ws.mp_lb.n, s);
      ws.mp_epoch = 0;
    }
    if (++ws.mp_epoch >= 0xFFFF) {
      fill_u64(ws.mp_lb.p, 0, ws.mp_lb.n, s);
      ws.mp_epoch = 1;
    }
    ws.mp_ctr.reserve(2, s);
    ws.mp_total.reserve(1, s);
    fill_u32(ws.mp_ctr.p, 0, 2, s);
  }
  mp_tile_kernel<K, MODE><<<uint32_t(tiles), kBlock, 0, s>>>(
      a.cols.p, a.stride, a.n, b.cols.p, b.stride, b.n, ws.mp_split.p, out, so,
      MODE == 0 ? ws.mp_lb.p : nullptr, MODE == 0 ? ws.mp_ctr.p : nullptr, ws.mp_epoch,
      MODE == 0 ? ws.mp_total.p : nullptr, nullptr, nullptr, ws.ticket_order);
  SRD_CHECK(hipGetLastError());
  return tiles;
}

// Pack the rows each MODE 4 tile left at its first A row into consecutive
// output rows: block t copies tile_cnt[t] rows from tmp row split[t] to out
// row tile_off[t].
template <int K>
__global__ __launch_bounds__(kBlock) void mp_gather_kernel(const uint32_t* __restrict__ tmp,
                                                           uint64_t st,
                                                           const uint64_t* __restrict__ split,
                                                           const uint32_t* __restrict__ tile_cnt,
                                                           const uint64_t* __restrict__ tile_off,
                                                           uint32_t* __restrict__ out, uint64_t so) {
  const uint32_t t = blockIdx.x;
  const uint64_t src = split[t], dst = tile_off[t];
  const uint32_t cnt = tile_cnt[t];
  for (uint32_t r = threadIdx.x; r < cnt; r += kBlock)
#pragma unroll
    for (int c = 0; c < K; ++c) out[uint64_t(c) * so + dst + r] = tmp[uint64_t(c) * st + src + r];
}

// Gathered set difference: one merge pass leaves each tile's kept rows at
// its first A row of a scratch copy (MODE 4), a scan of the per-tile counts
// gives the output offsets, and mp_gather_kernel packs the rows.  Costs
// 2*|D|*w more traffic than the look-back variant but no tile waits on
// another.  The total goes to ws.mp_total like the other variants'.
template <int K>
static void set_difference_gathered(const Index& a, const Index& b, uint32_t* out, uint64_t so,
                                    Workspace& ws, hipStream_t s) {
  const uint64_t tiles = (a.n + b.n + kMpTile - 1) / kMpTile;
  ws.mp_split.reserve(tiles + 1, s);
  mp_partition_kernel<K><<<uint32_t((tiles + 1 + kBlock - 1) / kBlock), kBlock, 0, s>>>(
      a.cols.p, a.stride, a.n, b.cols.p, b.stride, b.n, tiles, ws.mp_split.p);
  SRD_CHECK(hipGetLastError());
  const uint64_t ts = round_stride(a.n);
  DevBuf<uint32_t> tmp(ts * K, s);
  DevBuf<uint32_t> cnt(tiles, s);
  DevBuf<uint64_t> off(tiles + 1, s);
  mp_tile_kernel<K, 4><<<uint32_t(tiles), kBlock, 0, s>>>(
      a.cols.p, a.stride, a.n, b.cols.p, b.stride, b.n, ws.mp_split.p, tmp.p, ts, nullptr,
      nullptr, 0, nullptr, cnt.p, nullptr, false);
  SRD_CHECK(hipGetLastError());
  scan_impl<uint32_t>(cnt.p, off.p, tiles, s);
  mp_gather_kernel<K><<<uint32_t(tiles), kBlock, 0, s>>>(tmp.p, ts, ws.mp_split.p, cnt.p, off.p,
                                                         out, so);
  SRD_CHECK(hipGetLastError());
  ws.mp_total.reserve(1, s);
  ws.mp_ctr.reserve(2, s);
  fill_u32(ws.mp_ctr.p, 0, 2, s);
  SRD_CHECK(hipMemcpyAsync(ws.mp_total.p, off.p + tiles, 8, hipMemcpyDeviceToDevice, s));
}

// Two-pass set difference: count the kept rows per tile, scan, write.  The
// total goes to ws.mp_total like the single-pass variant's.
template <int K>
static void set_difference_two_pass(const Index& a, const Index& b, uint32_t* out, uint64_t so,
                                    Workspace& ws, hipStream_t s) {
  const uint64_t tiles = (a.n + b.n + kMpTile - 1) / kMpTile;
  ws.mp_split.reserve(tiles + 1, s);
  mp_partition_kernel<K><<<uint32_t((tiles + 1 + kBlock - 1) / kBlock), kBlock, 0, s>>>(
      a.cols.p, a.stride, a.n, b.cols.p, b.stride, b.n, tiles, ws.mp_split.p);
  SRD_CHECK(hipGetLastError());
  DevBuf<uint32_t> cnt(tiles, s);
  DevBuf<uint64_t> off(tiles + 1, s);
  mp_tile_kernel<K, 2><<<uint32_t(tiles), kBlock, 0, s>>>(
      a.cols.p, a.stride, a.n, b.cols.p, b.stride, b.n, ws.mp_split.p, out, so, nullptr, nullptr,
      0, nullptr, cnt.p, nullptr, false);
  SRD_CHECK(hipGetLastError());
  scan_impl<uint32_t>(cnt.p, off.p, tiles, s);
  mp_tile_kernel<K, 3><<<uint32_t(tiles), kBlock, 0, s>>>(
      a.cols.p, a.stride, a.n, b.cols.p, b.stride, b.n, ws.mp_split.p, out, so, nullptr, nullptr,
      0, nullptr, nullptr, off.p, false);
  SRD_CHECK(hipGetLastError());
  ws.mp_total.reserve(1, s);
  ws.mp_ctr.reserve(2, s);
  fill_u32(ws.mp_ctr.p, 0, 2, s);
  SRD_CHECK(hipMemcpyAsync(ws.mp_total.p, off.p + tiles, 8, hipMemcpyDeviceToDevice, s));
}

template <int K>
__global__ __launch_bounds__(kBlock) void compact_kernel(const uint32_t* __restrict__ a,
                                                         uint64_t sa, uint64_t na,
                                                         const uint8_t* __restrict__ flags,
                                                         const uint64_t* __restrict__ toff,
                                                         uint32_t* __restrict__ out,
                                                         uint64_t so) {
  __shared__ uint32_t s_cnt[kWavesPerBlock];
  const uint32_t w = threadIdx.x / kWave, l = lane_id();
  const uint64_t base = uint64_t(blockIdx.x) * kTile;
  uint64_t m[kItems];
  uint32_t cnt = 0;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    uint64_t idx = base + tile_pos(w, i, l);
    m[i] = ballot(idx < na && flags[idx]);
    cnt += __popcll(m[i]);
  }
  if (l == 0) s_cnt[w] = cnt;
  __syncthreads();
  uint64_t p = toff[blockIdx.x];
  for (uint32_t j = 0; j < w; ++j) p += s_cnt[j];
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    if ((m[i] >> l) & 1) {
      uint64_t idx = base + tile_pos(w, i, l);
      uint64_t o = p + popc_below(m[i]);
#pragma unroll
      for (int c = 0; c < K; ++c) out[uint64_t(c) * so + o] = a[uint64_t(c) * sa + idx];
    }
    p += __popcll(m[i]);
  }
}

template <int K>
static void set_difference_k(const Index& a, const Index& b, Index& out, Workspace& ws,
                             hipStream_t s) {
  if (b.n <= 8 * a.n) {
    // dense: one merge-path pass over NEW and FULL writes the kept rows
    out = Index();
    out.arity = K;
    out.col_max = a.col_max;
    out.stride = round_stride(a.n);
    out.cols.alloc(out.stride * K, s);
    {
      TimeScope t("setdiff.merge_path", s, double(a.n + b.n) * 4 * K);
      // SRDATALOG_SETDIFF = gather | lookback | twopass (read per call so
      // tests can switch variants within one process).  Default: the
      // gathered form for large inputs; below 2^22 rows the single-pass
      // look-back form, whose two launches cost less than the gathered
      // form's scan, gather and scratch allocation
      const char* v = std::getenv("SRDATALOG_SETDIFF");
      const std::string variant =
          v ? v
            : (std::getenv("SRDATALOG_SETDIFF_TWO_PASS")
                   ? "twopass"
                   : (a.n + b.n < (uint64_t(1) << 22) ? "lookback" : "gather"));
      if (variant == "twopass")
        set_difference_two_pass<K>(a, b, out.cols.p, out.stride, ws, s);
      else if (variant == "lookback")
        merge_path<K, 0>(a, b, out.cols.p, out.stride, ws, s);
      else
        set_difference_gathered<K>(a, b, out.cols.p, out.stride, ws, s);
    }
    uint64_t* h64 = reinterpret_cast<uint64_t*>(pinned_words(ws, 4));
    SRD_CHECK(hipMemcpyAsync(h64, ws.mp_total.p, 8, hipMemcpyDeviceToHost, s));
    SRD_CHECK(hipMemcpyAsync(h64 + 1, ws.mp_ctr.p, 8, hipMemcpyDeviceToHost, s));
    SRD_CHECK(hipStreamSynchronize(s));
    if (uint32_t(h64[1] >> 32) != 0) fail("set_difference: tile look-back did not resolve");
    out.n = h64[0];
    if (out.n == 0) {
      out = Index();
      out.arity = K;
      out.col_max = a.col_max;
      return;
    }
    build_roots(out, ws, s);
    return;
  }
  // sparse (FULL much larger than NEW): each tile of NEW probes only the
  // window of FULL its key range covers
  const uint64_t tiles = (a.n + kTile - 1) / kTile;
  ws.flags.reserve(a.n, s);
  DevBuf<uint32_t> tc(tiles, s);
  DevBuf<uint64_t> to(tiles + 1, s);
  {
    TimeScope t("setdiff.probe", s, double(a.n + b.n) * 4 * K + double(a.n));
    window_kernel<K, 0><<<uint32_t(tiles), kBlock, 0, s>>>(a.cols.p, a.stride, a.n, b.cols.p,
                                                           b.stride, b.n, ws.flags.p, tc.p,
                                                           nullptr, 0);
    SRD_CHECK(hipGetLastError());
  }
  scan_impl<uint32_t>(tc.p, to.p, tiles, s);
  uint64_t* h64 = reinterpret_cast<uint64_t*>(pinned_words(ws, 2));
  SRD_CHECK(hipMemcpyAsync(h64, to.p + tiles, 8, hipMemcpyDeviceToHost, s));
  SRD_CHECK(hipStreamSynchronize(s));
  const uint64_t d = h64[0];
  out = Index();
  out.arity = K;
  out.n = d;
  out.stride = round_stride(d);
  out.col_max = a.col_max;
  if (d == 0) return;
  out.cols.alloc(out.stride * K, s);
  {
    TimeScope t("setdiff.compact", s, double(a.n) * (1 + 4 * K) + double(d) * 4 * K);
    compact_kernel<K><<<uint32_t(tiles), kBlock, 0, s>>>(a.cols.p, a.stride, a.n, ws.flags.p,
                                                         to.p, out.cols.p, out.stride);
    SRD_CHECK(hipGetLastError());
  }
  build_roots(out, ws, s);
}

void copy_index(const Index& a, Index& out, hipStream_t) {
  if (&a == &out) return;
  out.arity = a.arity;
  out.n = a.n;
  out.stride = a.stride;
  out.r = a.r;
  out.col_max = a.col_max;
  out.cols = a.cols;
  out.root_vals = a.root_vals;
  out.root_offs = a.root_offs;
}

void set_difference(const Index& a, const Index& b, Index& out, Workspace& ws, hipStream_t s) {
  if (a.arity != b.arity && b.n) fail("set_difference: arity mismatch");
  if (a.n == 0) {
    out = Index();
    out.arity = a.arity;
    out.col_max = a.col_max;
    return;
  }
  if (b.n == 0) {
    copy_index(a, out, s);
    return;
  }
  switch (a.arity) {
    case 1: set_difference_k<1>(a, b, out, ws, s); break;
    case 2: set_difference_k<2>(a, b, out, ws, s); break;
    case 3: set_difference_k<3>(a, b, out, ws, s); break;
    case 4: set_difference_k<4>(a, b, out, ws, s); break;
    default: fail("set_difference: arity > 4 not supported");
  }
}

template <int K>
static void merge_k(const Index& a, const Index& b, Index& out, Workspace& ws, hipStream_t s) {
  out = Index();
  out.arity = K;
  out.n = a.n + b.n;
  out.stride = round_stride(out.n);
  out.cols.alloc(out.stride * K, s);
  out.col_max.assign(K, 0);
  for (int c = 0; c < K; ++c)
    out.col_max[c] = std::max(a.col_max.size() == K ? a.col_max[c] : 0u,
                              b.col_max.size() == K ? b.col_max[c] : 0u);
  {
    TimeScope t("merge.merge_path", s, double(out.n) * 8 * K);
    merge_path<K, 1>(a, b, out.cols.p, out.stride, ws, s);
  }
  build_roots(out, ws, s);
}

void merge_disjoint(const Index& a, const Index& b, Index& out, Workspace& ws, hipStream_t s) {
  if (a.n == 0) {
    copy_index(b, out, s);
    return;
  }
  if (b.n == 0) {
    copy_index(a, out, s);
    return;
  }
  if (a.arity != b.arity) fail("merge_disjoint: arity mismatch");
  if (a.n + b.n >= (uint64_t(1) << 32)) fail("merge_disjoint: more than 2^32-1 rows");
  switch (a.arity) {
    case 1: merge_k<1>(a, b, out, ws, s); break;
    case 2: merge_k<2>(a, b, out, ws, s); break;
    case 3: merge_k<3>(a, b, out, ws, s); break;
    case 4: merge_k<4>(a, b, out, ws, s); break;
    default: fail("merge_disjoint: arity > 4 not supported");
  }
}


// ============================================== device-sized primitives ====
// For device-resident iterations: row counts are read from device words,
// launch shapes come from host capacities, nothing is read back and
// nothing is allocated; blocks past the counted rows exit at once.

uint64_t scan_dev_words(uint64_t cap) { return std::max<uint64_t>(scan_noalloc_words(cap), 1); }

template <class T>
static void scan_dev(const T* in, uint64_t* out, uint64_t cap, const uint64_t* n_dev, uint64_t* tmp,
                     hipStream_t s) {
  if (cap <= kTile) {
    scan_tile_kernel<T><<<1, kBlock, 0, s>>>(in, cap, nullptr, out, n_dev);
    SRD_CHECK(hipGetLastError());
    return;
  }
  const uint64_t tiles = (cap + kTile - 1) / kTile;
  uint64_t* partial = tmp;
  uint64_t* carry = tmp + tiles;
  scan_reduce_kernel<T><<<uint32_t(tiles), kBlock, 0, s>>>(in, cap, partial, n_dev);
  SRD_CHECK(hipGetLastError());
  // every tile past the counted rows contributes 0: the partials scan at
  // the capacity's tile count
  scan_noalloc<uint64_t>(partial, carry, tiles, tmp + 2 * tiles + 1, s);
  scan_tile_kernel<T><<<uint32_t(tiles), kBlock, 0, s>>>(in, cap, carry, out, n_dev);
  SRD_CHECK(hipGetLastError());
}

void exclusive_scan_u32_dev(const uint32_t* in, uint64_t* out, uint64_t cap, const uint64_t* n_dev,
                            uint64_t* tmp, hipStream_t s) {
  scan_dev<uint32_t>(in, out, cap, n_dev, tmp, s);
}

void exclusive_scan_u64_dev(const uint64_t* in, uint64_t* out, uint64_t cap, const uint64_t* n_dev,
                            uint64_t* tmp, hipStream_t s) {
  scan_dev<uint64_t>(in, out, cap, n_dev, tmp, s);
}

__global__ void copy_word_kernel(const uint64_t* __restrict__ src, uint64_t* __restrict__ dst) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = *src;
}

uint64_t mp_tiles_for(uint64_t rows) { return (rows + kMpTile - 1) / kMpTile; }

template <int K>
static void set_difference_dev_k(const uint32_t* a, uint64_t sa, const uint32_t* b, uint64_t sb,
                                 uint64_t cap_ab, const uint64_t* nab, uint32_t* out, uint64_t so,
                                 const SetDiffDevScratch& x, uint64_t* total_out, hipStream_t s) {
  const uint64_t tiles = std::max<uint64_t>(mp_tiles_for(cap_ab), 1);
  mp_partition_kernel<K><<<uint32_t((tiles + 1 + kBlock - 1) / kBlock), kBlock, 0, s>>>(
      a, sa, 0, b, sb, 0, tiles, x.split, nab);
  SRD_CHECK(hipGetLastError());
  mp_tile_kernel<K, 4><<<uint32_t(tiles), kBlock, 0, s>>>(a, sa, 0, b, sb, 0, x.split, x.tmp,
                                                          x.tmp_stride, nullptr, nullptr, 0, nullptr,
                                                          x.cnt, nullptr, false, nab);
  SRD_CHECK(hipGetLastError());
  scan_dev<uint32_t>(x.cnt, x.off, tiles, nullptr, x.scan_tmp, s);
  mp_gather_kernel<K><<<uint32_t(tiles), kBlock, 0, s>>>(x.tmp, x.tmp_stride, x.split, x.cnt, x.off,
                                                         out, so);
  SRD_CHECK(hipGetLastError());
  copy_word_kernel<<<1, 64, 0, s>>>(x.off + tiles, total_out);
  SRD_CHECK(hipGetLastError());
}

void set_difference_dev(uint32_t k, const uint32_t* a, uint64_t sa, const uint32_t* b, uint64_t sb,
                        uint64_t cap_ab, const uint64_t* nab, uint32_t* out, uint64_t so,
                        const SetDiffDevScratch& x, uint64_t* total_out, hipStream_t s) {
  switch (k) {
    case 1: set_difference_dev_k<1>(a, sa, b, sb, cap_ab, nab, out, so, x, total_out, s); break;
    case 2: set_difference_dev_k<2>(a, sa, b, sb, cap_ab, nab, out, so, x, total_out, s); break;
    case 3: set_difference_dev_k<3>(a, sa, b, sb, cap_ab, nab, out, so, x, total_out, s); break;
    case 4: set_difference_dev_k<4>(a, sa, b, sb, cap_ab, nab, out, so, x, total_out, s); break;
    default: fail("set_difference_dev: arity must be in [1, 4]");
  }
}

template <int K>
static void merge_dev_k(const uint32_t* a, uint64_t sa, const uint32_t* b, uint64_t sb,
                        uint64_t cap_ab, const uint64_t* nab, uint32_t* out, uint64_t so,
                        uint64_t* split, hipStream_t s) {
  const uint64_t tiles = std::max<uint64_t>(mp_tiles_for(cap_ab), 1);
  mp_partition_kernel<K><<<uint32_t((tiles + 1 + kBlock - 1) / kBlock), kBlock, 0, s>>>(
      a, sa, 0, b, sb, 0, tiles, split, nab);
  SRD_CHECK(hipGetLastError());
  mp_tile_kernel<K, 1><<<uint32_t(tiles), kBlock, 0, s>>>(a, sa, 0, b, sb, 0, split, out, so,
                                                          nullptr, nullptr, 0, nullptr, nullptr,
                                                          nullptr, false, nab);
  SRD_CHECK(hipGetLastError());
}

void merge_dev(uint32_t k, const uint32_t* a, uint64_t sa, const uint32_t* b, uint64_t sb,
               uint64_t cap_ab, const uint64_t* nab, uint32_t* out, uint64_t so, uint64_t* split,
               hipStream_t s) {
  switch (k) {
    case 1: merge_dev_k<1>(a, sa, b, sb, cap_ab, nab, out, so, split, s); break;
    case 2: merge_dev_k<2>(a, sa, b, sb, cap_ab, nab, out, so, split, s); break;
    case 3: merge_dev_k<3>(a, sa, b, sb, cap_ab, nab, out, so, split, s); break;
    case 4: merge_dev_k<4>(a, sa, b, sb, cap_ab, nab, out, so, split, s); break;
    default: fail("merge_dev: arity must be in [1, 4]");
  }
}

__global__ void roots_dev_finish_kernel(const uint64_t* __restrict__ n_dev,
                                        const uint64_t* __restrict__ total,
                                        uint32_t* __restrict__ root_offs, uint64_t* __restrict__ r_out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint64_t r = *total;
  root_offs[r] = uint32_t(*n_dev);
  *r_out = r;
}

void build_roots_dev(const uint32_t* c0, uint64_t cap, const uint64_t* n_dev, uint32_t* pos,
                     uint32_t* tile_r, uint64_t* tile_off, uint64_t* scan_tmp, uint32_t* root_vals,
                     uint32_t* root_offs, uint64_t* r_out, hipStream_t s) {
  const uint64_t tiles = std::max<uint64_t>((cap + kTile - 1) / kTile, 1);
  root_mark_kernel<<<uint32_t(tiles), kBlock, 0, s>>>(c0, cap, pos, tile_r, n_dev);
  SRD_CHECK(hipGetLastError());
  scan_dev<uint32_t>(tile_r, tile_off, tiles, nullptr, scan_tmp, s);
  root_gather_kernel<<<uint32_t(tiles), kBlock, 0, s>>>(c0, pos, tile_r, tile_off, root_vals,
                                                        root_offs);
  SRD_CHECK(hipGetLastError());
  roots_dev_finish_kernel<<<1, 64, 0, s>>>(n_dev, tile_off + tiles, root_offs, r_out);
  SRD_CHECK(hipGetLastError());
}

uint64_t roots_dev_tiles(uint64_t cap) { return std::max<uint64_t>((cap + kTile - 1) / kTile, 1); }


// ====================================== device-resident iteration steps ====
// One-thread kernels running the bookkeeping steps of srd_devloop.h.
__global__ void dl_iter_begin_kernel(DevCtl* c, uint32_t nrels) { dl_iter_begin_step(c, nrels); }
__global__ void dl_bind_kernel(DevCtl* c, SrdDyn* d, DlBind b) { dl_bind_step(c, d, b); }
__global__ void dl_plan_kernel(DevCtl* c, SrdDyn* d, const uint64_t* ws, DlOut o) {
  dl_plan_step(c, d, ws, o);
}
__global__ void dl_out_kernel(DevCtl* c, SrdDyn* d, const uint64_t* off, DlOut o) {
  dl_out_step(c, d, off, o);
}
__global__ void dl_check_kernel(DevCtl* c, uint32_t nrels, uint64_t cap_full) {
  dl_check_step(c, nrels, cap_full);
}
__global__ void dl_sd_prep_kernel(DevCtl* c, uint32_t r) { dl_sd_prep_step(c, r); }
__global__ void dl_sd_fin_kernel(DevCtl* c, uint32_t r) { dl_sd_fin_step(c, r); }
__global__ void dl_mg_prep_kernel(DevCtl* c, uint32_t r) { dl_mg_prep_step(c, r); }
__global__ void dl_mg_fin_kernel(DevCtl* c, uint32_t r) { dl_mg_fin_step(c, r); }
__global__ void dl_iter_end_kernel(DevCtl* c, DlEnd e) { dl_iter_end_step(c, e); }

__global__ __launch_bounds__(kBlock) void dl_inst_kernel(const DevCtl* c, uint32_t r,
                                                         const uint32_t* __restrict__ col0,
                                                         const uint32_t* __restrict__ bounds,
                                                         uint32_t n_inst, uint64_t* __restrict__ out) {
  if (c->skip) return;
  const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
  if (g >= n_inst) return;
  const uint32_t n = uint32_t(c->delta_n[r]);
  const uint32_t lo = n ? lower_bound(col0, 0, n, bounds[g]) : 0;
  const uint32_t hi = n ? lower_bound(col0, lo, n, bounds[g + 1]) : 0;
  out[c->iter * n_inst + g] = hi - lo;
}

#define SRD_DL_LAUNCH(kern, ...)                  \
  do {                                            \
    kern<<<1, 64, 0, s>>>(__VA_ARGS__);           \
    SRD_CHECK(hipGetLastError());                 \
  } while (0)

void dl_iter_begin(DevCtl* c, uint32_t nrels, hipStream_t s) { SRD_DL_LAUNCH(dl_iter_begin_kernel, c, nrels); }
void dl_bind(DevCtl* c, SrdDyn* d, const DlBind& b, hipStream_t s) { SRD_DL_LAUNCH(dl_bind_kernel, c, d, b); }
void dl_plan(DevCtl* c, SrdDyn* d, const uint64_t* work_start, const DlOut& o, hipStream_t s) {
  SRD_DL_LAUNCH(dl_plan_kernel, c, d, work_start, o);
}
void dl_out(DevCtl* c, SrdDyn* d, const uint64_t* offsets, const DlOut& o, hipStream_t s) {
  SRD_DL_LAUNCH(dl_out_kernel, c, d, offsets, o);
}
void dl_check(DevCtl* c, uint32_t nrels, uint64_t cap_full, hipStream_t s) {
  SRD_DL_LAUNCH(dl_check_kernel, c, nrels, cap_full);
}
void dl_sd_prep(DevCtl* c, uint32_t r, hipStream_t s) { SRD_DL_LAUNCH(dl_sd_prep_kernel, c, r); }
void dl_sd_fin(DevCtl* c, uint32_t r, hipStream_t s) { SRD_DL_LAUNCH(dl_sd_fin_kernel, c, r); }
void dl_mg_prep(DevCtl* c, uint32_t r, hipStream_t s) { SRD_DL_LAUNCH(dl_mg_prep_kernel, c, r); }
void dl_mg_fin(DevCtl* c, uint32_t r, hipStream_t s) { SRD_DL_LAUNCH(dl_mg_fin_kernel, c, r); }
void dl_iter_end(DevCtl* c, const DlEnd& e, hipStream_t s) { SRD_DL_LAUNCH(dl_iter_end_kernel, c, e); }
void dl_inst(const DevCtl* c, uint32_t r, const uint32_t* col0, const uint32_t* bounds,
             uint32_t n_inst, uint64_t* out, hipStream_t s) {
  if (n_inst == 0) return;
  dl_inst_kernel<<<(n_inst + kBlock - 1) / kBlock, kBlock, 0, s>>>(c, r, col0, bounds, n_inst, out);
  SRD_CHECK(hipGetLastError());
}

}  // namespace srd
