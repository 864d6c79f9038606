// CPU implementations of the runtime primitives declared in srd_runtime.h,
// used only to exercise the engine's host logic on a machine without a GPU.
// Semantics are the definitions (lexicographic sort + dedup, set
// difference, disjoint union), independent of the device kernels.
#include <algorithm>
#include <cstdio>
#include <stdexcept>
#include <vector>

#include "srd_runtime.h"

namespace srd {

void fail(const std::string& msg) { throw std::runtime_error(msg); }
void check(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(what);
}

static PoolStats g_stats;
void pool_init(int) {}
void* pool_alloc(uint64_t bytes, hipStream_t) {
  g_stats.in_use += bytes;
  g_stats.peak = std::max(g_stats.peak, g_stats.in_use);
  return std::malloc(bytes ? bytes : 1);
}
void pool_free(void* p, uint64_t bytes, hipStream_t) {
  g_stats.in_use -= bytes;
  std::free(p);
}
void pool_retire_stream(hipStream_t) {}
PoolStats pool_stats() { return g_stats; }
hipStream_t stream_acquire() { return (void*)1; }
void stream_release(hipStream_t) {}
void* pinned_acquire(uint64_t bytes, uint64_t* got) {
  *got = bytes;
  return std::malloc(bytes ? bytes : 1);
}
void pinned_release(void* p, uint64_t) { std::free(p); }
static int g_running = 0;
bool run_enter() { return g_running++ == 0; }
void run_exit() { --g_running; }
void pool_set_budget(uint64_t b) { g_stats.budget = b; }
void pool_reset_peak() { g_stats.peak = g_stats.in_use; }

Workspace::~Workspace() { std::free(pinned); }
uint32_t* pinned_words(Workspace& ws, uint64_t words) {
  if (words > ws.pinned_words) {
    std::free(ws.pinned);
    ws.pinned = static_cast<uint32_t*>(std::malloc(words * 4));
    ws.pinned_words = words;
  }
  return ws.pinned;
}

using Row = std::vector<uint32_t>;

static std::vector<Row> rows_of(const Index& x) {
  std::vector<Row> out(x.n, Row(x.arity));
  for (uint64_t i = 0; i < x.n; ++i)
    for (uint32_t c = 0; c < x.arity; ++c) out[i][c] = x.col(c)[i];
  return out;
}

static void fill_index(Index& out, uint32_t k, const std::vector<Row>& rows, hipStream_t s) {
  out = Index();
  out.arity = k;
  out.n = rows.size();
  out.stride = round_stride(out.n);
  out.col_max.assign(k, 0);
  if (!out.n) return;
  out.cols.alloc(out.stride * k, s);
  for (uint64_t i = 0; i < out.n; ++i)
    for (uint32_t c = 0; c < k; ++c) {
      out.col(c)[i] = rows[i][c];
      out.col_max[c] = std::max(out.col_max[c], rows[i][c]);
    }
  Workspace ws;
  build_roots(out, ws, s);
}

void build_roots(Index& x, Workspace&, hipStream_t s) {
  x.root_vals.release();
  x.root_offs.release();
  x.r = 0;
  if (!x.n) return;
  std::vector<uint32_t> v, o;
  for (uint64_t i = 0; i < x.n; ++i)
    if (i == 0 || x.col(0)[i] != x.col(0)[i - 1]) v.push_back(x.col(0)[i]), o.push_back(uint32_t(i));
  o.push_back(uint32_t(x.n));
  x.r = uint32_t(v.size());
  x.root_vals.alloc(v.size(), s);
  x.root_offs.alloc(o.size(), s);
  std::copy(v.begin(), v.end(), x.root_vals.p);
  std::copy(o.begin(), o.end(), x.root_offs.p);
}

void build_index(const uint32_t* const* in, uint32_t k, uint64_t n, Index& out, Workspace&,
                 hipStream_t s, const std::vector<uint32_t>*, const DevBuf<uint32_t>*, uint64_t,
                 bool presorted) {
  std::vector<Row> rows(n, Row(k));
  for (uint64_t i = 0; i < n; ++i)
    for (uint32_t c = 0; c < k; ++c) rows[i][c] = in[c][i];
  if (presorted)  // the caller's promise: strictly increasing rows
    for (uint64_t i = 1; i < n; ++i)
      if (!(rows[i - 1] < rows[i])) throw std::runtime_error("build_index: presorted rows out of order");
  std::sort(rows.begin(), rows.end());
  rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
  fill_index(out, k, rows, s);
}

void build_index_from(const Index& src, const std::vector<uint32_t>& perm, Index& out,
                      Workspace& ws, hipStream_t s) {
  std::vector<const uint32_t*> in(src.arity);
  for (uint32_t c = 0; c < src.arity; ++c) in[c] = src.n ? src.col(perm[c]) : nullptr;
  build_index(in.data(), src.arity, src.n, out, ws, s, nullptr, nullptr, 0, false);
}

void set_difference(const Index& a, const Index& b, Index& out, Workspace&, hipStream_t s) {
  auto ra = rows_of(a), rb = rows_of(b);
  std::vector<Row> d;
  std::set_difference(ra.begin(), ra.end(), rb.begin(), rb.end(), std::back_inserter(d));
  fill_index(out, a.arity, d, s);
}

void merge_disjoint(const Index& a, const Index& b, Index& out, Workspace&, hipStream_t s) {
  auto ra = rows_of(a), rb = rows_of(b);
  std::vector<Row> m;
  std::merge(ra.begin(), ra.end(), rb.begin(), rb.end(), std::back_inserter(m));
  fill_index(out, a.arity ? a.arity : b.arity, m, s);
}

void copy_index(const Index& a, Index& out, hipStream_t) { out = a; }

void exclusive_scan_u32(const uint32_t* in, uint64_t* out, uint64_t n, Workspace&, hipStream_t) {
  uint64_t t = 0;
  for (uint64_t i = 0; i < n; ++i) out[i] = t, t += in[i];
  out[n] = t;
}
void exclusive_scan_u64(const uint64_t* in, uint64_t* out, uint64_t n, Workspace&, hipStream_t) {
  uint64_t t = 0;
  for (uint64_t i = 0; i < n; ++i) out[i] = t, t += in[i];
  out[n] = t;
}
void fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t) { std::fill(p, p + n, v); }
void fill_u64(uint64_t* p, uint64_t v, uint64_t n, hipStream_t) { std::fill(p, p + n, v); }

void export_rows(const Index& x, uint32_t* host_out, const uint32_t* bounds, uint32_t n_bounds,
                 Workspace&, hipStream_t) {
  for (uint64_t i = 0; i < x.n; ++i) {
    uint32_t base = 0;
    if (n_bounds) {
      const uint32_t v = x.cols.p[i];
      const uint32_t* g = std::upper_bound(bounds, bounds + n_bounds, v);
      base = g == bounds ? 0u : *(g - 1);
    }
    for (uint32_t c = 0; c < x.arity; ++c) host_out[i * x.arity + c] = x.cols.p[c * x.stride + i] - base;
  }
}

void slice_bounds(const uint64_t* pref, uint64_t n, uint64_t w, uint64_t S, uint64_t* out,
                  hipStream_t) {
  const uint64_t q = w / S, rem = w % S;
  for (uint64_t s = 0; s <= S; ++s) {
    const uint64_t target = s * q + std::min(s, rem);
    out[s] = uint64_t(std::lower_bound(pref, pref + n + 1, target) - pref);
  }
}

void build_heavy_bitmaps(const Index& x, uint32_t min_len, uint64_t max_bytes, Workspace&,
                         hipStream_t s) {
  if (!x.heavy) x.heavy = std::make_shared<HeavyBits>();
  HeavyBits& h = *x.heavy;
  h.n = 0;
  h.built_for = x.cols.p;
  h.built_n = x.n;
  if (x.arity != 2 || x.r == 0 || x.col_max.size() < 2 || max_bytes == 0) return;
  const uint32_t words = (x.col_max[1] >> 5) + 1;
  for (;;) {
    std::vector<uint32_t> st, en;
    for (uint32_t g = 0; g < x.r; ++g)
      if (x.root_offs.p[g + 1] - x.root_offs.p[g] >= min_len)
        st.push_back(x.root_offs.p[g]), en.push_back(x.root_offs.p[g + 1]);
    if (st.empty()) return;
    if (uint64_t(st.size()) * words * 4 <= max_bytes) {
      const uint32_t n = uint32_t(st.size());
      h.starts.alloc(n, s);
      h.bits.alloc(uint64_t(n) * words, s);
      std::fill(h.bits.p, h.bits.p + uint64_t(n) * words, 0u);
      for (uint32_t i = 0; i < n; ++i) {
        h.starts.p[i] = st[i];
        for (uint32_t row = st[i]; row < en[i]; ++row) {
          const uint32_t v = x.col(1)[row];
          h.bits.p[uint64_t(i) * words + (v >> 5)] |= 1u << (v & 31u);
        }
      }
      h.n = n;
      h.words = words;
      h.min_len = min_len;
      return;
    }
    if (min_len >= (1u << 30)) return;
    min_len *= 2;
  }
}

void build_root_pos(const Index& x, uint64_t max_entries, hipStream_t s) {
  if (!x.rootpos) x.rootpos = std::make_shared<RootPos>();
  RootPos& m = *x.rootpos;
  m.ok = false;
  m.built_for = x.root_vals.p;
  m.built_r = x.r;
  if (x.r == 0 || x.col_max.empty()) return;
  const uint64_t entries = uint64_t(x.col_max[0]) + 1;
  if (entries > max_entries) return;
  for (uint32_t g = 0; g < x.r; ++g)
    if (x.root_vals.p[g] > x.col_max[0]) throw std::runtime_error("build_root_pos: bound below a root value");
  m.pos.reserve(entries, s);
  std::fill(m.pos.p, m.pos.p + entries, ~0u);
  for (uint32_t g = 0; g < x.r; ++g) m.pos.p[x.root_vals.p[g]] = g;
  m.max = uint32_t(entries - 1);
  m.ok = true;
}

void instance_counts(const uint32_t* col0, uint64_t n, const uint32_t* bounds, uint32_t n_inst,
                     uint64_t* out, hipStream_t) {
  for (uint32_t g = 0; g < n_inst; ++g) {
    const uint32_t* lo = col0 ? std::lower_bound(col0, col0 + n, bounds[g]) : nullptr;
    const uint32_t* hi = col0 ? std::lower_bound(lo, col0 + n, bounds[g + 1]) : nullptr;
    out[g] = uint64_t(hi - lo);
  }
}

void outside_counts(const uint32_t* col0, uint64_t n, const uint32_t* bounds, uint32_t n_inst,
                    uint64_t* out, hipStream_t) {
  out[0] = col0 ? uint64_t(std::lower_bound(col0, col0 + n, bounds[0]) - col0) : 0;
  out[1] = col0 ? uint64_t(col0 + n - std::lower_bound(col0, col0 + n, bounds[n_inst])) : 0;
}

void gather_segments(const uint64_t* recs, const uint32_t* rec_ctr, uint64_t rec_cap,
                     const uint32_t* stage, uint64_t stage_stride, const uint64_t* offsets,
                     const GatherDest* dests, uint32_t n_dests, uint64_t, hipStream_t) {
  uint64_t n = std::min<uint64_t>(*rec_ctr, rec_cap);
  for (uint64_t r = 0; r < n; ++r) {
    uint64_t slice = recs[3 * r], src = recs[3 * r + 1], lc = recs[3 * r + 2];
    uint64_t cnt = lc & 0xFFFFFFFFu, local = lc >> 32;
    for (uint32_t d = 0; d < n_dests; ++d)
      for (uint32_t c = 0; c < dests[d].k; ++c)
        for (uint64_t j = 0; j < cnt; ++j)
          dests[d].cols[c * dests[d].stride + dests[d].base + offsets[slice] + local + j] =
              stage[(uint64_t(d) * 8 + c) * stage_stride + src + j];
  }
}

void timing_enable(bool) {}
bool timing_enabled() { return false; }
void timing_begin(const char*, hipStream_t) {}
void timing_end(hipStream_t, double) {}
std::vector<KernelTime> timing_collect() { return {}; }
void timing_reset() {}

}  // namespace srd
