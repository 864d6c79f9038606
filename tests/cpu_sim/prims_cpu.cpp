// CPU implementations of the runtime primitives declared in srd_runtime.h,
// used only to exercise the engine's host logic on a machine without a GPU.
// Semantics are the definitions (lexicographic sort + dedup, set
// difference, disjoint union), independent of the device kernels.
#include <algorithm>
#include <cstdio>
#include <stdexcept>
#include <vector>

#include "srd_devloop.h"
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
std::recursive_mutex& capture_mutex() {
  static std::recursive_mutex m;
  return m;
}
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
uint64_t dense_extract(uint32_t* bits, uint32_t* sum, uint64_t words, uint32_t k, uint32_t shift,
                       DevBuf<uint32_t>& out, uint64_t& stride, Workspace&, hipStream_t s) {
  std::fill(sum, sum + dense_sum_words(words), 0u);
  if (words == 0 || words % kDenseBlockWords) fail("dense_extract: bitmap size not whole blocks");
  if (k == 0 || k > 4 || shift == 0 || uint64_t(shift) * k > 63)
    fail("dense_extract: key layout out of range");
  uint64_t n = 0;
  for (uint64_t w = 0; w < words; ++w) n += uint64_t(__builtin_popcount(bits[w]));
  stride = round_stride(std::max<uint64_t>(n, 1));
  out.alloc(stride * k, s);
  const uint64_t mask = (uint64_t(1) << shift) - 1;
  uint64_t pos = 0;
  for (uint64_t w = 0; w < words; ++w) {
    uint32_t x = bits[w];
    bits[w] = 0;
    while (x) {
      const uint64_t key = (w << 5) | uint64_t(__builtin_ctz(x));
      x &= x - 1;
      for (uint32_t c = 0; c < k; ++c)
        out.p[uint64_t(c) * stride + pos] = uint32_t((key >> (shift * (k - 1 - c))) & mask);
      ++pos;
    }
  }
  return n;
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


// ---- device-sized primitives (counts read from "device" words, which are
// host memory here)
uint64_t scan_dev_words(uint64_t cap) { return cap + 1; }
uint64_t mp_tiles_for(uint64_t rows) { return (rows + 4095) / 4096; }
uint64_t roots_dev_tiles(uint64_t cap) { return std::max<uint64_t>((cap + 4095) / 4096, 1); }
uint64_t devsort_scratch_bytes(uint64_t cap) { return 64 + cap; }
void devsort_clear_lookback(void*, uint64_t, hipStream_t) {}

void exclusive_scan_u32_dev(const uint32_t* in, uint64_t* out, uint64_t cap, const uint64_t* n_dev,
                            uint64_t*, hipStream_t) {
  const uint64_t n = n_dev ? std::min(cap, *n_dev) : cap;
  uint64_t t = 0;
  for (uint64_t i = 0; i < n; ++i) out[i] = t, t += in[i];
  out[n] = t;
}
void exclusive_scan_u64_dev(const uint64_t* in, uint64_t* out, uint64_t cap, const uint64_t* n_dev,
                            uint64_t*, hipStream_t) {
  const uint64_t n = n_dev ? std::min(cap, *n_dev) : cap;
  uint64_t t = 0;
  for (uint64_t i = 0; i < n; ++i) out[i] = t, t += in[i];
  out[n] = t;
}

static std::vector<Row> rows_cols(const uint32_t* const* in, uint32_t k, uint64_t n) {
  std::vector<Row> rows(n, Row(k));
  for (uint64_t i = 0; i < n; ++i)
    for (uint32_t c = 0; c < k; ++c) rows[i][c] = in[c][i];
  return rows;
}
static std::vector<Row> rows_strided(const uint32_t* a, uint64_t sa, uint32_t k, uint64_t n) {
  std::vector<Row> rows(n, Row(k));
  for (uint64_t i = 0; i < n; ++i)
    for (uint32_t c = 0; c < k; ++c) rows[i][c] = a[c * sa + i];
  return rows;
}
static void put_rows(const std::vector<Row>& rows, uint32_t k, uint32_t* out, uint64_t so) {
  for (uint64_t i = 0; i < rows.size(); ++i)
    for (uint32_t c = 0; c < k; ++c) out[c * so + i] = rows[i][c];
}

void build_index_dev(const uint32_t* const* in, uint32_t k, uint64_t cap, const uint64_t* n_dev,
                     uint64_t n_host, uint32_t, void*, uint32_t* out_cols, uint64_t out_stride,
                     uint32_t* root_vals, uint32_t* root_offs, uint64_t* counts_out, bool,
                     hipStream_t, bool) {
  uint64_t n = n_dev ? *n_dev : n_host;
  uint64_t err = 0;
  if (n > cap) n = cap, err = 2;
  auto rows = rows_cols(in, k, n);
  std::sort(rows.begin(), rows.end());
  rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
  put_rows(rows, k, out_cols, out_stride);
  uint64_t r = 0;
  for (uint64_t i = 0; i < rows.size(); ++i)
    if (i == 0 || rows[i][0] != rows[i - 1][0]) root_vals[r] = rows[i][0], root_offs[r] = uint32_t(i), ++r;
  root_offs[r] = uint32_t(rows.size());
  if (counts_out) counts_out[0] = rows.size(), counts_out[1] = r, counts_out[2] = err;
}

void set_difference_dev(uint32_t k, const uint32_t* a, uint64_t sa, const uint32_t* b, uint64_t sb,
                        uint64_t cap_ab, const uint64_t* nab, uint32_t* out, uint64_t so,
                        const SetDiffDevScratch&, uint64_t* total_out, hipStream_t) {
  if (nab[0] + nab[1] > cap_ab) fail("set_difference_dev: rows above the capacity");
  auto ra = rows_strided(a, sa, k, nab[0]), rb = rows_strided(b, sb, k, nab[1]);
  std::vector<Row> d;
  std::set_difference(ra.begin(), ra.end(), rb.begin(), rb.end(), std::back_inserter(d));
  put_rows(d, k, out, so);
  *total_out = d.size();
}

void merge_dev(uint32_t k, const uint32_t* a, uint64_t sa, const uint32_t* b, uint64_t sb,
               uint64_t cap_ab, const uint64_t* nab, uint32_t* out, uint64_t so, uint64_t*,
               hipStream_t) {
  if (nab[0] + nab[1] > cap_ab) fail("merge_dev: rows above the capacity");
  auto ra = rows_strided(a, sa, k, nab[0]), rb = rows_strided(b, sb, k, nab[1]);
  std::vector<Row> m;
  std::merge(ra.begin(), ra.end(), rb.begin(), rb.end(), std::back_inserter(m));
  for (uint64_t i = 1; i < m.size(); ++i)
    if (!(m[i - 1] < m[i])) fail("merge_dev: inputs not disjoint");
  put_rows(m, k, out, so);
}

void build_roots_dev(const uint32_t* c0, uint64_t cap, const uint64_t* n_dev, uint32_t*, uint32_t*,
                     uint64_t*, uint64_t*, uint32_t* root_vals, uint32_t* root_offs, uint64_t* r_out,
                     hipStream_t) {
  const uint64_t n = std::min(cap, *n_dev);
  uint64_t r = 0;
  for (uint64_t i = 0; i < n; ++i)
    if (i == 0 || c0[i] != c0[i - 1]) root_vals[r] = c0[i], root_offs[r] = uint32_t(i), ++r;
  root_offs[r] = uint32_t(n);
  *r_out = r;
}

// ---- device-resident iteration steps, run directly
void dl_iter_begin(DevCtl* c, uint32_t nrels, hipStream_t) { dl_iter_begin_step(c, nrels); }
void dl_bind(DevCtl* c, SrdDyn* d, const DlBind& b, hipStream_t) { dl_bind_step(c, d, b); }
void dl_plan(DevCtl* c, SrdDyn* d, const uint64_t* ws, const DlOut& o, hipStream_t) {
  dl_plan_step(c, d, ws, o);
}
void dl_out(DevCtl* c, SrdDyn* d, const uint64_t* off, const DlOut& o, hipStream_t) {
  dl_out_step(c, d, off, o);
}
void dl_check(DevCtl* c, uint32_t nrels, uint64_t cap_full, hipStream_t) {
  dl_check_step(c, nrels, cap_full);
}
void dl_sd_prep(DevCtl* c, uint32_t r, hipStream_t) { dl_sd_prep_step(c, r); }
void dl_sd_fin(DevCtl* c, uint32_t r, hipStream_t) { dl_sd_fin_step(c, r); }
void dl_mg_prep(DevCtl* c, uint32_t r, hipStream_t) { dl_mg_prep_step(c, r); }
void dl_mg_fin(DevCtl* c, uint32_t r, hipStream_t) { dl_mg_fin_step(c, r); }
void dl_iter_end(DevCtl* c, const DlEnd& e, hipStream_t) { dl_iter_end_step(c, e); }
void dl_inst(const DevCtl* c, uint32_t r, const uint32_t* col0, const uint32_t* bounds,
             uint32_t n_inst, uint64_t* out, hipStream_t) {
  if (c->skip) return;
  const uint64_t n = c->delta_n[r];
  for (uint32_t g = 0; g < n_inst; ++g) {
    const uint32_t* lo = std::lower_bound(col0, col0 + n, bounds[g]);
    const uint32_t* hi = std::lower_bound(lo, col0 + n, bounds[g + 1]);
    out[c->iter * n_inst + g] = uint64_t(hi - lo);
  }
}

}  // namespace srd
