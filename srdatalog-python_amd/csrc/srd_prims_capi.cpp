// C ABI over the set primitives (declared in include/srdatalog_prims.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/srdatalog_prims.h"
#include "srd_runtime.h"

namespace {

thread_local std::string g_err;

}  // namespace

namespace srd {
// Message returned by srd_last_error() on this thread (other C-ABI sources
// of the library report through it).
void capi_set_error(const std::string& msg) { g_err = msg; }
}  // namespace srd

namespace {

struct Ctx {
  hipStream_t s = nullptr;
  srd::Workspace ws;
  Ctx() { s = srd::stream_acquire(); }
};

Ctx& ctx() {
  static Ctx* c = nullptr;
  if (!c) {
    srd::pool_init(0);
    c = new Ctx();
  }
  return *c;
}

template <class F>
int guard(F&& f) {
  srd::RunScope rs;
  try {
    ctx().ws.ticket_order = !rs.alone;
    f();
    g_err.clear();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  } catch (...) {
    g_err = "unknown error";
    return 2;
  }
}

// Upload row-major rows into k device columns (stride = round_stride(n)).
srd::DevBuf<uint32_t> upload_cols(const uint32_t* rows, uint64_t n, uint32_t k, uint64_t& stride,
                                  hipStream_t s) {
  stride = srd::round_stride(n);
  std::vector<uint32_t> t(stride * k, 0);
  for (uint64_t i = 0; i < n; ++i)
    for (uint32_t c = 0; c < k; ++c) t[c * stride + i] = rows[i * k + c];
  srd::DevBuf<uint32_t> d(stride * k ? stride * k : 1, s);
  if (!t.empty()) SRD_CHECK(hipMemcpyAsync(d.p, t.data(), t.size() * 4, hipMemcpyHostToDevice, s));
  SRD_CHECK(hipStreamSynchronize(s));
  return d;
}

void download_index(const srd::Index& x, uint32_t* out_rows, hipStream_t s) {
  if (x.n == 0) return;
  std::vector<uint32_t> t(x.stride * x.arity);
  SRD_CHECK(hipMemcpyAsync(t.data(), x.cols.p, t.size() * 4, hipMemcpyDeviceToHost, s));
  SRD_CHECK(hipStreamSynchronize(s));
  for (uint64_t i = 0; i < x.n; ++i)
    for (uint32_t c = 0; c < x.arity; ++c) out_rows[i * x.arity + c] = t[c * x.stride + i];
}

srd::Index index_from_sorted(const uint32_t* rows, uint64_t n, uint32_t k, Ctx& c) {
  srd::Index x;
  x.arity = k;
  x.n = n;
  x.col_max.assign(k, 0);
  for (uint64_t i = 0; i < n; ++i)
    for (uint32_t j = 0; j < k; ++j) x.col_max[j] = std::max(x.col_max[j], rows[i * k + j]);
  if (n == 0) return x;
  x.cols = upload_cols(rows, n, k, x.stride, c.s);
  srd::build_roots(x, c.ws, c.s);
  return x;
}

}  // namespace

extern "C" {

const char* srd_last_error(void) { return g_err.c_str(); }

int srd_prim_init(int device) {
  return guard([&] {
    srd::pool_init(device);
    ctx();
  });
}

int srd_prim_sort_dedup(const uint32_t* rows, uint64_t n, uint32_t k, uint32_t* out_rows,
                        uint64_t* out_n, uint32_t* root_vals, uint32_t* root_offs,
                        uint64_t* out_r) {
  return guard([&] {
    Ctx& c = ctx();
    uint64_t stride = 0;
    auto cols = upload_cols(rows, n, k, stride, c.s);
    std::vector<const uint32_t*> in(k);
    for (uint32_t j = 0; j < k; ++j) in[j] = cols.p + j * stride;
    srd::Index x;
    srd::build_index(in.data(), k, n, x, c.ws, c.s);
    download_index(x, out_rows, c.s);
    *out_n = x.n;
    if (out_r) *out_r = x.r;
    if (x.r && root_vals)
      SRD_CHECK(hipMemcpyAsync(root_vals, x.root_vals.p, uint64_t(x.r) * 4, hipMemcpyDeviceToHost, c.s));
    if (x.r && root_offs)
      SRD_CHECK(hipMemcpyAsync(root_offs, x.root_offs.p, uint64_t(x.r + 1) * 4,
                               hipMemcpyDeviceToHost, c.s));
    SRD_CHECK(hipStreamSynchronize(c.s));
  });
}

int srd_prim_set_difference(const uint32_t* a, uint64_t na, const uint32_t* b, uint64_t nb,
                            uint32_t k, uint32_t* out, uint64_t* out_n) {
  return guard([&] {
    Ctx& c = ctx();
    srd::Index xa = index_from_sorted(a, na, k, c), xb = index_from_sorted(b, nb, k, c), xo;
    srd::set_difference(xa, xb, xo, c.ws, c.s);
    download_index(xo, out, c.s);
    *out_n = xo.n;
  });
}

int srd_prim_merge(const uint32_t* a, uint64_t na, const uint32_t* b, uint64_t nb, uint32_t k,
                   uint32_t* out, uint64_t* out_n) {
  return guard([&] {
    Ctx& c = ctx();
    srd::Index xa = index_from_sorted(a, na, k, c), xb = index_from_sorted(b, nb, k, c), xo;
    srd::merge_disjoint(xa, xb, xo, c.ws, c.s);
    download_index(xo, out, c.s);
    *out_n = xo.n;
  });
}

int srd_prim_exclusive_scan(const uint32_t* in, uint64_t n, uint64_t* out) {
  return guard([&] {
    Ctx& c = ctx();
    srd::DevBuf<uint32_t> din(n ? n : 1, c.s);
    srd::DevBuf<uint64_t> dout(n + 1, c.s);
    if (n) SRD_CHECK(hipMemcpyAsync(din.p, in, n * 4, hipMemcpyHostToDevice, c.s));
    srd::exclusive_scan_u32(din.p, dout.p, n, c.ws, c.s);
    SRD_CHECK(hipMemcpyAsync(out, dout.p, (n + 1) * 8, hipMemcpyDeviceToHost, c.s));
    SRD_CHECK(hipStreamSynchronize(c.s));
  });
}

int srd_prim_time_sort_dedup(const uint32_t* rows, uint64_t n, uint32_t k, uint32_t reps,
                             double* ms_per_rep) {
  return guard([&] {
    Ctx& c = ctx();
    uint64_t stride = 0;
    auto cols = upload_cols(rows, n, k, stride, c.s);
    std::vector<const uint32_t*> in(k);
    for (uint32_t j = 0; j < k; ++j) in[j] = cols.p + j * stride;
    srd::Index x;
    srd::build_index(in.data(), k, n, x, c.ws, c.s);  // warm-up
    hipEvent_t e0, e1;
    SRD_CHECK(hipEventCreate(&e0));
    SRD_CHECK(hipEventCreate(&e1));
    SRD_CHECK(hipEventRecord(e0, c.s));
    for (uint32_t r = 0; r < reps; ++r) srd::build_index(in.data(), k, n, x, c.ws, c.s);
    SRD_CHECK(hipEventRecord(e1, c.s));
    SRD_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    SRD_CHECK(hipEventElapsedTime(&ms, e0, e1));
    *ms_per_rep = reps ? ms / reps : 0.0;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  });
}

int srd_prim_bench(int op, uint64_t na, uint64_t nb, uint32_t k, uint32_t reps,
                   double* ms_per_rep, uint64_t* out_n) {
  return guard([&] {
    Ctx& c = ctx();
    if (k < 1 || k > 4) throw std::runtime_error("srd_prim_bench: arity must be 1..4");
    const uint32_t fb = 20;                       // width of columns 1..k-1
    const uint32_t low = fb * (k - 1);
    auto sorted_index = [&](uint64_t n, uint64_t mul, uint64_t off) {
      srd::Index x;
      x.arity = k;
      x.n = n;
      x.col_max.assign(k, k > 1 ? (1u << fb) - 1 : 0u);
      if (n == 0) return x;
      const unsigned __int128 top = (unsigned __int128)mul * (n - 1) + off;
      if ((top >> low) > 0xffffffffu) throw std::runtime_error("srd_prim_bench: keys overflow u32");
      x.col_max[0] = uint32_t(uint64_t(top >> low));
      x.stride = srd::round_stride(n);
      x.cols.alloc(x.stride * k, c.s);
      srd::fill_pattern(x.cols.p, x.stride, n, k, fb, mul, off, 0, c.s);
      srd::build_roots(x, c.ws, c.s);
      return x;
    };
    srd::Index a, b, o;
    srd::DevBuf<uint32_t> raw;
    std::vector<const uint32_t*> in(k);
    uint64_t stride = srd::round_stride(na);
    if (op == 0) {
      raw.alloc(stride * k + 1, c.s);
      const uint32_t bits = k == 1 ? 32 : low + 20;  // column 0 spans 2^20 values
      srd::fill_pattern(raw.p, stride, na, k, fb, 0, 0, uint64_t(1) << bits, c.s);
      for (uint32_t j = 0; j < k; ++j) in[j] = raw.p + j * stride;
    } else if (op == 1) {
      a = sorted_index(na, 2, 0);                 // NEW  = {2i}
      b = sorted_index(nb, 3, 0);                 // FULL = {3j}
    } else if (op == 2) {
      a = sorted_index(na, 3, 0);                 // FULL  = {3i}
      b = sorted_index(nb, 3, 1);                 // DELTA = {3j+1}, disjoint
    } else {
      throw std::runtime_error("srd_prim_bench: op must be 0 (sort+dedup), 1 (set-difference) "
                               "or 2 (merge)");
    }
    auto once = [&] {
      if (op == 0) srd::build_index(in.data(), k, na, o, c.ws, c.s);
      else if (op == 1) srd::set_difference(a, b, o, c.ws, c.s);
      else srd::merge_disjoint(a, b, o, c.ws, c.s);
    };
    once();  // warm-up: pool growth, first touch
    SRD_CHECK(hipStreamSynchronize(c.s));
    hipEvent_t e0, e1;
    SRD_CHECK(hipEventCreate(&e0));
    SRD_CHECK(hipEventCreate(&e1));
    SRD_CHECK(hipEventRecord(e0, c.s));
    for (uint32_t r = 0; r < reps; ++r) once();
    SRD_CHECK(hipEventRecord(e1, c.s));
    SRD_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    SRD_CHECK(hipEventElapsedTime(&ms, e0, e1));
    *ms_per_rep = reps ? ms / reps : 0.0;
    if (out_n) *out_n = o.n;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  });
}

int srd_calib_copy(uint64_t bytes, int reps, double* ms_per_rep) {
  return guard([&] {
    Ctx& c = ctx();
    srd::calib_copy(bytes, reps, c.s, ms_per_rep);
  });
}

int srd_pool_stats(uint64_t* out, uint32_t cap) {
  return guard([&] {
    const srd::PoolStats st = srd::pool_stats();
    const uint64_t v[9] = {st.in_use,        st.peak,     st.budget, st.allocs,
                           st.driver_allocs, st.reserved, st.cached, st.trims,
                           st.cross_stream};
    for (uint32_t i = 0; i < cap && i < 9; ++i) out[i] = v[i];
  });
}

int srd_timing_enable(int on) {
  srd::timing_enable(on != 0);
  return 0;
}

uint64_t srd_timing_report(char* buf, uint64_t cap) {
  std::string js = "[";
  try {
    auto v = srd::timing_collect();
    for (size_t i = 0; i < v.size(); ++i) {
      char tmp[512];
      std::snprintf(tmp, sizeof(tmp), "%s{\"name\": \"%s\", \"launches\": %llu, \"ms\": %.6f, \"bytes\": %.1f}",
                    i ? ", " : "", v[i].name.c_str(), (unsigned long long)v[i].launches, v[i].ms,
                    v[i].bytes);
      js += tmp;
    }
  } catch (const std::exception& e) {
    g_err = e.what();
  }
  js += "]";
  if (buf && cap) {
    size_t n = std::min<size_t>(js.size(), cap - 1);
    std::memcpy(buf, js.data(), n);
    buf[n] = 0;
  }
  return js.size();
}

int srd_timing_reset(void) {
  return guard([&] { srd::timing_reset(); });
}

}  // extern "C"

