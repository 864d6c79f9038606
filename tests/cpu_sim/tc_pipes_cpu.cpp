// CPU versions of the transitive-closure pipelines (EdgeLoad, TCBase,
// TCRec_D0) following the root / count / write protocol of srd_pipeline.h,
// so the engine's host logic can be driven end to end without a GPU.
#include <algorithm>
#include <cstring>

#include "srd_engine.h"
#include "srd_pipeline.h"

namespace {

const uint32_t* colp(const srd::IndexView& v, int c) { return v.cols + uint64_t(c) * v.stride; }

srd::Rng cpu_root_find(const srd::IndexView& v, uint32_t key) {
  const uint32_t* e = v.root_vals + v.r;
  const uint32_t* it = std::lower_bound(v.root_vals, e, key);
  if (it == e || *it != key) return {0, 0};
  uint32_t g = uint32_t(it - v.root_vals);
  return {v.root_offs[g], v.root_offs[g + 1]};
}

// ---- scans: key_work[0] = rows; items are rows; swap = write (y, x)
template <bool SWAP>
void scan_root(const SrdLaunchArgs* a, hipStream_t) {
  a->key_work[0] = a->views.v[a->drv_view].n;
  a->key_aux[0] = 0;
}

template <bool SWAP, bool WRITE>
void scan_run(const SrdLaunchArgs* a, hipStream_t) {
  const srd::IndexView& v = a->views.v[0];
  for (uint64_t s = 0; s < a->n_slices; ++s) {
    uint64_t lo = s * a->grain, hi = std::min<uint64_t>(lo + a->grain, a->total_work);
    if (!WRITE) {
      a->counts[s] = uint32_t(hi - lo);
      continue;
    }
    uint64_t o = a->offsets[s];
    for (uint64_t i = lo; i < hi; ++i, ++o) {
      uint32_t x = colp(v, 0)[i], y = colp(v, 1)[i];
      a->out[0].cols[a->out[0].base + o] = SWAP ? y : x;
      a->out[0].cols[a->out[0].stride + a->out[0].base + o] = SWAP ? x : y;
    }
  }
}

// ---- TCRec_D0: root y over Path(1 0) DELTA (view 0) and Edge(0 1) (view 1)
void rec_root(const SrdLaunchArgs* a, hipStream_t) {
  const srd::IndexView& D = a->views.v[a->drv_view];
  for (uint32_t k = 0; k < a->n_root; ++k) {
    uint32_t key = D.root_vals[k];
    srd::Rng r[2];
    for (int s = 0; s < 2; ++s)
      r[s] = (a->drv_view == uint32_t(s)) ? srd::Rng{D.root_offs[k], D.root_offs[k + 1]}
                                           : cpu_root_find(a->views.v[s], key);
    uint64_t w = (r[0].e > r[0].b && r[1].e > r[1].b) ? uint64_t(r[0].e - r[0].b) * (r[1].e - r[1].b) : 0;
    for (int s = 0; s < 2; ++s) a->key_rng[k * 4 + 2 * s] = r[s].b, a->key_rng[k * 4 + 2 * s + 1] = r[s].e;
    a->key_aux[k] = 0;
    a->key_work[k] = w;
  }
}

template <bool WRITE>
void rec_run(const SrdLaunchArgs* a, hipStream_t) {
  for (uint64_t s = 0; s < a->n_slices; ++s) {
    uint64_t lo = s * a->grain, hi = std::min<uint64_t>(lo + a->grain, a->total_work);
    uint64_t o = WRITE ? a->offsets[s] : 0;
    uint32_t cnt = 0;
    for (uint64_t it = lo; it < hi; ++it) {
      uint32_t k = uint32_t(std::upper_bound(a->work_start, a->work_start + a->n_root + 1, it) -
                            a->work_start) - 1;
      uint64_t li = it - a->work_start[k];
      uint32_t b0 = a->key_rng[k * 4], e0 = a->key_rng[k * 4 + 1];
      uint32_t b1 = a->key_rng[k * 4 + 2], e1 = a->key_rng[k * 4 + 3];
      uint32_t d1 = e1 - b1;
      uint32_t x = colp(a->views.v[0], 1)[b0 + li / d1];
      uint32_t z = colp(a->views.v[1], 1)[b1 + li % d1];
      (void)e0;
      if (WRITE) {
        a->out[0].cols[a->out[0].base + o] = z;
        a->out[0].cols[a->out[0].stride + a->out[0].base + o] = x;
        ++o;
      }
      ++cnt;
    }
    if (!WRITE) a->counts[s] = cnt;
  }
}

const SrdPipelineDesc k_pipes[] = {
  {"EdgeLoad", 1, {{"ArcInput", 2, {0, 1}, SRD_FULL}}, SRD_ROOT_SCAN, 1, {0}, 1,
   {{"Edge", 2, {0, 1}}}, 0, 4096, &scan_root<false>, &scan_run<false, false>, &scan_run<false, true>},
  {"TCBase", 1, {{"Edge", 2, {0, 1}, SRD_FULL}}, SRD_ROOT_SCAN, 1, {0}, 1,
   {{"Path", 2, {1, 0}}}, 0, 4096, &scan_root<true>, &scan_run<true, false>, &scan_run<true, true>},
  {"TCRec_D0", 2, {{"Path", 2, {1, 0}, SRD_DELTA}, {"Edge", 2, {0, 1}, SRD_FULL}}, SRD_ROOT_JOIN, 2,
   {0, 1}, 1, {{"Path", 2, {1, 0}}}, 0, 7, &rec_root, &rec_run<false>, &rec_run<true>},
};

}  // namespace

extern "C" srd_engine* cpu_tc_engine(const char* program_text) {
  return srd_engine_create(program_text, k_pipes, 3);
}
