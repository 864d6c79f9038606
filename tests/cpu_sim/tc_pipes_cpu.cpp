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

// Emission shared by the modes: MODE 0 count, 1 write at the slice offset,
// 2 stage into SRD_SEG-row segments + records, 3 direct (row = item).
struct Emit {
  const SrdLaunchArgs* a;
  int mode;
  uint64_t slice = 0, opos = 0, seg = ~0ull, item = 0;
  uint32_t cnt = 0, used = 0, local = 0;
  void close() {
    if (seg == ~0ull || used == 0) return;
    uint32_t idx = (*a->rec_ctr)++;
    if (idx >= a->rec_cap) { *a->overflow = 1; return; }
    a->recs[3 * idx] = slice;
    a->recs[3 * idx + 1] = seg;
    a->recs[3 * idx + 2] = (uint64_t(local) << 32) | used;
  }
  void put(uint32_t c0, uint32_t c1) {
    if (mode == 1) {
      a->out[0].cols[a->out[0].base + opos] = c0;
      a->out[0].cols[a->out[0].stride + a->out[0].base + opos] = c1;
    } else if (mode == 3) {
      a->out[0].cols[a->out[0].base + item] = c0;
      a->out[0].cols[a->out[0].stride + a->out[0].base + item] = c1;
    } else if (mode == 2) {
      if (seg == ~0ull || used + 1 > SRD_SEG) {
        close();
        uint64_t s = *a->stage_ctr;
        *a->stage_ctr += SRD_SEG;
        seg = (s + SRD_SEG > a->stage_cap) ? (*a->overflow = 1, ~0ull) : s;
        used = 0;
        local = cnt;
      }
      if (seg != ~0ull) {
        a->stage[seg + used] = c0;
        a->stage[a->stage_cap + seg + used] = c1;
      }
      ++used;
    }
    ++opos, ++cnt;
  }
  void end() {
    if (mode == 2) close();
    if (mode == 0 || mode == 2) a->counts[slice] = cnt;
  }
};

// ---- scans: key_work[0] = rows; items are rows; swap = write (y, x)
template <bool SWAP>
void scan_root(const SrdLaunchArgs* a, hipStream_t) {
  a->key_work[0] = a->views.v[a->drv_view].n;
  a->key_aux[0] = 0;
}

template <bool SWAP, int MODE>
void scan_run(const SrdLaunchArgs* a, hipStream_t) {
  const srd::IndexView& v = a->views.v[0];
  // slices [slice_lo, n_slices), as the device kernels take them
  for (uint64_t s = a->slice_lo; s < a->n_slices; ++s) {
    uint64_t lo = a->slice_start ? a->slice_start[s] : s * a->grain;
    uint64_t hi = a->slice_start ? a->slice_start[s + 1] : std::min<uint64_t>(lo + a->grain, a->total_work);
    Emit em{a, MODE};
    em.slice = s;
    em.opos = MODE == 1 ? a->offsets[s] : 0;
    for (uint64_t i = lo; i < hi; ++i) {
      uint32_t x = colp(v, 0)[i], y = colp(v, 1)[i];
      em.item = i;
      em.put(SWAP ? y : x, SWAP ? x : y);
    }
    em.end();
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

template <int MODE>
void rec_run(const SrdLaunchArgs* a, hipStream_t) {
  // slices [slice_lo, n_slices), as the device kernels take them
  for (uint64_t s = a->slice_lo; s < a->n_slices; ++s) {
    uint64_t lo = a->slice_start ? a->slice_start[s] : s * a->grain;
    uint64_t hi = a->slice_start ? a->slice_start[s + 1] : std::min<uint64_t>(lo + a->grain, a->total_work);
    Emit em{a, MODE};
    em.slice = s;
    em.opos = MODE == 1 ? a->offsets[s] : 0;
    for (uint64_t it = lo; it < hi; ++it) {
      uint32_t k = uint32_t(std::upper_bound(a->work_start, a->work_start + a->n_root + 1, it) -
                            a->work_start) - 1;
      uint64_t li = it - a->work_start[k];
      uint32_t b0 = a->key_rng[k * 4];
      uint32_t b1 = a->key_rng[k * 4 + 2], e1 = a->key_rng[k * 4 + 3];
      uint32_t d1 = e1 - b1;
      uint32_t x = colp(a->views.v[0], 1)[b0 + li / d1];
      uint32_t z = colp(a->views.v[1], 1)[b1 + li % d1];
      em.item = it;
      em.put(z, x);
    }
    em.end();
  }
}

// weigh mode stand-in: any positive cost per item is a valid slicing, so a
// hash of the item exercises uneven slices (empty ones included)
void rec_weigh(const SrdLaunchArgs* a, hipStream_t) {
  for (uint64_t it = 0; it < a->total_work; ++it) {
    uint64_t h = (it + 1) * 0x9E3779B97F4A7C15ull;
    h ^= h >> 29;
    a->item_cost[it] = 1 + ((h % 7 == 0) ? h % 5000 : h % 3);
  }
}

// Device-sized launches: the sizes, views and output bases come from the
// SrdDyn the engine's bind / plan / out steps filled.
template <SrdLaunchFn F>
void dyn_launch(const SrdLaunchArgs* a, hipStream_t s) {
  const SrdDyn& d = *a->dyn;
  if (d.skip) return;
  SrdLaunchArgs b = *a;
  b.views = d.views;
  b.n_root = d.n_root;
  b.drv_view = d.drv_view;
  b.total_work = d.total_work;
  b.n_slices = d.n_slices;
  for (int k = 0; k < SRD_MAX_DESTS; ++k) b.out[k].base = d.out_base[k];
  b.dyn = nullptr;
  F(&b, s);
}

const SrdPipelineDesc k_pipes[] = {
  {"EdgeLoad", 1, {{"ArcInput", 2, {0, 1}, SRD_FULL}}, SRD_ROOT_SCAN, 1, {0}, 1,
   {{"Edge", 2, {0, 1}}}, 0, 4096, 1, &scan_root<false>, &scan_run<false, 0>, &scan_run<false, 1>,
   &scan_run<false, 2>, &scan_run<false, 3>},
  {"TCBase", 1, {{"Edge", 2, {0, 1}, SRD_FULL}}, SRD_ROOT_SCAN, 1, {0}, 1,
   {{"Path", 2, {1, 0}}}, 0, 4096, 1, &scan_root<true>, &scan_run<true, 0>, &scan_run<true, 1>,
   &scan_run<true, 2>, &scan_run<true, 3>},
  {"TCRec_D0", 2, {{"Path", 2, {1, 0}, SRD_DELTA}, {"Edge", 2, {0, 1}, SRD_FULL}}, SRD_ROOT_JOIN, 2,
   {0, 1}, 1, {{"Path", 2, {1, 0}}}, 0, 7, 1, &rec_root, &rec_run<0>, &rec_run<1>, &rec_run<2>,
   &rec_run<3>, &rec_weigh, 0, nullptr, &dyn_launch<&rec_root>, &dyn_launch<&rec_run<0>>,
   &dyn_launch<&rec_run<1>>, &dyn_launch<&rec_run<3>>},
};

}  // namespace

extern "C" srd_engine* cpu_tc_engine(const char* program_text) {
  return srd_engine_create(program_text, k_pipes, 3);
}
