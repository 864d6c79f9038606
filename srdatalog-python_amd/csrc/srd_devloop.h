// Device-resident fixpoint iteration: the control block and the single-
// thread bookkeeping steps that run between the launches of one iteration.
//
// A recursive step whose iteration the device can size by itself runs
// without host read-backs: every buffer is reserved at a capacity, the row
// counts live in a DevCtl in device memory, each pipeline's launches read
// their views and sizes from a device SrdDyn (srd_pipeline.h) that the bind
// step fills, and the steps below decide the sizes, output bases and
// convergence on the device.  A step that finds a capacity too small raises
// `overflow` with the size it needed and sets `skip`, which turns every
// later launch of the iteration (and of every iteration after it) into a
// no-op; the host then grows the buffers and runs that iteration again.
//
// The step functions are plain host/device code: the GPU build runs each in
// a one-thread kernel (srd_prims.hip), the host-only test build calls them
// directly.  The reference drives the same loop from the host
// (src/srdatalog/codegen/jit/orchestrator_jit.py:600-739) with a blocking
// size read per rule and phase (codegen/jit/complete_runner.py:1072-1112).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "srd_device.h"
#include "srd_pipeline.h"

namespace srd {

constexpr int kDevRels = 4;   // relations a device-resident step may update
constexpr int kDevPipes = 8;  // pipelines per iteration
enum : uint32_t { kOvfNew = 1, kOvfFull = 2, kOvfKeys = 4, kOvfSlices = 8, kOvfIters = 16 };

struct DevCtl {
  uint32_t skip, done, overflow, err;
  uint64_t iter, max_iter, iter_cap;
  uint64_t need_new, need_full, need_keys, need_slices;
  uint64_t join_rows;                 // tuples the joins emitted (before deduplication)
  uint64_t new_n[kDevRels];           // raw NEW rows written this iteration
  uint64_t newc[kDevRels][4];         // sorted NEW: distinct rows, column-0 values, error bits
  uint64_t full_n[kDevRels];
  uint64_t delta_n[kDevRels];
  uint64_t delta_r[kDevRels];
  uint64_t sd_nab[kDevRels][2];       // set-difference input sizes (0 when skipping)
  uint64_t sd_total[kDevRels];
  uint64_t mg_nab[kDevRels][2];       // merge input sizes (0 when skipping)
  uint64_t keys[kDevPipes];           // root keys of each pipeline this iteration
  uint64_t slices[kDevPipes];
};

struct DlBind {
  int32_t vrel[SRD_MAX_VIEWS];   // step relation whose DELTA the view reads, -1: static view
  uint32_t root_view[SRD_MAX_VIEWS];
  uint32_t n_views, root_kind, n_root_src, pipe;
  uint64_t cap_keys;
};

struct DlOut {
  int32_t rel[SRD_MAX_DESTS];    // step relation of each destination
  uint32_t n_dest, pipe, grain, direct;
  uint64_t cap_slices, cap_new;
};

struct DlEnd {
  uint64_t* deltas[kDevRels];    // per relation: |DELTA| of every iteration
  uint32_t nrels;
};

#define SRD_HD __host__ __device__

SRD_HD inline void dl_raise(DevCtl* c, uint32_t bit, uint64_t* need, uint64_t v) {
  c->overflow |= bit;
  c->skip = 1;
  if (v > *need) *need = v;
}

// Start of an iteration: NEW is empty; skip once converged, overflowed,
// failed or at max_iter; the per-iteration delta record must have room.
SRD_HD inline void dl_iter_begin_step(DevCtl* c, uint32_t nrels) {
  for (uint32_t r = 0; r < nrels; ++r) c->new_n[r] = 0;
  c->skip = (c->done || c->overflow || c->err || c->iter >= c->max_iter) ? 1u : 0u;
  if (!c->skip && c->iter >= c->iter_cap) {
    c->overflow |= kOvfIters;
    c->skip = 1;
  }
}

// Sizes of the DELTA views, the driver (the root source with the fewest
// distinct keys, as the host-driven pipeline picks it) and the key count.
SRD_HD inline void dl_bind_step(DevCtl* c, SrdDyn* d, const DlBind& b) {
  for (uint32_t v = 0; v < b.n_views; ++v)
    if (b.vrel[v] >= 0) {
      d->views.v[v].n = uint32_t(c->delta_n[b.vrel[v]]);
      d->views.v[v].r = uint32_t(c->delta_r[b.vrel[v]]);
    }
  uint32_t drv = b.root_view[0], nr;
  if (b.root_kind == SRD_ROOT_JOIN) {
    for (uint32_t i = 0; i < b.n_root_src; ++i) {
      const uint32_t v = b.root_view[i];
      if (d->views.v[v].r < d->views.v[drv].r) drv = v;
    }
    nr = d->views.v[drv].r;
  } else {
    nr = d->views.v[drv].n ? 1u : 0u;
  }
  if (!c->skip && nr > b.cap_keys) dl_raise(c, kOvfKeys, &c->need_keys, nr);
  d->skip = c->skip;
  if (c->skip) nr = 0;
  d->drv_view = drv;
  d->n_root = nr;
  d->total_work = 0;
  d->n_slices = 0;
  c->keys[b.pipe] = nr;
}

// Output bases of a launch's `rows` rows in NEW.
SRD_HD inline void dl_reserve(DevCtl* c, SrdDyn* d, const DlOut& o, uint64_t rows) {
  for (uint32_t k = 0; k < o.n_dest; ++k) {
    const uint64_t base = c->new_n[o.rel[k]];
    if (base + rows > o.cap_new) {
      dl_raise(c, kOvfNew, &c->need_new, base + rows);
      d->skip = 1;
      return;
    }
    d->out_base[k] = base;
    c->new_n[o.rel[k]] = base + rows;
  }
  c->join_rows += rows;
}

// After the root launch and the scan of its per-key work: total work and
// slice count; a direct launch (one row per item) reserves its rows here.
SRD_HD inline void dl_plan_step(DevCtl* c, SrdDyn* d, const uint64_t* work_start, const DlOut& o) {
  if (c->skip) {
    d->skip = 1;
    d->n_slices = 0;
    c->slices[o.pipe] = 0;
    return;
  }
  const uint64_t tw = work_start[d->n_root];
  const uint64_t ns = (tw + o.grain - 1) / o.grain;
  d->total_work = tw;
  if (ns > o.cap_slices) {
    dl_raise(c, kOvfSlices, &c->need_slices, ns);
    d->skip = 1;
    d->n_slices = 0;
    c->slices[o.pipe] = 0;
    return;
  }
  d->n_slices = ns;
  c->slices[o.pipe] = ns;
  if (o.direct) dl_reserve(c, d, o, tw);
}

// After the count launch and the scan of its per-slice counts.
SRD_HD inline void dl_out_step(DevCtl* c, SrdDyn* d, const uint64_t* offsets, const DlOut& o) {
  if (c->skip) {
    d->skip = 1;
    return;
  }
  dl_reserve(c, d, o, offsets[d->n_slices]);
}

// After the last NEW sort of the iteration, before any DELTA or FULL is
// written: FULL must be able to hold FULL u NEW.
SRD_HD inline void dl_check_step(DevCtl* c, uint32_t nrels, uint64_t cap_full) {
  if (c->skip) return;
  for (uint32_t r = 0; r < nrels; ++r) {
    if (c->newc[r][2]) {  // the device-planned sort reported an error
      c->err |= 1u;
      c->skip = 1;
      return;
    }
    if (c->full_n[r] + c->newc[r][0] > cap_full)
      dl_raise(c, kOvfFull, &c->need_full, c->full_n[r] + c->newc[r][0]);
  }
}

SRD_HD inline void dl_sd_prep_step(DevCtl* c, uint32_t r) {
  c->sd_nab[r][0] = c->skip ? 0 : c->newc[r][0];
  c->sd_nab[r][1] = c->skip ? 0 : c->full_n[r];
  c->sd_total[r] = 0;
}

SRD_HD inline void dl_sd_fin_step(DevCtl* c, uint32_t r) {
  if (!c->skip) c->delta_n[r] = c->sd_total[r];
}

SRD_HD inline void dl_mg_prep_step(DevCtl* c, uint32_t r) {
  c->mg_nab[r][0] = c->skip ? 0 : c->full_n[r];
  c->mg_nab[r][1] = c->skip ? 0 : c->delta_n[r];
}

SRD_HD inline void dl_mg_fin_step(DevCtl* c, uint32_t r) {
  if (!c->skip) c->full_n[r] += c->delta_n[r];
}

// End of an iteration: record every |DELTA|, count the iteration, and stop
// once the summed DELTA is empty (orchestrator_jit.py:659-695).
SRD_HD inline void dl_iter_end_step(DevCtl* c, const DlEnd& e) {
  if (c->skip) return;
  uint64_t sum = 0;
  for (uint32_t r = 0; r < e.nrels; ++r) {
    e.deltas[r][c->iter] = c->delta_n[r];
    sum += c->delta_n[r];
  }
  c->iter += 1;
  if (sum == 0) c->done = 1;
}

// Enqueue the steps (one thread each) on stream s.
void dl_iter_begin(DevCtl* c, uint32_t nrels, hipStream_t s);
void dl_bind(DevCtl* c, SrdDyn* d, const DlBind& b, hipStream_t s);
void dl_plan(DevCtl* c, SrdDyn* d, const uint64_t* work_start, const DlOut& o, hipStream_t s);
void dl_out(DevCtl* c, SrdDyn* d, const uint64_t* offsets, const DlOut& o, hipStream_t s);
void dl_check(DevCtl* c, uint32_t nrels, uint64_t cap_full, hipStream_t s);
void dl_sd_prep(DevCtl* c, uint32_t r, hipStream_t s);
void dl_sd_fin(DevCtl* c, uint32_t r, hipStream_t s);
void dl_mg_prep(DevCtl* c, uint32_t r, hipStream_t s);
void dl_mg_fin(DevCtl* c, uint32_t r, hipStream_t s);
void dl_iter_end(DevCtl* c, const DlEnd& e, hipStream_t s);
// Per-instance |DELTA| of a packed batch (two searches per instance on the
// sorted column 0 of relation r's DELTA) into row c->iter of out.
void dl_inst(const DevCtl* c, uint32_t r, const uint32_t* col0, const uint32_t* bounds,
             uint32_t n_inst, uint64_t* out, hipStream_t s);

}  // namespace srd
