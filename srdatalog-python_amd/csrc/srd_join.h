// Device helpers used by the generated per-rule join kernels.
//
// Execution model of a generated kernel (see srd_pipeline.h for the launch
// protocol): a wave64 owns one slice, a contiguous range of the flattened
// level-1 items (possibly spanning keys).  Items are processed 64 at a time,
// one per lane.  If the rule has a second nested level, the lanes that
// produced a binding publish it in LDS and the wave walks the concatenation
// of their level-2 iteration spaces 64 items at a time (each position's
// owner found from per-batch start flags, a ballot and a popcount), so a
// lane with a huge level-2 range does not serialize the wave.  Every emitted
// tuple is one lane's result for one item; ballot + mbcnt give its position,
// so emission never diverges.
//
// Counterparts in the reference: the warp-cooperative WCOJ body emitted by
// src/srdatalog/codegen/jit/pipeline.py:360 / root.py / instructions.py over
// the search primitives of runtime/generalized_datalog/gpu/search.h and
// IntersectView (gpu/device_sorted_array_index.ipp:1284-1487).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "srd_device.h"
#include "srd_pipeline.h"

namespace srdj {

using srd::IndexView;
using srd::Rng;

__device__ __forceinline__ const uint32_t* colp(const IndexView& v, int c) {
  return v.cols + uint64_t(c) * v.stride;
}

// find_slot, find_slot_from: srd_device.h
using srd::find_slot;
using srd::find_slot_from;

// Equal range of the value at driver position pos inside [b,e) of column
// col (values within [b,e) sorted); `last` = column is the index's last.
__device__ __forceinline__ Rng run_at(const uint32_t* __restrict__ col, uint32_t pos, uint32_t e,
                                      uint32_t v, bool last) {
  if (last) return Rng{pos, pos + 1};
  return Rng{pos, srd::upper_bound_la(col, pos + 1, e, v)};
}

// Persistent-wave slice dispenser.  Every lane executes the atomic (lane 0
// adds 1, the others 0) and the wave takes lane 0's result with v_readlane,
// a scalar.  There must be no lane-dependent branch around the atomic: with
// `if (lane == 0) s = atomicAdd(...)` the compiler threads the lanes that
// skip the branch (whose s is the constant 0) straight back to the
// broadcast, forming a second loop in which lanes 1..63 re-read lane 0's
// stale register and never fetch a new slice.
// Slices [slice_lo, n_slices) are split into SRD_XCDS ranges; a wave takes
// the next slice of its own XCD's range, then of the following ranges in
// turn; it returns n_slices when every range is exhausted.
__device__ __forceinline__ uint64_t next_slice(const SrdLaunchArgs& A, uint64_t n) {
  const uint64_t lo = A.slice_lo;
  if (n <= lo) return n;
  const uint64_t span = n - lo;
  const uint32_t nr = A.n_ranges ? A.n_ranges : 1u;
  const uint32_t home = blockIdx.x % nr;
  for (uint32_t k = 0; k < nr; ++k) {
    const uint32_t x = (home + k) % nr;
    const uint64_t b = lo + span * x / nr, e = lo + span * (x + 1) / nr;
    if (b >= e) continue;
    const uint32_t t = atomicAdd(A.work_ctr + x * SRD_CTR_STRIDE, srd::lane_id() == 0 ? 1u : 0u);
    const uint64_t s = b + __builtin_amdgcn_readlane(t, 0);
    if (s < e) return s;
  }
  return n;
}

// Fused mode: open a staging segment of SRD_SEG rows for the wave (all lanes
// take part; lane 0's bump is broadcast with v_readlane).  Returns ~0 when
// staging is exhausted (the overflow flag is raised; the slice keeps
// counting, and the host then re-runs the write pass).
__device__ __forceinline__ uint64_t seg_open(const SrdLaunchArgs& A) {
  const unsigned long long r =
      atomicAdd(A.stage_ctr, srd::lane_id() == 0 ? (unsigned long long)SRD_SEG : 0ull);
  const uint64_t s = srd::read_lane_u64(r, 0);
  const bool ovf = s + SRD_SEG > A.stage_cap;  // wave-uniform
  if (ovf) atomicOr(A.overflow, srd::lane_id() == 0 ? 1u : 0u);
  return ovf ? ~uint64_t(0) : s;
}

// Fused mode: log segment [start, start+used) as rows local..local+used of
// `slice`'s output.
__device__ __forceinline__ void seg_close(const SrdLaunchArgs& A, uint64_t slice, uint64_t start,
                                          uint32_t used, uint32_t local) {
  if (start == ~uint64_t(0) || used == 0) return;
  const uint32_t r = atomicAdd(A.rec_ctr, srd::lane_id() == 0 ? 1u : 0u);
  const uint32_t idx = __builtin_amdgcn_readlane(r, 0);
  const bool ovf = idx >= A.rec_cap;  // wave-uniform
  if (ovf) atomicOr(A.overflow, srd::lane_id() == 0 ? 1u : 0u);
  if (!ovf && srd::lane_id() == 0) {
    A.recs[3ull * idx] = slice;
    A.recs[3ull * idx + 1] = start;
    A.recs[3ull * idx + 2] = (uint64_t(local) << 32) | used;
  }
}

// Cooperative lower bound: for every lane with `on`, the first position p in
// [b, e) of col with col[p] >= v (e if none).  Consecutive `on` lanes with the
// same `seg` form a run; the lanes of a run must share [b, e) and hold
// non-decreasing v (a run is one owner's candidates, read in column order).
// The first lane of each run searches lower_bound(v_first) and the last one
// upper_bound(v_last) over the whole range; every other lane of the run then
// searches only the window between those two positions, where its answer
// lies.  A divergent full-range search costs one cache-line lookup per lane
// and step; the window searches of one run touch a few shared lines instead.
// Must be called with all 64 lanes active (wave-uniform control flow).
__device__ __forceinline__ uint32_t seg_lower_bound(const uint32_t* __restrict__ col, uint32_t b,
                                                    uint32_t e, uint32_t v, bool on, uint32_t seg) {
  const uint32_t lane = srd::lane_id();
  const uint64_t onm = srd::ballot(on);
  const uint32_t seg_p = __shfl_up(seg, 1, 64);
  const uint32_t seg_n = __shfl_down(seg, 1, 64);
  const bool on_p = lane > 0 && ((onm >> (lane - 1)) & 1ull);
  const bool on_n = lane < 63 && ((onm >> (lane + 1)) & 1ull);
  const bool first = on && !(on_p && seg_p == seg);
  const bool last = on && !(on_n && seg_n == seg);
  // run ends: first lanes lower_bound(v), last-only lanes upper_bound(v)
  const bool upper = last && !first;
  uint32_t lo = b;
  uint32_t n = (first || last) ? e - b : 0u;
  while (n > 0) {
    const uint32_t half = n >> 1;
    const uint32_t mid = lo + half;
    const uint32_t x = col[mid];
    if (upper ? x <= v : x < v) {
      lo = mid + 1;
      n -= half + 1;
    } else {
      n = half;
    }
  }
  const uint64_t fm = srd::ballot(first);
  const uint64_t lm = srd::ballot(last);
  const uint64_t le = lane == 63 ? ~0ull : ((2ull << lane) - 1);
  const uint64_t below = fm & le;
  const uint32_t src_f = below ? 63u - uint32_t(__clzll((long long)below)) : lane;
  const uint64_t above = lm >> lane;
  const uint32_t src_l = above ? lane + uint32_t(__ffsll((unsigned long long)above)) - 1u : lane;
  const uint32_t w_lo = __shfl(lo, int(src_f), 64);
  const uint32_t w_hi = __shfl(lo, int(src_l), 64);
  if (!on || first) return lo;
  // w_lo <= answer <= w_hi for every lane strictly inside or at the end of a run
  uint32_t l2 = w_lo;
  uint32_t n2 = w_hi - w_lo;
  while (n2 > 0) {
    const uint32_t half = n2 >> 1;
    const uint32_t mid = l2 + half;
    if (col[mid] < v) {
      l2 = mid + 1;
      n2 -= half + 1;
    } else {
      n2 = half;
    }
  }
  return l2;
}

// Whether index view v holds a row whose first sizeof...(key) columns equal
// key (negated atoms: the rule's binding survives only when this is false).
template <class... T>
__device__ __forceinline__ bool has_prefix(const IndexView& v, T... key) {
  constexpr int n = int(sizeof...(T));
  const uint32_t k[n] = {uint32_t(key)...};
  if (v.n == 0) return false;
  Rng r = srd::root_find(v, k[0]);
#pragma unroll
  for (int d = 1; d < n; ++d) {
    if (r.e <= r.b) return false;
    r = srd::child_find(v, d, r, k[d], d == int(v.arity) - 1);
  }
  return r.e > r.b;
}

// Bitmap slot of the heavy key whose depth-1 range of view v starts at row
// b, or ~0u when that key has no bitmap.
__device__ __forceinline__ uint32_t heavy_slot(const IndexView& v, uint32_t b) {
  if (v.n_hv == 0) return ~0u;
  const uint32_t p = srd::lower_bound(v.hv_start, 0, v.n_hv, b);
  return (p < v.n_hv && v.hv_start[p] == b) ? p : ~0u;
}

// Whether value x is in the range whose bitmap is `slot` of view v.
__device__ __forceinline__ bool heavy_has(const IndexView& v, uint32_t slot, uint32_t x) {
  const uint32_t w = x >> 5;
  return w < v.bm_words && ((v.bm[uint64_t(slot) * v.bm_words + w] >> (x & 31u)) & 1u);
}

// Wave-uniform copy of a 64-bit value held by `lane`.
__device__ __forceinline__ uint64_t uniform_from(uint64_t v, int lane) {
  uint32_t lo = __builtin_amdgcn_readlane(uint32_t(v), lane);
  uint32_t hi = __builtin_amdgcn_readlane(uint32_t(v >> 32), lane);
  return (uint64_t(hi) << 32) | lo;
}

}  // namespace srdj
