// Device-side building blocks shared by the runtime primitives and the
// generated per-rule join kernels: wave64 intrinsics, sorted-column search,
// and the POD index view handed to kernels.
//
// Index layout in HBM (one relation version, one column order):
//   cols       k u32 columns, column c at cols + c*stride, rows sorted
//              lexicographically and duplicate-free;
//   root_vals  the r distinct values of column 0, ascending;
//   root_offs  r+1 row offsets: rows with column 0 == root_vals[g] are
//              [root_offs[g], root_offs[g+1]).
// This is the sorted-array column index of the reference
// (runtime/generalized_datalog/gpu/device_sorted_array_index.h:619-643)
// plus a CSR offset array so a depth-0 key lookup is one search over the
// distinct keys instead of two over the whole column.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace srd {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;

struct IndexView {
  const uint32_t* cols;
  const uint32_t* root_vals;
  const uint32_t* root_offs;
  uint64_t stride;
  uint32_t n;      // rows
  uint32_t r;      // distinct column-0 values
  uint32_t arity;
  uint32_t bm_min;  // ranges shorter than this have no bitmap
  // Membership bitmaps of the column-1 values of heavy keys (two-column
  // indexes a join searches at depth 1): bitmap i covers the rows starting
  // at hv_start[i] (ascending), bm_words u32 each; n_hv = 0 when absent.
  const uint32_t* bm;
  const uint32_t* hv_start;
  uint32_t n_hv;
  uint32_t bm_words;
  // Dense map of the root values: root_pos[x] = g when root_vals[g] == x,
  // ~0u when x is not a root value, for x <= root_max; null when absent.
  const uint32_t* root_pos;
  uint32_t root_max;
  uint32_t pad_;
};

struct Rng {
  uint32_t b, e;
};

__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Number of set bits of m strictly below this lane.
__device__ __forceinline__ uint32_t popc_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Value of v in `lane`, as a wave-uniform (scalar) result; v_readlane does
// not depend on EXEC, unlike a __shfl through ds_bpermute.
__device__ __forceinline__ uint64_t read_lane_u64(uint64_t v, int lane) {
  uint32_t lo = __builtin_amdgcn_readlane(uint32_t(v), lane);
  uint32_t hi = __builtin_amdgcn_readlane(uint32_t(v >> 32), lane);
  return (uint64_t(hi) << 32) | lo;
}

// Adds, for every valid lane, 1 to h[d] (d this lane's bin), with one LDS
// atomic per run of equal bins in consecutive lanes instead of one per lane:
// keys a join wrote side by side share their high digits, and 64 lanes
// incrementing one LDS word serialise.  The whole wave must call it, and the
// valid lanes must be a prefix of the wave (the tail of a range).
__device__ __forceinline__ void hist_add_runs(uint32_t* h, uint32_t d, bool valid) {
  const uint32_t l = lane_id();
  const uint32_t prev = __shfl_up(d, 1, kWave);  // lane 0 reads its own
  const uint64_t vm = ballot(valid);
  const bool start = valid && (l == 0 || prev != d);
  const uint64_t sm = ballot(start);
  if (start) {
    // the run ends at the next start or at the first invalid lane
    const uint64_t ends = (sm | ~vm) & ~((uint64_t(2) << l) - 1);
    const uint32_t e = ends ? uint32_t(__builtin_ctzll(ends)) : uint32_t(kWave);
    atomicAdd(&h[d], e - l);
  }
}

template <class T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  const uint32_t l = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    T t = __shfl_up(v, o, kWave);
    if (l >= uint32_t(o)) v += t;
  }
  return v;
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

template <class T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    T t = __shfl_xor(v, o, kWave);
    v = t > v ? t : v;
  }
  return v;
}

// Block-wide exclusive scan for NW waves (the block's size); returns this
// thread's exclusive prefix and writes the block total to *total.  `wsum`
// must hold NW entries of shared memory.
template <class T, int NW = kWavesPerBlock>
__device__ __forceinline__ T block_excl_scan(T v, T* wsum, T* total) {
  const uint32_t l = lane_id();
  const uint32_t w = threadIdx.x / kWave;
  T inc = wave_incl_scan(v);
  if (l == kWave - 1) wsum[w] = inc;
  __syncthreads();
  T base = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < NW; ++i) {
    T s = wsum[i];
    if (uint32_t(i) < w) base += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return base + inc - v;
}

// First position p in [lo,hi) with a[p] >= v (hi if none).
__device__ __forceinline__ uint32_t lower_bound(const uint32_t* __restrict__ a, uint32_t lo,
                                                uint32_t hi, uint32_t v) {
  uint32_t n = hi - lo;
  while (n > 0) {
    uint32_t half = n >> 1;
    uint32_t mid = lo + half;
    if (a[mid] < v) {
      lo = mid + 1;
      n -= half + 1;
    } else {
      n = half;
    }
  }
  return lo;
}

// First position p in [lo,hi) with a[p] > v (hi if none).
__device__ __forceinline__ uint32_t upper_bound(const uint32_t* __restrict__ a, uint32_t lo,
                                                uint32_t hi, uint32_t v) {
  uint32_t n = hi - lo;
  while (n > 0) {
    uint32_t half = n >> 1;
    uint32_t mid = lo + half;
    if (a[mid] <= v) {
      lo = mid + 1;
      n -= half + 1;
    } else {
      n = half;
    }
  }
  return lo;
}

// lower_bound / upper_bound that take two halving steps per memory round
// trip: with the midpoint it also loads the midpoints of both halves, so the
// next step's value is already there whichever way the comparison goes.
// Latency-bound searches (the join's) then wait for half as many dependent
// loads, at 1.5x the loads.  Both steps are written as selects, so lanes
// that go different ways do not split the wave into two branch paths.
// (Three steps per trip, 7 loads, measured slower on the triangle join: the
// extra loads cost more than the saved trips.)
template <bool UPPER>
__device__ __forceinline__ uint32_t bound_la(const uint32_t* __restrict__ a, uint32_t lo,
                                             uint32_t hi, uint32_t v) {
  uint32_t n = hi - lo;
  while (n > 1) {
    const uint32_t half = n >> 1;
    const uint32_t mid = lo + half;
    const uint32_t nl = half, nr = n - half - 1;  // nl >= 1
    const uint32_t ml = lo + (nl >> 1);           // < mid
    const uint32_t mr = mid + 1 + (nr >> 1);      // < hi when nr > 0
    const uint32_t x = a[mid];
    const uint32_t xl = a[ml];
    const uint32_t xr = a[nr ? mr : mid];
    const bool right = UPPER ? x <= v : x < v;
    const uint32_t lo1 = right ? mid + 1 : lo;    // after the first step
    const uint32_t n1 = right ? nr : nl;
    const uint32_t m2 = right ? mr : ml;          // = lo1 + (n1 >> 1)
    const uint32_t y = right ? xr : xl;
    const bool right2 = n1 > 0 && (UPPER ? y <= v : y < v);
    const uint32_t h2 = n1 >> 1;
    lo = right2 ? m2 + 1 : lo1;
    n = n1 == 0 ? 0u : (right2 ? n1 - h2 - 1 : h2);
  }
  if (n == 1 && (UPPER ? a[lo] <= v : a[lo] < v)) ++lo;
  return lo;
}

__device__ __forceinline__ uint32_t lower_bound_la(const uint32_t* __restrict__ a, uint32_t lo,
                                                   uint32_t hi, uint32_t v) {
  return bound_la<false>(a, lo, hi, v);
}

__device__ __forceinline__ uint32_t upper_bound_la(const uint32_t* __restrict__ a, uint32_t lo,
                                                   uint32_t hi, uint32_t v) {
  return bound_la<true>(a, lo, hi, v);
}

// Largest k in [0, n) with start[k] <= x (start has n+1 ascending entries).
__device__ __forceinline__ uint32_t find_slot(const uint64_t* __restrict__ start, uint32_t n,
                                              uint64_t x) {
  uint32_t lo = 0, hi = n;  // answer in [lo, hi)
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (start[mid] <= x) lo = mid;
    else hi = mid;
  }
  return lo;
}

// find_slot for a search that starts from a known slot: the largest k in
// [lo, n) with start[k] <= x, given start[lo] <= x.  It gallops from lo
// (lo+1, lo+2, lo+4, ...) before bisecting, so an answer d slots away costs
// about 2*log2(d) loads; consecutive batches of a slice move a few slots.
__device__ __forceinline__ uint32_t find_slot_from(const uint64_t* __restrict__ start, uint32_t lo,
                                                   uint32_t n, uint64_t x) {
  uint32_t step = 1;
  uint32_t hi = lo + 1;
  while (hi < n && start[hi] <= x) {
    lo = hi;
    step <<= 1;
    hi = (n - lo > step) ? lo + step : n;
  }
  if (hi > n) hi = n;
  while (hi - lo > 1) {  // answer in [lo, hi)
    const uint32_t mid = (lo + hi) >> 1;
    if (start[mid] <= x) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ const uint32_t* col(const IndexView& v, int c) {
  return v.cols + uint64_t(c) * v.stride;
}

// Position of key among the root values through the dense map (~0u when
// absent); v.root_pos must be set.
__device__ __forceinline__ uint32_t root_pos_of(const IndexView& v, uint32_t key) {
  return key <= v.root_max ? v.root_pos[key] : ~0u;
}

// Rows whose column 0 equals key (empty range if absent), via root arrays.
__device__ __forceinline__ Rng root_find(const IndexView& v, uint32_t key) {
  uint32_t g = lower_bound_la(v.root_vals, 0, v.r, key);
  if (g < v.r && v.root_vals[g] == key) return Rng{v.root_offs[g], v.root_offs[g + 1]};
  return Rng{0, 0};
}

// Rows within r whose column `depth` equals key; `last` says column `depth`
// is the index's last column, where values within a fixed prefix are unique.
__device__ __forceinline__ Rng child_find(const IndexView& v, int depth, Rng r, uint32_t key,
                                          bool last) {
  const uint32_t* c = col(v, depth);
  uint32_t lo = lower_bound_la(c, r.b, r.e, key);
  if (lo >= r.e || c[lo] != key) return Rng{0, 0};
  if (last) return Rng{lo, lo + 1};
  return Rng{lo, upper_bound_la(c, lo + 1, r.e, key)};
}

// Lexicographic comparison of row i of a against row j of b over k columns.
__device__ __forceinline__ int cmp_rows(const uint32_t* a, uint64_t sa, uint64_t i,
                                        const uint32_t* b, uint64_t sb, uint64_t j, int k) {
  for (int c = 0; c < k; ++c) {
    uint32_t x = a[uint64_t(c) * sa + i], y = b[uint64_t(c) * sb + j];
    if (x != y) return x < y ? -1 : 1;
  }
  return 0;
}

}  // namespace srd
