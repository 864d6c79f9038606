// Host-side API of the MI355X semi-naive runtime: device memory pool,
// owned sorted indexes and the stream-ordered set primitives the
// fixpoint loop is built from (sort+dedup, set difference, merge, scan).
//
// Counterparts in the reference runtime (runtime/generalized_datalog/):
//   Pool            gpu/device_array_rmm_impl.h:53-160 (RMM pool)
//   Index           gpu/device_sorted_array_index.h:337-643
//   build_index     gpu/device_sorted_array_index.ipp:263-553 + gpu/sort.h:297
//                   + gpu/deduplication.h:178-499
//   set_difference  gpu/device_sorted_array_index.ipp:966-1140
//   merge_disjoint  gpu/device_sorted_array_index.ipp:709-850
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "srd_device.h"

namespace srd {

[[noreturn]] void fail(const std::string& msg);
void check(hipError_t e, const char* what);
#define SRD_CHECK(x) ::srd::check((x), #x)

// ------------------------------------------------------------------ pool --
// Caching device allocator over hipMalloc: freed blocks are kept on a free
// list tagged with the stream that freed them and reused by later
// allocations on that stream (or on any stream once it is retired), so
// steady-state iterations and later sessions do not call into the driver.
// `budget` (bytes, 0 = unlimited) makes allocations beyond it throw, which
// is how admission control sees memory pressure.
struct PoolStats {
  uint64_t in_use = 0;         // bytes requested by live buffers
  uint64_t peak = 0;
  uint64_t budget = 0;
  uint64_t allocs = 0;         // allocations served
  uint64_t driver_allocs = 0;  // of which went to hipMalloc
  uint64_t reserved = 0;       // bytes obtained from hipMalloc so far
  uint64_t cached = 0;         // bytes held on the free list
  uint64_t trims = 0;          // times cached memory was handed back to the driver
  uint64_t cross_stream = 0;   // allocations served from another stream's free range
};

void pool_init(int device);
void* pool_alloc(uint64_t bytes, hipStream_t s);
void pool_free(void* p, uint64_t bytes, hipStream_t s);
// Synchronise s and let every block it freed serve any stream.
void pool_retire_stream(hipStream_t s);
// Declare s a stream that lives until the process ends: when an allocation
// on another stream finds no free range of its own, a range s freed serves
// it after that stream is made to wait for s's work so far (an event).
void pool_register_stream(hipStream_t s);
PoolStats pool_stats();
void pool_set_budget(uint64_t bytes);
void pool_reset_peak();

// Process-wide reuse of streams (default priority, non-blocking) and of
// pinned host blocks across engine sessions.
hipStream_t stream_acquire();
void stream_release(hipStream_t s);  // retires s in the pool first
void* pinned_acquire(uint64_t bytes, uint64_t* got);
void pinned_release(void* p, uint64_t bytes);

// Held while a stream capture is open (HIP graph capture of a device-
// resident iteration) and around the calls that would invalidate another
// thread's open capture: hipDeviceSynchronize / hipFree when the pool hands
// memory back, hipHostMalloc, hipMemGetInfo.  Engines of one process run on
// several host threads (serving), so one thread's capture must not overlap
// those calls in another.
std::recursive_mutex& capture_mutex();

// Runs in flight: run_enter() returns true when no other run is in flight.
// Only such a run may take inter-tile look-back order from the block index;
// runs that overlap another take their tile order from a ticket, so a tile
// never waits on one that has not been dispatched (see onesweep_kernel).
bool run_enter();
void run_exit();
struct RunScope {
  bool alone;
  RunScope() : alone(run_enter()) {}
  ~RunScope() { run_exit(); }
};

// Device buffer allocated from the pool.  Copies share the allocation (the
// last owner returns it to the pool on its stream): indexes are immutable
// once built, so FULL / DELTA / NEW can share columns without copying.
template <class T>
struct DevBuf {
  T* p = nullptr;
  uint64_t n = 0;
  hipStream_t s = nullptr;
  std::shared_ptr<void> owner;
  DevBuf() = default;
  DevBuf(uint64_t count, hipStream_t st) { alloc(count, st); }
  void alloc(uint64_t count, hipStream_t st) {
    release();
    s = st;
    n = count;
    if (count) {
      p = static_cast<T*>(pool_alloc(count * sizeof(T), st));
      const uint64_t bytes = count * sizeof(T);
      owner = std::shared_ptr<void>(p, [bytes, st](void* q) { pool_free(q, bytes, st); });
    }
  }
  // Grow to at least `count` elements (contents not preserved).
  void reserve(uint64_t count, hipStream_t st) {
    if (count > n) alloc(count + count / 4, st);
  }
  void release() {
    owner.reset();
    p = nullptr;
    n = 0;
  }
};

// ----------------------------------------------------------------- index --
// Sorted, duplicate-free k-column index with root (distinct column-0)
// arrays; see srd_device.h for the layout.
// Column-1 membership bitmaps of the keys of a two-column index whose
// ranges are long (built on demand for joins that search them; see
// build_heavy_bitmaps).
struct HeavyBits {
  DevBuf<uint32_t> bits;     // n x words
  DevBuf<uint32_t> starts;   // first row of each heavy key's range, ascending
  uint32_t n = 0, words = 0;
  uint32_t min_len = 0;                 // shortest range that has a bitmap
  const uint32_t* built_for = nullptr;  // cols buffer and row count they describe
  uint64_t built_n = 0;
};

// Dense map value -> position among an index's root values.
struct RootPos {
  DevBuf<uint32_t> pos;  // max + 1 entries
  uint32_t max = 0;
  bool ok = false;
  const uint32_t* built_for = nullptr;  // root_vals buffer and count it describes
  uint32_t built_r = 0;
};

struct Index {
  uint32_t arity = 0;
  uint64_t n = 0;
  uint64_t stride = 0;
  uint32_t r = 0;
  DevBuf<uint32_t> cols;       // arity * stride
  DevBuf<uint32_t> root_vals;  // r
  DevBuf<uint32_t> root_offs;  // r + 1
  std::vector<uint32_t> col_max;  // per-column maximum (0 when empty)
  mutable std::shared_ptr<HeavyBits> heavy;  // derived data, shared by copies
  mutable std::shared_ptr<RootPos> rootpos;

  IndexView view() const {
    IndexView v;
    v.cols = cols.p;
    v.root_vals = root_vals.p;
    v.root_offs = root_offs.p;
    v.stride = stride;
    v.n = uint32_t(n);
    v.r = r;
    v.arity = arity;
    const bool hb = heavy && heavy->n && heavy->built_for == cols.p && heavy->built_n == n;
    v.bm = hb ? heavy->bits.p : nullptr;
    v.hv_start = hb ? heavy->starts.p : nullptr;
    v.n_hv = hb ? heavy->n : 0;
    v.bm_words = hb ? heavy->words : 0;
    v.bm_min = hb ? heavy->min_len : ~0u;
    const bool rp = rootpos && rootpos->ok && rootpos->built_for == root_vals.p &&
                    rootpos->built_r == r;
    v.root_pos = rp ? rootpos->pos.p : nullptr;
    v.root_max = rp ? rootpos->max : 0;
    v.pad_ = 0;
    return v;
  }
  const uint32_t* col(uint32_t c) const { return cols.p + uint64_t(c) * stride; }
  uint32_t* col(uint32_t c) { return cols.p + uint64_t(c) * stride; }
};

inline uint64_t round_stride(uint64_t n) { return (n + 63) & ~uint64_t(63); }

// Reusable scratch for the primitives (grown on demand, kept across calls).
struct Workspace {
  DevBuf<uint64_t> keys[2][2];   // [ping/pong][word]
  DevBuf<uint64_t> lookback;     // onesweep inter-tile state
  DevBuf<uint32_t> small;        // histograms, counters, tile totals
  DevBuf<uint64_t> tile_scan;    // per-tile offsets
  DevBuf<uint8_t> flags;
  DevBuf<uint32_t> hostbuf_dev;  // staging for values read back
  uint32_t epoch = 0;            // lookback epoch (avoids per-pass clears)
  DevBuf<uint64_t> mp_lb;        // merge-path set-difference inter-tile state
  DevBuf<uint64_t> mp_split;     // merge-path tile split points
  DevBuf<uint32_t> mp_ctr;       // [tile ticket, error flag]
  DevBuf<uint64_t> mp_total;     // rows the set difference kept
  uint32_t mp_epoch = 0;
  // look-back kernels take tile order from a ticket (other runs may share
  // the GPU) instead of the block id; set per run from RunScope
  bool ticket_order = false;
  uint32_t* pinned = nullptr;    // pinned host words for read-backs
  uint64_t pinned_words = 0;
  uint8_t* stage_host = nullptr; // two pinned chunks for exports to pageable memory
  uint64_t stage_host_bytes = 0;
  ~Workspace();
};

uint32_t* pinned_words(Workspace& ws, uint64_t words);

// Sort + dedup the k input columns (column c at in[c], n rows, taken in
// the given order) into `out` (replaced).  `col_max` may give known upper
// bounds per column (empty: computed on the device, together with an
// order check).  Rows already strictly increasing skip the sort entirely;
// if `adopt` holds those same columns (column c at adopt->p + c*stride) the
// index takes the buffer over without copying.
// presorted: the caller guarantees the rows are strictly increasing and
// col_max holds per-column upper bounds; only the root arrays are built.
void build_index(const uint32_t* const* in, uint32_t k, uint64_t n, Index& out, Workspace& ws,
                 hipStream_t s, const std::vector<uint32_t>* col_max = nullptr,
                 const DevBuf<uint32_t>* adopt = nullptr, uint64_t adopt_stride = 0,
                 bool presorted = false);

// Same, with the source columns permuted: out column c = src column perm[c].
void build_index_from(const Index& src, const std::vector<uint32_t>& perm, Index& out,
                      Workspace& ws, hipStream_t s);

// Device-planned sort + dedup: build_index without host read-backs or
// allocations.  Rows: n_dev (device word, may be null) or n_host, at most
// `cap` (launch shapes come from cap).  w = key words (1 suffices for k <= 2;
// 2 for k <= 4 columns of any width).  scratch: devsort_scratch_bytes(cap)
// bytes, 256-byte aligned; its layout depends on cap alone, so a scratch
// reused across calls must always be given the same cap (its look-back
// words would otherwise overlap earlier calls' keys).  Outputs: out_cols (k columns, out_stride apart,
// cap rows), root_vals (cap), root_offs (cap + 1) and, if counts_out is not
// null, counts_out[0..2] = distinct rows, distinct column-0 values, error
// bits (1: a look-back did not resolve, 2: rows above cap, 4: key wider
// than w words).  The look-back epochs are kept in the scratch itself and
// advanced by a kernel, so a captured sort can be replayed; ticket: take tile order
// from tickets (required when other work may share the GPU); timed: record
// the launches in the kernel timing (false: no events are created).
uint64_t devsort_scratch_bytes(uint64_t cap);
void build_index_dev(const uint32_t* const* in, uint32_t k, uint64_t cap, const uint64_t* n_dev,
                     uint64_t n_host, uint32_t w, void* scratch, uint32_t* out_cols,
                     uint64_t out_stride, uint32_t* root_vals, uint32_t* root_offs,
                     uint64_t* counts_out, bool ticket, hipStream_t s, bool timed = true);
// Zero the look-back words and epoch of a devsort scratch (once, before its
// first use).
void devsort_clear_lookback(void* scratch, uint64_t cap, hipStream_t s);

// ------------------------------------------------ device-sized primitives --
// For device-resident iterations (a recursive step whose iterations are
// enqueued or captured without host read-backs): row counts are read from
// device words, launch shapes come from host capacities; nothing is read
// back or allocated.  Scratch sizes: scan_dev_words(entries) u64 words for a
// scan, mp_tiles_for(rows) merge-path tiles, roots_dev_tiles(rows) tiles.
uint64_t scan_dev_words(uint64_t cap);
// out[0..n] = exclusive prefix sums of in[0..n), n = *n_dev <= cap (n_dev
// null: n = cap), out[n] = total.
void exclusive_scan_u32_dev(const uint32_t* in, uint64_t* out, uint64_t cap, const uint64_t* n_dev,
                            uint64_t* tmp, hipStream_t s);
void exclusive_scan_u64_dev(const uint64_t* in, uint64_t* out, uint64_t cap, const uint64_t* n_dev,
                            uint64_t* tmp, hipStream_t s);
uint64_t mp_tiles_for(uint64_t rows);
uint64_t roots_dev_tiles(uint64_t cap);
struct SetDiffDevScratch {
  uint64_t* split;      // tiles + 1
  uint32_t* tmp;        // k columns of tmp_stride >= |a| rows
  uint64_t tmp_stride;
  uint32_t* cnt;        // tiles
  uint64_t* off;        // tiles + 1
  uint64_t* scan_tmp;   // scan_dev_words(tiles)
};
// out = a - b for sorted duplicate-free k-column inputs (k <= 4); nab =
// device words {|a|, |b|} with |a| + |b| <= cap_ab (tiles =
// mp_tiles_for(cap_ab)); *total_out = rows kept.
void set_difference_dev(uint32_t k, const uint32_t* a, uint64_t sa, const uint32_t* b, uint64_t sb,
                        uint64_t cap_ab, const uint64_t* nab, uint32_t* out, uint64_t so,
                        const SetDiffDevScratch& x, uint64_t* total_out, hipStream_t s);
// out = a u b for disjoint inputs, |a| + |b| (device words nab) <= cap_ab;
// split: mp_tiles_for(cap_ab) + 1 words.
void merge_dev(uint32_t k, const uint32_t* a, uint64_t sa, const uint32_t* b, uint64_t sb,
               uint64_t cap_ab, const uint64_t* nab, uint32_t* out, uint64_t so, uint64_t* split,
               hipStream_t s);
// Root arrays of the *n_dev sorted rows whose column 0 is c0 (capacity
// cap): *r_out = distinct values, root_offs[*r_out] = *n_dev.  pos: cap
// words; tile_r: roots_dev_tiles(cap); tile_off: tiles + 1;
// scan_tmp: scan_dev_words(tiles).
void build_roots_dev(const uint32_t* c0, uint64_t cap, const uint64_t* n_dev, uint32_t* pos,
                     uint32_t* tile_r, uint64_t* tile_off, uint64_t* scan_tmp, uint32_t* root_vals,
                     uint32_t* root_offs, uint64_t* r_out, hipStream_t s);

// out = a - b (both sorted and duplicate-free, same arity and column order).
void set_difference(const Index& a, const Index& b, Index& out, Workspace& ws, hipStream_t s);

// out = a u b for disjoint sorted inputs.
void merge_disjoint(const Index& a, const Index& b, Index& out, Workspace& ws, hipStream_t s);

// (Re)build root_vals / root_offs of an index whose columns are sorted.
void build_roots(Index& x, Workspace& ws, hipStream_t s);

// Copy (shares the immutable column / root buffers).
void copy_index(const Index& a, Index& out, hipStream_t s);

// out[0..n] = exclusive prefix sums of in[0..n) with out[n] = total.
void exclusive_scan_u32(const uint32_t* in, uint64_t* out, uint64_t n, Workspace& ws,
                        hipStream_t s);
void exclusive_scan_u64(const uint64_t* in, uint64_t* out, uint64_t n, Workspace& ws,
                        hipStream_t s);

// Dense-bitmap NEW (the joins' mark mode): bit `key` of the bitmap holds
// tuple (v_0 .. v_k-1) with key = sum(v_c << shift*(k-1-c)).  words must be
// a multiple of kDenseBlockWords (the allocation is rounded up to it and the
// tail stays clear); summary bit b (of words / kDenseBlockWords, in
// dense_sum_words u32) is set when block b (2^18 keys) holds a set bit.
constexpr uint64_t kDenseBlockWords = 8192;
inline uint64_t dense_sum_words(uint64_t words) { return (words / kDenseBlockWords + 31) / 32; }
// The tuples of the set bits, in key order (the index order: sorted and
// duplicate-free), written to `out` (allocated here, column c at
// out.p + c * stride, stride returned); only blocks marked in the summary
// are read, and the bitmap and summary are left clear.  Returns the number
// of tuples.
uint64_t dense_extract(uint32_t* bits, uint32_t* sum, uint64_t words, uint32_t k, uint32_t shift,
                       DevBuf<uint32_t>& out, uint64_t& stride, Workspace& ws, hipStream_t s);

// Fill n words with value (a kernel, safe under stream capture).
void fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t s);
void fill_u64(uint64_t* p, uint64_t value, uint64_t n, hipStream_t s);

// Move staged join segments to their final rows: record i (3 u64:
// slice, source row, local_offset<<32 | count) copies `count` rows of the k
// staging columns (stride stage_stride, dest d's columns starting at column
// d*8) to row out_base + offsets[slice] + local_offset of each dest.
struct GatherDest {
  uint32_t* cols;
  uint64_t stride;
  uint64_t base;
  uint32_t k;
};
void gather_segments(const uint64_t* recs, const uint32_t* rec_ctr, uint64_t rec_cap,
                     const uint32_t* stage, uint64_t stage_stride, const uint64_t* offsets,
                     const GatherDest* dests, uint32_t n_dests, uint64_t total, hipStream_t s);

// Streaming copy of `bytes` (reps launches) for bandwidth / counter calibration.
void calib_copy(uint64_t bytes, int reps, hipStream_t s, double* ms_per_rep);

// Copy index x to host memory as row-major rows (n x arity u32).  With an
// instance table (dev_bounds, n_bounds ascending column-0 range starts) every
// value of a row is rebased by its instance's start.
void export_rows(const Index& x, uint32_t* host_out, const uint32_t* dev_bounds,
                 uint32_t n_bounds, Workspace& ws, hipStream_t s);

// out[g] = rows of the column-0-sorted column col0 (n rows) with value in
// [bounds[g], bounds[g+1]), g < n_inst (bounds has n_inst+1 entries).
void instance_counts(const uint32_t* col0, uint64_t n, const uint32_t* bounds, uint32_t n_inst,
                     uint64_t* out, hipStream_t s);
// out[0] = rows of col0 below bounds[0], out[1] = rows at or above
// bounds[n_inst] (rows that belong to no instance).
void outside_counts(const uint32_t* col0, uint64_t n, const uint32_t* bounds, uint32_t n_inst,
                    uint64_t* out, hipStream_t s);

// Membership bitmaps for the keys of two-column index x whose ranges hold at
// least min_len rows (min_len doubled until they fit max_bytes): bitmap i
// has bit v set when row (key_i, v) exists.  Fills x.heavy.
void build_heavy_bitmaps(const Index& x, uint32_t min_len, uint64_t max_bytes, Workspace& ws,
                         hipStream_t s);

// Dense root-value map of x (x.rootpos) when its column-0 bound (col_max[0])
// gives at most max_entries entries; otherwise the map is marked absent.
void build_root_pos(const Index& x, uint64_t max_entries, hipStream_t s);

// Cost-balanced slice bounds: pref holds n+1 ascending item-cost prefix sums
// (pref[n] = w, every item costing at least 1); out[s] (s = 0..S) = first
// item i with pref[i] >= s*w/S, so out[0] = 0, out[S] = n and slice s =
// [out[s], out[s+1]) carries about w/S of the cost.
void slice_bounds(const uint64_t* pref, uint64_t n, uint64_t w, uint64_t S, uint64_t* out,
                  hipStream_t s);

// Synthetic columns (k fields of one 64-bit key per row, low k-1 fields
// fbits wide): key = mul*i + off when range == 0 (ascending, distinct for
// mul > 0), else a hash of i+off reduced mod range.
void fill_pattern(uint32_t* cols, uint64_t stride, uint64_t n, uint32_t k, uint32_t fbits,
                  uint64_t mul, uint64_t off, uint64_t range, hipStream_t s);

// Kernel timing hooks: when enabled, every primitive kernel launch is
// bracketed by events and its duration accumulated under its name.
struct KernelTime {
  std::string name;
  uint64_t launches = 0;
  double ms = 0;
  double bytes = 0;  // algorithmic bytes attributed by the caller
};
void timing_enable(bool on);
bool timing_enabled();
void timing_begin(const char* name, hipStream_t s);
void timing_end(hipStream_t s, double alg_bytes = 0);
std::vector<KernelTime> timing_collect();
void timing_reset();

}  // namespace srd
