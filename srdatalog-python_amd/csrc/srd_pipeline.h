// Contract between the runtime's MIR executor and the per-rule join kernels
// that the Python emitter generates for a program.
//
// A pipeline (one `execute-pipeline` MIR node) runs as:
//   root   one thread per distinct key of the driver root source: look the
//          key up in every root source, record the ranges, and size the
//          first nested level's iteration space (key_work);
//   then the per-key work is flattened into one global item space (an
//   exclusive scan of key_work) cut into slices -- `grain` items each, or,
//   after a weigh launch, of about equal level-2 cost -- that persistent
//   waves pull from per-XCD counters, so a slice may span many small keys
//   or a fraction of one huge key.  The traversal runs in one of these modes:
//   direct (3)  the rule emits exactly one tuple per item (last level is a
//               filter-free cartesian product or scan): tuple i goes to row
//               base+i of NEW, one pass, no counting;
//   fused  (2)  one pass; each slice appends its tuples to segments of a
//               staging buffer (bump-allocated with an atomic) and logs
//               (slice, local offset, source, count) records; after the
//               scan of the per-slice counts a gather kernel moves every
//               segment to its slice's offset in NEW;
//   count  (0)  counts only (count-only rules; fallback when staging
//               overflows);
//   write  (1)  re-runs the traversal writing at the scanned offsets;
//   probe  (5)  traverses like count but writes nothing except the number
//               of searches, root-map lookups and bitmap tests it made
//               (probe_ctr); run once, untimed, to report join probes;
//   weigh  (4)  rules with two or more nested levels: writes each level-1
//               item's cost (its level-2 candidates + SRD_ITEM_COST) so the
//               runtime can cut slices of equal cost instead of equal item
//               counts (slice_start); a few items of two hub keys can carry
//               more level-2 work than whole slices of ordinary ones.
// Output order depends only on the item order, never on which wave ran a
// slice, so every mode yields the same NEW rows in the same order.
//
// Reference counterpart: the JitRunner_<rule> phases setup / launch_count /
// scan_and_resize / launch_materialize of
// src/srdatalog/codegen/jit/complete_runner.py:929-1523 and the block-group
// work partition of complete_runner.py:377-736.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "srd_device.h"

#define SRD_MAX_VIEWS 8
#define SRD_MAX_COLS 8
#define SRD_MAX_DESTS 2

enum SrdVersion : uint32_t { SRD_FULL = 0, SRD_DELTA = 1, SRD_NEW = 2 };
enum SrdRootKind : uint32_t { SRD_ROOT_JOIN = 0, SRD_ROOT_SCAN = 1 };

struct SrdViewSpec {
  const char* rel;
  uint32_t ncols;
  uint32_t cols[SRD_MAX_COLS];  // index column order
  uint32_t ver;                 // SrdVersion
  uint32_t want_bitmaps;        // a nested level searches this two-column view at depth 1
  uint32_t want_root_pos;       // a nested level looks keys up among this view's root values
};

struct SrdDestSpec {
  const char* rel;
  uint32_t arity;
  uint32_t cols[SRD_MAX_COLS];  // dedup index (column order written to NEW)
};

struct SrdViews {
  srd::IndexView v[SRD_MAX_VIEWS];
};

struct SrdOut {
  uint32_t* cols;   // NEW buffer, column c at cols + c*stride
  uint64_t stride;
  uint64_t base;    // first row this launch writes
};

struct SrdDyn;

struct SrdLaunchArgs {
  SrdViews views;
  uint32_t n_root;             // driver root keys (join) or 1 (scan)
  uint32_t drv_view;           // view whose root_vals enumerate the keys
  uint32_t grain;              // level-1 items per slice
  uint32_t grid_blocks;        // persistent blocks for count / write
  uint64_t n_slices;           // ceil(total work / grain)
  uint64_t total_work;
  uint32_t* key_rng;           // per key: 2 u32 per root slot
  uint32_t* key_aux;           // per key: level-1 driver choice
  uint64_t* key_work;          // per key: level-1 item count
  uint64_t* work_start;        // exclusive scan of key_work, n_root+1 entries
  uint32_t* counts;            // per slice tuple count
  uint64_t* offsets;           // exclusive scan of counts
  uint32_t* work_ctr;          // SRD_XCDS slice counters, SRD_CTR_STRIDE words apart (zeroed per launch)
  SrdOut out[SRD_MAX_DESTS];
  // fused mode staging (column c of dest d at stage + (d*SRD_MAX_COLS + c)*stage_cap)
  uint32_t* stage;
  uint64_t stage_cap;          // tuples
  unsigned long long* stage_ctr;   // bump allocator (tuples)
  uint64_t* recs;              // 3 u64 per segment record
  uint64_t rec_cap;
  uint32_t* rec_ctr;
  uint32_t* overflow;          // set when staging or the record log ran out
  // cost-balanced slicing: slice s covers items [slice_start[s], slice_start[s+1])
  // (nullptr: [s*grain, min((s+1)*grain, total_work)))
  const uint64_t* slice_start;
  uint64_t* item_cost;         // weigh mode output, one entry per item
  uint64_t slice_lo;           // the launch takes slices [slice_lo, n_slices)
  uint32_t n_ranges;           // slice ranges, 1..SRD_XCDS (0 is taken as 1)
  uint32_t window_min;         // nested column joins narrow the driving range of items with
                               // at least this many candidates (0: never)
  // probe runs only (else nullptr): the root launch and the probe mode add
  // their lower/upper-bound searches, root-map lookups and bitmap tests
  unsigned long long* probe_ctr;
  // mark mode (6) only: the destination's NEW as a dense bitmap over its
  // key space, a tuple (v0 .. vk-1) in index order at bit
  // sum(v_c << mark_shift * (k-1-c)); mark_sum has a bit per block of
  // 2^18 keys, set when a bit of the block is set, so extraction reads only
  // marked blocks
  uint32_t* mark_bits;
  uint32_t* mark_sum;
  uint32_t mark_shift;
  uint32_t mark_pad_;
  // device-sized launches (the *_dyn entry points) only: views, key and
  // slice counts and output bases are read from this device struct, which
  // kernels of the same iteration fill in, instead of the fields above
  const SrdDyn* dyn;
};

// Device-resident part of a device-sized launch.  skip != 0 turns every
// launch reading it into a no-op (an iteration after convergence, or one
// abandoned because a buffer would overflow).
struct SrdDyn {
  SrdViews views;
  uint64_t total_work;
  uint64_t n_slices;
  uint64_t out_base[SRD_MAX_DESTS];
  uint32_t n_root;
  uint32_t drv_view;
  uint32_t skip;
  uint32_t pad_;
};

// Slices are split into SRD_XCDS contiguous ranges, one per XCD (workgroups
// are dispatched to the XCDs round-robin, so block b runs on XCD b % 8):
// the waves of one XCD work through neighbouring slices, whose keys (and
// their index ranges and bitmaps) then stay in that XCD's L2, and take
// slices of the other ranges once their own is done.
#define SRD_XCDS 8u
#define SRD_CTR_STRIDE 32u     // one 128-byte line per counter: atomics on one line serialise

#define SRD_ITEM_COST 4u       // level-1 cost of an item, in level-2 candidates
// Two-source cartesian products at a nested last level with at least this
// many tuples, and a source of at least SRD_CART_TILE_DIM rows, are walked in
// register tiles (one 64-row tile per lane register, the other broadcast a
// row at a time) instead of the flattened per-tuple decode.
#ifndef SRD_CART_TILE_MIN
#define SRD_CART_TILE_MIN 1024u
#endif
#ifndef SRD_CART_TILE_DIM
#define SRD_CART_TILE_DIM 48u
#endif
#define SRD_BM_MIN 512u        // default shortest range given a heavy-key bitmap

#define SRD_SEG 2048u          // staging segment size (tuples)

typedef void (*SrdLaunchFn)(const SrdLaunchArgs* args, hipStream_t stream);

struct SrdPipelineDesc {
  const char* rule;
  uint32_t n_views;
  SrdViewSpec views[SRD_MAX_VIEWS];
  uint32_t root_kind;                // SrdRootKind
  uint32_t n_root_src;               // root sources (join) / 1 (scan)
  uint32_t root_view[SRD_MAX_VIEWS];
  uint32_t n_dest;
  SrdDestSpec dest[SRD_MAX_DESTS];
  uint32_t count_only;               // rule carries .with_count()
  uint32_t default_grain;
  uint32_t direct_ok;                // one tuple per level-1 item (mode 3 valid)
  SrdLaunchFn root;
  SrdLaunchFn count;
  SrdLaunchFn write;
  SrdLaunchFn fused;
  SrdLaunchFn direct;
  SrdLaunchFn weigh;                 // nullptr unless the rule has >= 2 nested levels
  // The rule writes its single destination's rows strictly increasing in
  // the destination's index order: every level is a column join binding the
  // next index column, so tuples come out in (root key, level 1, level 2 ...)
  // order with no repeats, in every mode.
  uint32_t sorted_output;
  SrdLaunchFn probe;                 // mode 5 (search counting)
  // device-sized variants (SrdLaunchArgs::dyn set; grid_blocks fixed);
  // nullptr outside recursive steps, and direct_dyn unless direct_ok
  SrdLaunchFn root_dyn;
  SrdLaunchFn count_dyn;
  SrdLaunchFn write_dyn;
  SrdLaunchFn direct_dyn;
  // mode 6: set the bit of every derived tuple in the destination's dense
  // bitmap instead of writing rows; nullptr unless the rule has a single
  // destination whose terms are all variables
  SrdLaunchFn mark;
};

#ifdef __cplusplus
extern "C" {
#endif

struct srd_engine;

// Engine entry points used by the generated extern "C" shim of a program.
// program_text: relation declarations followed by the MIR program in the
// reference's S-expression form.
struct srd_engine* srd_engine_create(const char* program_text, const SrdPipelineDesc* pipelines,
                                     uint32_t n_pipelines);
void srd_engine_destroy(struct srd_engine* e);
const char* srd_engine_last_error(struct srd_engine* e);

#ifdef __cplusplus
}
#endif
