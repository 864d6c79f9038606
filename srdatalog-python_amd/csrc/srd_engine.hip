// MIR executor of the MI355X runtime.
//
// A program arrives as text: relation declarations followed by the MIR
// program in the reference's S-expression form (the format of the
// reference's tests/fixtures/**/*.mir.sexpr goldens), plus a table of
// generated per-rule pipeline kernels (srd_pipeline.h).  The engine keeps
// every relation version (FULL / DELTA / NEW) as sorted u32 column indexes
// in HBM and executes the MIR ops with the device primitives of
// srd_prims.hip.
//
// Semantics follow the reference host orchestration:
//   recursive step loop   src/srdatalog/codegen/jit/orchestrator_jit.py:600-739
//                         (stop at iter>0 when the summed canonical DELTA
//                         size is 0)
//   non-recursive step    orchestrator_jit.py:747-855 (one pass)
//   iteration-0 DELTA     reads FULL (complete_runner.py:834-838)
//   rebuild-index         gpu_mir_helpers.h:203 (sort+dedup NEW)
//   compute-delta-index   gpu_mir_helpers.h:87 (DELTA = NEW - FULL)
//   merge-index           gpu_mir_helpers.h:135 (FULL u= DELTA)
//   rebuild-index-from-index  gpu_mir_helpers.h:250
//   clear-relation        gpu_mir_helpers.h:307
//   count-only rules      orchestrator_jit.py:156-172 (count printed, no
//                         materialisation)
//   CSV input             runtime/io.h:178-395 (delimiter auto-detect)
#include <fcntl.h>
#include <hip/hip_runtime.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <iostream>
#include <map>
#include <set>
#include <memory>
#include <string>
#include <vector>

#include "srd_devloop.h"
#include "srd_engine.h"
#include "srd_pipeline.h"
#include "srd_runtime.h"

namespace srd {

// ============================================================ S-expr ====
struct SNode {
  bool atom = false;
  bool quoted = false;
  std::string text;
  std::vector<SNode> kids;

  bool is_sym(const char* s) const { return atom && !quoted && text == s; }
  const SNode* kw(const char* key) const {  // value after :key
    for (size_t i = 0; i + 1 < kids.size(); ++i)
      if (kids[i].atom && !kids[i].quoted && kids[i].text == key) return &kids[i + 1];
    return nullptr;
  }
  std::string head() const { return (!atom && !kids.empty() && kids[0].atom) ? kids[0].text : ""; }
};

class SParser {
 public:
  explicit SParser(const std::string& t) : s_(t) {}
  std::vector<SNode> parse_all() {
    std::vector<SNode> out;
    skip();
    while (i_ < s_.size()) {
      out.push_back(parse());
      skip();
    }
    return out;
  }

 private:
  const std::string& s_;
  size_t i_ = 0;
  void skip() {
    while (i_ < s_.size()) {
      if (isspace((unsigned char)s_[i_])) {
        ++i_;
      } else if (s_[i_] == ';') {
        while (i_ < s_.size() && s_[i_] != '\n') ++i_;
      } else {
        break;
      }
    }
  }
  SNode parse() {
    skip();
    if (i_ >= s_.size()) fail("mir: unexpected end of program text");
    SNode n;
    if (s_[i_] == '(') {
      ++i_;
      skip();
      while (i_ < s_.size() && s_[i_] != ')') {
        n.kids.push_back(parse());
        skip();
      }
      if (i_ >= s_.size()) fail("mir: unbalanced parenthesis");
      ++i_;
      return n;
    }
    if (s_[i_] == ')') fail("mir: unexpected ')'");
    n.atom = true;
    if (s_[i_] == '"') {
      n.quoted = true;
      ++i_;
      while (i_ < s_.size() && s_[i_] != '"') {
        if (s_[i_] == '\\' && i_ + 1 < s_.size()) ++i_;
        n.text.push_back(s_[i_++]);
      }
      ++i_;
      return n;
    }
    while (i_ < s_.size() && !isspace((unsigned char)s_[i_]) && s_[i_] != '(' && s_[i_] != ')')
      n.text.push_back(s_[i_++]);
    return n;
  }
};

static std::vector<uint32_t> int_list(const SNode* n) {
  std::vector<uint32_t> v;
  if (!n) return v;
  if (n->atom) {
    v.push_back(uint32_t(std::stoul(n->text)));
    return v;
  }
  for (auto& k : n->kids) v.push_back(uint32_t(std::stoul(k.text)));
  return v;
}

static uint32_t parse_ver(const SNode* n) {
  if (!n) fail("mir: missing :ver");
  if (n->text == "FULL") return SRD_FULL;
  if (n->text == "DELTA") return SRD_DELTA;
  if (n->text == "NEW") return SRD_NEW;
  fail("mir: unknown version " + n->text);
}

// ============================================================= model ====
enum class OpKind {
  Pipeline,
  Parallel,
  RebuildIndex,
  ClearRelation,
  CheckSize,
  ComputeDelta,
  MergeIndex,
  RebuildFromIndex,
  FlatView,
  Reconstruct,
  Hook,
};

struct Op {
  OpKind kind;
  int rel = -1;
  std::vector<uint32_t> index, index2;
  uint32_t ver = SRD_FULL;
  std::string rule;
  const SrdPipelineDesc* pipe = nullptr;
  std::vector<Op> kids;
};

struct Step {
  bool recursive = false;
  std::vector<Op> ops;
  std::vector<int> canonical_rels;  // relations with compute-delta in the plan
  // relations of a recursive step whose FULL no pipeline of the step reads:
  // their canonical FULL is kept as sorted segments while the step loops
  std::vector<int> segmented_rels;
};

using Order = std::vector<uint32_t>;

struct Versioned {
  Index idx;
  uint64_t version = 0;
};

struct Relation {
  std::string name;
  uint32_t arity = 0;
  std::string input_file;
  bool print_size = false;
  Order canonical;
  bool canonical_set = false;
  // FULL: indexes with the FULL version they reflect.
  std::map<Order, Versioned> full;
  uint64_t full_version = 0;   // deltas merged into the canonical FULL index
  uint64_t delta_seq = 0;      // version FULL reaches once the current DELTA is merged
  std::map<Order, Index> delta;
  std::map<Order, Index> newv;
  // NEW raw rows, columns in raw_order
  DevBuf<uint32_t> raw;
  uint64_t raw_n = 0, raw_cap = 0;
  Order raw_order;
  // raw rows written by one sorted-output pipeline (strictly increasing in
  // raw_order), with per-column upper bounds: the index adopts them as is
  bool raw_sorted = false;
  std::vector<uint32_t> raw_bound;
  // input data (EDB)
  std::vector<uint32_t> host_rows;  // row-major
  uint64_t host_n = 0;
  bool has_input = false;
  DevBuf<uint32_t> dev_input;       // columns, stride dev_stride
  uint64_t dev_stride = 0;
  bool dev_input_valid = false;
  std::vector<uint32_t> input_max;
  // results
  std::vector<uint64_t> deltas;     // per-iteration |DELTA| of the last step computing it
  int deltas_step = -1;
  // with an instance table: per iteration, |DELTA| rows of every instance
  std::vector<DevBuf<uint64_t>> inst_deltas;
  // segmented canonical FULL: full[canonical] is the oldest segment, segs
  // the younger ones (pairwise disjoint, sizes roughly halving)
  std::vector<Index> segs;
  bool segmented = false;
  // NEW rows of pipelines too large to materialise at once, written in
  // slice chunks, each sorted and deduplicated (all in new_parts_order)
  std::vector<Index> new_parts;
  Order new_parts_order;
  // FULL in relation column order for export, valid while export_version
  // equals full_version (the size query and the copy share one re-sort)
  Index export_idx;
  uint64_t export_version = ~uint64_t(0);
  uint64_t count = 0;               // count-only result
  bool counted = false;
  // NEW as a bitmap over the relation's key space while this iteration's
  // pipelines mark it (their mode 6): tuple (v_0 .. v_k-1) in dense_order is
  // bit sum(v_c << dense_shift*(k-1-c)).  new_index, or a pipeline that
  // writes rows instead, turns the set bits into a NEW part; the bitmap is
  // kept (clear) for the next iteration.
  DevBuf<uint32_t> dense_bits;
  DevBuf<uint32_t> dense_sum;  // a bit per block of the bitmap holding a set bit
  uint64_t dense_words = 0;
  uint32_t dense_shift = 0;
  uint32_t dense_bound = 0;
  bool dense_on = false;
  Order dense_order;
};

// Device scratch of one pipeline run (root keys, slices, fused staging,
// weighed slices, probe counter), grown on demand and kept across runs.
struct PipeScratch {
  DevBuf<uint32_t> key_rng, key_aux, counts_buf, work_ctr;
  DevBuf<uint64_t> key_work, work_start, offsets;
  DevBuf<uint32_t> stage, overflow;
  DevBuf<uint64_t> recs, stage_ctr;
  DevBuf<uint64_t> item_cost, item_pref, slice_start;
  DevBuf<uint64_t> probe_buf;
};

}  // namespace srd

struct srd_engine {
  std::vector<srd::Relation> rels;
  std::map<std::string, int> rel_by_name;
  std::vector<srd::Step> steps;
  std::map<std::string, const SrdPipelineDesc*> pipes;
  hipStream_t s = nullptr;
  srd::Workspace ws;
  std::string err;
  bool quiet = false;
  bool ran = false;
  uint64_t iterations = 0;         // iterations of the last recursive step
  // resumable run (srd_engine_run_begin): the program's last step, its only
  // recursive one, advanced a few iterations per call
  bool resumable = false;
  size_t res_step = 0;
  uint64_t res_it = 0;
  double run_seconds = 0;
  std::vector<std::pair<std::string, uint64_t>> counts;  // rule -> count (count-only)
  uint32_t grain_override = 0;
  // per-pipeline scratch (one per member of the widest parallel group),
  // grown on demand, and the streams, fences and pinned words of groups
  std::vector<std::unique_ptr<srd::PipeScratch>> scratch;
  std::vector<hipStream_t> group_s;
  std::vector<hipEvent_t> fence_events;
  uint64_t* run_pins = nullptr;
  uint64_t run_pins_bytes = 0;
  bool group_streams = true;       // SRDATALOG_GROUP_STREAMS=0: a group's members run in order
  // fused-mode staging is sized from the largest output each rule produced
  std::map<std::string, uint64_t> stage_hint;
  bool two_pass = false;           // SRDATALOG_JOIN_MODE=twopass: count + write only
  bool no_direct = false;          // SRDATALOG_JOIN_MODE=fused: never the direct mode
  uint64_t stage_cap_env = 0;      // SRDATALOG_STAGE_CAP: fixed staging rows (tests overflow)
  uint64_t stage_max_bytes = uint64_t(16) << 30;  // SRDATALOG_STAGE_MAX_BYTES
  uint64_t new_chunk_rows = 0;     // SRDATALOG_NEW_CHUNK_ROWS: fixed NEW chunk (tests)
  // cost-balanced slicing of rules with >= 2 nested levels
  bool no_weigh = false;           // SRDATALOG_NO_WEIGH=1: uniform slices only
  // fewer items: uniform slices.  Low on purpose: a rule with few level-1
  // items can still carry most of an iteration's work (Andersen's Store rule
  // at small scales: ~13 uniform slices for seconds of join)
  uint64_t weigh_min = 64;         // SRDATALOG_WEIGH_MIN
  // smaller slices keep fewer keys (and their bitmaps) active per XCD at a
  // time; 8192 candidates and up to 48 slices per wave measured best on the
  // triangle join (16384/24: +0.8 ms, 2048/192: +4.5 ms)
  uint64_t weigh_target = 16384;   // SRDATALOG_WEIGH_TARGET: cost (candidates) per slice
  uint64_t slices_per_wave = 48;   // SRDATALOG_SLICES_PER_WAVE: cap of the weighed slices
  uint64_t min_slices_per_wave = 4;  // SRDATALOG_MIN_SLICES_PER_WAVE: floor of the weighed slices
  uint32_t xcd_ranges = SRD_XCDS;  // SRDATALOG_XCD_RANGES: slice ranges of the join launches
  // SRDATALOG_WINDOW_MIN: a nested column join narrows an item's driving
  // range to the values the other sources can match (between the largest
  // first and the smallest last value of their ranges) when the item has at
  // least this many candidates, two searches per item (0: never)
  uint32_t window_min = 64;
  // SRDATALOG_DENSE_MAX_BITS: a destination whose key space (arity x the bit
  // width of the largest value the engine holds) has at most this many keys
  // takes its NEW as a bitmap the joins mark, instead of raw rows sorted to
  // remove repeats (0: never).  2^36 keys are 8 GiB of HBM per relation,
  // cleared once per engine; extraction reads only the blocks of 2^18 keys
  // that were marked.  All of an engine's bitmaps together stay within an
  // eighth of the device's memory.
  uint64_t dense_max_bits = uint64_t(1) << 36;
  uint64_t dense_bytes = 0;        // bitmaps allocated so far
  uint64_t mark_chunk_slices = 0;  // SRDATALOG_MARK_CHUNK_SLICES (0: one launch per traversal)
  uint32_t weigh_ranges = SRD_XCDS;  // SRDATALOG_WEIGH_RANGES: of the weigh launch
  // column-1 bitmaps of heavy keys for views nested levels search
  uint32_t bitmap_min = SRD_BM_MIN;        // SRDATALOG_BITMAP_MIN
  bool root_pos = true;                    // SRDATALOG_ROOT_POS=0: search root values instead
  // SRDATALOG_BITMAP_BYTES (0: none).  The bitmaps are probed at random, so
  // their total is held to a few hundred MB, the size of the last-level
  // cache: on the R-MAT scale-20 triangle join 308 MB of bitmaps (threshold
  // 1024 rows) ran 8 % faster than 571 MB (512 rows).
  uint64_t bitmap_bytes = uint64_t(384) << 20;
  uint64_t join_probes = 0;        // tuples emitted by joins (pre-dedup)
  // probe runs (srdatalog_count_probes): every pipeline also counts its
  // searches, root-map lookups and bitmap tests in an extra, untimed-by-the-
  // caller launch of the probe mode
  bool count_probes = false;
  uint64_t join_searches = 0;
  bool trace = false;              // SRDATALOG_TRACE: one stderr line per op
  bool op_profile = false;         // SRDATALOG_OP_PROFILE: host time per op, printed per run
  std::map<std::string, std::pair<double, uint64_t>> op_times;
  // instance table of a packed batch: n_inst+1 ascending column-0 bounds
  srd::DevBuf<uint32_t> inst_bounds;
  uint32_t n_inst = 0;
  // device-resident iterations of eligible recursive steps
  bool devloop = true;             // SRDATALOG_DEVLOOP=0: host-driven loop everywhere
  bool devloop_graph = true;       // SRDATALOG_DEVLOOP_GRAPH=0: enqueue without capturing
  uint64_t devloop_batch = 4;      // SRDATALOG_DEVLOOP_BATCH: iterations per control read
  uint64_t devloop_caps = 0;       // SRDATALOG_DEVLOOP_CAPS: initial capacities (tests)
  // SRDATALOG_DEVLOOP_MAX_BYTES: a loop that would reserve more hands its
  // state to the host loop (default: a quarter of the device memory)
  uint64_t devloop_max_bytes = 0;
  // SRDATALOG_DEVLOOP_MAX_ROWS: a step whose FULL or NEW outgrows this many
  // rows continues on the host loop.  Its iterations then run long enough
  // that the host loop's read-backs cost little, while growing the device
  // loop's capacities costs a re-run of the iteration and fresh buffers.
  uint64_t devloop_max_rows = uint64_t(1) << 24;
};

namespace srd {

// =========================================================== parsing ====
static int rel_id(srd_engine& e, const std::string& name) {
  auto it = e.rel_by_name.find(name);
  if (it == e.rel_by_name.end()) fail("mir: unknown relation " + name);
  return it->second;
}

static void parse_index_ref(srd_engine& e, const SNode* n, int& rel, Order& idx) {
  if (!n || n->atom || n->kids.empty()) fail("mir: bad :index form");
  rel = rel_id(e, n->kids[0].text);
  idx.clear();
  for (size_t i = 1; i < n->kids.size(); ++i) idx.push_back(uint32_t(std::stoul(n->kids[i].text)));
}

static Op parse_op(srd_engine& e, const SNode& n) {
  Op op;
  std::string h = n.head();
  if (h == "execute-pipeline") {
    op.kind = OpKind::Pipeline;
    const SNode* r = n.kw(":rule");
    if (!r) fail("mir: execute-pipeline without :rule");
    op.rule = r->text;
    auto it = e.pipes.find(op.rule);
    if (it == e.pipes.end()) fail("mir: no compiled kernel for rule " + op.rule);
    op.pipe = it->second;
  } else if (h == "parallel-group") {
    op.kind = OpKind::Parallel;
    for (size_t i = 1; i < n.kids.size(); ++i)
      if (!n.kids[i].atom) op.kids.push_back(parse_op(e, n.kids[i]));
  } else if (h == "rebuild-index") {
    op.kind = OpKind::RebuildIndex;
    parse_index_ref(e, n.kw(":index"), op.rel, op.index);
    op.ver = parse_ver(n.kw(":ver"));
  } else if (h == "clear-relation") {
    op.kind = OpKind::ClearRelation;
    op.rel = rel_id(e, n.kw(":schema")->text);
    op.ver = parse_ver(n.kw(":ver"));
  } else if (h == "check-size") {
    op.kind = OpKind::CheckSize;
    op.rel = rel_id(e, n.kw(":schema")->text);
  } else if (h == "compute-delta-index") {
    op.kind = OpKind::ComputeDelta;
    op.rel = rel_id(e, n.kw(":schema")->text);
    op.index = int_list(n.kw(":canonical-index"));
  } else if (h == "merge-index") {
    op.kind = OpKind::MergeIndex;
    parse_index_ref(e, n.kw(":index"), op.rel, op.index);
  } else if (h == "rebuild-index-from-index") {
    op.kind = OpKind::RebuildFromIndex;
    int r2;
    parse_index_ref(e, n.kw(":source"), op.rel, op.index);
    parse_index_ref(e, n.kw(":target"), r2, op.index2);
    if (r2 != op.rel) fail("mir: rebuild-index-from-index across relations");
    op.ver = parse_ver(n.kw(":ver"));
  } else if (h == "create-flat-view") {
    op.kind = OpKind::FlatView;
    parse_index_ref(e, n.kw(":index"), op.rel, op.index);
    op.ver = parse_ver(n.kw(":ver"));
  } else if (h == "post-stratum-reconstruct-intern-cols") {
    op.kind = OpKind::Reconstruct;
    op.rel = rel_id(e, n.kw(":rel")->text);
    op.index = int_list(n.kw(":canonical-index"));
  } else if (h == "inject-cpp-hook") {
    op.kind = OpKind::Hook;
  } else {
    fail("mir: unsupported op " + h);
  }
  return op;
}

static void collect_canonical(srd_engine& e, const std::vector<Op>& ops, Step& st) {
  for (auto& op : ops) {
    if (op.kind == OpKind::ComputeDelta) {
      if (std::find(st.canonical_rels.begin(), st.canonical_rels.end(), op.rel) ==
          st.canonical_rels.end())
        st.canonical_rels.push_back(op.rel);
      Relation& r = e.rels[op.rel];
      if (!r.canonical_set) {
        r.canonical = op.index;
        r.canonical_set = true;
      }
    }
    if (op.kind == OpKind::Parallel) collect_canonical(e, op.kids, st);
  }
}

// FULL readers of a step: pipeline views at FULL, and merges into a
// non-canonical order.  A relation with compute-delta and no such reader can
// keep its canonical FULL segmented while the step loops.
static void full_readers(srd_engine& e, const std::vector<Op>& ops, std::vector<int>& rd) {
  for (auto& op : ops) {
    if (op.kind == OpKind::Pipeline && op.pipe) {
      for (uint32_t i = 0; i < op.pipe->n_views; ++i)
        if (op.pipe->views[i].ver == SRD_FULL) rd.push_back(rel_id(e, op.pipe->views[i].rel));
    }
    if (op.kind == OpKind::MergeIndex && op.index != e.rels[op.rel].canonical) rd.push_back(op.rel);
    if (op.kind == OpKind::Parallel) full_readers(e, op.kids, rd);
  }
}

static void collect_segmented(srd_engine& e, Step& st) {
  std::vector<int> rd;
  full_readers(e, st.ops, rd);
  for (int ri : st.canonical_rels)
    if (std::find(rd.begin(), rd.end(), ri) == rd.end()) st.segmented_rels.push_back(ri);
}

static void parse_program(srd_engine& e, const std::string& text) {
  auto top = SParser(text).parse_all();
  for (auto& n : top) {
    if (n.head() == "relations") {
      for (size_t i = 1; i < n.kids.size(); ++i) {
        const SNode& d = n.kids[i];
        Relation r;
        r.name = d.kw(":name")->text;
        r.arity = uint32_t(std::stoul(d.kw(":arity")->text));
        if (auto* f = d.kw(":input-file")) r.input_file = f->text;
        if (auto* p = d.kw(":print-size")) r.print_size = p->text == "true";
        for (uint32_t c = 0; c < r.arity; ++c) r.canonical.push_back(c);
        e.rel_by_name[r.name] = int(e.rels.size());
        e.rels.push_back(std::move(r));
      }
    }
  }
  for (auto& n : top) {
    if (n.head() != "program") continue;
    for (size_t i = 1; i < n.kids.size(); ++i) {
      const SNode& sn = n.kids[i];
      if (sn.head() != "step") fail("mir: expected (step ...)");
      Step st;
      if (auto* r = sn.kw(":recursive")) st.recursive = r->text == "true";
      for (size_t j = 1; j < sn.kids.size(); ++j) {
        const SNode& body = sn.kids[j];
        if (body.atom) continue;
        if (body.head() == "fixpoint-plan") {
          for (size_t k = 1; k < body.kids.size(); ++k)
            if (!body.kids[k].atom) st.ops.push_back(parse_op(e, body.kids[k]));
        } else {
          st.ops.push_back(parse_op(e, body));
        }
      }
      collect_canonical(e, st.ops, st);
      if (st.recursive && !std::getenv("SRDATALOG_NO_SEGMENTS")) collect_segmented(e, st);
      e.steps.push_back(std::move(st));
    }
  }
}

// ========================================================= relations ====
static Order perm_between(const Order& from, const Order& to) {
  // out column c of `to` order = column perm[c] of `from` order
  Order p(to.size());
  for (size_t c = 0; c < to.size(); ++c) {
    auto it = std::find(from.begin(), from.end(), to[c]);
    if (it == from.end()) fail("index orders are not permutations of each other");
    p[c] = uint32_t(it - from.begin());
  }
  return p;
}

static Index empty_index(uint32_t arity) {
  Index x;
  x.arity = arity;
  x.col_max.assign(arity, 0);
  return x;
}

static void upload_input(srd_engine& e, Relation& r) {
  if (r.dev_input_valid || !r.has_input) return;
  const uint64_t n = r.host_n, k = r.arity;
  r.dev_stride = round_stride(n);
  std::vector<uint32_t> t(r.dev_stride * k, 0);
  r.input_max.assign(k, 0);
  for (uint64_t i = 0; i < n; ++i)
    for (uint64_t c = 0; c < k; ++c) {
      uint32_t v = r.host_rows[i * k + c];
      t[c * r.dev_stride + i] = v;
      r.input_max[c] = std::max(r.input_max[c], v);
    }
  r.dev_input.alloc(std::max<uint64_t>(t.size(), 1), e.s);
  if (!t.empty())
    SRD_CHECK(hipMemcpyAsync(r.dev_input.p, t.data(), t.size() * 4, hipMemcpyHostToDevice, e.s));
  SRD_CHECK(hipStreamSynchronize(e.s));
  r.dev_input_valid = true;
}

// FULL index in `order`, current with the relation's FULL version.
// Merge a segmented canonical FULL back into one index (smallest segments
// first).
static void materialize_full(srd_engine& e, Relation& r) {
  if (r.segs.empty()) return;
  Index acc = r.segs.back();
  for (size_t i = r.segs.size() - 1; i-- > 0;) {
    Index m;
    merge_disjoint(r.segs[i], acc, m, e.ws, e.s);
    acc = std::move(m);
  }
  r.segs.clear();
  auto& base = r.full[r.canonical];
  Index m;
  merge_disjoint(base.idx, acc, m, e.ws, e.s);
  base.idx = std::move(m);
}

static const Index& full_index(srd_engine& e, Relation& r, const Order& order) {
  if (!r.segs.empty()) materialize_full(e, r);
  auto it = r.full.find(order);
  if (it != r.full.end() && it->second.version == r.full_version) return it->second.idx;
  Versioned v;
  v.version = r.full_version;
  auto cit = r.full.find(r.canonical);
  if (order != r.canonical && cit != r.full.end() && cit->second.version == r.full_version) {
    build_index_from(cit->second.idx, perm_between(r.canonical, order), v.idx, e.ws, e.s);
  } else if (r.has_input && r.full_version == 0) {
    upload_input(e, r);
    std::vector<const uint32_t*> in(r.arity);
    bool identity = true;
    for (uint32_t c = 0; c < r.arity; ++c) {
      in[c] = r.dev_input.p + uint64_t(order[c]) * r.dev_stride;
      identity = identity && order[c] == c;
    }
    // Input already sorted and duplicate-free in this order is adopted as is
    // (the resident input is immutable, so the index may share it).
    if (r.host_n)
      build_index(in.data(), r.arity, r.host_n, v.idx, e.ws, e.s, nullptr,
                  identity ? &r.dev_input : nullptr, r.dev_stride);
    else
      v.idx = empty_index(r.arity);
  } else if (r.full_version == 0 && !r.has_input) {
    v.idx = empty_index(r.arity);
  } else {
    fail("relation " + r.name + ": FULL index requested before its canonical index exists");
  }
  auto& slot = r.full[order];
  slot = std::move(v);
  return slot.idx;
}

static const Index& delta_index(srd_engine& e, Relation& r, const Order& order) {
  auto it = r.delta.find(order);
  if (it != r.delta.end()) return it->second;
  auto cit = r.delta.find(r.canonical);
  Index x;
  if (cit != r.delta.end())
    build_index_from(cit->second, perm_between(r.canonical, order), x, e.ws, e.s);
  else
    x = empty_index(r.arity);
  auto& slot = r.delta[order];
  slot = std::move(x);
  return slot;
}

static void union_into(srd_engine& e, Index& a, const Index& b);
static void dense_flush(srd_engine& e, Relation& r);

static const Index& new_index(srd_engine& e, Relation& r, const Order& order) {
  auto it = r.newv.find(order);
  if (it != r.newv.end()) return it->second;
  dense_flush(e, r);  // marked tuples join the NEW parts
  Index x;
  if (!r.new_parts.empty()) {
    // chunked pipelines' parts, united with the rows other pipelines wrote
    const Order po = r.new_parts_order;
    Index acc;
    if (r.raw_n > 0) {
      Order p = perm_between(r.raw_order, po);
      std::vector<const uint32_t*> in(r.arity);
      for (uint32_t c = 0; c < r.arity; ++c) in[c] = r.raw.p + uint64_t(p[c]) * r.raw_cap;
      build_index(in.data(), r.arity, r.raw_n, acc, e.ws, e.s);
    } else {
      acc = empty_index(r.arity);
    }
    for (auto& part : r.new_parts) union_into(e, acc, part);
    r.new_parts.clear();
    r.raw.release();
    r.raw_n = r.raw_cap = 0;
    r.raw_sorted = false;
    if (order == po) {
      x = std::move(acc);
    } else {
      r.newv[po] = acc;
      build_index_from(acc, perm_between(po, order), x, e.ws, e.s);
    }
  } else if (r.raw_n > 0) {
    Order p = perm_between(r.raw_order, order);
    std::vector<const uint32_t*> in(r.arity);
    for (uint32_t c = 0; c < r.arity; ++c) in[c] = r.raw.p + uint64_t(p[c]) * r.raw_cap;
    // NEW rows that a join emitted already in index order are adopted
    const bool pre = r.raw_sorted && r.raw_order == order && r.raw_bound.size() == r.arity;
    build_index(in.data(), r.arity, r.raw_n, x, e.ws, e.s, pre ? &r.raw_bound : nullptr,
                r.raw_order == order ? &r.raw : nullptr, r.raw_cap, pre);
  } else if (!r.newv.empty()) {
    auto& any = *r.newv.begin();
    build_index_from(any.second, perm_between(any.first, order), x, e.ws, e.s);
  } else {
    x = empty_index(r.arity);
  }
  auto& slot = r.newv[order];
  slot = std::move(x);
  return slot;
}

// Make room for `extra` more NEW rows (columns in `order`); returns base row.
static uint64_t reserve_new(srd_engine& e, Relation& r, const Order& order, uint64_t extra) {
  if (r.raw_n == 0 && r.raw_order != order) r.raw_order = order;
  if (r.raw_order != order) fail("relation " + r.name + ": pipelines insert with different dedup indexes");
  uint64_t need = r.raw_n + extra;
  if (need > r.raw_cap) {
    uint64_t cap = std::max<uint64_t>(round_stride(need + need / 2), 1024);
    DevBuf<uint32_t> nb(cap * r.arity, e.s);
    if (r.raw_n)
      for (uint32_t c = 0; c < r.arity; ++c)
        SRD_CHECK(hipMemcpyAsync(nb.p + uint64_t(c) * cap, r.raw.p + uint64_t(c) * r.raw_cap,
                                 r.raw_n * 4, hipMemcpyDeviceToDevice, e.s));
    r.raw = std::move(nb);
    r.raw_cap = cap;
  }
  uint64_t base = r.raw_n;
  r.raw_n = need;
  return base;
}

static void clear_new(Relation& r) {
  if (r.dense_on) {  // marks nobody read: clear them for the next iteration
    fill_u32(r.dense_bits.p, 0, r.dense_words, r.dense_bits.s);
    fill_u32(r.dense_sum.p, 0, dense_sum_words(r.dense_words), r.dense_sum.s);
    r.dense_on = false;
  }
  r.newv.clear();
  r.raw.release();
  r.raw_n = r.raw_cap = 0;
  r.raw_sorted = false;
  r.new_parts.clear();
}

// a := a u b for sorted duplicate-free indexes of one order.
static void union_into(srd_engine& e, Index& a, const Index& b) {
  if (b.n == 0) return;
  if (a.n == 0) {
    copy_index(b, a, e.s);
    return;
  }
  Index d, m;
  set_difference(b, a, d, e.ws, e.s);
  if (d.n == 0) return;
  merge_disjoint(a, d, m, e.ws, e.s);
  a = std::move(m);
}

// Add a deduplicated chunk of NEW rows; the two youngest parts are united
// while within a factor 2 of each other, so few parts are kept.
static void add_new_part(srd_engine& e, Relation& r, const Order& order, Index&& part) {
  if (r.new_parts.empty()) r.new_parts_order = order;
  if (order != r.new_parts_order) fail("relation " + r.name + ": NEW chunks in different orders");
  r.new_parts.push_back(std::move(part));
  while (r.new_parts.size() >= 2) {
    Index& older = r.new_parts[r.new_parts.size() - 2];
    if (older.n > 2 * r.new_parts.back().n) break;
    union_into(e, older, r.new_parts.back());
    r.new_parts.pop_back();
  }
}

// The marked tuples of a dense NEW become a NEW part (already in order and
// without repeats: the index adopts them); the bitmap is left clear.
static void dense_flush(srd_engine& e, Relation& r) {
  if (!r.dense_on) return;
  r.dense_on = false;
  DevBuf<uint32_t> buf;
  uint64_t stride = 0;
  const uint64_t n = dense_extract(r.dense_bits.p, r.dense_sum.p, r.dense_words, r.arity,
                                   r.dense_shift, buf, stride, e.ws, e.s);
  if (n == 0) return;
  std::vector<const uint32_t*> in(r.arity);
  for (uint32_t c = 0; c < r.arity; ++c) in[c] = buf.p + uint64_t(c) * stride;
  const std::vector<uint32_t> bound(r.arity, r.dense_bound);
  Index part;
  build_index(in.data(), r.arity, n, part, e.ws, e.s, &bound, &buf, stride, /*presorted=*/true);
  add_new_part(e, r, r.dense_order, std::move(part));
}

// Turn the relation's raw NEW rows into a deduplicated NEW part, so that
// raw rows never exceed one chunk (several pipelines may write one NEW).
static void flush_raw(srd_engine& e, Relation& r) {
  if (r.raw_n == 0) return;
  std::vector<const uint32_t*> in(r.arity);
  for (uint32_t c = 0; c < r.arity; ++c) in[c] = r.raw.p + uint64_t(c) * r.raw_cap;
  Index part;
  build_index(in.data(), r.arity, r.raw_n, part, e.ws, e.s);
  const Order order = r.raw_order;
  r.raw.release();
  r.raw_n = r.raw_cap = 0;
  r.raw_sorted = false;
  add_new_part(e, r, order, std::move(part));
}

// ============================================================ pipelines ====
static uint64_t read_u64(srd_engine& e, const uint64_t* dev) {
  uint64_t* h = reinterpret_cast<uint64_t*>(pinned_words(e.ws, 2));
  SRD_CHECK(hipMemcpyAsync(h, dev, 8, hipMemcpyDeviceToHost, e.s));
  SRD_CHECK(hipStreamSynchronize(e.s));
  return h[0];
}

// Staging columns of a fused launch: destination k's columns start at 8*k.
static uint32_t stage_cols(const SrdPipelineDesc* d) {
  uint32_t kmax = 0;
  for (uint32_t k = 0; k < d->n_dest; ++k) kmax = std::max(kmax, 8 * k + d->dest[k].arity);
  return kmax;
}

// One pipeline in flight.  A single pipeline runs its phases on the
// engine's stream; the pipelines of a parallel group run each phase on
// their own streams with their own scratch and meet only where the host
// needs their sizes -- one wait per phase for the whole group instead of
// one per pipeline and phase (the reference's ParallelGroup dispatch,
// codegen/jit/orchestrator_jit.py:236-474, over a StreamPool,
// runtime/generalized_datalog/gpu/runtime/stream_pool.h:18-89).
struct PipeRun {
  const SrdPipelineDesc* d = nullptr;
  PipeScratch* x = nullptr;
  hipStream_t s = nullptr;
  SrdLaunchArgs a;
  std::vector<uint32_t> view_max[SRD_MAX_VIEWS];
  uint64_t* host = nullptr;   // 4 pinned words for this pipeline's read-backs
  uint64_t total = 0;         // tuples emitted
  uint64_t slices = 0;
  uint64_t chunk_rows = 0;
  uint64_t rec_cap = 0;
  uint64_t weigh_items = 0;   // > 0: slices are being weighed
  bool dense = false;         // the destination's NEW is a bitmap: the traversal marks it
  int mode = 0;               // 0 nothing to run, 1 direct, 2 count (+ write), 3 fused, 4 mark
  bool ovf = false;
};

// Read-back of one u64 on the run's stream, collected after the phase's wait.
static void pr_fetch(PipeRun& R, int slot, const void* dev, size_t bytes) {
  SRD_CHECK(hipMemcpyAsync(R.host + slot, dev, bytes, hipMemcpyDeviceToHost, R.s));
}

// Size the fused mode's staging buffer for a launch of `slices` slices: the
// expected output (the rule's largest so far, +25%) plus one segment per
// wave, which may stay partly filled, clamped to stage_max_bytes.  A launch
// that still overflows falls back to the write pass over the counts it
// produced.  Returns false, leaving the two-pass mode to run, when the
// buffer cannot be allocated.
static bool stage_ready(srd_engine& e, PipeRun& R) {
  const SrdPipelineDesc* d = R.d;
  const uint32_t kmax = stage_cols(d);
  uint64_t cap = std::max<uint64_t>(e.stage_hint[d->rule], uint64_t(1) << 24);
  // a wave keeps one segment open across its slices: at most one partly
  // filled segment per wave of the persistent grid
  cap = std::max<uint64_t>(cap, std::min<uint64_t>(R.slices, 4 * 2048) * SRD_SEG + (uint64_t(1) << 20));
  cap = std::min<uint64_t>(cap, e.stage_max_bytes / (4ull * kmax));
  if (e.stage_cap_env) cap = std::max<uint64_t>(e.stage_cap_env, 1);
  if (R.x->stage.n >= cap * kmax) return true;
  R.x->stage.release();
  try {
    R.x->stage.alloc(cap * kmax, R.s);
  } catch (const std::exception& ex) {
    if (e.trace) {
      std::fprintf(stderr, "[srd]   %s: no staging (%s): two-pass\n", d->rule, ex.what());
      std::fflush(stderr);
    }
    return false;
  }
  return true;
}

// Phase 1 (host, engine stream): resolve the views (iteration 0 reads FULL
// wherever the rule reads DELTA), build the bitmaps and root maps they ask
// for, pick the driver.
static void pr_views(srd_engine& e, PipeRun& R, uint64_t iteration) {
  const SrdPipelineDesc* d = R.d;
  std::memset(&R.a, 0, sizeof(R.a));
  for (uint32_t v = 0; v < d->n_views; ++v) {
    const SrdViewSpec& vs = d->views[v];
    Relation& r = e.rels[rel_id(e, vs.rel)];
    Order order(vs.cols, vs.cols + vs.ncols);
    const Index* x;
    if (vs.ver == SRD_FULL || (vs.ver == SRD_DELTA && iteration == 0))
      x = &full_index(e, r, order);
    else if (vs.ver == SRD_DELTA)
      x = &delta_index(e, r, order);
    else
      x = &new_index(e, r, order);
    R.view_max[v] = x->col_max;
    if (vs.want_bitmaps && x->arity == 2 && e.bitmap_bytes) {
      const HeavyBits* h = x->heavy.get();
      if (!h || h->built_for != x->cols.p || h->built_n != x->n) {
        timing_begin("join.bitmaps", e.s);
        build_heavy_bitmaps(*x, e.bitmap_min, e.bitmap_bytes, e.ws, e.s);
        timing_end(e.s);
        if (e.trace) {
          std::fprintf(stderr, "[srd]   %s view %u: %u heavy-key bitmaps of %u words\n", d->rule, v,
                       x->heavy->n, x->heavy->words);
          std::fflush(stderr);
        }
      }
    }
    if (vs.want_root_pos && e.root_pos) {
      const RootPos* m = x->rootpos.get();
      if (!m || m->built_for != x->root_vals.p || m->built_r != x->r) {
        timing_begin("join.root_pos", e.s);
        build_root_pos(*x, std::max<uint64_t>(8ull * x->r, uint64_t(1) << 20), e.s);
        timing_end(e.s);
      }
    }
    R.a.views.v[v] = x->view();
  }
  R.a.n_ranges = e.xcd_ranges;
  R.a.window_min = e.window_min;
  R.a.grain = e.grain_override ? e.grain_override : d->default_grain;
  if (d->root_kind == SRD_ROOT_JOIN) {
    uint32_t drv = d->root_view[0];
    for (uint32_t i = 0; i < d->n_root_src; ++i) {
      uint32_t v = d->root_view[i];
      if (R.a.views.v[v].r < R.a.views.v[drv].r) drv = v;
    }
    R.a.drv_view = drv;
    R.a.n_root = R.a.views.v[drv].r;
  } else {
    R.a.drv_view = d->root_view[0];
    R.a.n_root = R.a.views.v[R.a.drv_view].n ? 1 : 0;
  }
}

static void pr_launch(srd_engine& e, PipeRun& R, SrdLaunchFn fn, const char* phase) {
  fill_u32(R.a.work_ctr, 0, SRD_XCDS * SRD_CTR_STRIDE, R.s);
  std::string nm = std::string("join.") + phase + ":" + R.d->rule;
  timing_begin(nm.c_str(), R.s);
  fn(&R.a, R.s);
  SRD_CHECK(hipGetLastError());
  timing_end(R.s);
}

// Phase 2 (run stream): the root launch and the scan of the per-key work;
// the work total is fetched.
static void pr_root(srd_engine& e, PipeRun& R) {
  const SrdPipelineDesc* d = R.d;
  PipeScratch& x = *R.x;
  const uint64_t nk = R.a.n_root;
  x.key_rng.reserve(nk * 2 * SRD_MAX_VIEWS, R.s);
  x.key_aux.reserve(nk, R.s);
  x.key_work.reserve(nk, R.s);
  x.work_start.reserve(nk + 1, R.s);
  x.work_ctr.reserve(SRD_XCDS * SRD_CTR_STRIDE, R.s);
  if (e.count_probes) {
    x.probe_buf.reserve(1, R.s);
    fill_u64(x.probe_buf.p, 0, 1, R.s);
    R.a.probe_ctr = reinterpret_cast<unsigned long long*>(x.probe_buf.p);
  }
  R.a.key_rng = x.key_rng.p;
  R.a.key_aux = x.key_aux.p;
  R.a.key_work = x.key_work.p;
  R.a.work_start = x.work_start.p;
  R.a.work_ctr = x.work_ctr.p;
  {
    std::string nm = std::string("join.root:") + d->rule;
    timing_begin(nm.c_str(), R.s);
    d->root(&R.a, R.s);
    SRD_CHECK(hipGetLastError());
    timing_end(R.s);
  }
  exclusive_scan_u64(R.a.key_work, R.a.work_start, nk, e.ws, R.s);
  pr_fetch(R, 0, R.a.work_start + nk, 8);
}

// Phase 3 (host + run stream): uniform slices, and for a rule with >= 2
// nested levels the weigh launch -- every level-1 item's level-2 candidates
// + SRD_ITEM_COST, prefix-summed -- whose total is fetched.
static void pr_slices(srd_engine& e, PipeRun& R) {
  const SrdPipelineDesc* d = R.d;
  R.a.total_work = R.host[0];
  R.a.n_slices = (R.a.total_work + R.a.grain - 1) / R.a.grain;
  const bool direct = d->direct_ok && !d->count_only && !e.two_pass && !e.no_direct;
  R.weigh_items = 0;
  if (!direct && d->weigh && !e.no_weigh && R.a.total_work >= e.weigh_min &&
      R.a.total_work <= (uint64_t(1) << 28)) {
    PipeScratch& x = *R.x;
    const uint64_t n = R.a.total_work;
    x.item_cost.reserve(n, R.s);
    x.item_pref.reserve(n + 1, R.s);
    R.a.item_cost = x.item_cost.p;
    R.a.n_ranges = e.weigh_ranges;
    R.a.grid_blocks = uint32_t(std::min<uint64_t>((R.a.n_slices + 3) / 4, 2048));
    pr_launch(e, R, d->weigh, "weigh");
    exclusive_scan_u64(R.a.item_cost, x.item_pref.p, n, e.ws, R.s);
    pr_fetch(R, 1, x.item_pref.p + n, 8);
    R.weigh_items = n;
  }
}

// Replace the uniform slices (grain items each) of a weighed rule by slices
// of about equal cost: slice s is cut at the item where the cost prefix
// crosses s/S of the total.  An item is never split, so one item's
// candidates bound a slice's cost.
static void pr_weighed(srd_engine& e, PipeRun& R) {
  if (!R.weigh_items) return;
  PipeScratch& x = *R.x;
  const uint64_t n = R.weigh_items, w = R.host[1];
  // up to e.slices_per_wave slices per persistent wave of a full grid (4 x 2048), never
  // fewer than the uniform slicing had, never more than one per item
  // and at least e.min_slices_per_wave per wave, so that a mid-sized join
  // whose cost / weigh_target is below the wave count still spreads
  const uint64_t cap = std::max<uint64_t>(R.a.n_slices, e.slices_per_wave * 4 * 2048);
  const uint64_t floor_s = e.min_slices_per_wave * 4 * 2048;
  uint64_t S = std::min<uint64_t>(
      std::max<uint64_t>(std::max<uint64_t>(R.a.n_slices, w / e.weigh_target), floor_s), cap);
  S = std::max<uint64_t>(std::min<uint64_t>(S, n), 1);
  x.slice_start.reserve(S + 1, R.s);
  slice_bounds(x.item_pref.p, n, w, S, x.slice_start.p, R.s);
  R.a.slice_start = x.slice_start.p;
  R.a.n_slices = S;
  R.a.item_cost = nullptr;
  R.a.n_ranges = e.xcd_ranges;
  if (e.trace) {
    std::fprintf(stderr, "[srd]   %s: weighed %llu items, cost %llu, %llu slices\n", R.d->rule,
                 (unsigned long long)n, (unsigned long long)w, (unsigned long long)S);
    std::fflush(stderr);
  }
}

// Phase 4 (run stream): the traversal that sizes the output -- nothing for
// the direct mode (one row per item), the count launch, or the fused launch
// into staging -- with its total (and the staging overflow flag) fetched.
static void pr_traverse(srd_engine& e, PipeRun& R) {
  const SrdPipelineDesc* d = R.d;
  PipeScratch& x = *R.x;
  if (e.trace) {
    std::fprintf(stderr, "[srd]   %s: keys %u drv %u work %llu slices %llu\n", d->rule, R.a.n_root,
                 R.a.drv_view, (unsigned long long)R.a.total_work, (unsigned long long)R.a.n_slices);
    std::fflush(stderr);
  }
  R.slices = R.a.n_slices;
  R.mode = 0;
  if (R.slices == 0) return;
  R.a.grid_blocks = uint32_t(std::min<uint64_t>((R.slices + 3) / 4, 2048));
  if (R.dense) {
    // every derived tuple sets its bit in the destination's bitmap: nothing
    // to count, size or write.  With mark_chunk_slices the launch is split
    // into slice ranges, each waited for and traced when tracing (a
    // traversal of minutes then reports as it goes).
    R.mode = 4;
    const uint64_t step = e.mark_chunk_slices;
    if (!step || R.slices <= step) {
      pr_launch(e, R, d->mark, "mark");
      return;
    }
    for (uint64_t s0 = 0; s0 < R.slices; s0 += step) {
      R.a.slice_lo = s0;
      R.a.n_slices = std::min(s0 + step, R.slices);
      pr_launch(e, R, d->mark, "mark");
      if (e.trace) {
        SRD_CHECK(hipStreamSynchronize(R.s));
        std::fprintf(stderr, "[srd]   %s: marked slices [%llu, %llu) of %llu\n", d->rule,
                     (unsigned long long)s0, (unsigned long long)R.a.n_slices,
                     (unsigned long long)R.slices);
        std::fflush(stderr);
      }
    }
    R.a.slice_lo = 0;
    R.a.n_slices = R.slices;
    return;
  }
  x.counts_buf.reserve(R.slices, R.s);
  x.offsets.reserve(R.slices + 1, R.s);
  R.a.counts = x.counts_buf.p;
  R.a.offsets = x.offsets.p;
  // Rows of the write (direct) pass, materialised at once when they fit
  // one chunk, else in slice ranges whose rows are each sorted and
  // deduplicated into a NEW part (a join's output before deduplication
  // can be far larger than HBM).
  uint32_t amax = 1;
  for (uint32_t k = 0; k < d->n_dest; ++k) amax = std::max(amax, d->dest[k].arity);
  R.chunk_rows = e.new_chunk_rows ? e.new_chunk_rows
                                  : std::min<uint64_t>(e.stage_max_bytes / (4ull * amax), uint64_t(1) << 31);
  const bool direct = d->direct_ok && !d->count_only && !e.two_pass && !e.no_direct;
  if (direct) {
    // one tuple per item: rows land at base + item, nothing to count
    R.mode = 1;
    R.total = R.a.total_work;
    return;
  }
  if (d->count_only || e.two_pass || !stage_ready(e, R)) {
    R.mode = 2;
    pr_launch(e, R, d->count, "count");
    exclusive_scan_u32(R.a.counts, R.a.offsets, R.slices, e.ws, R.s);
    pr_fetch(R, 2, R.a.offsets + R.slices, 8);
    return;
  }
  // fused: one traversal into staging segments, then gather.  Every slice
  // may hold one partly filled segment, so the capacity wanted is the
  // expected output plus a segment per slice; it is clamped to
  // stage_max_bytes, and a launch that overflows falls back to the write
  // pass over the counts it produced.
  R.mode = 3;
  const uint32_t kmax = stage_cols(d);
  uint64_t cap = x.stage.n / kmax;
  if (e.stage_cap_env) cap = std::min<uint64_t>(cap, e.stage_cap_env);
  R.rec_cap = R.slices + cap / (SRD_SEG - 64) + 64;
  x.recs.reserve(3 * R.rec_cap, R.s);
  x.stage_ctr.reserve(2, R.s);
  x.overflow.reserve(2 * SRD_CTR_STRIDE, R.s);  // flag and record counter on own lines
  fill_u64(x.stage_ctr.p, 0, 2, R.s);
  fill_u32(x.overflow.p, 0, 2 * SRD_CTR_STRIDE, R.s);
  R.a.stage = x.stage.p;
  R.a.stage_cap = cap;
  R.a.stage_ctr = reinterpret_cast<unsigned long long*>(x.stage_ctr.p);
  R.a.recs = x.recs.p;
  R.a.rec_cap = R.rec_cap;
  R.a.rec_ctr = x.overflow.p + SRD_CTR_STRIDE;
  R.a.overflow = x.overflow.p;
  pr_launch(e, R, d->fused, "fused");
  exclusive_scan_u32(R.a.counts, R.a.offsets, R.slices, e.ws, R.s);
  pr_fetch(R, 2, R.a.offsets + R.slices, 8);
  pr_fetch(R, 3, x.overflow.p, 4);
}

// After phase 4's wait: the totals, and whether the rows fit one chunk and
// (fused) staging.  Returns true when the rows are written by the simple
// paths of phase 5 (the chunked and overflow paths run sequentially after).
static bool pr_sized(srd_engine& e, PipeRun& R) {
  const SrdPipelineDesc* d = R.d;
  if (R.mode == 2 || R.mode == 3) R.total = R.host[2];
  R.ovf = R.mode == 3 && uint32_t(R.host[3]) != 0;
  if (R.mode == 3)
    e.stage_hint[d->rule] = std::max<uint64_t>(
        e.stage_hint[d->rule], R.total + R.total / 4 + std::min<uint64_t>(R.slices, 4 * 2048) * SRD_SEG);
  if (R.mode == 0 || R.total == 0 || d->count_only) return false;
  return R.total <= R.chunk_rows && !R.ovf;
}

// Reserve the NEW rows of the destinations (host, engine stream).
static void pr_bind(srd_engine& e, PipeRun& R, uint64_t n_rows) {
  const SrdPipelineDesc* d = R.d;
  for (uint32_t k = 0; k < d->n_dest; ++k) {
    const SrdDestSpec& ds = d->dest[k];
    Relation& r = e.rels[rel_id(e, ds.rel)];
    dense_flush(e, r);  // rows marked by earlier pipelines become a NEW part
    if (r.raw_n > 0 && r.raw_n + n_rows > R.chunk_rows) flush_raw(e, r);
    // rows of one sorted-output pipeline into an empty NEW stay an index;
    // bounds: the largest value of any column the join reads
    r.raw_sorted = r.raw_n == 0 && d->sorted_output && d->n_dest == 1;
    if (r.raw_sorted) {
      uint32_t bound = 0;
      for (uint32_t v = 0; v < d->n_views; ++v)
        for (uint32_t m : R.view_max[v]) bound = std::max(bound, m);
      r.raw_bound.assign(ds.arity, bound);
    }
    Order order(ds.cols, ds.cols + ds.arity);
    uint64_t base = reserve_new(e, r, order, n_rows);
    R.a.out[k].cols = r.raw.p;
    R.a.out[k].stride = r.raw_cap;
    R.a.out[k].base = base;
  }
}

// Phase 5 (run stream): write the bound rows -- the direct launch, or the
// gather of the fused mode's staging.
static void pr_write(srd_engine& e, PipeRun& R) {
  const SrdPipelineDesc* d = R.d;
  if (R.mode == 1) {
    pr_launch(e, R, d->direct, "direct");
  } else if (R.mode == 2) {
    pr_launch(e, R, d->write, "write");
  } else if (R.mode == 3) {
    GatherDest gd[SRD_MAX_DESTS];
    for (uint32_t k = 0; k < d->n_dest; ++k)
      gd[k] = GatherDest{R.a.out[k].cols, R.a.out[k].stride, R.a.out[k].base, d->dest[k].arity};
    gather_segments(R.a.recs, R.a.rec_ctr, R.rec_cap, R.a.stage, R.a.stage_cap, R.a.offsets, gd,
                    d->n_dest, R.total, R.s);
  }
}

// The rows of a run that exceeds one chunk (written in slice ranges, each
// sorted and deduplicated into a NEW part) or whose fused launch overflowed
// staging (the write pass over its counts).  Engine stream.
static void pr_write_slow(srd_engine& e, PipeRun& R) {
  const SrdPipelineDesc* d = R.d;
  const bool by_items = R.mode == 1;
  SrdLaunchFn fn = R.mode == 1 ? d->direct : d->write;
  const char* phase = R.mode == 1 ? "direct" : "write";
  const uint64_t slices = R.slices;
  if (R.total <= R.chunk_rows) {
    pr_bind(e, R, R.total);
    pr_launch(e, R, fn, phase);
    return;
  }
  std::vector<uint64_t> off;
  if (!by_items) {
    off.resize(slices + 1);
    SRD_CHECK(hipMemcpyAsync(off.data(), R.a.offsets, (slices + 1) * 8, hipMemcpyDeviceToHost, R.s));
    SRD_CHECK(hipStreamSynchronize(R.s));
  }
  auto start_of = [&](uint64_t sl) {
    return by_items ? std::min<uint64_t>(sl * R.a.grain, R.a.total_work) : off[sl];
  };
  std::vector<DevBuf<uint32_t>> bufs(d->n_dest);
  uint64_t s0 = 0;
  while (s0 < slices) {
    uint64_t s1 = s0 + 1;  // at least one slice, even when it alone is larger
    while (s1 < slices && start_of(s1 + 1) - start_of(s0) <= R.chunk_rows) ++s1;
    const uint64_t rows = start_of(s1) - start_of(s0);
    if (rows > 0) {
      const uint64_t stride = round_stride(rows);
      for (uint32_t k = 0; k < d->n_dest; ++k) {
        bufs[k].reserve(stride * d->dest[k].arity, R.s);
        R.a.out[k].cols = bufs[k].p;
        R.a.out[k].stride = stride;
        R.a.out[k].base = uint64_t(0) - start_of(s0);  // row = position - chunk start
      }
      R.a.slice_lo = s0;
      R.a.n_slices = s1;
      pr_launch(e, R, fn, phase);
      for (uint32_t k = 0; k < d->n_dest; ++k) {
        const SrdDestSpec& ds = d->dest[k];
        Relation& r = e.rels[rel_id(e, ds.rel)];
        std::vector<const uint32_t*> in(ds.arity);
        for (uint32_t c = 0; c < ds.arity; ++c) in[c] = bufs[k].p + uint64_t(c) * stride;
        Index part;
        build_index(in.data(), ds.arity, rows, part, e.ws, R.s);
        add_new_part(e, r, Order(ds.cols, ds.cols + ds.arity), std::move(part));
      }
      if (e.trace) {
        std::fprintf(stderr, "[srd]   %s: chunk of slices [%llu, %llu): %llu rows\n", d->rule,
                     (unsigned long long)s0, (unsigned long long)s1, (unsigned long long)rows);
        std::fflush(stderr);
      }
    }
    s0 = s1;
  }
  R.a.slice_lo = 0;
  R.a.n_slices = slices;
  if (e.trace) {
    std::fprintf(stderr, "[srd]   %s: %llu rows written in chunks of <= %llu\n", d->rule,
                 (unsigned long long)R.total, (unsigned long long)R.chunk_rows);
    std::fflush(stderr);
  }
}

// Count-only bookkeeping, probes and trace of a finished run.
static void pr_finish(srd_engine& e, PipeRun& R) {
  const SrdPipelineDesc* d = R.d;
  if (e.trace && R.mode != 0) {
    std::fprintf(stderr, "[srd]   %s: %llu tuples%s\n", d->rule, (unsigned long long)R.total,
                 R.ovf ? " (staging overflow: write pass)" : "");
    std::fflush(stderr);
  }
  e.join_probes += R.total;
  if (d->count_only) {
    e.counts.emplace_back(d->rule, R.total);
    for (uint32_t k = 0; k < d->n_dest; ++k) {
      Relation& r = e.rels[rel_id(e, d->dest[k].rel)];
      r.count += R.total;
      r.counted = true;
    }
    if (!e.quiet) std::cout << " >>>>>>>>>>>>>>>>> " << d->rule << " count: " << R.total << std::endl;
  }
}

// Wait for every run's stream (the phase's read-backs have landed).
static void pr_wait(std::vector<PipeRun>& runs) {
  for (auto& R : runs) SRD_CHECK(hipStreamSynchronize(R.s));
}

// Make every stream of `to` wait for the work enqueued so far on `from`.
static void pr_fence(srd_engine& e, hipStream_t from, const std::vector<PipeRun>& runs, bool into_runs) {
  if (runs.size() == 1 && runs[0].s == e.s) return;
  if (into_runs) {
    hipEvent_t ev = e.fence_events[0];
    SRD_CHECK(hipEventRecord(ev, from));
    for (auto& R : runs) SRD_CHECK(hipStreamWaitEvent(R.s, ev, 0));
  } else {
    for (size_t i = 0; i < runs.size(); ++i) {
      hipEvent_t ev = e.fence_events[1 + i];
      SRD_CHECK(hipEventRecord(ev, runs[i].s));
      SRD_CHECK(hipStreamWaitEvent(from, ev, 0));
    }
  }
}

// Largest value of any FULL or DELTA index or input of the engine (index
// maxima are upper bounds): no join of the current iteration reads a value
// beyond it, save from a NEW view, whose maxima the callers add.
static uint32_t value_bound(const srd_engine& e) {
  uint32_t b = 0;
  auto take = [&](const std::vector<uint32_t>& mx) {
    for (uint32_t m : mx) b = std::max(b, m);
  };
  for (const auto& r : e.rels) {
    for (const auto& kv : r.full) take(kv.second.idx.col_max);
    for (const auto& sg : r.segs) take(sg.col_max);
    for (const auto& kv : r.delta) take(kv.second.col_max);
    take(r.input_max);
  }
  return b;
}

// After the runs' views are prepared and before any of them launches: the
// destinations whose key space fits dense_max_bits take their NEW as a
// bitmap that the runs writing them mark (mode 6).  A destination some run
// writes rows into (a rule with several destinations, constants in its
// head, count-only or sorted output) keeps raw rows for all of them.  The
// key layout covers every value the runs can emit -- the engine's value
// bound and their views' maxima -- and a bitmap marked earlier in the
// iteration under a narrower layout is first turned into a NEW part.
static void pr_dense(srd_engine& e, std::vector<PipeRun*>& runs) {
  if (!e.dense_max_bits) return;
  std::set<int> rows;
  std::map<int, uint32_t> need;
  for (PipeRun* R : runs) {
    const SrdPipelineDesc* d = R->d;
    const bool can = d->mark && d->n_dest == 1 && !d->count_only && !d->sorted_output;
    for (uint32_t k = 0; k < d->n_dest; ++k) {
      const int ri = rel_id(e, d->dest[k].rel);
      if (!can) {
        rows.insert(ri);
        continue;
      }
      uint32_t b = 0;
      for (uint32_t v = 0; v < d->n_views; ++v)
        for (uint32_t m : R->view_max[v]) b = std::max(b, m);
      need[ri] = std::max(need[ri], b);
    }
  }
  if (need.empty()) return;
  const uint32_t gb = value_bound(e);
  for (auto& kv : need) {
    if (rows.count(kv.first)) continue;
    Relation& r = e.rels[kv.first];
    const uint32_t bound = std::max(kv.second, gb);
    const uint32_t sh = std::max<uint32_t>(32 - __builtin_clz(bound | 1u), 1);
    const uint64_t bits = uint64_t(sh) * r.arity;
    if (r.dense_on && r.dense_shift < sh) dense_flush(e, r);
    if (!r.dense_on) {
      if (bits > 40 || (uint64_t(1) << bits) > e.dense_max_bits) continue;
      // rows already written this iteration stay (they join the parts);
      // an index already built from them does not take more
      if (!r.newv.empty()) continue;
      const SrdDestSpec* ds = nullptr;
      for (PipeRun* R : runs)
        if (R->d->n_dest == 1 && rel_id(e, R->d->dest[0].rel) == kv.first) ds = &R->d->dest[0];
      const Order order(ds->cols, ds->cols + ds->arity);
      if (!r.new_parts.empty() && r.new_parts_order != order) continue;
      if (r.raw_n > 0 && r.raw_order != order) continue;
      const uint64_t words =
          std::max<uint64_t>(((uint64_t(1) << bits) / 32 + kDenseBlockWords - 1) / kDenseBlockWords,
                             1) * kDenseBlockWords;
      if (r.dense_bits.n < words) {
        size_t free_b = 0, total_b = 0;
        {
          std::lock_guard<std::recursive_mutex> lk(capture_mutex());
          SRD_CHECK(hipMemGetInfo(&free_b, &total_b));
        }
        if (e.dense_bytes - r.dense_bits.n * 4 + words * 4 > total_b / 8) continue;
        e.dense_bytes = e.dense_bytes - r.dense_bits.n * 4 + words * 4;
        r.dense_bits.alloc(words, e.s);
        fill_u32(r.dense_bits.p, 0, words, e.s);
        r.dense_sum.alloc(dense_sum_words(words), e.s);
        fill_u32(r.dense_sum.p, 0, dense_sum_words(words), e.s);
      }
      r.dense_words = words;
      r.dense_shift = sh;
      r.dense_bound = bound;
      r.dense_order = order;
      r.dense_on = true;
      if (e.trace) {
        std::fprintf(stderr, "[srd]   %s: NEW as a bitmap of 2^%llu keys (%u bits per column)\n",
                     r.name.c_str(), (unsigned long long)bits, sh);
        std::fflush(stderr);
      }
    }
    // a layout kept from earlier in the iteration is wide enough; its bound
    // (the extracted part's column maxima) must still cover these runs
    r.dense_bound = std::max(r.dense_bound, bound);
    for (PipeRun* R : runs)
      if (R->d->mark && R->d->n_dest == 1 && rel_id(e, R->d->dest[0].rel) == kv.first) {
        R->dense = true;
        R->a.mark_bits = r.dense_bits.p;
        R->a.mark_sum = r.dense_sum.p;
        R->a.mark_shift = r.dense_shift;
      }
  }
}

// Run pipelines -- one, or the members of a parallel group -- phase by
// phase.  A group's members each get a stream and scratch of their own (a
// member counting probes or needing the chunked or overflow path finishes
// on the engine's stream, after the others).
static void run_pipelines(srd_engine& e, const std::vector<const SrdPipelineDesc*>& ds,
                          uint64_t iteration) {
  const size_t P = ds.size();
  const bool group = P > 1 && e.group_streams;
  while (e.scratch.size() < P) e.scratch.emplace_back(new PipeScratch());
  if (group) {
    while (e.group_s.size() < P) e.group_s.push_back(stream_acquire());
    while (e.fence_events.size() < P + 1) {
      hipEvent_t ev;
      SRD_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      e.fence_events.push_back(ev);
    }
  } else if (e.fence_events.empty()) {
    hipEvent_t ev;
    SRD_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    e.fence_events.push_back(ev);
  }
  // pinned words of the runs' read-backs (a block of their own: the sorts a
  // run may start reallocate the workspace's words)
  if (e.run_pins_bytes < 64 * P) {
    pinned_release(e.run_pins, e.run_pins_bytes);
    e.run_pins = nullptr;
    e.run_pins_bytes = 0;
    e.run_pins = static_cast<uint64_t*>(pinned_acquire(64 * std::max<size_t>(P, 8), &e.run_pins_bytes));
  }
  uint64_t* pins = e.run_pins;
  if (!group) {
    // members one after the other on the engine's stream, as a single pipeline
    for (size_t i = 0; i < P; ++i) {
      std::vector<PipeRun> one(1);
      PipeRun& R = one[0];
      R.d = ds[i];
      R.x = e.scratch[0].get();
      R.s = e.s;
      R.host = pins;
      pr_views(e, R, iteration);
      {
        std::vector<PipeRun*> me{&R};
        pr_dense(e, me);
      }
      if (R.a.n_root > 0) {
        pr_root(e, R);
        pr_wait(one);
        pr_slices(e, R);
        if (R.weigh_items) pr_wait(one);
        pr_weighed(e, R);
        pr_traverse(e, R);
        if (R.mode == 2 || R.mode == 3) pr_wait(one);
        if (pr_sized(e, R)) {
          pr_bind(e, R, R.total);
          pr_write(e, R);
        } else if (R.mode != 0 && R.total > 0 && !R.d->count_only) {
          pr_write_slow(e, R);
        }
        if (e.count_probes && R.slices > 0 && R.d->probe) pr_launch(e, R, R.d->probe, "probe");
        if (e.count_probes) {
          pr_fetch(R, 0, R.x->probe_buf.p, 8);
          pr_wait(one);
          e.join_searches += R.host[0];
        }
      }
      pr_finish(e, R);
    }
    return;
  }
  std::vector<PipeRun> runs(P);
  for (size_t i = 0; i < P; ++i) {
    runs[i].d = ds[i];
    runs[i].x = e.scratch[i].get();
    runs[i].s = e.group_s[i];
    runs[i].host = pins + 8 * i;
    pr_views(e, runs[i], iteration);
  }
  {
    std::vector<PipeRun*> all;
    for (auto& R : runs) all.push_back(&R);
    pr_dense(e, all);
  }
  // every member reads indexes built on the engine's stream (and the
  // bitmaps cleared there)
  pr_fence(e, e.s, runs, true);
  std::vector<PipeRun*> live;
  for (auto& R : runs)
    if (R.a.n_root > 0) {
      pr_root(e, R);
      live.push_back(&R);
    }
  pr_wait(runs);
  bool weighed = false;
  for (PipeRun* R : live) {
    pr_slices(e, *R);
    weighed = weighed || R->weigh_items;
  }
  if (weighed) pr_wait(runs);
  for (PipeRun* R : live) {
    pr_weighed(e, *R);
    pr_traverse(e, *R);
  }
  pr_wait(runs);
  // sizes known: reserve every member's NEW rows on the engine's stream (the
  // members are idle), then write them concurrently
  std::vector<PipeRun*> fast, slow;
  for (PipeRun* R : live) {
    if (pr_sized(e, *R)) fast.push_back(R);
    else if (R->mode != 0 && R->total > 0 && !R->d->count_only) slow.push_back(R);
  }
  // The members' rows go to NEW one after the other.  A destination whose
  // rows would exceed a chunk takes them sequentially (the chunked path
  // flushes written rows, and these are not written yet).
  std::map<int, uint64_t> need;
  for (PipeRun* R : fast)
    for (uint32_t k = 0; k < R->d->n_dest; ++k) need[rel_id(e, R->d->dest[k].rel)] += R->total;
  std::vector<PipeRun*> bound;
  for (PipeRun* R : fast) {
    bool fits = true;
    for (uint32_t k = 0; k < R->d->n_dest; ++k) {
      const int ri = rel_id(e, R->d->dest[k].rel);
      fits = fits && e.rels[ri].raw_n + need[ri] <= R->chunk_rows;
    }
    (fits ? bound : slow).push_back(R);
  }
  for (PipeRun* R : bound) pr_bind(e, *R, R->total);
  // a later reservation may have moved a relation's NEW buffer (rows kept):
  // point every member at the final one
  for (PipeRun* R : bound)
    for (uint32_t k = 0; k < R->d->n_dest; ++k) {
      const Relation& r = e.rels[rel_id(e, R->d->dest[k].rel)];
      R->a.out[k].cols = r.raw.p;
      R->a.out[k].stride = r.raw_cap;
    }
  fast = bound;
  pr_fence(e, e.s, runs, true);
  for (PipeRun* R : fast) pr_write(e, *R);
  if (e.count_probes)
    for (PipeRun* R : live)
      if (R->slices > 0 && R->d->probe) pr_launch(e, *R, R->d->probe, "probe");
  pr_fence(e, e.s, runs, false);
  // the chunked / overflow members, one after the other
  for (PipeRun* R : slow) {
    R->s = e.s;
    pr_write_slow(e, *R);
  }
  if (e.count_probes) {
    for (PipeRun* R : live) {
      R->s = e.s;
      pr_fetch(*R, 0, R->x->probe_buf.p, 8);
    }
    SRD_CHECK(hipStreamSynchronize(e.s));
    for (PipeRun* R : live) e.join_searches += R->host[0];
  }
  for (auto& R : runs) pr_finish(e, R);
}

// =========================================================== MIR ops ====
static void exec_ops(srd_engine& e, const std::vector<Op>& ops, uint64_t iteration, int step_no);

static const char* op_name(OpKind k) {
  switch (k) {
    case OpKind::Pipeline: return "execute-pipeline";
    case OpKind::Parallel: return "parallel-group";
    case OpKind::RebuildIndex: return "rebuild-index";
    case OpKind::ClearRelation: return "clear-relation";
    case OpKind::CheckSize: return "check-size";
    case OpKind::ComputeDelta: return "compute-delta-index";
    case OpKind::MergeIndex: return "merge-index";
    case OpKind::RebuildFromIndex: return "rebuild-index-from-index";
    case OpKind::FlatView: return "create-flat-view";
    case OpKind::Reconstruct: return "reconstruct";
    case OpKind::Hook: return "hook";
  }
  return "?";
}

static void exec_op_body(srd_engine& e, const Op& op, uint64_t iteration, int step_no);

// With SRDATALOG_OP_PROFILE set, the host wall time of every op (launches,
// and the waits of the ops that read sizes back) is summed per op and rule
// and printed at the end of srdatalog_run.
static void exec_op(srd_engine& e, const Op& op, uint64_t iteration, int step_no) {
  if (!e.op_profile) return exec_op_body(e, op, iteration, step_no);
  const auto t0 = std::chrono::steady_clock::now();
  exec_op_body(e, op, iteration, step_no);
  const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::string key = std::string(op_name(op.kind)) + " ";
  if (op.kind == OpKind::Pipeline) key += op.rule;
  else if (op.kind == OpKind::Parallel)
    for (auto& k : op.kids) key += k.rule + " ";
  else if (op.rel >= 0) key += e.rels[op.rel].name;
  auto& slot = e.op_times[key];
  slot.first += dt;
  slot.second += 1;
}

// Merge the two youngest parts of a segmented FULL while the older is
// within 2x of the younger (the oldest part is full[canonical]).
static void merge_young_segments(srd_engine& e, Relation& r) {
  auto it = r.full.find(r.canonical);
  while (!r.segs.empty()) {
    const size_t m = r.segs.size();
    Index& older = m >= 2 ? r.segs[m - 2] : it->second.idx;
    if (older.n > 2 * r.segs.back().n && m < 24) break;
    Index mg;
    merge_disjoint(older, r.segs.back(), mg, e.ws, e.s);
    older = std::move(mg);
    r.segs.pop_back();
  }
}

static void exec_op_body(srd_engine& e, const Op& op, uint64_t iteration, int step_no) {
  if (e.trace) {
    std::fprintf(stderr, "[srd] step %d iter %llu %s %s\n", step_no, (unsigned long long)iteration,
                 op_name(op.kind), op.kind == OpKind::Pipeline ? op.rule.c_str()
                                   : (op.rel >= 0 ? e.rels[op.rel].name.c_str() : ""));
    std::fflush(stderr);
  }
  switch (op.kind) {
    case OpKind::Pipeline:
      run_pipelines(e, {op.pipe}, iteration);
      break;
    case OpKind::Parallel: {
      std::vector<const SrdPipelineDesc*> ds;
      for (auto& k : op.kids)
        if (k.kind == OpKind::Pipeline) ds.push_back(k.pipe);
      if (ds.size() == op.kids.size()) run_pipelines(e, ds, iteration);
      else exec_ops(e, op.kids, iteration, step_no);
      break;
    }
    case OpKind::RebuildIndex: {
      Relation& r = e.rels[op.rel];
      if (op.ver == SRD_NEW) {
        r.newv.erase(op.index);
        new_index(e, r, op.index);
        // the sorted index now owns the NEW rows
        r.raw.release();
        r.raw_n = r.raw_cap = 0;
      } else if (op.ver == SRD_DELTA) {
        r.delta.erase(op.index);
        delta_index(e, r, op.index);
      } else {
        full_index(e, r, op.index);
      }
      break;
    }
    case OpKind::ClearRelation: {
      Relation& r = e.rels[op.rel];
      if (op.ver == SRD_NEW) clear_new(r);
      else if (op.ver == SRD_DELTA) r.delta.clear();
      else {
        r.full.clear();
        r.full_version = r.delta_seq = 0;
      }
      break;
    }
    case OpKind::CheckSize:
      break;
    case OpKind::ComputeDelta: {
      Relation& r = e.rels[op.rel];
      const Index& nw = new_index(e, r, op.index);
      Index d;
      uint64_t full_n = 0;
      if (r.segmented && op.index == r.canonical && !r.segs.empty()) {
        // NEW minus every segment of FULL (each pass reads only the windows
        // of the segment that NEW's tiles cover)
        auto bit = r.full.find(r.canonical);
        copy_index(nw, d, e.s);
        std::vector<const Index*> parts;
        if (bit != r.full.end()) parts.push_back(&bit->second.idx);
        for (auto& sg : r.segs) parts.push_back(&sg);
        for (const Index* part : parts) {
          full_n += part->n;
          if (d.n == 0 || part->n == 0) continue;
          Index t;
          set_difference(d, *part, t, e.ws, e.s);
          d = std::move(t);
        }
      } else {
        const Index& fl = full_index(e, r, op.index);
        full_n = fl.n;
        if (fl.n == 0)
          copy_index(nw, d, e.s);
        else
          set_difference(nw, fl, d, e.ws, e.s);
      }
      r.delta.clear();
      if (r.deltas_step != step_no) {
        r.deltas.clear();
        r.inst_deltas.clear();
        r.deltas_step = step_no;
      }
      r.deltas.push_back(d.n);
      if (e.n_inst) {
        DevBuf<uint64_t> per(e.n_inst, e.s);
        instance_counts(d.n ? d.cols.p : nullptr, d.n, e.inst_bounds.p, e.n_inst, per.p, e.s);
        r.inst_deltas.push_back(per);
      }
      if (e.trace) {
        std::fprintf(stderr, "[srd]   delta %s = %llu (new %llu, full %llu, segments %zu)\n",
                     r.name.c_str(), (unsigned long long)d.n, (unsigned long long)nw.n,
                     (unsigned long long)full_n, r.segs.size() + 1);
        std::fflush(stderr);
      }
      r.delta[op.index] = std::move(d);
      r.delta_seq = r.full_version + 1;
      break;
    }
    case OpKind::MergeIndex: {
      Relation& r = e.rels[op.rel];
      const uint64_t target = r.delta_seq;
      if (target == 0) break;  // nothing computed yet
      auto it = r.full.find(op.index);
      if (it != r.full.end() && it->second.version == target) break;
      if (r.segmented && op.index == r.canonical) {
        // append DELTA as the youngest segment; merge the two youngest while
        // they are within a factor 2 (the oldest is full[canonical])
        if (it == r.full.end() || it->second.version != target - 1) {
          full_index(e, r, op.index);
          it = r.full.find(op.index);
        }
        const Index& dl = delta_index(e, r, op.index);
        if (dl.n > 0) {
          Index c;
          copy_index(dl, c, e.s);
          r.segs.push_back(std::move(c));
        }
        merge_young_segments(e, r);
        it->second.version = target;
        r.full_version = target;
        break;
      }
      const Index& dl = delta_index(e, r, op.index);
      Versioned v;
      v.version = target;
      if (it != r.full.end() && it->second.version == target - 1) {
        merge_disjoint(it->second.idx, dl, v.idx, e.ws, e.s);
      } else if (op.index != r.canonical) {
        auto cit = r.full.find(r.canonical);
        if (cit != r.full.end() && cit->second.version == target) {
          build_index_from(cit->second.idx, perm_between(r.canonical, op.index), v.idx, e.ws, e.s);
        } else {
          const Index& base = full_index(e, r, op.index);  // at target - 1
          merge_disjoint(base, dl, v.idx, e.ws, e.s);
        }
      } else {
        const Index& base = full_index(e, r, op.index);
        merge_disjoint(base, dl, v.idx, e.ws, e.s);
      }
      r.full[op.index] = std::move(v);
      if (op.index == r.canonical) r.full_version = target;
      break;
    }
    case OpKind::RebuildFromIndex: {
      Relation& r = e.rels[op.rel];
      Index x;
      if (op.ver == SRD_DELTA) {
        const Index& src = delta_index(e, r, op.index);
        build_index_from(src, perm_between(op.index, op.index2), x, e.ws, e.s);
        r.delta[op.index2] = std::move(x);
      } else if (op.ver == SRD_NEW) {
        const Index& src = new_index(e, r, op.index);
        build_index_from(src, perm_between(op.index, op.index2), x, e.ws, e.s);
        r.newv[op.index2] = std::move(x);
      } else {
        full_index(e, r, op.index2);
      }
      break;
    }
    case OpKind::FlatView: {
      // A split rule's temp relation (written by its first pipeline, read
      // by its second, no maintenance of its own): the rows become its NEW
      // index in the given order and, as the same index, its FULL one --
      // the second pipeline scans it as NEW in a recursive step and as FULL
      // in a non-recursive one (gpu_mir_helpers.h:436-466 builds the index
      // the scan iterates).
      Relation& r = e.rels[op.rel];
      r.newv.erase(op.index);
      const Index& nw = new_index(e, r, op.index);
      Versioned v;
      copy_index(nw, v.idx, e.s);
      r.full.clear();
      r.segs.clear();
      r.full_version += 1;
      r.delta_seq = r.full_version;
      v.version = r.full_version;
      r.full[op.index] = std::move(v);
      break;
    }
    case OpKind::Reconstruct:
    case OpKind::Hook:
      break;
  }
}

static void exec_ops(srd_engine& e, const std::vector<Op>& ops, uint64_t iteration, int step_no) {
  for (auto& op : ops) exec_op(e, op, iteration, step_no);
}

static void reset_state(srd_engine& e) {
  for (auto& r : e.rels) {
    r.full.clear();
    r.delta.clear();
    clear_new(r);
    r.full_version = r.delta_seq = 0;
    r.deltas.clear();
    r.inst_deltas.clear();
    r.segs.clear();
    r.segmented = false;
    r.export_idx = Index();
    r.export_version = ~uint64_t(0);
    r.deltas_step = -1;
    r.count = 0;
    r.counted = false;
  }
  e.counts.clear();
  e.iterations = 0;
  e.join_probes = 0;
  e.join_searches = 0;
}

}  // namespace srd

// ================================================= device-resident loop ====
// A recursive step whose iteration the device can size by itself runs
// without host read-backs: every buffer is reserved at a capacity, row
// counts live in a device control block (DevCtl), each pipeline's launches
// read their views and sizes from a device SrdDyn that a bind kernel fills,
// NEW is sorted by the device-planned radix sort, and DELTA = NEW - FULL and
// FULL u DELTA are merge-path kernels sized from device counts.  One
// iteration is captured into a HIP graph -- two, since FULL alternates
// between two buffers -- and replayed; the host reads the control block
// every few iterations.  The reference drives the same loop from the host
// (codegen/jit/orchestrator_jit.py:600-739) with a blocking size read per
// rule and phase (codegen/jit/complete_runner.py:1072-1112).
//
// Iteration 0 reads FULL wherever a rule reads DELTA
// (complete_runner.py:834-838); the loop starts with DELTA := FULL, which
// is the same computation.  Capacities are guesses: a kernel that finds one
// too small raises `overflow` with the size it needed and sets `skip`, which
// turns every later launch into a no-op.  All capacity checks of an
// iteration come before its first write of DELTA or FULL, so the host grows
// the buffers and runs the abandoned iteration again.
namespace srd {

struct DevRelState {
  int rel = -1;
  uint32_t k = 0;
  Order canon;          // index order of NEW, DELTA and FULL
  Order raw_order;      // column order the pipelines write NEW in
  DevBuf<uint32_t> raw, nsort, nroot_v, nroot_o, delta, droot_v, droot_o, full[2];
  DevBuf<uint8_t> sort_scratch;
  DevBuf<uint64_t> sd_split, sd_off, sd_scan, mg_split, rt_off, rt_scan, deltas, inst;
  DevBuf<uint32_t> sd_cnt, rt_cnt;
  uint32_t bound = 0;   // upper bound of every column value (for col_max)
};

struct DevPipeState {
  const SrdPipelineDesc* d = nullptr;
  SrdLaunchArgs a;
  SrdDyn dyn_host;
  DevBuf<uint8_t> dyn;
  DevBuf<uint32_t> key_rng, key_aux, work_ctr, counts;
  DevBuf<uint64_t> key_work, work_start, offsets, scan_keys, scan_slices;
  DlBind bind;
  DlOut out;
  int32_t vrel[SRD_MAX_VIEWS];
  bool direct = false;
};

struct DevLoop {
  std::vector<DevRelState> rels;
  std::vector<DevPipeState> pipes;
  DevBuf<uint8_t> ctl_buf;
  DevCtl* ctl = nullptr;
  uint64_t cap_new = 0, cap_full = 0, cap_keys = 0, cap_slices = 0, iter_cap = 0;
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  ~DevLoop() { drop_graphs(); }
  void drop_graphs() {
    for (auto& g : exec)
      if (g) {
        (void)hipGraphExecDestroy(g);
        g = nullptr;
      }
  }
};

}  // namespace srd

// ------------------------------------------------------- host side ----
namespace srd {

static void flatten_ops(const std::vector<Op>& ops, std::vector<const Op*>& out) {
  for (auto& op : ops) {
    if (op.kind == OpKind::Parallel) flatten_ops(op.kids, out);
    else out.push_back(&op);
  }
}

// Whether the device can size every op of the step: linear rules over the
// step's DELTAs (canonical order) and FULLs of relations the step does not
// write, every relation rebuilding its NEW once before any DELTA or FULL
// update, at most kDevRels relations of arity <= 4 and kDevPipes pipelines.
static bool devloop_eligible(srd_engine& e, const Step& st, std::string& why) {
  std::vector<const Op*> ops;
  flatten_ops(st.ops, ops);
  const std::vector<int>& rels = st.canonical_rels;
  auto in_step = [&](int r) { return std::find(rels.begin(), rels.end(), r) != rels.end(); };
  if (rels.empty() || rels.size() > size_t(kDevRels)) return why = "relation count", false;
  for (int ri : rels)
    if (e.rels[ri].arity == 0 || e.rels[ri].arity > 4) return why = "arity", false;
  bool updated = false;
  size_t npipes = 0;
  std::map<int, int> rebuilds;
  std::map<int, Order> raw_order;
  for (const Op* op : ops) {
    switch (op->kind) {
      case OpKind::Pipeline: {
        if (updated) return why = "pipeline after a DELTA/FULL update", false;
        const SrdPipelineDesc* d = op->pipe;
        if (d->count_only || !d->root_dyn || !d->count_dyn || !d->write_dyn)
          return why = "pipeline without device-sized launches", false;
        for (uint32_t v = 0; v < d->n_views; ++v) {
          const SrdViewSpec& vs = d->views[v];
          const int ri = rel_id(e, vs.rel);
          Order order(vs.cols, vs.cols + vs.ncols);
          if (in_step(ri)) {
            if (vs.ver != SRD_DELTA || order != e.rels[ri].canonical)
              return why = "view of a step relation other than its canonical DELTA", false;
          } else if (vs.ver != SRD_FULL) {
            return why = "DELTA/NEW view of a relation outside the step", false;
          }
        }
        for (uint32_t k = 0; k < d->n_dest; ++k) {
          const int ri = rel_id(e, d->dest[k].rel);
          if (!in_step(ri)) return why = "destination outside the step", false;
          Order order(d->dest[k].cols, d->dest[k].cols + d->dest[k].arity);
          auto it = raw_order.find(ri);
          if (it != raw_order.end() && it->second != order)
            return why = "destinations in different orders", false;
          raw_order[ri] = order;
        }
        ++npipes;
        break;
      }
      case OpKind::RebuildIndex:
        if (op->ver != SRD_NEW || !in_step(op->rel) || op->index != e.rels[op->rel].canonical)
          return why = "rebuild-index other than NEW in canonical order", false;
        if (updated) return why = "rebuild after a DELTA/FULL update", false;
        rebuilds[op->rel]++;
        break;
      case OpKind::ComputeDelta:
      case OpKind::MergeIndex:
        if (!in_step(op->rel) || op->index != e.rels[op->rel].canonical)
          return why = "delta/merge other than canonical", false;
        updated = true;
        break;
      case OpKind::ClearRelation:
      case OpKind::CheckSize:
        break;
      default:
        return why = std::string("op ") + op_name(op->kind), false;
    }
  }
  if (npipes == 0 || npipes > size_t(kDevPipes)) return why = "pipeline count", false;
  for (int ri : rels)
    if (rebuilds[ri] != 1 || !raw_order.count(ri)) return why = "NEW not rebuilt exactly once", false;
  return true;
}

struct DevCaps {
  uint64_t new_rows, full_rows, keys, slices, iters;
};

// IndexView of a static (FULL, not written by the step) view, with the
// bitmaps and root map the pipeline asks for, as run_pipeline prepares them.
static IndexView static_view(srd_engine& e, const SrdViewSpec& vs, std::vector<uint32_t>& vmax) {
  Relation& r = e.rels[rel_id(e, vs.rel)];
  Order order(vs.cols, vs.cols + vs.ncols);
  const Index& x = full_index(e, r, order);
  vmax = x.col_max;
  if (vs.want_bitmaps && x.arity == 2 && e.bitmap_bytes) {
    const HeavyBits* h = x.heavy.get();
    if (!h || h->built_for != x.cols.p || h->built_n != x.n)
      build_heavy_bitmaps(x, e.bitmap_min, e.bitmap_bytes, e.ws, e.s);
  }
  if (vs.want_root_pos && e.root_pos) {
    const RootPos* m = x.rootpos.get();
    if (!m || m->built_for != x.root_vals.p || m->built_r != x.r)
      build_root_pos(x, std::max<uint64_t>(8ull * x.r, uint64_t(1) << 20), e.s);
  }
  return x.view();
}

static void copy_cols(uint32_t* dst, uint64_t ds, const uint32_t* src, uint64_t ss, uint32_t k,
                      uint64_t n, hipStream_t s) {
  if (n == 0) return;
  for (uint32_t c = 0; c < k; ++c)
    SRD_CHECK(hipMemcpyAsync(dst + uint64_t(c) * ds, src + uint64_t(c) * ss, n * 4,
                             hipMemcpyDeviceToDevice, s));
}

// Allocate every buffer of the loop at the given capacities and fill the
// launch templates.  The iteration's state (FULL, DELTA with its root
// arrays, recorded deltas) is then copied in by the caller.
static void devloop_alloc(srd_engine& e, const Step& st, DevLoop& L, const DevCaps& cp) {
  hipStream_t s = e.s;
  L.cap_new = round_stride(cp.new_rows);
  L.cap_full = round_stride(cp.full_rows);
  L.cap_keys = cp.keys;
  L.cap_slices = cp.slices;
  L.iter_cap = cp.iters;
  const uint64_t cn = L.cap_new, cf = L.cap_full;
  const uint64_t sd_tiles = std::max<uint64_t>(mp_tiles_for(cn + cf), 1);
  const uint64_t mg_tiles = std::max<uint64_t>(mp_tiles_for(cf), 1);
  const uint64_t rt_tiles = roots_dev_tiles(cn);
  for (auto& R : L.rels) {
    const uint32_t k = R.k;
    R.raw.alloc(k * cn, s);
    R.nsort.alloc(k * cn, s);
    R.nroot_v.alloc(cn, s);
    R.nroot_o.alloc(cn + 1, s);
    R.sort_scratch.alloc(devsort_scratch_bytes(cn), s);
    devsort_clear_lookback(R.sort_scratch.p, cn, s);
    R.delta.alloc(k * cn, s);
    R.droot_v.alloc(cn, s);
    R.droot_o.alloc(cn + 1, s);
    R.full[0].alloc(k * cf, s);
    R.full[1].alloc(k * cf, s);
    R.sd_split.alloc(sd_tiles + 1, s);
    // the set difference stages its kept rows in the raw NEW columns and the
    // root build its key positions in the sorted NEW's root values: both are
    // dead once NEW is sorted
    R.sd_cnt.alloc(sd_tiles, s);
    R.sd_off.alloc(sd_tiles + 1, s);
    R.sd_scan.alloc(scan_dev_words(sd_tiles), s);
    R.mg_split.alloc(mg_tiles + 1, s);
    R.rt_cnt.alloc(rt_tiles, s);
    R.rt_off.alloc(rt_tiles + 1, s);
    R.rt_scan.alloc(scan_dev_words(rt_tiles), s);
    R.deltas.alloc(L.iter_cap, s);
    if (e.n_inst) R.inst.alloc(L.iter_cap * e.n_inst, s);
  }
  auto slot_of = [&](int ri) {
    for (size_t j = 0; j < L.rels.size(); ++j)
      if (L.rels[j].rel == ri) return int(j);
    return -1;
  };
  for (size_t pi = 0; pi < L.pipes.size(); ++pi) {
    DevPipeState& P = L.pipes[pi];
    const SrdPipelineDesc* d = P.d;
    P.key_rng.alloc(L.cap_keys * 2 * SRD_MAX_VIEWS, s);
    P.key_aux.alloc(L.cap_keys, s);
    P.key_work.alloc(L.cap_keys, s);
    P.work_start.alloc(L.cap_keys + 1, s);
    P.scan_keys.alloc(scan_dev_words(L.cap_keys), s);
    P.counts.alloc(L.cap_slices, s);
    P.offsets.alloc(L.cap_slices + 1, s);
    P.scan_slices.alloc(scan_dev_words(L.cap_slices), s);
    P.work_ctr.alloc(SRD_XCDS * SRD_CTR_STRIDE, s);
    P.dyn.alloc(sizeof(SrdDyn), s);
    std::memset(&P.a, 0, sizeof(P.a));
    P.a.grain = e.grain_override ? e.grain_override : d->default_grain;
    P.a.grid_blocks = 2048;
    P.a.n_ranges = e.xcd_ranges;
    P.a.window_min = e.window_min;
    P.a.key_rng = P.key_rng.p;
    P.a.key_aux = P.key_aux.p;
    P.a.key_work = P.key_work.p;
    P.a.work_start = P.work_start.p;
    P.a.counts = P.counts.p;
    P.a.offsets = P.offsets.p;
    P.a.work_ctr = P.work_ctr.p;
    P.a.dyn = reinterpret_cast<const SrdDyn*>(P.dyn.p);
    for (uint32_t k = 0; k < d->n_dest; ++k) {
      const int j = slot_of(rel_id(e, d->dest[k].rel));
      P.a.out[k].cols = L.rels[j].raw.p;
      P.a.out[k].stride = cn;
      P.a.out[k].base = 0;
      P.out.rel[k] = j;
    }
    P.out.n_dest = d->n_dest;
    P.out.pipe = uint32_t(pi);
    P.out.grain = P.a.grain;
    P.out.direct = P.direct ? 1u : 0u;
    P.out.cap_slices = L.cap_slices;
    P.out.cap_new = cn;
    P.bind.cap_keys = L.cap_keys;
    for (uint32_t v = 0; v < d->n_views; ++v) {
      if (P.vrel[v] < 0) continue;  // static: prepared once in dyn_host
      const DevRelState& R = L.rels[P.vrel[v]];
      IndexView iv;
      std::memset(&iv, 0, sizeof(iv));
      iv.cols = R.delta.p;
      iv.root_vals = R.droot_v.p;
      iv.root_offs = R.droot_o.p;
      iv.stride = cn;
      iv.arity = R.k;
      iv.bm_min = ~0u;
      P.dyn_host.views.v[v] = iv;
    }
    P.dyn_host.skip = 0;
    SRD_CHECK(hipMemcpyAsync(P.dyn.p, &P.dyn_host, sizeof(SrdDyn), hipMemcpyHostToDevice, s));
  }
  (void)st;
}

// Enqueue one iteration reading FULL from full[parity] (merged into the
// other buffer).  No allocation, no host synchronisation: it may be captured.
static void devloop_enqueue(srd_engine& e, const Step& st, DevLoop& L, int parity) {
  hipStream_t s = e.s;
  DevCtl* c = L.ctl;
  const uint32_t nrels = uint32_t(L.rels.size());
  const uint64_t cn = L.cap_new, cf = L.cap_full;
  auto slot_of = [&](int ri) {
    for (size_t j = 0; j < L.rels.size(); ++j)
      if (L.rels[j].rel == ri) return uint32_t(j);
    fail("device loop: relation outside the step");
  };
  std::vector<const Op*> ops;
  flatten_ops(st.ops, ops);
  size_t last_rebuild = 0;
  for (size_t i = 0; i < ops.size(); ++i)
    if (ops[i]->kind == OpKind::RebuildIndex) last_rebuild = i;
  dl_iter_begin(c, nrels, s);
  size_t pi = 0;
  for (size_t i = 0; i < ops.size(); ++i) {
    const Op* op = ops[i];
    switch (op->kind) {
      case OpKind::Pipeline: {
        DevPipeState& P = L.pipes[pi];
        const SrdPipelineDesc* d = P.d;
        SrdDyn* dy = reinterpret_cast<SrdDyn*>(P.dyn.p);
        dl_bind(c, dy, P.bind, s);
        d->root_dyn(&P.a, s);
        SRD_CHECK(hipGetLastError());
        exclusive_scan_u64_dev(P.key_work.p, P.work_start.p, L.cap_keys, &c->keys[pi], P.scan_keys.p, s);
        dl_plan(c, dy, P.work_start.p, P.out, s);
        fill_u32(P.work_ctr.p, 0, SRD_XCDS * SRD_CTR_STRIDE, s);
        if (P.direct) {
          d->direct_dyn(&P.a, s);
          SRD_CHECK(hipGetLastError());
        } else {
          d->count_dyn(&P.a, s);
          SRD_CHECK(hipGetLastError());
          exclusive_scan_u32_dev(P.counts.p, P.offsets.p, L.cap_slices, &c->slices[pi],
                                 P.scan_slices.p, s);
          dl_out(c, dy, P.offsets.p, P.out, s);
          fill_u32(P.work_ctr.p, 0, SRD_XCDS * SRD_CTR_STRIDE, s);
          d->write_dyn(&P.a, s);
          SRD_CHECK(hipGetLastError());
        }
        ++pi;
        break;
      }
      case OpKind::RebuildIndex: {
        const uint32_t j = slot_of(op->rel);
        DevRelState& R = L.rels[j];
        const Order p = perm_between(R.raw_order, R.canon);
        std::vector<const uint32_t*> in(R.k);
        for (uint32_t c2 = 0; c2 < R.k; ++c2) in[c2] = R.raw.p + uint64_t(p[c2]) * cn;
        build_index_dev(in.data(), R.k, cn, &c->new_n[j], 0, R.k <= 2 ? 1u : 2u,
                        R.sort_scratch.p, R.nsort.p, cn, R.nroot_v.p, R.nroot_o.p, &c->newc[j][0],
                        /*ticket=*/true, s, /*timed=*/false);
        if (i == last_rebuild) {
          dl_check(c, nrels, cf, s);
        }
        break;
      }
      case OpKind::ComputeDelta: {
        const uint32_t j = slot_of(op->rel);
        DevRelState& R = L.rels[j];
        dl_sd_prep(c, j, s);
        SetDiffDevScratch x{R.sd_split.p, R.raw.p, cn, R.sd_cnt.p, R.sd_off.p, R.sd_scan.p};
        set_difference_dev(R.k, R.nsort.p, cn, R.full[parity].p, cf, cn + cf, c->sd_nab[j], R.delta.p,
                           cn, x, &c->sd_total[j], s);
        dl_sd_fin(c, j, s);
        build_roots_dev(R.delta.p, cn, &c->delta_n[j], R.nroot_v.p, R.rt_cnt.p, R.rt_off.p,
                        R.rt_scan.p, R.droot_v.p, R.droot_o.p, &c->delta_r[j], s);
        if (e.n_inst) {
          dl_inst(c, j, R.delta.p, e.inst_bounds.p, e.n_inst, R.inst.p, s);
        }
        break;
      }
      case OpKind::MergeIndex: {
        const uint32_t j = slot_of(op->rel);
        DevRelState& R = L.rels[j];
        dl_mg_prep(c, j, s);
        merge_dev(R.k, R.full[parity].p, cf, R.delta.p, cn, cf, c->mg_nab[j], R.full[parity ^ 1].p,
                  cf, R.mg_split.p, s);
        dl_mg_fin(c, j, s);
        break;
      }
      default:
        break;
    }
  }
  DlEnd end{};
  end.nrels = nrels;
  for (uint32_t j = 0; j < nrels; ++j) end.deltas[j] = L.rels[j].deltas.p;
  dl_iter_end(c, end, s);
}

static void devloop_capture(srd_engine& e, const Step& st, DevLoop& L) {
  L.drop_graphs();
  std::lock_guard<std::recursive_mutex> cg(capture_mutex());
  for (int p = 0; p < 2; ++p) {
    hipGraph_t g = nullptr;
    SRD_CHECK(hipStreamBeginCapture(e.s, hipStreamCaptureModeRelaxed));
    try {
      devloop_enqueue(e, st, L, p);
    } catch (...) {
      (void)hipStreamEndCapture(e.s, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    SRD_CHECK(hipStreamEndCapture(e.s, &g));
    hipError_t err = hipGraphInstantiate(&L.exec[p], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    SRD_CHECK(err);
  }
}

}  // namespace srd

namespace srd {

// Device bytes of a loop at capacities cp (what devloop_alloc reserves).
static uint64_t devloop_bytes(const DevLoop& L, const DevCaps& cp) {
  const uint64_t cn = round_stride(cp.new_rows), cf = round_stride(cp.full_rows);
  uint64_t b = 0;
  for (auto& R : L.rels) {
    b += uint64_t(R.k) * 4 * (3 * cn + 2 * cf);   // raw (also set-difference scratch), sorted NEW, DELTA, FULL x2
    b += 4 * 4 * cn;                              // root arrays of sorted NEW (also root scratch) and DELTA
    b += devsort_scratch_bytes(cn);
    b += cp.iters * 8 * 2;
  }
  b += L.pipes.size() * (cp.keys * (4 * 2 * SRD_MAX_VIEWS + 4 + 8 + 8) + cp.slices * 12);
  return b;
}

// Copy the loop state (FULL in full[parity], DELTA and its root arrays,
// the recorded per-iteration sizes) from `from` into the fresh buffers of `to`.
static void devloop_copy_state(srd_engine& e, const DevLoop& from, int parity, DevLoop& to,
                               const DevCtl& h) {
  hipStream_t s = e.s;
  for (size_t j = 0; j < from.rels.size(); ++j) {
    const DevRelState& A = from.rels[j];
    DevRelState& B = to.rels[j];
    copy_cols(B.full[0].p, to.cap_full, A.full[parity].p, from.cap_full, A.k, h.full_n[j], s);
    copy_cols(B.delta.p, to.cap_new, A.delta.p, from.cap_new, A.k, h.delta_n[j], s);
    if (h.delta_n[j]) {
      SRD_CHECK(hipMemcpyAsync(B.droot_v.p, A.droot_v.p, h.delta_r[j] * 4, hipMemcpyDeviceToDevice, s));
      SRD_CHECK(hipMemcpyAsync(B.droot_o.p, A.droot_o.p, (h.delta_r[j] + 1) * 4,
                               hipMemcpyDeviceToDevice, s));
    }
    if (h.iter) {
      SRD_CHECK(hipMemcpyAsync(B.deltas.p, A.deltas.p, h.iter * 8, hipMemcpyDeviceToDevice, s));
      if (e.n_inst)
        SRD_CHECK(hipMemcpyAsync(B.inst.p, A.inst.p, h.iter * e.n_inst * 8, hipMemcpyDeviceToDevice, s));
    }
  }
}

// Run the step's iterations on the device.  Returns the iterations run;
// *converged tells whether the fixpoint was reached (else the host loop
// continues from the returned iteration with the state handed back).
static uint64_t run_step_device(srd_engine& e, const Step& st, int step_no, uint64_t max_iters,
                                bool* converged) {
  hipStream_t s = e.s;
  *converged = false;
  std::vector<const Op*> ops;
  flatten_ops(st.ops, ops);
  auto L = std::make_unique<DevLoop>();
  // relations and their entry state: FULL as built by the base step
  uint64_t full_max = 0;
  std::vector<const Index*> full0;
  for (int ri : st.canonical_rels) {
    Relation& r = e.rels[ri];
    DevRelState R;
    R.rel = ri;
    R.k = r.arity;
    R.canon = r.canonical;
    const Index& f = full_index(e, r, r.canonical);
    full0.push_back(&f);
    full_max = std::max<uint64_t>(full_max, f.n);
    for (uint32_t m : f.col_max) R.bound = std::max(R.bound, m);
    L->rels.push_back(std::move(R));
  }
  auto slot_of = [&](int ri) {
    for (size_t j = 0; j < L->rels.size(); ++j)
      if (L->rels[j].rel == ri) return int(j);
    return -1;
  };
  uint32_t vbound = 0;
  for (const Op* op : ops) {
    if (op->kind != OpKind::Pipeline) continue;
    DevPipeState P;
    P.d = op->pipe;
    const SrdPipelineDesc* d = P.d;
    P.direct = d->direct_ok && d->direct_dyn && !e.two_pass;
    std::memset(&P.dyn_host, 0, sizeof(P.dyn_host));
    std::memset(&P.bind, 0, sizeof(P.bind));
    std::memset(&P.out, 0, sizeof(P.out));
    for (uint32_t v = 0; v < d->n_views; ++v) {
      const int j = slot_of(rel_id(e, d->views[v].rel));
      P.vrel[v] = j;
      P.bind.vrel[v] = j;
      if (j < 0) {
        std::vector<uint32_t> vm;
        P.dyn_host.views.v[v] = static_view(e, d->views[v], vm);
        for (uint32_t m : vm) vbound = std::max(vbound, m);
      }
    }
    P.bind.n_views = d->n_views;
    P.bind.root_kind = d->root_kind;
    P.bind.n_root_src = d->root_kind == SRD_ROOT_JOIN ? d->n_root_src : 1;
    for (uint32_t i = 0; i < SRD_MAX_VIEWS; ++i) P.bind.root_view[i] = d->root_view[i];
    P.bind.pipe = uint32_t(L->pipes.size());
    for (uint32_t k = 0; k < d->n_dest; ++k) {
      const int j = slot_of(rel_id(e, d->dest[k].rel));
      L->rels[j].raw_order = Order(d->dest[k].cols, d->dest[k].cols + d->dest[k].arity);
    }
    L->pipes.push_back(std::move(P));
  }
  for (auto& R : L->rels) R.bound = std::max(R.bound, vbound);

  DevCaps cp;
  cp.full_rows = std::max<uint64_t>(uint64_t(1) << 20, 4 * full_max);
  cp.new_rows = std::max<uint64_t>(uint64_t(1) << 20, 2 * full_max);
  cp.keys = uint64_t(1) << 20;
  cp.slices = uint64_t(1) << 16;
  cp.iters = 1024;
  if (e.devloop_caps) {
    cp.full_rows = cp.new_rows = cp.keys = cp.slices = e.devloop_caps;
    cp.iters = std::max<uint64_t>(e.devloop_caps, 2);
  }
  // DELTA starts as FULL: both must hold the base step's FULL
  cp.full_rows = std::max<uint64_t>(cp.full_rows, std::max<uint64_t>(full_max, 1));
  cp.new_rows = std::max<uint64_t>(cp.new_rows, std::max<uint64_t>(full_max, 1));
  if (full_max > e.devloop_max_rows) {
    if (e.trace) {
      std::fprintf(stderr, "[srd]   device loop: FULL above %llu rows: host loop\n",
                   (unsigned long long)e.devloop_max_rows);
      std::fflush(stderr);
    }
    return 0;
  }
  if (devloop_bytes(*L, cp) > e.devloop_max_bytes) {
    if (e.trace) {
      std::fprintf(stderr, "[srd]   device loop: initial capacities above the memory budget: host loop\n");
      std::fflush(stderr);
    }
    return 0;
  }
  try {
    devloop_alloc(e, st, *L, cp);
  } catch (const std::exception& ex) {
    if (e.trace) {
      std::fprintf(stderr, "[srd]   device loop: %s: host loop\n", ex.what());
      std::fflush(stderr);
    }
    return 0;
  }

  // DELTA := FULL (iteration 0 reads FULL where the rules read DELTA)
  DevCtl h;
  std::memset(&h, 0, sizeof(h));
  h.max_iter = max_iters;
  h.iter_cap = L->iter_cap;
  for (size_t j = 0; j < L->rels.size(); ++j) {
    DevRelState& R = L->rels[j];
    const Index& f = *full0[j];
    if (f.n > L->cap_full || f.n > L->cap_new) fail("device loop: initial capacities below FULL");
    copy_cols(R.full[0].p, L->cap_full, f.cols.p, f.stride, R.k, f.n, s);
    copy_cols(R.delta.p, L->cap_new, f.cols.p, f.stride, R.k, f.n, s);
    if (f.n) {
      SRD_CHECK(hipMemcpyAsync(R.droot_v.p, f.root_vals.p, uint64_t(f.r) * 4, hipMemcpyDeviceToDevice, s));
      SRD_CHECK(hipMemcpyAsync(R.droot_o.p, f.root_offs.p, (uint64_t(f.r) + 1) * 4,
                               hipMemcpyDeviceToDevice, s));
    }
    h.full_n[j] = h.delta_n[j] = f.n;
    h.delta_r[j] = f.r;
  }
  L->ctl_buf.alloc(sizeof(DevCtl), s);
  L->ctl = reinterpret_cast<DevCtl*>(L->ctl_buf.p);
  uint64_t pin_bytes = 0;
  DevCtl* hp = static_cast<DevCtl*>(pinned_acquire(sizeof(DevCtl), &pin_bytes));
  std::unique_ptr<DevCtl, std::function<void(DevCtl*)>> pin_guard(
      hp, [pin_bytes](DevCtl* q) { pinned_release(q, pin_bytes); });
  *hp = h;
  SRD_CHECK(hipMemcpyAsync(L->ctl, hp, sizeof(DevCtl), hipMemcpyHostToDevice, s));
  SRD_CHECK(hipStreamSynchronize(s));
  if (e.devloop_graph) devloop_capture(e, st, *L);

  int parity = 0;
  uint64_t done_iters = 0;
  const uint64_t batch = e.devloop_batch;
  for (;;) {
    for (uint64_t b = 0; b < batch; ++b) {
      const int p = int((parity + b) & 1);
      if (e.devloop_graph) SRD_CHECK(hipGraphLaunch(L->exec[p], s));
      else devloop_enqueue(e, st, *L, p);
    }
    SRD_CHECK(hipMemcpyAsync(hp, L->ctl, sizeof(DevCtl), hipMemcpyDeviceToHost, s));
    SRD_CHECK(hipStreamSynchronize(s));
    h = *hp;
    parity ^= int((h.iter - done_iters) & 1);
    done_iters = h.iter;
    if (e.trace) {
      std::fprintf(stderr, "[srd]   device loop: %llu iterations, overflow %u, done %u\n",
                   (unsigned long long)h.iter, h.overflow, h.done);
      std::fflush(stderr);
    }
    if (h.err) fail("device loop: a device-planned sort reported an error");
    if (h.overflow) {
      if (h.need_new > e.devloop_max_rows || h.need_full > e.devloop_max_rows) {
        if (e.trace) {
          std::fprintf(stderr, "[srd]   device loop: NEW or FULL above %llu rows: host loop from "
                       "iteration %llu\n", (unsigned long long)e.devloop_max_rows,
                       (unsigned long long)h.iter);
          std::fflush(stderr);
        }
        break;
      }
      DevCaps nc{L->cap_new, L->cap_full, L->cap_keys, L->cap_slices, L->iter_cap};
      auto grow = [](uint64_t cur, uint64_t need) {
        return std::max<uint64_t>(2 * cur, need + need / 4);
      };
      if (h.overflow & kOvfNew) {
        nc.new_rows = grow(nc.new_rows, h.need_new);
        // FULL u NEW follows: grow FULL with NEW instead of re-running the
        // iteration once more when the union outgrows it
        uint64_t fmax = 0;
        for (size_t j = 0; j < L->rels.size(); ++j) fmax = std::max(fmax, h.full_n[j]);
        nc.full_rows = std::max(nc.full_rows, fmax + h.need_new);
      }
      if (h.overflow & kOvfFull) nc.full_rows = grow(nc.full_rows, h.need_full);
      if (h.overflow & kOvfKeys) nc.keys = grow(nc.keys, h.need_keys);
      if (h.overflow & kOvfSlices) nc.slices = grow(nc.slices, h.need_slices);
      if (h.overflow & kOvfIters) nc.iters = 2 * nc.iters;
      if (e.trace) {
        std::fprintf(stderr, "[srd]   device loop: grow (overflow %u): new %llu full %llu keys %llu "
                     "slices %llu iters %llu\n", h.overflow, (unsigned long long)nc.new_rows,
                     (unsigned long long)nc.full_rows, (unsigned long long)nc.keys,
                     (unsigned long long)nc.slices, (unsigned long long)nc.iters);
        std::fflush(stderr);
      }
      if (devloop_bytes(*L, nc) > e.devloop_max_bytes) {
        if (e.trace) {
          std::fprintf(stderr, "[srd]   device loop: growth above the memory budget: host loop from "
                       "iteration %llu\n", (unsigned long long)h.iter);
          std::fflush(stderr);
        }
        break;
      }
      auto L2 = std::make_unique<DevLoop>();
      for (auto& R : L->rels) {
        DevRelState B;
        B.rel = R.rel;
        B.k = R.k;
        B.canon = R.canon;
        B.raw_order = R.raw_order;
        B.bound = R.bound;
        L2->rels.push_back(std::move(B));
      }
      for (auto& P : L->pipes) {
        DevPipeState Q;
        Q.d = P.d;
        Q.direct = P.direct;
        Q.dyn_host = P.dyn_host;
        Q.bind = P.bind;
        std::memcpy(Q.vrel, P.vrel, sizeof(Q.vrel));
        L2->pipes.push_back(std::move(Q));
      }
      try {
        devloop_alloc(e, st, *L2, nc);
      } catch (const std::exception& ex) {
        if (e.trace) {
          std::fprintf(stderr, "[srd]   device loop: %s: host loop from iteration %llu\n", ex.what(),
                       (unsigned long long)h.iter);
          std::fflush(stderr);
        }
        break;
      }
      devloop_copy_state(e, *L, parity, *L2, h);
      L2->ctl_buf = L->ctl_buf;
      L2->ctl = L->ctl;
      h.overflow = 0;
      h.skip = 0;
      h.need_new = h.need_full = h.need_keys = h.need_slices = 0;
      h.iter_cap = L2->iter_cap;
      *hp = h;
      SRD_CHECK(hipMemcpyAsync(L2->ctl, hp, sizeof(DevCtl), hipMemcpyHostToDevice, s));
      SRD_CHECK(hipStreamSynchronize(s));
      L = std::move(L2);
      parity = 0;
      if (e.devloop_graph) devloop_capture(e, st, *L);
      continue;
    }
    if (h.done || h.iter >= max_iters) break;
  }

  // hand the state back to the relations
  std::vector<uint64_t> dl(h.iter);
  for (size_t j = 0; j < L->rels.size(); ++j) {
    DevRelState& R = L->rels[j];
    Relation& r = e.rels[R.rel];
    if (h.iter) {
      SRD_CHECK(hipMemcpyAsync(dl.data(), R.deltas.p, h.iter * 8, hipMemcpyDeviceToHost, s));
      SRD_CHECK(hipStreamSynchronize(s));
    }
    if (r.deltas_step != step_no) {
      r.deltas.clear();
      r.inst_deltas.clear();
      r.deltas_step = step_no;
    }
    for (uint64_t i = 0; i < h.iter; ++i) {
      r.deltas.push_back(dl[i]);
      if (e.n_inst) {
        DevBuf<uint64_t> per;
        per.p = R.inst.p + i * e.n_inst;
        per.n = e.n_inst;
        per.s = s;
        per.owner = R.inst.owner;
        r.inst_deltas.push_back(per);
      }
    }
    Index f;
    f.arity = R.k;
    f.n = h.full_n[j];
    f.stride = L->cap_full;
    f.cols = R.full[parity];
    f.col_max.assign(R.k, f.n ? R.bound : 0);
    build_roots(f, e.ws, e.s);
    Index d;
    d.arity = R.k;
    d.n = h.delta_n[j];
    d.r = uint32_t(h.delta_r[j]);
    d.stride = L->cap_new;
    d.col_max.assign(R.k, d.n ? R.bound : 0);
    if (d.n) {
      d.cols = R.delta;
      d.root_vals = R.droot_v;
      d.root_offs = R.droot_o;
    }
    r.segs.clear();
    r.full.clear();
    r.full_version += h.iter;
    r.delta_seq = r.full_version;
    Versioned v;
    v.idx = std::move(f);
    v.version = r.full_version;
    r.full[r.canonical] = std::move(v);
    r.delta.clear();
    r.delta[r.canonical] = std::move(d);
    clear_new(r);
  }
  e.join_probes += h.join_rows;
  *converged = h.done != 0 || h.iter >= max_iters;
  return h.iter;
}

}  // namespace srd

namespace srd {

static void run_program(srd_engine& e, uint64_t max_iters) {
  reset_state(e);
  for (auto& r : e.rels) upload_input(e, r);
  auto t0 = std::chrono::steady_clock::now();
  for (size_t si = 0; si < e.steps.size(); ++si) {
    const Step& st = e.steps[si];
    auto ts = std::chrono::steady_clock::now();
    if (!st.recursive) {
      exec_ops(e, st.ops, 0, int(si));
    } else {
      for (int ri : st.segmented_rels) e.rels[ri].segmented = true;
      uint64_t it = 0;
      bool converged = false;
      std::string why;
      if (e.devloop && !e.count_probes && devloop_eligible(e, st, why)) {
        it = run_step_device(e, st, int(si), max_iters, &converged);
        converged = converged || it >= max_iters;
      } else if (e.trace && e.devloop) {
        std::fprintf(stderr, "[srd] step %zu: host-driven loop (%s)\n", si, why.c_str());
        std::fflush(stderr);
      }
      for (; !converged && it < max_iters; ++it) {
        if (it > 0) {
          uint64_t tot = 0;
          for (int ri : st.canonical_rels) {
            Relation& r = e.rels[ri];
            auto d = r.delta.find(r.canonical);
            if (d != r.delta.end()) tot += d->second.n;
          }
          if (std::getenv("SRDATALOG_PRINT_DELTA"))
            std::cerr << "[iter " << it << "] delta=" << tot << std::endl;
          if (tot == 0) break;
        }
        exec_ops(e, st.ops, it, int(si));
      }
      for (int ri : st.segmented_rels) {
        Relation& r = e.rels[ri];
        materialize_full(e, r);
        r.segmented = false;
      }
      e.iterations = it;
    }
    SRD_CHECK(hipStreamSynchronize(e.s));
    if (!e.quiet && std::getenv("SRDATALOG_PRINT_STEPS")) {
      double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ts).count();
      std::cout << "[Step " << si << "] " << ms << " ms" << std::endl;
    }
  }
  SRD_CHECK(hipStreamSynchronize(e.s));
  e.run_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  e.ran = true;
  if (e.op_profile) {
    std::vector<std::pair<std::string, std::pair<double, uint64_t>>> v(e.op_times.begin(),
                                                                       e.op_times.end());
    std::sort(v.begin(), v.end(), [](auto& a, auto& b) { return a.second.first > b.second.first; });
    std::fprintf(stderr, "[srd] op profile of a %.3f s run (host ms, count):\n", e.run_seconds);
    for (auto& x : v)
      std::fprintf(stderr, "[srd]   %9.3f %7llu  %s\n", x.second.first * 1e3,
                   (unsigned long long)x.second.second, x.first.c_str());
    std::fflush(stderr);
    e.op_times.clear();
  }
  if (!e.quiet)
    for (auto& r : e.rels)
      if (r.print_size) {
        auto it = r.full.find(r.canonical);
        uint64_t n = it == r.full.end() ? 0 : it->second.idx.n;
        std::cout << " >>>>>>>>>>>>>>>>> " << r.name << " : " << n << std::endl;
      }
}

// ======================================================= resumable runs ====
// Iteration-granular serving.  A program whose last step is its only
// recursive one runs its earlier steps once (run_begin) and then a few
// iterations per call (run_iters).  Between two calls, absorb() unites into
// it the state another engine of the same program reached after its own
// run_begin -- the base-step results of new instances -- and erase_ranges()
// drops the rows of finished instances.  Instances occupy disjoint value
// ranges, every column of an instance's rows lies in its range, so in any
// lexicographic index an instance's rows are contiguous and never join with
// another instance's: absorbed rows enter FULL and DELTA of the running
// step as they would enter a fresh run at its iteration 0 (which reads FULL
// for DELTA), and each instance's rows and per-iteration deltas are those of
// a run of that instance alone, shifted by the iteration it was absorbed at.

static uint64_t delta_total(srd_engine& e, const Step& st) {
  uint64_t tot = 0;
  for (int ri : st.canonical_rels) {
    Relation& r = e.rels[ri];
    auto d = r.delta.find(r.canonical);
    if (d != r.delta.end()) tot += d->second.n;
  }
  return tot;
}

// A step that computes nothing (post-stratum reconstructs, hooks, checks).
static bool step_is_noop(const Step& st) {
  for (const Op& op : st.ops)
    if (op.kind != OpKind::Reconstruct && op.kind != OpKind::Hook && op.kind != OpKind::CheckSize)
      return false;
  return true;
}

static void run_begin(srd_engine& e) {
  int rs = -1;
  for (size_t si = 0; si < e.steps.size(); ++si)
    if (e.steps[si].recursive) {
      if (rs >= 0) fail("srdatalog_run_begin: the program has more than one recursive step");
      rs = int(si);
    }
  bool last = rs >= 0;
  for (size_t si = size_t(std::max(rs, 0)) + 1; last && si < e.steps.size(); ++si)
    last = step_is_noop(e.steps[si]);
  if (!last)
    fail("srdatalog_run_begin: the program's only recursive step must be its last computing step");
  reset_state(e);
  for (auto& r : e.rels) upload_input(e, r);
  for (int si = 0; si < rs; ++si) exec_ops(e, e.steps[si].ops, 0, si);
  // FULL of relations no join of the step reads stays log-structured, so an
  // iteration (and an absorb) costs what it adds, not the size of FULL
  for (int ri : e.steps[rs].segmented_rels) e.rels[ri].segmented = true;
  SRD_CHECK(hipStreamSynchronize(e.s));
  e.res_step = size_t(rs);
  e.res_it = 0;
  e.resumable = true;
  e.iterations = 0;
  e.ran = true;
}

static uint64_t run_iters(srd_engine& e, uint64_t k, bool* converged) {
  if (!e.resumable) fail("srdatalog_run_iterations: no run was begun");
  const Step& st = e.steps[e.res_step];
  uint64_t done = 0;
  *converged = false;
  for (; done < k; ++done) {
    if (e.res_it > 0 && delta_total(e, st) == 0) {
      *converged = true;
      break;
    }
    exec_ops(e, st.ops, e.res_it, int(e.res_step));
    e.res_it += 1;
  }
  if (!*converged && e.res_it > 0 && delta_total(e, st) == 0) *converged = true;
  SRD_CHECK(hipStreamSynchronize(e.s));
  e.iterations = e.res_it;
  return done;
}

static bool in_step(const Step& st, int ri) {
  return std::find(st.canonical_rels.begin(), st.canonical_rels.end(), ri) != st.canonical_rels.end();
}

// The disjoint sorted parts FULL consists of (canonical order): the oldest
// and the segments of a segmented FULL, else the one canonical index.
static std::vector<Index*> full_parts(srd_engine& e, Relation& r) {
  std::vector<Index*> parts;
  if (!r.canonical_set || (r.full.empty() && !r.has_input)) return parts;
  auto it = r.full.find(r.canonical);
  if (!r.segs.empty() && it != r.full.end()) {
    parts.push_back(&it->second.idx);
    for (auto& sg : r.segs) parts.push_back(&sg);
  } else {
    parts.push_back(const_cast<Index*>(&full_index(e, r, r.canonical)));
  }
  return parts;
}

static void absorb(srd_engine& d, srd_engine& s) {
  if (&d == &s) fail("srdatalog_absorb: an engine cannot absorb itself");
  if (!d.resumable || !s.resumable || s.res_it != 0)
    fail("srdatalog_absorb: both engines need a begun run, the absorbed one at iteration 0");
  if (d.rels.size() != s.rels.size()) fail("srdatalog_absorb: engines of different programs");
  SRD_CHECK(hipStreamSynchronize(s.s));
  const Step& st = d.steps[d.res_step];
  for (size_t i = 0; i < d.rels.size(); ++i) {
    Relation& rs = s.rels[i];
    Relation& rd = d.rels[i];
    if (rs.name != rd.name || rs.arity != rd.arity)
      fail("srdatalog_absorb: engines of different programs");
    if (!rs.canonical_set || (rs.full.empty() && !rs.has_input)) continue;
    const Index& add = full_index(s, rs, rs.canonical);
    if (add.n == 0) continue;
    if (rd.segmented && rd.full.count(rd.canonical)) {
      // the youngest segment; the two youngest merge while within 2x
      Index c;
      copy_index(add, c, d.s);
      rd.segs.push_back(std::move(c));
      merge_young_segments(d, rd);
      Versioned base = std::move(rd.full[rd.canonical]);
      rd.full.clear();
      rd.full_version += 1;
      rd.delta_seq = rd.full_version;
      base.version = rd.full_version;
      rd.full[rd.canonical] = std::move(base);
    } else {
      Versioned v;
      const Index& cur = full_index(d, rd, rd.canonical);
      if (cur.n == 0) copy_index(add, v.idx, d.s);
      else merge_disjoint(cur, add, v.idx, d.ws, d.s);
      rd.full.clear();
      rd.segs.clear();
      rd.full_version += 1;
      rd.delta_seq = rd.full_version;
      v.version = rd.full_version;
      rd.full[rd.canonical] = std::move(v);
    }
    if (in_step(st, int(i))) {
      auto it = rd.delta.find(rd.canonical);
      Index nd;
      if (it == rd.delta.end() || it->second.n == 0) copy_index(add, nd, d.s);
      else merge_disjoint(it->second, add, nd, d.ws, d.s);
      rd.delta.clear();
      rd.delta[rd.canonical] = std::move(nd);
    }
  }
  SRD_CHECK(hipStreamSynchronize(d.s));
}

// Rows of x (sorted on its first column) with first-column value in each of
// the m ranges [lohi[2g], lohi[2g+1]) (ascending, disjoint): first row and
// count of each.
static std::vector<std::pair<uint64_t, uint64_t>> value_ranges(srd_engine& e, const Index& x,
                                                               const uint32_t* lohi, uint32_t m) {
  std::vector<std::pair<uint64_t, uint64_t>> out(m, {0, 0});
  if (x.n == 0 || m == 0) return out;
  // bounds 0, lo0, hi0, lo1, ...: slot 0 counts the rows before the first
  // range, then ranges and gaps alternate
  std::vector<uint32_t> b(2ull * m + 1);
  b[0] = 0;
  for (uint64_t i = 0; i < 2ull * m; ++i) b[i + 1] = lohi[i];
  DevBuf<uint32_t> db(b.size(), e.s);
  DevBuf<uint64_t> per(2ull * m, e.s);
  SRD_CHECK(hipMemcpyAsync(db.p, b.data(), b.size() * 4, hipMemcpyHostToDevice, e.s));
  instance_counts(x.cols.p, x.n, db.p, 2 * m, per.p, e.s);
  std::vector<uint64_t> h(2ull * m);
  SRD_CHECK(hipMemcpyAsync(h.data(), per.p, h.size() * 8, hipMemcpyDeviceToHost, e.s));
  SRD_CHECK(hipStreamSynchronize(e.s));
  uint64_t pos = 0;
  for (uint32_t g = 0; g < m; ++g) {
    pos += h[2ull * g];
    out[g] = {pos, h[2ull * g + 1]};
    pos += h[2ull * g + 1];
  }
  return out;
}

static std::pair<uint64_t, uint64_t> value_range(srd_engine& e, const Index& x, uint32_t lo,
                                                 uint32_t hi) {
  if (lo >= hi) return {0, 0};
  const uint32_t lh[2] = {lo, hi};
  return value_ranges(e, x, lh, 1)[0];
}

// x without the row ranges rm (ascending, disjoint; first row, count).
static void erase_rows(srd_engine& e, const Index& x,
                       const std::vector<std::pair<uint64_t, uint64_t>>& rm, Index& out) {
  uint64_t gone = 0;
  for (auto& [a, c] : rm) gone += c;
  const uint64_t n2 = x.n - gone;
  if (n2 == 0) {
    out = empty_index(x.arity);
    return;
  }
  DevBuf<uint32_t> cols(uint64_t(x.arity) * n2, e.s);
  std::vector<const uint32_t*> in(x.arity);
  for (uint32_t c = 0; c < x.arity; ++c) {
    const uint32_t* src = x.cols.p + uint64_t(c) * x.stride;
    uint32_t* dst = cols.p + uint64_t(c) * n2;
    uint64_t from = 0, to = 0;
    for (size_t i = 0; i <= rm.size(); ++i) {
      const uint64_t end = i < rm.size() ? rm[i].first : x.n;
      if (end > from)
        SRD_CHECK(hipMemcpyAsync(dst + to, src + from, (end - from) * 4, hipMemcpyDeviceToDevice, e.s));
      to += end - from;
      if (i < rm.size()) from = rm[i].first + rm[i].second;
    }
    in[c] = dst;
  }
  build_index(in.data(), x.arity, n2, out, e.ws, e.s, nullptr, &cols, n2, /*presorted=*/true);
}

static void erase_ranges(srd_engine& e, const uint32_t* lohi, uint32_t m) {
  if (!e.resumable) fail("srdatalog_erase_ranges: no run was begun");
  for (uint64_t i = 1; i < 2ull * m; ++i)
    if (lohi[i] < lohi[i - 1]) fail("srdatalog_erase_ranges: ranges must be ascending and disjoint");
  if (m == 0) return;
  auto any = [](const std::vector<std::pair<uint64_t, uint64_t>>& v) {
    for (auto& p : v)
      if (p.second) return true;
    return false;
  };
  for (auto& r : e.rels) {
    bool changed = false;
    for (Index* part : full_parts(e, r)) {
      auto rm = value_ranges(e, *part, lohi, m);
      if (!any(rm)) continue;
      Index kept;
      erase_rows(e, *part, rm, kept);
      *part = std::move(kept);
      changed = true;
    }
    if (changed) {
      // FULL in other column orders follows the canonical parts again
      Versioned base = std::move(r.full[r.canonical]);
      r.full.clear();
      r.full_version += 1;
      r.delta_seq = r.full_version;
      base.version = r.full_version;
      r.full[r.canonical] = std::move(base);
    }
    auto it = r.delta.find(r.canonical);
    if (it != r.delta.end()) {
      auto drm = value_ranges(e, it->second, lohi, m);
      if (any(drm)) {
        Index nd;
        erase_rows(e, it->second, drm, nd);
        r.delta.clear();
        r.delta[r.canonical] = std::move(nd);
      }
    }
  }
  SRD_CHECK(hipStreamSynchronize(e.s));
}

// =============================================================== CSV ====
static void load_csv_file(Relation& r, const std::string& path) {
  int fd = open(path.c_str(), O_RDONLY);
  if (fd < 0) fail("could not open file: " + path);
  struct stat sb;
  if (fstat(fd, &sb) != 0) {
    close(fd);
    fail("could not stat: " + path);
  }
  size_t size = size_t(sb.st_size);
  r.host_rows.clear();
  r.host_n = 0;
  r.has_input = true;
  r.dev_input_valid = false;
  if (size == 0) {
    close(fd);
    return;
  }
  char* m = (char*)mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
  close(fd);
  if (m == MAP_FAILED) fail("could not mmap: " + path);
  const char* p = m;
  const char* end = m + size;
  // delimiter from the first line: tab, else comma, else runs of spaces
  const char* le = p;
  while (le < end && *le != '\n') ++le;
  char delim = ' ';
  bool tabs = std::find(p, le, '\t') != le, commas = std::find(p, le, ',') != le;
  if (tabs) delim = '\t';
  else if (commas) delim = ',';
  r.host_rows.reserve(size / 8);
  std::vector<uint32_t> row(r.arity);
  while (p < end) {
    while (p < end && (*p == '\n' || *p == '\r')) ++p;
    if (p >= end) break;
    bool ok = true;
    for (uint32_t c = 0; c < r.arity; ++c) {
      while (p < end && (*p == ' ' || *p == '\t') && *p != delim) ++p;
      if (delim == ' ' || delim == '\t')
        while (p < end && (*p == ' ' || *p == '\t')) ++p;
      bool neg = false;
      if (p < end && *p == '-') {
        neg = true;
        ++p;
      }
      uint64_t v = 0;
      const char* st = p;
      while (p < end && *p >= '0' && *p <= '9') v = v * 10 + uint64_t(*p++ - '0');
      if (p == st) ok = false;
      row[c] = neg ? uint32_t(-int64_t(v)) : uint32_t(v);
      while (p < end && (*p == ' ' || *p == '\t') && delim != ' ' && delim != '\t') ++p;
      if (p < end && *p == delim) ++p;
    }
    while (p < end && *p != '\n') ++p;
    if (p < end) ++p;
    if (ok) {
      r.host_rows.insert(r.host_rows.end(), row.begin(), row.end());
      r.host_n++;
    }
  }
  munmap(m, size);
}

}  // namespace srd

// ============================================================ C ABI ====
using namespace srd;

template <class F>
static int eguard(srd_engine* e, int code_on_error, F&& f) {
  if (!e) return 1;
  RunScope rs;
  e->ws.ticket_order = !rs.alone;
  try {
    f();
    e->err.clear();
    return 0;
  } catch (const std::exception& ex) {
    e->err = ex.what();
    std::cerr << "srdatalog: " << ex.what() << std::endl;
    return code_on_error;
  } catch (...) {
    e->err = "unknown error";
    return code_on_error;
  }
}

extern "C" {

srd_engine* srd_engine_create(const char* text, const SrdPipelineDesc* pipes, uint32_t n) {
  auto* e = new srd_engine();
  try {
    pool_init(0);
    e->s = stream_acquire();
    const char* tr = std::getenv("SRDATALOG_TRACE");
    e->trace = tr && *tr && *tr != '0';
    const char* op = std::getenv("SRDATALOG_OP_PROFILE");
    e->op_profile = op && *op && *op != '0';
    const char* q = std::getenv("SRDATALOG_QUIET");
    e->quiet = q && *q && *q != '0';
    if (const char* g = std::getenv("SRDATALOG_GRAIN")) e->grain_override = uint32_t(std::atoi(g));
    if (const char* c = std::getenv("SRDATALOG_STAGE_CAP")) e->stage_cap_env = std::strtoull(c, nullptr, 10);
    if (const char* c = std::getenv("SRDATALOG_NEW_CHUNK_ROWS"))
      e->new_chunk_rows = std::strtoull(c, nullptr, 10);
    if (const char* c = std::getenv("SRDATALOG_NO_WEIGH")) e->no_weigh = *c && *c != '0';
    if (const char* c = std::getenv("SRDATALOG_XCD_RANGES"))
      e->xcd_ranges = std::min<uint32_t>(std::max<uint32_t>(uint32_t(std::atoi(c)), 1), SRD_XCDS);
    if (const char* c = std::getenv("SRDATALOG_MARK_CHUNK_SLICES"))
      e->mark_chunk_slices = std::strtoull(c, nullptr, 10);
    if (const char* c = std::getenv("SRDATALOG_DENSE_MAX_BITS"))
      e->dense_max_bits = std::strtoull(c, nullptr, 10);
    if (const char* c = std::getenv("SRDATALOG_WINDOW_MIN"))
      e->window_min = uint32_t(std::strtoul(c, nullptr, 10));
    if (const char* c = std::getenv("SRDATALOG_WEIGH_RANGES"))
      e->weigh_ranges = std::min<uint32_t>(std::max<uint32_t>(uint32_t(std::atoi(c)), 1), SRD_XCDS);
    if (const char* c = std::getenv("SRDATALOG_BITMAP_MIN"))
      e->bitmap_min = std::max<uint32_t>(uint32_t(std::strtoul(c, nullptr, 10)), 16);
    if (const char* c = std::getenv("SRDATALOG_BITMAP_BYTES")) e->bitmap_bytes = std::strtoull(c, nullptr, 10);
    if (const char* c = std::getenv("SRDATALOG_ROOT_POS")) e->root_pos = !(*c == '0' && c[1] == 0);
    if (const char* c = std::getenv("SRDATALOG_WEIGH_MIN")) e->weigh_min = std::strtoull(c, nullptr, 10);
    if (const char* c = std::getenv("SRDATALOG_SLICES_PER_WAVE"))
      e->slices_per_wave = std::max<uint64_t>(std::strtoull(c, nullptr, 10), 1);
    if (const char* c = std::getenv("SRDATALOG_WEIGH_TARGET"))
      e->weigh_target = std::max<uint64_t>(std::strtoull(c, nullptr, 10), 1);
    if (const char* c = std::getenv("SRDATALOG_MIN_SLICES_PER_WAVE"))
      e->min_slices_per_wave = std::strtoull(c, nullptr, 10);
    if (const char* c = std::getenv("SRDATALOG_STAGE_MAX_BYTES"))
      e->stage_max_bytes = std::max<uint64_t>(std::strtoull(c, nullptr, 10), uint64_t(1) << 20);
    if (const char* c = std::getenv("SRDATALOG_DEVLOOP")) e->devloop = !(*c == '0' && c[1] == 0);
    if (const char* c = std::getenv("SRDATALOG_GROUP_STREAMS"))
      e->group_streams = !(*c == '0' && c[1] == 0);
    if (const char* c = std::getenv("SRDATALOG_DEVLOOP_GRAPH"))
      e->devloop_graph = !(*c == '0' && c[1] == 0);
    if (const char* c = std::getenv("SRDATALOG_DEVLOOP_BATCH"))
      e->devloop_batch = std::max<uint64_t>(std::strtoull(c, nullptr, 10), 1);
    if (const char* c = std::getenv("SRDATALOG_DEVLOOP_CAPS"))
      e->devloop_caps = std::strtoull(c, nullptr, 10);
    if (const char* c = std::getenv("SRDATALOG_DEVLOOP_MAX_ROWS"))
      e->devloop_max_rows = std::strtoull(c, nullptr, 10);
    {
      size_t fr = 0, tot = 0;
      std::lock_guard<std::recursive_mutex> cg(capture_mutex());
      e->devloop_max_bytes = hipMemGetInfo(&fr, &tot) == hipSuccess ? uint64_t(tot) / 4 : uint64_t(32) << 30;
      if (const char* c = std::getenv("SRDATALOG_DEVLOOP_MAX_BYTES"))
        e->devloop_max_bytes = std::strtoull(c, nullptr, 10);
    }
    if (const char* m = std::getenv("SRDATALOG_JOIN_MODE")) {
      e->two_pass = std::string(m) == "twopass";
      e->no_direct = std::string(m) == "fused";
    }
    for (uint32_t i = 0; i < n; ++i) e->pipes[pipes[i].rule] = &pipes[i];
    parse_program(*e, text);
  } catch (const std::exception& ex) {
    std::cerr << "srdatalog_init: " << ex.what() << std::endl;
    e->err = ex.what();
  }
  return e;
}

void srd_engine_destroy(srd_engine* e) {
  if (!e) return;
  hipStream_t s = e->s;
  std::vector<hipStream_t> gs = e->group_s;
  {
    std::unique_ptr<srd_engine> hold(e);
    for (auto& r : e->rels) {
      r.full.clear();
      r.delta.clear();
      clear_new(r);
      r.dev_input.release();
    }
    e->scratch.clear();
    for (hipEvent_t ev : e->fence_events) (void)hipEventDestroy(ev);
    pinned_release(e->run_pins, e->run_pins_bytes);
  }
  // the streams (and the blocks freed on them) serve the next engine
  for (hipStream_t g : gs) stream_release(g);
  stream_release(s);
}

const char* srd_engine_last_error(srd_engine* e) { return e ? e->err.c_str() : "no engine"; }

int srd_engine_ok(srd_engine* e) { return e && e->err.empty() ? 1 : 0; }

int srd_engine_load_csv(srd_engine* e, const char* rel, const char* path) {
  if (!e || !rel || !path) return 1;
  auto it = e->rel_by_name.find(rel);
  if (it == e->rel_by_name.end() || e->rels[it->second].input_file.empty()) {
    std::cerr << "srdatalog_load_csv: unknown or non-input relation " << rel << std::endl;
    return 2;
  }
  return eguard(e, 3, [&] { load_csv_file(e->rels[it->second], path); });
}

int srd_engine_load_all(srd_engine* e, const char* dir) {
  if (!e || !dir) return 1;
  return eguard(e, 2, [&] {
    for (auto& r : e->rels)
      if (!r.input_file.empty()) load_csv_file(r, std::string(dir) + "/" + r.input_file);
  });
}

int srd_engine_load_rows(srd_engine* e, const char* rel, const uint32_t* rows, uint64_t n,
                         uint32_t arity) {
  if (!e || !rel || (n && !rows)) return 1;
  auto it = e->rel_by_name.find(rel);
  if (it == e->rel_by_name.end()) return 2;
  Relation& r = e->rels[it->second];
  if (arity != r.arity) return 2;
  return eguard(e, 3, [&] {
    r.host_rows.assign(rows, rows + n * arity);
    r.host_n = n;
    r.has_input = true;
    r.dev_input_valid = false;
  });
}

int srd_engine_prepare(srd_engine* e) {
  return eguard(e, 2, [&] {
    for (auto& r : e->rels) upload_input(*e, r);
  });
}

int srd_engine_run(srd_engine* e, unsigned long long max_iters) {
  if (!e) return 1;
  if (!e->err.empty() && e->steps.empty()) return 2;
  bool any_input = false;
  for (auto& r : e->rels) any_input |= r.has_input;
  bool any_decl = false;
  for (auto& r : e->rels) any_decl |= !r.input_file.empty();
  if (any_decl && !any_input) return 1;
  return eguard(e, 2, [&] { run_program(*e, max_iters ? max_iters : uint64_t(2147483647)); });
}

unsigned long long srd_engine_size(srd_engine* e, const char* rel) {
  if (!e || !rel) return 0;
  auto it = e->rel_by_name.find(rel);
  if (it == e->rel_by_name.end()) return 0;
  Relation& r = e->rels[it->second];
  if (r.counted) return r.count;
  auto f = r.full.find(r.canonical);
  if (f != r.full.end()) return f->second.idx.n;
  if (!e->ran && r.has_input) return r.host_n;
  return 0;
}

// FULL of r in relation column order (0..k-1), cached per FULL version.
static const Index* export_index(srd_engine& e, Relation& r) {
  auto f = r.full.find(r.canonical);
  if (f == r.full.end() || f->second.idx.n == 0) return nullptr;
  Order ident(r.arity);
  for (uint32_t c = 0; c < r.arity; ++c) ident[c] = c;
  if (r.canonical == ident) return &f->second.idx;
  if (r.export_version != r.full_version) {
    build_index_from(f->second.idx, perm_between(r.canonical, ident), r.export_idx, e.ws, e.s);
    r.export_version = r.full_version;
  }
  return &r.export_idx;
}

int srd_engine_export(srd_engine* e, const char* rel, uint32_t* out, uint64_t cap_rows,
                      uint64_t* n_rows) {
  if (!e || !rel || !n_rows) return 1;
  auto it = e->rel_by_name.find(rel);
  if (it == e->rel_by_name.end()) return 2;
  Relation& r = e->rels[it->second];
  return eguard(e, 3, [&] {
    *n_rows = 0;
    const Index* x = export_index(*e, r);
    if (!x) return;
    *n_rows = x->n;
    if (!out) return;
    if (cap_rows < x->n) fail("srdatalog_export: buffer too small");
    export_rows(*x, out, nullptr, 0, e->ws, e->s);
  });
}

int srd_engine_export_instances(srd_engine* e, const char* rel, uint32_t* out, uint64_t cap_rows,
                                uint64_t* n_rows, uint64_t* cuts) {
  if (!e || !rel || !n_rows) return 1;
  auto it = e->rel_by_name.find(rel);
  if (it == e->rel_by_name.end()) return 2;
  Relation& r = e->rels[it->second];
  return eguard(e, 3, [&] {
    if (!e->n_inst) fail("srdatalog_export_instances: no instance table");
    *n_rows = 0;
    const Index* x = export_index(*e, r);
    if (cuts) std::fill(cuts, cuts + e->n_inst + 1, 0);
    if (!x) return;
    *n_rows = x->n;
    if (!out) return;
    if (cap_rows < x->n) fail("srdatalog_export_instances: buffer too small");
    if (cuts) {
      DevBuf<uint64_t> per(e->n_inst + 2, e->s);
      instance_counts(x->cols.p, x->n, e->inst_bounds.p, e->n_inst, per.p, e->s);
      outside_counts(x->cols.p, x->n, e->inst_bounds.p, e->n_inst, per.p + e->n_inst, e->s);
      std::vector<uint64_t> h(e->n_inst + 2);
      SRD_CHECK(hipMemcpyAsync(h.data(), per.p, (e->n_inst + 2) * 8, hipMemcpyDeviceToHost, e->s));
      SRD_CHECK(hipStreamSynchronize(e->s));
      // Rows below the first bound belong to no instance and precede cut 0;
      // a row at or above the last bound would be rebased into the wrong
      // range by the export, so it is an error.
      if (h[e->n_inst + 1] != 0)
        fail("srdatalog_export_instances: " + std::to_string(h[e->n_inst + 1]) +
             " rows lie at or above the last instance bound");
      uint64_t acc = h[e->n_inst];
      for (uint32_t g = 0; g < e->n_inst; ++g) {
        cuts[g] = acc;
        acc += h[g];
      }
      cuts[e->n_inst] = acc;
    }
    export_rows(*x, out, e->inst_bounds.p, e->n_inst + 1, e->ws, e->s);
  });
}

unsigned long long srd_engine_num_iterations(srd_engine* e, const char* rel) {
  if (!e || !rel) return 0;
  auto it = e->rel_by_name.find(rel);
  if (it == e->rel_by_name.end()) return 0;
  return e->rels[it->second].deltas.size();
}

int srd_engine_delta_sizes(srd_engine* e, const char* rel, unsigned long long* out,
                           unsigned long long cap) {
  if (!e || !rel) return -1;
  auto it = e->rel_by_name.find(rel);
  if (it == e->rel_by_name.end()) return -1;
  auto& d = e->rels[it->second].deltas;
  for (uint64_t i = 0; i < d.size() && i < cap; ++i) out[i] = d[i];
  return int(d.size());
}

int srd_engine_shutdown(srd_engine* e) {
  if (!e) return 0;
  for (auto& r : e->rels) {
    r.full.clear();
    r.delta.clear();
    clear_new(r);
    r.host_rows.clear();
    r.host_rows.shrink_to_fit();
    r.host_n = 0;
    r.has_input = false;
    r.dev_input.release();
    r.dev_input_valid = false;
  }
  e->ran = false;
  return 0;
}

int srd_engine_run_begin(srd_engine* e) {
  if (!e) return 1;
  return eguard(e, 2, [&] { run_begin(*e); });
}

int srd_engine_run_iterations(srd_engine* e, unsigned long long k, unsigned long long* done,
                              int* converged) {
  if (!e) return 1;
  return eguard(e, 2, [&] {
    bool c = false;
    const uint64_t d = run_iters(*e, k, &c);
    if (done) *done = d;
    if (converged) *converged = c ? 1 : 0;
  });
}

int srd_engine_absorb(srd_engine* dst, srd_engine* src) {
  if (!dst || !src) return 1;
  return eguard(dst, 2, [&] { absorb(*dst, *src); });
}

int srd_engine_range_counts(srd_engine* e, const char* rel, int delta, const uint32_t* lohi,
                            uint32_t m, unsigned long long* out) {
  if (!e || !rel || (m && (!lohi || !out))) return 1;
  auto it = e->rel_by_name.find(rel);
  if (it == e->rel_by_name.end()) return 2;
  Relation& r = e->rels[it->second];
  return eguard(e, 3, [&] {
    for (uint32_t g = 0; g < m; ++g) out[g] = 0;
    if (m == 0) return;
    for (uint64_t i = 1; i < 2ull * m; ++i)
      if (lohi[i] < lohi[i - 1]) fail("srdatalog_range_counts: ranges must be ascending and disjoint");
    std::vector<Index*> parts;
    if (delta) {
      auto d = r.delta.find(r.canonical);
      if (d != r.delta.end()) parts.push_back(&d->second);
    } else {
      parts = full_parts(*e, r);
    }
    for (Index* x : parts) {
      auto v = value_ranges(*e, *x, lohi, m);
      for (uint32_t g = 0; g < m; ++g) out[g] += v[g].second;
    }
  });
}

int srd_engine_export_ranges(srd_engine* e, const char* rel, const uint32_t* lohi, uint32_t m,
                             int rebase, uint32_t* out, uint64_t cap_rows, uint64_t* n_rows,
                             uint64_t* cuts) {
  if (!e || !rel || !n_rows || (m && !lohi)) return 1;
  auto it = e->rel_by_name.find(rel);
  if (it == e->rel_by_name.end()) return 2;
  Relation& r = e->rels[it->second];
  return eguard(e, 3, [&] {
    for (uint64_t i = 1; i < 2ull * m; ++i)
      if (lohi[i] < lohi[i - 1]) fail("srdatalog_export_ranges: ranges must be ascending and disjoint");
    *n_rows = 0;
    if (cuts) std::fill(cuts, cuts + m + 1, 0);
    if (m == 0) return;
    // the ranges' rows of every part of FULL (canonical order), gathered,
    // then sorted in relation column order: only these rows are touched,
    // and the ranges come out one after the other
    std::vector<std::pair<const Index*, std::vector<std::pair<uint64_t, uint64_t>>>> hits;
    std::vector<uint64_t> per(m, 0);
    uint64_t tot = 0;
    for (Index* part : full_parts(*e, r)) {
      auto v = value_ranges(*e, *part, lohi, m);
      for (uint32_t g = 0; g < m; ++g) {
        per[g] += v[g].second;
        tot += v[g].second;
      }
      hits.push_back({part, std::move(v)});
    }
    *n_rows = tot;
    if (cuts)
      for (uint32_t g = 0; g < m; ++g) cuts[g + 1] = cuts[g] + per[g];
    if (!out || tot == 0) return;
    if (cap_rows < tot) fail("srdatalog_export_ranges: buffer too small");
    const uint32_t k = r.arity;
    DevBuf<uint32_t> cols(uint64_t(k) * tot, e->s);
    uint64_t at = 0;
    for (auto& [part, v] : hits)
      for (auto& [a, cnt] : v) {
        if (!cnt) continue;
        for (uint32_t c = 0; c < k; ++c)
          SRD_CHECK(hipMemcpyAsync(cols.p + uint64_t(c) * tot + at,
                                   part->cols.p + uint64_t(c) * part->stride + a, cnt * 4,
                                   hipMemcpyDeviceToDevice, e->s));
        at += cnt;
      }
    // relation column j is canonical column c with canonical[c] == j
    std::vector<const uint32_t*> in(k);
    for (uint32_t c = 0; c < k; ++c) in[r.canonical[c]] = cols.p + uint64_t(c) * tot;
    Index x;
    build_index(in.data(), k, tot, x, e->ws, e->s);
    if (!rebase) {
      export_rows(x, out, nullptr, 0, e->ws, e->s);
      return;
    }
    // rebased on the device: the 2m bounds lo0, hi0, lo1, ... make ranges
    // and gaps alternate; a row of range g loses lo_g from every value
    DevBuf<uint32_t> b(2ull * m, e->s);
    SRD_CHECK(hipMemcpyAsync(b.p, lohi, 8ull * m, hipMemcpyHostToDevice, e->s));
    export_rows(x, out, b.p, 2 * m, e->ws, e->s);
  });
}

int srd_engine_erase_ranges(srd_engine* e, const uint32_t* lohi, uint32_t m) {
  if (!e || (m && !lohi)) return 1;
  return eguard(e, 2, [&] { erase_ranges(*e, lohi, m); });
}

int srd_engine_set_instances(srd_engine* e, const uint32_t* bounds, uint64_t n_bounds) {
  if (!e) return 1;
  return eguard(e, 2, [&] {
    if (n_bounds == 1 || n_bounds > (uint64_t(1) << 31)) fail("set_instances: need 0 or >= 2 bounds");
    for (uint64_t i = 1; i < n_bounds; ++i)
      if (bounds[i] < bounds[i - 1]) fail("set_instances: bounds must be ascending");
    e->n_inst = 0;
    e->inst_bounds.release();
    if (n_bounds == 0) return;
    e->inst_bounds.alloc(n_bounds, e->s);
    SRD_CHECK(hipMemcpyAsync(e->inst_bounds.p, bounds, n_bounds * 4, hipMemcpyHostToDevice, e->s));
    SRD_CHECK(hipStreamSynchronize(e->s));
    e->n_inst = uint32_t(n_bounds - 1);
  });
}

long long srd_engine_instance_delta_sizes(srd_engine* e, const char* rel, unsigned long long* out,
                                          unsigned long long cap_iters) {
  if (!e || !rel) return -1;
  auto it = e->rel_by_name.find(rel);
  if (it == e->rel_by_name.end()) return -1;
  auto& v = e->rels[it->second].inst_deltas;
  if (out && cap_iters) {
    const int rc = eguard(e, 2, [&] {
      for (size_t i = 0; i < v.size() && i < cap_iters; ++i)
        SRD_CHECK(hipMemcpyAsync(out + i * e->n_inst, v[i].p, uint64_t(e->n_inst) * 8,
                                 hipMemcpyDeviceToHost, e->s));
      SRD_CHECK(hipStreamSynchronize(e->s));
    });
    if (rc != 0) return -1;
  }
  return (long long)v.size();
}

double srd_engine_run_seconds(srd_engine* e) { return e ? e->run_seconds : 0.0; }

unsigned long long srd_engine_join_tuples(srd_engine* e) { return e ? e->join_probes : 0; }

void srd_engine_count_probes(srd_engine* e, int on) {
  if (e) e->count_probes = on != 0;
}

unsigned long long srd_engine_join_searches(srd_engine* e) { return e ? e->join_searches : 0; }

void srd_engine_set_quiet(srd_engine* e, int quiet) {
  if (e) e->quiet = quiet != 0;
}

void srd_engine_set_grain(srd_engine* e, unsigned grain) {
  if (e) e->grain_override = grain;
}

}  // extern "C"
