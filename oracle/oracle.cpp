// CPU oracle for the semi-naive fixpoint path.  TEST INFRASTRUCTURE ONLY:
// tests/, the smoke check and the bench's cpu_baseline leg use it as the
// checker; the product never links, loads or calls it.
//
// It restates, per benchmark program, the reference's execution semantics:
//
//  * strata and variants as in the reference MIR goldens
//    (tests/fixtures/tc.mir.sexpr, tests/fixtures/integration/{sg,andersen,
//    lsqb_triangle_count}.mir.sexpr): base rules run in their own
//    non-recursive step before the recursive step;
//  * iteration 0 of a recursive step reads FULL wherever a variant reads
//    DELTA (src/srdatalog/codegen/jit/complete_runner.py:834-838);
//  * every non-delta atom reads FULL, which already contains the previous
//    delta (src/srdatalog/hir/semi_naive.py:24-60);
//  * maintenance order per iteration: NEW = sort+dedup(join output),
//    DELTA = NEW - FULL, FULL = FULL u DELTA
//    (src/srdatalog/hir/lower.py:487-565;
//     runtime/generalized_datalog/gpu/runtime/gpu_mir_helpers.h:87-232);
//  * the loop stops when, at iter > 0, the summed canonical DELTA size is 0
//    (src/srdatalog/codegen/jit/orchestrator_jit.py:659-695).
//
// Relations are sorted, duplicate-free vectors of packed keys (arity 2:
// u64 = c0<<32|c1 in index-column order; arity 3: u128).  Joins run over
// root keys with OpenMP; sorting uses the libstdc++ parallel mode.
//
// Pinning (tests/test_oracle.py): brute-force naive Jacobi iteration on tiny
// inputs for every program, all-pairs BFS distance histograms for TC deltas,
// and scipy sparse A.(A@A) for triangle counts.

#include <omp.h>
#include <parallel/algorithm>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace {

using u32 = uint32_t;
using u64 = uint64_t;
using u128 = unsigned __int128;
using Keys = std::vector<u64>;

inline u64 pack(u32 hi, u32 lo) { return (u64(hi) << 32) | lo; }
inline u32 hi32(u64 k) { return u32(k >> 32); }
inline u32 lo32(u64 k) { return u32(k); }

template <class K>
void sort_unique(std::vector<K>& v) {
  if (v.size() > (1u << 16)) __gnu_parallel::sort(v.begin(), v.end());
  else std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
}

// Keys of a binary relation in the given column order: (0 1) or (1 0).
Keys make_index(const u32* rows, u64 n, bool swap) {
  Keys k(n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < int64_t(n); ++i) {
    u32 a = rows[2 * i], b = rows[2 * i + 1];
    k[i] = swap ? pack(b, a) : pack(a, b);
  }
  sort_unique(k);
  return k;
}

Keys swap_index(const Keys& k) {
  Keys out(k.size());
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < int64_t(k.size()); ++i) out[i] = pack(lo32(k[i]), hi32(k[i]));
  sort_unique(out);
  return out;
}

// CSR view over sorted keys: distinct first-column values and row offsets.
struct Groups {
  std::vector<u32> vals;
  std::vector<u64> offs;  // size vals.size()+1
  explicit Groups(const Keys& k) {
    for (u64 i = 0; i < k.size(); ++i)
      if (i == 0 || hi32(k[i]) != hi32(k[i - 1])) {
        vals.push_back(hi32(k[i]));
        offs.push_back(i);
      }
    offs.push_back(k.size());
  }
  // [begin,end) rows whose first column is v; empty if absent.
  std::pair<u64, u64> find(u32 v) const {
    auto it = std::lower_bound(vals.begin(), vals.end(), v);
    if (it == vals.end() || *it != v) return {0, 0};
    size_t g = it - vals.begin();
    return {offs[g], offs[g + 1]};
  }
};

// Thread-local output buffers, compacted when they grow, then concatenated.
template <class K>
struct Sink {
  std::vector<std::vector<K>> per;
  explicit Sink(int t) : per(t) {}
  std::vector<K>& mine() { return per[omp_get_thread_num()]; }
  static void maybe_compact(std::vector<K>& v) {
    if (v.size() >= (size_t(1) << 25)) {
      std::sort(v.begin(), v.end());
      v.erase(std::unique(v.begin(), v.end()), v.end());
    }
  }
  std::vector<K> collect() {
    size_t tot = 0;
    for (auto& v : per) tot += v.size();
    std::vector<K> out;
    out.reserve(tot);
    for (auto& v : per) {
      out.insert(out.end(), v.begin(), v.end());
      std::vector<K>().swap(v);
    }
    sort_unique(out);
    return out;
  }
};

Keys set_diff(const Keys& a, const Keys& b) {
  Keys out;
  out.reserve(a.size());
  std::set_difference(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(out));
  return out;
}

Keys set_union_disjoint(const Keys& a, const Keys& b) {
  Keys out(a.size() + b.size());
  std::merge(a.begin(), a.end(), b.begin(), b.end(), out.begin());
  return out;
}

struct Result {
  std::map<std::string, std::vector<u32>> rels;  // row-major, sorted by relation columns
  std::map<std::string, std::vector<u64>> deltas;  // per-iteration |DELTA| of the recursive step
  u64 iterations = 0;
  u64 count = 0;  // count-only rules
  std::map<std::string, u64> rule_counts;  // count-only head relation -> count
  double seconds = 0;
  std::string error;
};

void store_binary(Result& r, const char* name, const Keys& k, bool swapped) {
  Keys rows = swapped ? swap_index(k) : k;
  auto& out = r.rels[name];
  out.resize(2 * rows.size());
  for (u64 i = 0; i < rows.size(); ++i) {
    out[2 * i] = hi32(rows[i]);
    out[2 * i + 1] = lo32(rows[i]);
  }
}

// ---------------------------------------------------------------- TC ------
// Path(x,y) <= Edge(x,y);  Path(x,z) <= Path(x,y) & Edge(y,z).
// Path canonical index (1 0): keys (y,x).  Edge index (0 1): keys (y,z).
void run_tc(const u32* arc, u64 m, u64 max_iters, int threads, Result& r) {
  Keys edge = make_index(arc, m, false);        // EdgeLoad: Edge = dedup(ArcInput)
  Groups eg(edge);
  Keys full = swap_index(edge);                 // TCBase: Path(x,y), index (1 0)
  Keys delta = full;
  auto& dl = r.deltas["Path"];
  for (u64 it = 0; it < max_iters; ++it) {
    if (it > 0 && delta.empty()) break;
    const Keys& src = (it == 0) ? full : delta;  // iteration-0 DELTA -> FULL
    Groups dg(src);
    Sink<u64> sink(threads);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t g = 0; g < int64_t(dg.vals.size()); ++g) {
      u32 y = dg.vals[g];
      auto [eb, ee] = eg.find(y);
      if (eb == ee) continue;
      auto& out = sink.mine();
      for (u64 i = dg.offs[g]; i < dg.offs[g + 1]; ++i) {
        u32 x = lo32(src[i]);
        for (u64 j = eb; j < ee; ++j) out.push_back(pack(lo32(edge[j]), x));  // (z,x)
      }
      Sink<u64>::maybe_compact(out);
    }
    Keys nw = sink.collect();
    delta = set_diff(nw, full);
    full = set_union_disjoint(full, delta);
    dl.push_back(delta.size());
    r.iterations = it + 1;
  }
  store_binary(r, "Edge", edge, false);
  store_binary(r, "Path", full, true);
}

// ---------------------------------------------------------------- SG ------
// Sg(x,y) <= Arc(p,x) & Arc(p,y) & x != y;
// Sg(x,y) <= Arc(a,x) & Sg(a,b) & Arc(b,y).  Sg canonical index (0 1).
void run_sg(const u32* arc_rows, u64 m, u64 max_iters, int threads, Result& r) {
  Keys arc = make_index(arc_rows, m, false);
  Groups ag(arc);
  Sink<u64> base(threads);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t g = 0; g < int64_t(ag.vals.size()); ++g) {
    auto& out = base.mine();
    for (u64 i = ag.offs[g]; i < ag.offs[g + 1]; ++i)
      for (u64 j = ag.offs[g]; j < ag.offs[g + 1]; ++j) {
        u32 x = lo32(arc[i]), y = lo32(arc[j]);
        if (x != y) out.push_back(pack(x, y));
      }
    Sink<u64>::maybe_compact(out);
  }
  Keys full = base.collect();
  Keys delta = full;
  auto& dl = r.deltas["Sg"];
  for (u64 it = 0; it < max_iters; ++it) {
    if (it > 0 && delta.empty()) break;
    const Keys& src = (it == 0) ? full : delta;
    Groups dg(src);
    Sink<u64> sink(threads);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t g = 0; g < int64_t(dg.vals.size()); ++g) {
      u32 a = dg.vals[g];
      auto [xb, xe] = ag.find(a);
      if (xb == xe) continue;
      auto& out = sink.mine();
      for (u64 i = dg.offs[g]; i < dg.offs[g + 1]; ++i) {
        u32 b = lo32(src[i]);
        auto [yb, ye] = ag.find(b);
        for (u64 p = xb; p < xe; ++p)
          for (u64 q = yb; q < ye; ++q) out.push_back(pack(lo32(arc[p]), lo32(arc[q])));
      }
      Sink<u64>::maybe_compact(out);
    }
    Keys nw = sink.collect();
    delta = set_diff(nw, full);
    full = set_union_disjoint(full, delta);
    dl.push_back(delta.size());
    r.iterations = it + 1;
  }
  store_binary(r, "Sg", full, false);
}

// ---------------------------------------------------------- Andersen ------
// PointsTo(v,o) <= AddressOf(v,o);
// PointsTo(v,o) <= Assign(v,w) & PointsTo(w,o);
// PointsTo(x,o) <= Load(x,y) & PointsTo(y,a) & PointsTo(a,o);
// PointsTo(a,o) <= Store(x,y) & PointsTo(x,a) & PointsTo(y,o).
// Recursive step = the five delta variants of the reference MIR
// (Assign_D0, Load_D0, Load_D2, Store_D0, Store_D2).
void run_andersen(const u32* addr, u64 n_addr, const u32* assign, u64 n_assign,
                  const u32* load, u64 n_load, const u32* store, u64 n_store,
                  u64 max_iters, int threads, Result& r) {
  Keys assign_by_w = make_index(assign, n_assign, true);   // (w, v)
  Keys load_by_y = make_index(load, n_load, true);         // (y, x)
  Keys store_by_x = make_index(store, n_store, false);     // (x, y)
  Keys store_by_y = make_index(store, n_store, true);      // (y, x)
  Groups asg(assign_by_w), lg(load_by_y), sxg(store_by_x), syg(store_by_y);
  Keys full = make_index(addr, n_addr, false);             // PointsTo (0 1)
  Keys full_rev = swap_index(full);                        // PointsTo (1 0)
  Keys delta = full;
  auto& dl = r.deltas["PointsTo"];
  for (u64 it = 0; it < max_iters; ++it) {
    if (it > 0 && delta.empty()) break;
    const Keys& d = (it == 0) ? full : delta;
    Groups dg(d), fg(full), frg(full_rev);
    Sink<u64> sink(threads);
    // Assign_D0: dPT(w,o) & Assign(v,w) -> (v,o).
    // Load_D0:   dPT(y,a) & Load(x,y) & PT(a,o) -> (x,o).
    // Store_D0:  dPT(x,a) & Store(x,y) & PT(y,o) -> (a,o).
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t g = 0; g < int64_t(dg.vals.size()); ++g) {
      u32 k = dg.vals[g];
      auto& out = sink.mine();
      u64 db = dg.offs[g], de = dg.offs[g + 1];
      auto [ab, ae] = asg.find(k);
      for (u64 j = ab; j < ae; ++j)
        for (u64 i = db; i < de; ++i) out.push_back(pack(lo32(assign_by_w[j]), lo32(d[i])));
      auto [lb, le] = lg.find(k);
      if (lb != le)
        for (u64 i = db; i < de; ++i) {
          auto [pb, pe] = fg.find(lo32(d[i]));
          for (u64 j = lb; j < le; ++j)
            for (u64 p = pb; p < pe; ++p) out.push_back(pack(lo32(load_by_y[j]), lo32(full[p])));
        }
      auto [sb, se] = sxg.find(k);
      for (u64 j = sb; j < se; ++j) {
        auto [pb, pe] = fg.find(lo32(store_by_x[j]));
        for (u64 i = db; i < de; ++i)
          for (u64 p = pb; p < pe; ++p) out.push_back(pack(lo32(d[i]), lo32(full[p])));
      }
      Sink<u64>::maybe_compact(out);
    }
    // Load_D2:  PT(y,a) & Load(x,y) & dPT(a,o) -> (x,o).
    // Store_D2: PT(x,a) & Store(x,y) & dPT(y,o) -> (a,o).
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t g = 0; g < int64_t(dg.vals.size()); ++g) {
      u32 k = dg.vals[g];
      auto& out = sink.mine();
      u64 db = dg.offs[g], de = dg.offs[g + 1];
      auto [rb, re] = frg.find(k);  // PT(y, k): rows (k, y) of the (1 0) index
      for (u64 q = rb; q < re; ++q) {
        auto [lb, le] = lg.find(lo32(full_rev[q]));
        for (u64 j = lb; j < le; ++j)
          for (u64 i = db; i < de; ++i) out.push_back(pack(lo32(load_by_y[j]), lo32(d[i])));
      }
      auto [sb, se] = syg.find(k);  // Store(x, k): rows (k, x)
      for (u64 j = sb; j < se; ++j) {
        auto [pb, pe] = fg.find(lo32(store_by_y[j]));
        for (u64 p = pb; p < pe; ++p)
          for (u64 i = db; i < de; ++i) out.push_back(pack(lo32(full[p]), lo32(d[i])));
      }
      Sink<u64>::maybe_compact(out);
    }
    Keys nw = sink.collect();
    delta = set_diff(nw, full);
    full = set_union_disjoint(full, delta);
    full_rev = set_union_disjoint(full_rev, swap_index(delta));
    dl.push_back(delta.size());
    r.iterations = it + 1;
  }
  store_binary(r, "PointsTo", full, false);
}

// Order-independent digest of a tuple set (test infrastructure): per tuple
// h = m(m(m(t0) ^ t1) ^ t2 ...) with m the splitmix64 finaliser; the digest
// is (count, sum of h, sum of m(h + c)) mod 2^64.  Two sets with equal
// digests and equal counts are equal unless the 128-bit sums collide.
inline u64 mix64(u64 z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
constexpr u64 kDigestC = 0x632be59bd9b4e019ull;
struct Digest {
  u64 n = 0, s1 = 0, s2 = 0;
  void add(u64 h) {
    ++n;
    s1 += h;
    s2 += mix64(h + kDigestC);
  }
  void merge(const Digest& o) {
    n += o.n;
    s1 += o.s1;
    s2 += o.s2;
  }
};
inline u64 tuple_hash3(u32 a, u32 b, u32 c) { return mix64(mix64(mix64(a) ^ b) ^ c); }

// Root-key sample of the triangle programs: only keys a with a % stride == 0
// (1 = every key); set through orc_set_key_stride.
u32 g_key_stride = 1;

// ---------------------------------------------------------- Triangle ------
// Triangle(a,b,c) <= Knows(a,b) & Knows(b,c) & Knows(a,c), var order a,b,c.
// count_only: the rule carries .with_count(); only the count is produced.
// key_lo/key_hi bound the root keys a (for bounded baseline samples).
void run_triangle(const u32* rows, u64 m, bool count_only, u32 key_lo, u32 key_hi,
                  int threads, Result& r) {
  Keys kn = make_index(rows, m, false);
  Groups g(kn);
  std::vector<u64> counts(threads, 0);
  Sink<u128> sink(threads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t ga = 0; ga < int64_t(g.vals.size()); ++ga) {
    u32 a = g.vals[ga];
    if (a < key_lo || a >= key_hi || a % g_key_stride != 0) continue;
    u64 ab = g.offs[ga], ae = g.offs[ga + 1];
    u64 local = 0;
    auto& out = sink.mine();
    for (u64 i = ab; i < ae; ++i) {
      u32 b = lo32(kn[i]);
      auto [bb, be] = g.find(b);
      // c in Knows[b] intersect Knows[a], merge of two sorted ranges
      u64 p = bb, q = ab;
      while (p < be && q < ae) {
        u32 cb = lo32(kn[p]), ca = lo32(kn[q]);
        if (cb < ca) ++p;
        else if (ca < cb) ++q;
        else {
          ++local;
          if (!count_only) out.push_back((u128(a) << 64) | (u128(b) << 32) | cb);
          ++p, ++q;
        }
      }
    }
    counts[omp_get_thread_num()] += local;
  }
  for (u64 c : counts) r.count += c;
  store_binary(r, "Knows", kn, false);
  auto& dl = r.deltas["Triangle"];
  if (!count_only) {
    std::vector<u128> tri = sink.collect();
    auto& out = r.rels["Triangle"];
    out.resize(3 * tri.size());
    for (u64 i = 0; i < tri.size(); ++i) {
      out[3 * i] = u32(tri[i] >> 64);
      out[3 * i + 1] = u32(tri[i] >> 32);
      out[3 * i + 2] = u32(tri[i]);
    }
    dl.push_back(tri.size());
  } else {
    r.rels["Triangle"];
    dl.push_back(0);
  }
  r.iterations = 1;
}

// -------------------------------------------------------------- Reach -----
// Reach(y) <= Src(y);  Reach(z) <= Reach(y) & Edge(y,z).
void run_reach(const u32* edge_rows, u64 m, const u32* src, u64 n_src, u64 max_iters,
               int threads, Result& r) {
  Keys edge = make_index(edge_rows, m, false);
  Groups eg(edge);
  std::vector<u32> full(src, src + n_src);
  sort_unique(full);
  std::vector<u32> delta = full;
  auto& dl = r.deltas["Reach"];
  for (u64 it = 0; it < max_iters; ++it) {
    if (it > 0 && delta.empty()) break;
    const std::vector<u32>& d = (it == 0) ? full : delta;
    Sink<u32> sink(threads);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < int64_t(d.size()); ++i) {
      auto [b, e] = eg.find(d[i]);
      auto& out = sink.mine();
      for (u64 j = b; j < e; ++j) out.push_back(lo32(edge[j]));
    }
    std::vector<u32> nw = sink.collect();
    std::vector<u32> nd;
    std::set_difference(nw.begin(), nw.end(), full.begin(), full.end(), std::back_inserter(nd));
    std::vector<u32> merged(full.size() + nd.size());
    std::merge(full.begin(), full.end(), nd.begin(), nd.end(), merged.begin());
    full.swap(merged);
    delta.swap(nd);
    dl.push_back(delta.size());
    r.iterations = it + 1;
  }
  r.rels["Reach"] = full;
}

// ------------------------------------------------------------- Neg2hop ---
// LSQB Q9 (examples/lsqb_q9_neg2hop.py): Path(p1,p2,p3,t) <= Knows(p1,p2),
// Knows(p2,p3), HasInterest(p3,t), ~Knows(p1,p3), p1 != p3.  Knows and
// HasInterest are loaded from their inputs (KnowsLoad / InterestLoad).  One
// non-recursive step; count_only mirrors the rule's .with_count().
void run_neg2hop(const u32* knows_rows, u64 mk, const u32* hi_rows, u64 mh, bool count_only,
                 int threads, Result& r) {
  Keys kn = make_index(knows_rows, mk, false);     // (p1,p2) / (p2,p3) / (p1,p3)
  Keys hi = make_index(hi_rows, mh, false);        // (p3,t)
  Groups kg(kn), hg(hi);
  auto has_knows = [&](u32 a, u32 b) { return std::binary_search(kn.begin(), kn.end(), pack(a, b)); };
  std::vector<u64> counts(threads, 0);
  std::vector<std::vector<u128>> outs(threads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t g1 = 0; g1 < int64_t(kg.vals.size()); ++g1) {
    const u32 p1 = kg.vals[g1];
    u64 local = 0;
    auto& out = outs[omp_get_thread_num()];
    for (u64 i = kg.offs[g1]; i < kg.offs[g1 + 1]; ++i) {
      const u32 p2 = lo32(kn[i]);
      auto [b2, e2] = kg.find(p2);
      for (u64 j = b2; j < e2; ++j) {
        const u32 p3 = lo32(kn[j]);
        if (p1 == p3 || has_knows(p1, p3)) continue;
        auto [b3, e3] = hg.find(p3);
        for (u64 q = b3; q < e3; ++q) {
          ++local;
          if (!count_only)
            out.push_back((u128(p1) << 96) | (u128(p2) << 64) | (u128(p3) << 32) | lo32(hi[q]));
        }
      }
    }
    counts[omp_get_thread_num()] += local;
  }
  for (u64 c : counts) r.count += c;
  store_binary(r, "Knows", kn, false);
  store_binary(r, "HasInterest", hi, false);
  auto& dl = r.deltas["Path"];
  if (!count_only) {
    std::vector<u128> all;
    for (auto& v : outs) all.insert(all.end(), v.begin(), v.end());
    sort_unique(all);
    auto& out = r.rels["Path"];
    out.resize(4 * all.size());
    for (u64 i = 0; i < all.size(); ++i) {
      out[4 * i] = u32(all[i] >> 96);
      out[4 * i + 1] = u32(all[i] >> 64);
      out[4 * i + 2] = u32(all[i] >> 32);
      out[4 * i + 3] = u32(all[i]);
    }
    dl.push_back(all.size());
  } else {
    r.rels["Path"];
    dl.push_back(0);
  }
  r.iterations = 1;
}

// ---------------------------------------------------------------- Q7 ------
// LSQB Q7 (examples/lsqb_q7_optional.py), four count-only rules over messages
// m:  HasReply(m) <= ReplyOf(_, m);  HasLiker(m) <= Likes(_, m);
//   Case1(m,t,c,comment,liker) <= HasTag(m,t), HasCreator(m,c), ReplyOf(comment,m), Likes(liker,m)
//   Case2(m,t,c,comment) <= HasTag, HasCreator, ReplyOf(comment,m), ~HasLiker(m)
//   Case3(m,t,c,liker)   <= HasTag, HasCreator, Likes(liker,m), ~HasReply(m)
//   Case4(m,t,c)         <= HasTag, HasCreator, ~HasReply(m), ~HasLiker(m)
// Each count is the number of distinct head tuples, enumerated per m.
void run_q7(const u32* tag, u64 nt, const u32* cre, u64 nc, const u32* rep, u64 nr, const u32* lik,
            u64 nl, Result& r) {
  Keys ht = make_index(tag, nt, false);    // (m,t)
  Keys hc = make_index(cre, nc, false);    // (m,c)
  Keys ro = make_index(rep, nr, true);     // ReplyOf(comment,m) keyed (m,comment)
  Keys li = make_index(lik, nl, true);     // Likes(liker,m) keyed (m,liker)
  Groups gt(ht), gc(hc), gr(ro), gl(li);
  u64 c1 = 0, c2 = 0, c3 = 0, c4 = 0;
  for (size_t g = 0; g < gt.vals.size(); ++g) {
    const u32 m = gt.vals[g];
    const u64 t = gt.offs[g + 1] - gt.offs[g];
    auto [cb, ce] = gc.find(m);
    auto [rb, re] = gr.find(m);
    auto [lb, le] = gl.find(m);
    const u64 c = ce - cb, rr = re - rb, ll = le - lb;
    c1 += t * c * rr * ll;
    if (ll == 0) c2 += t * c * rr;
    if (rr == 0) c3 += t * c * ll;
    if (rr == 0 && ll == 0) c4 += t * c;
  }
  r.rule_counts["Case1"] = c1;
  r.rule_counts["Case2"] = c2;
  r.rule_counts["Case3"] = c3;
  r.rule_counts["Case4"] = c4;
  r.iterations = 1;
}

// ---------------------------------------------------------------- Q3 ------
// LSQB Q3 (examples/lsqb_q3_triangle.py): Triangle(p1,p2,p3) <= Knows(p1,p2),
// Knows(p2,p3), Knows(p3,p1), IsLocatedIn(p_i,c_i), IsPartOf(c_i,co) for one
// country co shared by the three cities.  The head drops c1..c3 and co, so
// the result is the set of directed 3-cycles whose persons share a country.
void run_q3(const u32* kr, u64 mk, const u32* lr, u64 ml, const u32* pr, u64 mp, int threads,
            Result& r) {
  Keys kn = make_index(kr, mk, false), loc = make_index(lr, ml, false),
       part = make_index(pr, mp, false);
  Groups kg(kn), lg(loc), pg(part);
  auto countries = [&](u32 p) {
    std::vector<u32> out;
    auto [b, e] = lg.find(p);
    for (u64 i = b; i < e; ++i) {
      auto [cb, ce] = pg.find(lo32(loc[i]));
      for (u64 j = cb; j < ce; ++j) out.push_back(lo32(part[j]));
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    return out;
  };
  std::vector<std::vector<u32>> cty(kg.vals.size());
  std::vector<std::vector<u128>> outs(threads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t g1 = 0; g1 < int64_t(kg.vals.size()); ++g1) {
    const u32 p1 = kg.vals[g1];
    auto& out = outs[omp_get_thread_num()];
    const std::vector<u32> c1 = countries(p1);
    if (c1.empty()) continue;
    for (u64 i = kg.offs[g1]; i < kg.offs[g1 + 1]; ++i) {
      const u32 p2 = lo32(kn[i]);
      auto [b2, e2] = kg.find(p2);
      if (b2 == e2) continue;
      std::vector<u32> c12;
      const std::vector<u32> c2 = countries(p2);
      std::set_intersection(c1.begin(), c1.end(), c2.begin(), c2.end(), std::back_inserter(c12));
      if (c12.empty()) continue;
      for (u64 j = b2; j < e2; ++j) {
        const u32 p3 = lo32(kn[j]);
        if (!std::binary_search(kn.begin(), kn.end(), pack(p3, p1))) continue;
        const std::vector<u32> c3 = countries(p3);
        std::vector<u32> all;
        std::set_intersection(c12.begin(), c12.end(), c3.begin(), c3.end(), std::back_inserter(all));
        if (!all.empty()) out.push_back((u128(p1) << 64) | (u128(p2) << 32) | p3);
      }
    }
  }
  std::vector<u128> tri;
  for (auto& v : outs) tri.insert(tri.end(), v.begin(), v.end());
  sort_unique(tri);
  auto& o = r.rels["Triangle"];
  o.resize(3 * tri.size());
  for (u64 i = 0; i < tri.size(); ++i) {
    o[3 * i] = u32(tri[i] >> 64);
    o[3 * i + 1] = u32(tri[i] >> 32);
    o[3 * i + 2] = u32(tri[i]);
  }
  r.deltas["Triangle"].push_back(tri.size());
  r.iterations = 1;
}

// Digest of the oriented triangle listing Tri(a,b,c) <= Knows(a,b),
// Knows(b,c), Knows(a,c) without materialising it: the same enumeration as
// run_triangle (every triangle is listed once, a < b < c on an oriented
// input, so the listing is a set), each tuple hashed as it is found.
Digest triangle_digest(const u32* rows, u64 m, int threads) {
  Keys kn = make_index(rows, m, false);
  Groups g(kn);
  std::vector<Digest> part(threads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t ga = 0; ga < int64_t(g.vals.size()); ++ga) {
    const u32 a = g.vals[ga];
    const u64 ab = g.offs[ga], ae = g.offs[ga + 1];
    Digest& d = part[omp_get_thread_num()];
    for (u64 i = ab; i < ae; ++i) {
      const u32 b = lo32(kn[i]);
      auto [bb, be] = g.find(b);
      u64 p = bb, q = ab;
      while (p < be && q < ae) {
        const u32 cb = lo32(kn[p]), ca = lo32(kn[q]);
        if (cb < ca) ++p;
        else if (ca < cb) ++q;
        else {
          d.add(tuple_hash3(a, b, cb));
          ++p, ++q;
        }
      }
    }
  }
  Digest tot;
  for (auto& d : part) tot.merge(d);
  return tot;
}

}  // namespace

// ------------------------------------------------------------- C ABI ------
extern "C" {

struct orc_input {
  const char* name;   // relation name as in the program
  const uint32_t* rows;  // row-major n x arity
  uint64_t n;
  uint32_t arity;
};

static const orc_input* find_in(const orc_input* in, int n, const char* name) {
  for (int i = 0; i < n; ++i)
    if (std::strcmp(in[i].name, name) == 0) return &in[i];
  return nullptr;
}

// program: "tc" | "sg" | "andersen" | "triangle" | "triangle_count" | "reach" |
// "neg2hop" | "neg2hop_count" | "q3" | "q7".
// Returns an opaque result (check orc_error) or null for an unknown program.
void* orc_run(const char* program, const orc_input* inputs, int n_inputs, uint64_t max_iters,
              int threads, uint32_t key_lo, uint32_t key_hi) {
  auto* r = new Result();
  if (threads <= 0) threads = omp_get_max_threads();
  omp_set_num_threads(threads);
  if (max_iters == 0) max_iters = ~uint64_t(0);
  std::string p(program);
  auto need = [&](const char* nm, uint32_t arity) -> const orc_input* {
    const orc_input* x = find_in(inputs, n_inputs, nm);
    if (!x || x->arity != arity) r->error = std::string("missing input ") + nm;
    return x;
  };
  auto t0 = std::chrono::steady_clock::now();
  if (p == "tc") {
    auto* a = need("ArcInput", 2);
    if (a) run_tc(a->rows, a->n, max_iters, threads, *r);
  } else if (p == "sg") {
    auto* a = need("Arc", 2);
    if (a) run_sg(a->rows, a->n, max_iters, threads, *r);
  } else if (p == "andersen") {
    auto *ad = need("AddressOf", 2), *as = need("Assign", 2), *ld = need("Load", 2),
         *st = need("Store", 2);
    if (ad && as && ld && st)
      run_andersen(ad->rows, ad->n, as->rows, as->n, ld->rows, ld->n, st->rows, st->n,
                   max_iters, threads, *r);
  } else if (p == "triangle" || p == "triangle_count") {
    auto* k = need("KnowsInput", 2);
    if (k) run_triangle(k->rows, k->n, p == "triangle_count", key_lo, key_hi ? key_hi : ~0u,
                        threads, *r);
  } else if (p == "reach") {
    auto *e = need("Edge", 2), *s = need("Src", 1);
    if (e && s) run_reach(e->rows, e->n, s->rows, s->n, max_iters, threads, *r);
  } else if (p == "q3") {
    auto *k = need("KnowsInput", 2), *l = need("IsLocatedInInput", 2), *pt = need("IsPartOfInput", 2);
    if (k && l && pt) run_q3(k->rows, k->n, l->rows, l->n, pt->rows, pt->n, threads, *r);
  } else if (p == "q7") {
    auto *t = need("HasTag", 2), *c = need("HasCreator", 2), *rp = need("ReplyOf", 2),
         *l = need("Likes", 2);
    if (t && c && rp && l) run_q7(t->rows, t->n, c->rows, c->n, rp->rows, rp->n, l->rows, l->n, *r);
  } else if (p == "neg2hop" || p == "neg2hop_count") {
    auto *k = need("KnowsInput", 2), *h = need("HasInterestInput", 2);
    if (k && h) run_neg2hop(k->rows, k->n, h->rows, h->n, p == "neg2hop_count", threads, *r);
  } else {
    delete r;
    return nullptr;
  }
  r->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return r;
}

const char* orc_error(void* h) {
  auto* r = static_cast<Result*>(h);
  return r->error.empty() ? nullptr : r->error.c_str();
}

// Number of 32-bit words of the sorted row-major tuple set (0 if absent).
uint64_t orc_words(void* h, const char* rel) {
  auto* r = static_cast<Result*>(h);
  auto it = r->rels.find(rel);
  return it == r->rels.end() ? 0 : it->second.size();
}

int orc_export(void* h, const char* rel, uint32_t* out, uint64_t cap_words) {
  auto* r = static_cast<Result*>(h);
  auto it = r->rels.find(rel);
  if (it == r->rels.end()) return 1;
  if (cap_words < it->second.size()) return 2;
  std::memcpy(out, it->second.data(), it->second.size() * sizeof(uint32_t));
  return 0;
}

uint64_t orc_num_deltas(void* h, const char* rel) {
  auto* r = static_cast<Result*>(h);
  auto it = r->deltas.find(rel);
  return it == r->deltas.end() ? 0 : it->second.size();
}

int orc_deltas(void* h, const char* rel, uint64_t* out, uint64_t cap) {
  auto* r = static_cast<Result*>(h);
  auto it = r->deltas.find(rel);
  if (it == r->deltas.end()) return 1;
  if (cap < it->second.size()) return 2;
  std::memcpy(out, it->second.data(), it->second.size() * sizeof(uint64_t));
  return 0;
}

uint64_t orc_iterations(void* h) { return static_cast<Result*>(h)->iterations; }
uint64_t orc_count(void* h) { return static_cast<Result*>(h)->count; }
// Count of a count-only rule's head relation (0 if absent).
uint64_t orc_rule_count(void* h, const char* rel) {
  auto* r = static_cast<Result*>(h);
  auto it = r->rule_counts.find(rel);
  return it == r->rule_counts.end() ? 0 : it->second;
}
double orc_seconds(void* h) { return static_cast<Result*>(h)->seconds; }
void orc_free(void* h) { delete static_cast<Result*>(h); }

// Root-key sample of the triangle programs (keys a with a % stride == 0).
void orc_set_key_stride(uint32_t stride) { g_key_stride = stride ? stride : 1; }

// Digest (count, sum1, sum2) of n row-major tuples of arity k (k <= 3; a
// shorter tuple hashes as if padded on the left by nothing: k = 2 hashes
// m(m(t0) ^ t1)).  The rows must be distinct for the digest to be a set's.
void orc_digest_rows(const uint32_t* rows, uint64_t n, uint32_t k, int threads, uint64_t* out) {
  if (threads <= 0) threads = omp_get_max_threads();
  std::vector<Digest> part(threads);
#pragma omp parallel for schedule(static) num_threads(threads)
  for (int64_t i = 0; i < int64_t(n); ++i) {
    const uint32_t* t = rows + uint64_t(i) * k;
    u64 h = mix64(t[0]);
    for (uint32_t c = 1; c < k; ++c) h = mix64(h ^ t[c]);
    part[omp_get_thread_num()].add(h);
  }
  Digest tot;
  for (auto& d : part) tot.merge(d);
  out[0] = tot.n;
  out[1] = tot.s1;
  out[2] = tot.s2;
}

// Digest of the full triangle listing of the KnowsInput rows (m pairs);
// *seconds = wall time of the enumeration (index build included).
void orc_triangle_digest(const uint32_t* rows, uint64_t m, int threads, uint64_t* out,
                         double* seconds) {
  if (threads <= 0) threads = omp_get_max_threads();
  omp_set_num_threads(threads);
  const auto t0 = std::chrono::steady_clock::now();
  const Digest d = triangle_digest(rows, m, threads);
  *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  out[0] = d.n;
  out[1] = d.s1;
  out[2] = d.s2;
}

}  // extern "C"
