"""Pin the C++ oracle (oracle/oracle.cpp) before trusting it.

Checks, none of which retype the oracle's own loops:
* brute force: pure-Python naive Jacobi iteration (oracle/brute.py) on tiny
  seeded inputs, for every program -- final sets AND per-iteration deltas;
* closed form (TC): |DELTA_i| equals the number of ordered pairs whose
  shortest-path distance is exactly i+2 (all-pairs BFS with scipy);
* closed form (triangles): sum(A .* (A @ A)) over the oriented adjacency;
* invariants: FULL = disjoint union of deltas, deltas end with 0.
"""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as csg

import oracle
from oracle import brute
from srdatalog_hip import datasets as ds


def as_set(rows):
  return set(map(tuple, np.asarray(rows).tolist()))


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
@pytest.mark.parametrize("n,m", [(12, 20), (40, 90), (80, 200)])
def test_tc_matches_brute(seed, n, m):
  e = ds.tc_edges(seed, n, m)
  r = oracle.run("tc", {"ArcInput": e})
  b, bd = brute.tc(e)
  assert as_set(r.relations["Path"]) == b["Path"]
  assert r.deltas["Path"] == bd["Path"]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_tc_deltas_are_bfs_distance_histogram(seed):
  n, m = 300, 900
  e = ds.tc_edges(seed, n, m)
  a = sp.csr_matrix((np.ones(m), (e[:, 0], e[:, 1])), shape=(n, n))
  dist = csg.shortest_path(a, method="D", unweighted=True)
  # Path after the base step = distance-1 pairs (and cycles give dist(x,x)
  # via walks; shortest_path puts 0 on the diagonal, so handle x->x walks
  # separately with the shortest cycle through x).
  d = dist.copy()
  np.fill_diagonal(d, np.inf)
  cyc = np.full(n, np.inf)
  for x in range(n):
    succ = e[e[:, 0] == x, 1]
    if succ.size:
      cyc[x] = 1 + dist[succ, x].min()
  d[np.arange(n), np.arange(n)] = cyc
  finite = d[np.isfinite(d)].astype(np.int64)
  hist = np.bincount(finite)
  r = oracle.run("tc", {"ArcInput": e})
  want = [int(hist[k]) if k < hist.size else 0 for k in range(2, 2 + len(r.deltas["Path"]))]
  assert r.deltas["Path"] == want
  assert r.deltas["Path"][-1] == 0
  assert sum(r.deltas["Path"]) + m == r.size("Path") == finite.size


@pytest.mark.parametrize("seed", [0, 1, 2, 5])
@pytest.mark.parametrize("n", [30, 120])
def test_sg_matches_brute(seed, n):
  e = ds.sg_forest_edges(np.random.default_rng(seed), n)
  r = oracle.run("sg", {"Arc": e})
  b, bd = brute.sg(e)
  assert as_set(r.relations["Sg"]) == b["Sg"]
  assert r.deltas["Sg"] == bd["Sg"]


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_andersen_matches_brute(seed):
  a = ds.andersen_inputs(seed, n_vars=300, n_objs=80, sizes=(120, 400, 120, 120))
  r = oracle.run("andersen", a)
  b, bd = brute.andersen(a["AddressOf"], a["Assign"], a["Load"], a["Store"])
  assert as_set(r.relations["PointsTo"]) == b["PointsTo"]
  assert r.deltas["PointsTo"] == bd["PointsTo"]


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("scale,m", [(6, 300), (9, 3000)])
def test_triangle_matches_brute_and_closed_form(seed, scale, m):
  e = ds.triangle_edges(seed, scale, m)
  r = oracle.run("triangle", {"KnowsInput": e})
  b, _ = brute.triangle(e)
  assert as_set(r.relations["Triangle"]) == b["Triangle"]
  n = 1 << scale
  a = sp.csr_matrix((np.ones(e.shape[0]), (e[:, 0], e[:, 1])), shape=(n, n))
  closed = int(a.multiply(a @ a).sum())
  assert r.count == closed == len(b["Triangle"])
  rc = oracle.run("triangle_count", {"KnowsInput": e})
  assert rc.count == closed and rc.relations["Triangle"].shape[0] == 0
  # Orientation low->high makes every listed triangle strictly increasing.
  t = r.relations["Triangle"]
  assert np.all(t[:, 0] < t[:, 1]) and np.all(t[:, 1] < t[:, 2])


@pytest.mark.parametrize("seed", [0, 1])
def test_reach_matches_brute_and_bfs(seed):
  n = 200
  e = ds.tc_edges(seed, n, 300)
  src = np.array([seed * 7 % n], dtype=np.uint32)
  r = oracle.run("reach", {"Edge": e, "Src": src})
  b, bd = brute.reach(e, src)
  assert set(r.relations["Reach"].ravel().tolist()) == b["Reach"]
  assert r.deltas["Reach"] == bd["Reach"]
  a = sp.csr_matrix((np.ones(e.shape[0]), (e[:, 0], e[:, 1])), shape=(n, n))
  lv = csg.breadth_first_order(a, int(src[0]), return_predecessors=False)
  assert len(lv) == r.size("Reach")


def test_max_iters_truncates():
  e = ds.tc_edges(0, 40, 90)
  full = oracle.run("tc", {"ArcInput": e})
  part = oracle.run("tc", {"ArcInput": e}, max_iters=2)
  assert part.deltas["Path"] == full.deltas["Path"][:2]
  assert part.iterations == 2


def test_sorted_unique_exports():
  e = ds.tc_edges(1, 50, 120)
  r = oracle.run("tc", {"ArcInput": e})
  p = r.relations["Path"].astype(np.uint64)
  k = (p[:, 0] << np.uint64(32)) | p[:, 1]
  assert np.all(k[1:] > k[:-1])


def _neg2hop_inputs(seed, n=60, m=300, tags=12, mi=150):
  rng = np.random.default_rng(seed)
  k = ds.tc_edges(seed, n, m)
  hi = np.unique(np.stack([rng.integers(0, n, mi), rng.integers(0, tags, mi)], 1).astype(np.uint32), axis=0)
  return {"KnowsInput": k, "HasInterestInput": hi}


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_neg2hop_matches_brute(seed):
  inp = _neg2hop_inputs(seed)
  r = oracle.run("neg2hop", inp)
  b, _ = brute.neg2hop(inp["KnowsInput"], inp["HasInterestInput"])
  assert as_set(r.relations["Path"]) == b["Path"]
  c = oracle.run("neg2hop_count", inp)
  assert c.count == len(b["Path"]) == r.count


def test_neg2hop_closed_form():
  """count = sum over 2-paths (p1,p2,p3) with p1 != p3 and no edge p1->p3
  of |tags(p3)|, with dense matrices: A2 = A@A masked by (1-A) and off-diagonal."""
  inp = _neg2hop_inputs(5, n=80, m=500, tags=9, mi=200)
  n = 80
  a = np.zeros((n, n), np.int64)
  e = inp["KnowsInput"]
  a[e[:, 0], e[:, 1]] = 1
  t = np.zeros(n, np.int64)
  np.add.at(t, inp["HasInterestInput"][:, 0], 1)
  a2 = a @ a
  mask = (1 - a) * (1 - np.eye(n, dtype=np.int64))
  want = int(((a2 * mask) @ t).sum())
  assert oracle.run("neg2hop_count", inp).count == want


def q7_inputs(seed, n_msg=40, n_tag=6, n_person=30, sizes=(80, 50, 60, 70)):
  rng = np.random.default_rng(seed)
  def pairs(n, a, b):
    return np.unique(np.stack([rng.integers(0, a, n), rng.integers(0, b, n)], 1).astype(np.uint32), axis=0)
  return {"HasTag": pairs(sizes[0], n_msg, n_tag), "HasCreator": pairs(sizes[1], n_msg, n_person),
          "ReplyOf": pairs(sizes[2], 1000, n_msg), "Likes": pairs(sizes[3], n_person, n_msg)}


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_q7_matches_brute(seed):
  inp = q7_inputs(seed)
  r = oracle.run("q7", inp)
  want = brute.q7(inp["HasTag"], inp["HasCreator"], inp["ReplyOf"], inp["Likes"])
  assert r.rule_counts == want
  assert all(v > 0 for v in want.values()) or seed > 0


def q3_inputs(seed, n=40, m=260, cities=10, countries=3, extra_loc=6):
  rng = np.random.default_rng(seed)
  k = ds.tc_edges(seed, n, m)
  loc = np.stack([np.arange(n), rng.integers(0, cities, n)], 1)
  more = np.stack([rng.integers(0, n, extra_loc), rng.integers(0, cities, extra_loc)], 1)
  loc = np.unique(np.concatenate([loc, more]).astype(np.uint32), axis=0)
  part = np.stack([np.arange(cities), rng.integers(0, countries, cities)], 1)
  part = np.unique(np.concatenate([part, [[0, countries - 1]]]).astype(np.uint32), axis=0)
  return {"KnowsInput": k, "IsLocatedInInput": loc, "IsPartOfInput": part}


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_q3_matches_brute(seed):
  inp = q3_inputs(seed)
  r = oracle.run("q3", inp)
  want = brute.q3(inp["KnowsInput"], inp["IsLocatedInInput"], inp["IsPartOfInput"])["Triangle"]
  assert as_set(r.relations["Triangle"]) == want
  assert len(want) > 0


def _py_digest(rows):
  """The tuple-set digest written out in Python integers (splitmix64
  finaliser chained over the columns; two sums mod 2^64)."""
  M = (1 << 64) - 1

  def mix(z):
    z = (z + 0x9e3779b97f4a7c15) & M
    z = ((z ^ (z >> 30)) * 0xbf58476d1ce4e5b9) & M
    z = ((z ^ (z >> 27)) * 0x94d049bb133111eb) & M
    return z ^ (z >> 31)
  s1 = s2 = 0
  for t in np.asarray(rows).tolist():
    h = mix(t[0])
    for v in t[1:]:
      h = mix(h ^ v)
    s1 = (s1 + h) & M
    s2 = (s2 + mix((h + 0x632be59bd9b4e019) & M)) & M
  return len(rows), s1, s2


def test_digest_pins():
  rng = np.random.default_rng(5)
  rows = np.unique(rng.integers(0, 1 << 20, size=(500, 3), dtype=np.uint32), axis=0)
  d = oracle.digest_rows(rows)
  assert d == _py_digest(rows)
  # order-independent, and sensitive to one changed value or a swapped pair
  assert oracle.digest_rows(rows[rng.permutation(rows.shape[0])]) == d
  bad = rows.copy()
  bad[7, 2] ^= 1
  assert oracle.digest_rows(bad) != d
  sw = rows.copy()
  sw[3, [1, 2]] = sw[3, [2, 1]]
  assert oracle.digest_rows(sw) != d
  assert oracle.digest_rows(rows[:-1]) != d


@pytest.mark.parametrize("seed,scale,m", [(0, 9, 3000), (1, 12, 60000)])
def test_triangle_digest_matches_listing(seed, scale, m):
  e = ds.triangle_edges(seed, scale, m)
  r = oracle.run("triangle", {"KnowsInput": e})
  d, _ = oracle.triangle_digest(e)
  assert d == oracle.digest_rows(r.relations["Triangle"])
  assert d[0] == r.count


def test_triangle_key_stride_sample():
  e = ds.triangle_edges(2, 10, 8000)
  full = oracle.run("triangle", {"KnowsInput": e}).relations["Triangle"]
  oracle.set_key_stride(4)
  try:
    part = oracle.run("triangle", {"KnowsInput": e}).relations["Triangle"]
  finally:
    oracle.set_key_stride(1)
  np.testing.assert_array_equal(part, full[full[:, 0] % 4 == 0])


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_split_rec_brute_is_blocked_bfs(seed):
  """The split rule's model R(x,z) <= R(x,y), !Filter(y), Edge(y,z) from
  Seed: (x, z) is in R iff some seed (x, y) reaches z along edges leaving
  only unblocked vertices, and it first appears in iteration d - 1, d the
  fewest such edges; checked by a per-x breadth-first search."""
  from collections import deque
  rng = np.random.default_rng(seed)
  n = 60
  edge = ds.tc_edges(seed, n, 150)
  sd = np.stack([rng.integers(0, n, 12), rng.integers(0, n, 12)], 1)
  filt = rng.choice(n, size=8 * seed, replace=False).reshape(-1, 1)
  rel, deltas = brute.split_rec(sd, filt, edge)
  deltas = deltas["R"]
  blocked = {int(v) for v in filt.ravel()}
  succ = {}
  for a, b in edge:
    succ.setdefault(int(a), []).append(int(b))
  dist_count, pairs = {}, set()
  for x in sorted({int(a) for a, _ in sd}):
    dist = {int(b): 0 for a, b in sd if int(a) == x}
    q = deque(dist)
    while q:
      u = q.popleft()
      if u in blocked:
        continue
      for w in succ.get(u, ()):
        if w not in dist:
          dist[w] = dist[u] + 1
          q.append(w)
    for z, d in dist.items():
      pairs.add((x, z))
      if d > 0:
        dist_count[d] = dist_count.get(d, 0) + 1
  assert rel["R"] == pairs
  want = [dist_count.get(i + 1, 0) for i in range(len(deltas))]
  assert deltas == want
  assert deltas[-1] == 0 and all(d > 0 for d in deltas[:-1])


def test_split_nonrec_brute_closed_form():
  rng = np.random.default_rng(5)
  a = rng.integers(0, 40, (200, 2))
  b = rng.choice(40, 10, replace=False).reshape(-1, 1)
  c = rng.integers(0, 40, (150, 2))
  rel, _ = brute.split_nonrec(a, b, c)
  want = {(int(x), int(z)) for x, y in a for xx, z in c if x == xx and int(y) not in set(b.ravel())}
  assert rel["Q"] == want
