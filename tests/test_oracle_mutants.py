"""The oracle's pins must be able to fail.

Each test plants one plausible mistake in a copy of the brute-force model
(oracle/brute.py) and checks that the independent pin used for that program
-- the all-pairs BFS distance histogram for TC deltas, the sparse closed
form for triangles, agreement with the C++ restatement for Andersen / SG --
rejects it, while the unmutated model passes the same check.
"""

import numpy as np
import scipy.sparse as sp
import scipy.sparse.csgraph as csg

import oracle
from oracle import brute
from srdatalog_hip import datasets as ds


def as_set(rows):
  return set(map(tuple, np.asarray(rows).tolist()))


def tc_bfs_deltas(e, n, count):
  a = sp.csr_matrix((np.ones(len(e)), (e[:, 0], e[:, 1])), shape=(n, n))
  dist = csg.shortest_path(a, method="D", unweighted=True)
  d = dist.copy()
  np.fill_diagonal(d, np.inf)
  for x in range(n):
    succ = e[e[:, 0] == x, 1]
    if succ.size:
      d[x, x] = 1 + dist[succ, x].min()
  hist = np.bincount(d[np.isfinite(d)].astype(np.int64))
  return [int(hist[k]) if k < hist.size else 0 for k in range(2, 2 + count)]


def tc_model(arc, *, iter0_reads_delta=False, delta_not_subtracted=False, max_iters=50):
  """Naive Jacobi TC; the flags plant one mistake each: iteration 0 reading
  an empty DELTA instead of FULL, or DELTA reported as T(F) instead of
  T(F) - F."""
  edge = {(int(a), int(b)) for a, b in arc}
  succ = {}
  for a, b in edge:
    succ.setdefault(a, set()).add(b)
  full, deltas = set(edge), []
  for it in range(max_iters):
    src = set() if (it == 0 and iter0_reads_delta) else full
    t = {(x, z) for (x, y) in src for z in succ.get(y, ())}
    new = t - full
    deltas.append(len(t) if delta_not_subtracted else len(new))
    full |= new
    if not new:
      break
  return full, deltas


def test_tc_pin_rejects_mutants():
  n, m = 120, 300
  e = ds.tc_edges(4, n, m)
  _, good = tc_model(e)
  assert good == brute.tc(e)[1]["Path"]
  assert good == tc_bfs_deltas(e, n, len(good))
  for kw in ({"iter0_reads_delta": True}, {"delta_not_subtracted": True}):
    _, bad = tc_model(e, **kw)
    assert bad != tc_bfs_deltas(e, n, len(bad)), kw


def test_triangle_pin_rejects_transposed_atom():
  e = ds.triangle_edges(1, 9, 3000)
  n = int(e.max()) + 1
  a = sp.csr_matrix((np.ones(len(e)), (e[:, 0], e[:, 1])), shape=(n, n))
  closed = int(a.multiply(a @ a).sum())
  kn = {(int(x), int(y)) for x, y in e}
  succ = {}
  for x, y in kn:
    succ.setdefault(x, set()).add(y)
  good = {(x, y, z) for (x, y) in kn for z in succ.get(y, ()) if z in succ.get(x, ())}
  # mistake: third atom Edge(z, x) instead of Edge(x, z)
  bad = {(x, y, z) for (x, y) in kn for z in succ.get(y, ()) if x in succ.get(z, ())}
  assert len(good) == closed > 0
  assert len(bad) != closed


def test_andersen_pin_rejects_dropped_rule():
  a = ds.andersen_inputs(2, n_vars=300, n_objs=80, sizes=(150, 500, 150, 150))
  r = oracle.run("andersen", a)
  good, gd = brute.andersen(a["AddressOf"], a["Assign"], a["Load"], a["Store"])
  assert as_set(r.relations["PointsTo"]) == good["PointsTo"]
  assert r.deltas["PointsTo"] == gd["PointsTo"]
  # mistake: the Store rule dropped (empty Store input is the same model)
  empty = np.zeros((0, 2), np.uint32)
  bad, bd = brute.andersen(a["AddressOf"], a["Assign"], a["Load"], empty)
  assert as_set(r.relations["PointsTo"]) != bad["PointsTo"] or r.deltas["PointsTo"] != bd["PointsTo"]


def test_sg_pin_rejects_missing_inequality():
  e = ds.sg_forest_edges(np.random.default_rng(5), 300)
  r = oracle.run("sg", {"Arc": e})
  good, _ = brute.sg(e)
  assert as_set(r.relations["Sg"]) == good["Sg"]
  # mistake: base rule without x != y
  arc = {(int(p), int(x)) for p, x in e}
  kids = {}
  for p, x in arc:
    kids.setdefault(p, set()).add(x)
  bad_base = {(x, y) for ks in kids.values() for x in ks for y in ks}
  assert not bad_base <= as_set(r.relations["Sg"])
