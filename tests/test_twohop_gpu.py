"""GPU parity of LSQB Q6 (the reference's examples/lsqb_q6_2hop.py).  The
listing program keeps the example's plan order, which the reference's
planner compiles to Path(p1,p2,p3,t) <= Knows(p2,p1), Knows(p2,p3),
HasInterest(p3,t), p1 != p3 (see programs.twohop_program): it is checked
against that model on small and mid inputs and, on a large input where
enumeration in Python is too slow, by its closed-form size plus the
validity of every exported row (rows distinct and all valid, count exact:
the set is exactly the answer).  The count program (p2-first order) counts
the rule as written, against its closed form.  Knows holds self-loops,
which p1 != p3 filters."""

import numpy as np
import pytest

import oracle.brute as brute
from srdatalog_hip import programs

pytestmark = pytest.mark.gpu


def inputs(seed, n, m, tags, mi):
  rng = np.random.default_rng(seed)
  k = np.unique(rng.integers(0, n, size=(m, 2)).astype(np.uint32), axis=0)
  h = np.unique(np.stack([rng.integers(0, n, mi), rng.integers(0, tags, mi)], 1)
                .astype(np.uint32), axis=0)
  return k, h


def run(name, k, h):
  lib = programs.load(name)
  assert lib.init() == 0
  try:
    lib.load_rows("KnowsInput", k)
    lib.load_rows("HasInterestInput", h)
    assert lib.run(0) == 0
    size = lib.size("Path")
    rows = lib.export("Path", 4) if name == "twohop" else None
  finally:
    lib.shutdown()
  return size, rows


def pair_keys(a, b):
  return (a.astype(np.uint64) << np.uint64(32)) | b.astype(np.uint64)


@pytest.mark.parametrize("seed,n,m,tags,mi", [(0, 40, 200, 6, 60), (1, 3000, 30000, 40, 5000)])
def test_twohop_matches_model(seed, n, m, tags, mi):
  k, h = inputs(seed, n, m, tags, mi)
  planned = brute.twohop_as_planned(k, h)
  want = np.array(sorted(planned), dtype=np.uint32).reshape(-1, 4)
  rels, _ = brute.twohop(k, h)
  print(f"twohop seed={seed} n={n} m={m}: |Path| planned {want.shape[0]}, "
        f"as written {len(rels['Path'])}", flush=True)
  size, got = run("twohop", k, h)
  assert size == want.shape[0]
  np.testing.assert_array_equal(got, want)


@pytest.mark.xfail(reason="p2-first TwoHopPath, which reads Knows(p1, p2) through a second "
                          "index (1 0), listed 4,503 tuples where the rule has 1,223 (the "
                          "count-only program counts the same 4,503): open")
@pytest.mark.parametrize("seed,n,m,tags,mi", [(0, 40, 200, 6, 60)])
def test_twohop_p2_first_listing_matches_rule(seed, n, m, tags, mi):
  """The p2-first order (Knows(p1, p2) read through index (1 0)) lists the
  rule as written."""
  k, h = inputs(seed, n, m, tags, mi)
  rels, _ = brute.twohop(k, h)
  want = np.array(sorted(rels["Path"]), dtype=np.uint32).reshape(-1, 4)
  size, got = run("twohop_p2", k, h)
  assert size == want.shape[0]
  np.testing.assert_array_equal(got, want)


@pytest.mark.xfail(reason="count-only TwoHopPath uses the p2-first plan, which lists 4,503 "
                          "tuples where the rule has 1,223 (see the listing test above): open")
@pytest.mark.parametrize("seed,n,m,tags,mi", [(0, 40, 200, 6, 60)])
def test_twohop_count_matches_rule(seed, n, m, tags, mi):
  k, h = inputs(seed, n, m, tags, mi)
  rels, _ = brute.twohop(k, h)
  size_c, _ = run("twohop_count", k, h)
  assert size_c == len(rels["Path"])


def test_twohop_large_closed_form_and_validity():
  n, m, tags, mi = 200000, 2000000, 500, 300000
  k, h = inputs(7, n, m, tags, mi)
  want = brute.twohop_as_planned_count_closed_form(k, h, n)
  want_c = brute.twohop_count_closed_form(k, h, n)
  print(f"twohop large n={n} m={k.shape[0]}: |Path| planned {want}, as written {want_c}",
        flush=True)
  size, got = run("twohop", k, h)
  assert size == want and got.shape == (want, 4)
  p1, p2, p3, t = (got[:, c] for c in range(4))
  assert not (p1 == p3).any()
  kk = np.sort(pair_keys(k[:, 0], k[:, 1]))
  hk = np.sort(pair_keys(h[:, 0], h[:, 1]))
  for keys, ref in ((pair_keys(p2, p1), kk), (pair_keys(p2, p3), kk), (pair_keys(p3, t), hk)):
    pos = np.minimum(np.searchsorted(ref, keys), ref.shape[0] - 1)
    assert (ref[pos] == keys).all()
  # strictly increasing rows: distinct, and in the index's order
  hi = pair_keys(p1, p2)
  lo = pair_keys(p3, t)
  step = (hi[1:] > hi[:-1]) | ((hi[1:] == hi[:-1]) & (lo[1:] > lo[:-1]))
  assert step.all()
