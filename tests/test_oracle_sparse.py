"""The sparse-matrix Andersen restatement (oracle/sparse.py) agrees with the
C++ oracle and the brute-force evaluator on small and mid-size inputs, so it
can check scales the C++ oracle cannot finish in seconds."""

import numpy as np
import pytest

import oracle
from oracle import brute, sparse
from srdatalog_hip import datasets as ds


def _inputs(div, seed=0):
  return ds.andersen_inputs(seed, n_vars=200_000 // div, n_objs=50_000 // div,
                            sizes=(100_000 // div, 400_000 // div, 100_000 // div,
                                   100_000 // div))


@pytest.mark.parametrize("div,seed", [(1000, 0), (300, 1), (100, 0), (50, 2)])
def test_sparse_matches_cpp_oracle(div, seed):
  a = _inputs(div, seed)
  rows, deltas = sparse.andersen(a["AddressOf"], a["Assign"], a["Load"], a["Store"])
  want = oracle.run("andersen", a)
  np.testing.assert_array_equal(rows, want.relations["PointsTo"])
  assert deltas == want.deltas["PointsTo"]
  assert sparse.tuple_hash(rows) == sparse.tuple_hash(want.relations["PointsTo"])


@pytest.mark.parametrize("seed", range(4))
def test_sparse_matches_brute(seed):
  rng = np.random.default_rng(seed)
  mk = lambda k, hi_a, hi_b: np.unique(np.stack([rng.integers(0, hi_a, k), rng.integers(0, hi_b, k)],
                                                1).astype(np.uint32), axis=0)
  addr, assign, load, store = mk(15, 30, 40), mk(25, 30, 30), mk(10, 30, 30), mk(10, 30, 30)
  rows, deltas = sparse.andersen(addr, assign, load, store)
  rel, d = brute.andersen(addr, assign, load, store)
  assert {tuple(map(int, r)) for r in rows} == rel["PointsTo"]
  assert deltas == d["PointsTo"]


def test_tuple_hash_is_a_set_hash():
  r = np.array([[1, 2], [3, 4], [0, 7]], np.uint32)
  assert sparse.tuple_hash(r) == sparse.tuple_hash(r[::-1])
  assert sparse.tuple_hash(r) != sparse.tuple_hash(r[:2])
  assert sparse.tuple_hash(r) != sparse.tuple_hash(np.array([[2, 1], [3, 4], [0, 7]], np.uint32))
