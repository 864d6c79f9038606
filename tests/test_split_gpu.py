"""GPU parity of split rules (a rule body cut by the SPLIT marker into a
pipeline writing a temp relation and a pipeline scanning it): the relation
and every per-iteration delta equal the brute-force model, and the split
plan gives exactly what the same rule gives unsplit."""

import numpy as np
import pytest

from oracle import brute
from srdatalog_hip import datasets as ds

pytestmark = pytest.mark.gpu


def run(prog, inputs):
  from srdatalog_hip import programs
  lib = programs.load(prog)
  assert lib.init() == 0
  for rel, rows in inputs.items():
    lib.load_rows(rel, rows)
  assert lib.run(0) == 0
  return lib


def as_rows(s, k):
  if not s:
    return np.zeros((0, k), dtype=np.uint32)
  return np.array(sorted(s), dtype=np.uint32).reshape(-1, k)


@pytest.mark.parametrize("n,m,seeds,blocked,seed", [(50, 120, 3, 5, 0), (400, 1200, 10, 40, 1),
                                                     (2000, 5000, 25, 0, 2), (2000, 5000, 25, 2000, 3),
                                                     (3000, 9000, 1, 300, 4)])
def test_split_rec(n, m, seeds, blocked, seed):
  rng = np.random.default_rng(seed)
  edge = ds.tc_edges(seed, n, m)
  sd = np.stack([rng.integers(0, n, seeds), rng.integers(0, n, seeds)], 1).astype(np.uint32)
  filt = rng.choice(n, size=min(blocked, n), replace=False).astype(np.uint32).reshape(-1, 1)
  inputs = {"Seed": sd, "Filter": filt, "Edge": edge}
  want_rel, want_d = brute.split_rec(sd, filt, edge)
  want = as_rows(want_rel["R"], 2)
  print(f"split_rec n={n} m={m} |R|={want.shape[0]} iterations={len(want_d['R'])}", flush=True)
  for prog in ("split_rec", "split_rec_plain"):
    lib = run(prog, inputs)
    try:
      got = lib.export("R", 2)
      assert got.shape == want.shape, prog
      np.testing.assert_array_equal(got, want)
      assert lib.delta_sizes("R") == want_d["R"], prog
    finally:
      lib.shutdown()


@pytest.mark.parametrize("n,seed", [(30, 0), (2000, 1), (20000, 2)])
def test_split_nonrec(n, seed):
  rng = np.random.default_rng(seed)
  a = np.stack([rng.integers(0, n, 3 * n), rng.integers(0, n, 3 * n)], 1).astype(np.uint32)
  b = rng.choice(n, size=n // 3, replace=False).astype(np.uint32).reshape(-1, 1)
  c = np.stack([rng.integers(0, n, 2 * n), rng.integers(0, n, 2 * n)], 1).astype(np.uint32)
  want_rel, _ = brute.split_nonrec(a, b, c)
  want = as_rows(want_rel["Q"], 2)
  lib = run("split_nonrec", {"A": a, "B": b, "C": c})
  try:
    got = lib.export("Q", 2)
    assert got.shape == want.shape
    np.testing.assert_array_equal(got, want)
  finally:
    lib.shutdown()
