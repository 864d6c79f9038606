"""Andersen points-to at 1/10 of the full workload (20,000 variables, 5,000
objects, AddressOf 10k / Assign 40k / Load 10k / Store 10k pairs): its joins
write about 1.2e11 tuples before deduplication, far beyond HBM, so this runs
the chunked NEW path (slice ranges deduplicated into NEW parts).  Checked
against the committed fixture tests/golden/andersen_scaled.json (per-iteration
|Delta|, |PointsTo| and the tuple-set hash of oracle/sparse.py, written by
tests/golden/gen_andersen_scaled.py), plus sortedness and uniqueness of the
exported rows."""

import json
import os

import numpy as np
import pytest

from oracle import sparse
from srdatalog_hip import datasets as ds

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "andersen_scaled.json")


@pytest.mark.parametrize("case", json.load(open(GOLD)), ids=lambda c: f"div{c['div']}-seed{c['seed']}")
def test_andersen_scaled(case):
  from srdatalog_hip import programs
  d = case["div"]
  a = ds.andersen_inputs(case["seed"], n_vars=200_000 // d, n_objs=50_000 // d,
                         sizes=(100_000 // d, 400_000 // d, 100_000 // d, 100_000 // d))
  print(f"andersen 1/{d} seed {case['seed']}", flush=True)
  lib = programs.load("andersen")
  assert lib.init() == 0
  try:
    for rel, rows in a.items():
      lib.load_rows(rel, rows)
    assert lib.run(0) == 0
    assert lib.delta_sizes("PointsTo") == case["deltas"]
    assert lib.size("PointsTo") == case["size"]
    rows = lib.export("PointsTo", 2)
    key = (rows[:, 0].astype(np.uint64) << np.uint64(32)) | rows[:, 1].astype(np.uint64)
    assert bool(np.all(key[1:] > key[:-1])), "rows not strictly sorted"
    assert str(sparse.tuple_hash(rows)) == case["hash"]
  finally:
    lib.shutdown()
