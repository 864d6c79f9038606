"""GPU parity of CSPA (the reference's examples/cspa.py): three mutually
recursive relations in one strongly connected component, against the
brute-force model's relations and per-iteration delta sizes."""

import numpy as np
import pytest

import oracle.brute as brute
from srdatalog_hip import programs

pytestmark = pytest.mark.gpu

ENGINES = [({}, "default"), ({"SRDATALOG_DEVLOOP": "0"}, "host-loop"),
           ({"SRDATALOG_DEVLOOP": "0", "SRDATALOG_DENSE_MAX_BITS": "0"}, "host-loop-rows")]


def inputs(seed, n, n_assign, n_deref):
  rng = np.random.default_rng(seed)
  pairs = lambda k: np.unique(rng.integers(0, n, size=(k, 2)).astype(np.uint32), axis=0)
  return {"AssignInput": pairs(n_assign), "DerefInput": pairs(n_deref)}


_models: dict = {}


def model(case, inp):
  """The brute-force result of one case, computed once for all engines
  (the larger case takes seconds in pure Python)."""
  if case not in _models:
    _models[case] = brute.cspa(inp["AssignInput"], inp["DerefInput"])
  return _models[case]


@pytest.mark.parametrize("env,label", ENGINES, ids=[e[1] for e in ENGINES])
@pytest.mark.parametrize("seed,n,n_assign,n_deref", [(0, 60, 80, 30), (1, 200, 250, 80)])
def test_cspa_matches_model(seed, n, n_assign, n_deref, env, label, monkeypatch):
  for k in ("SRDATALOG_DEVLOOP", "SRDATALOG_DENSE_MAX_BITS"):
    monkeypatch.delenv(k, raising=False)
  for k, v in env.items():
    monkeypatch.setenv(k, v)
  inp = inputs(seed, n, n_assign, n_deref)
  want, deltas = model((seed, n, n_assign, n_deref), inp)
  print(f"cspa seed={seed} n={n}: " + ", ".join(f"{k} {len(v)}" for k, v in want.items())
        + f", {len(deltas['ValueFlow'])} iterations", flush=True)
  lib = programs.load("cspa")
  assert lib.init() == 0
  try:
    for rel, rows in inp.items():
      lib.load_rows(rel, rows)
    assert lib.run(0) == 0
    for rel, rows in want.items():
      got = lib.export(rel, 2)
      exp = np.array(sorted(rows), dtype=np.uint32).reshape(-1, 2)
      assert got.shape == exp.shape, (rel, got.shape, exp.shape)
      np.testing.assert_array_equal(got, exp, err_msg=rel)
      assert lib.delta_sizes(rel) == deltas[rel], rel
  finally:
    lib.shutdown()
