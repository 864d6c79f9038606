"""GPU parity of Galen (the reference's examples/galen.py): OutP and OutQ
mutually recursive through three-atom joins with delta-specific plans,
against the brute-force model's relations and per-iteration delta sizes."""

import numpy as np
import pytest

import oracle.brute as brute
from srdatalog_hip import programs

pytestmark = pytest.mark.gpu

ENGINES = [({}, "default"), ({"SRDATALOG_DEVLOOP": "0"}, "host-loop"),
           ({"SRDATALOG_DEVLOOP": "0", "SRDATALOG_DENSE_MAX_BITS": "0"}, "host-loop-rows")]


def inputs(seed, n, roles, m):
  rng = np.random.default_rng(seed)

  def rel(k, cols):
    a = np.stack([rng.integers(0, hi, size=k) for hi in cols], 1).astype(np.uint32)
    return np.unique(a, axis=0)
  return {"PInput": rel(m, (n, n)), "QInput": rel(m, (n, roles, n)),
          "RInput": rel(m // 3, (roles, roles, roles)), "CInput": rel(m // 2, (n, n, n)),
          "UInput": rel(m // 2, (n, roles, n)), "SInput": rel(roles, (roles, roles))}


_models: dict = {}


@pytest.mark.parametrize("env,label", ENGINES, ids=[e[1] for e in ENGINES])
@pytest.mark.parametrize("seed,n,roles,m", [(0, 30, 4, 40), (1, 120, 6, 150)])
def test_galen_matches_model(seed, n, roles, m, env, label, monkeypatch):
  for k in ("SRDATALOG_DEVLOOP", "SRDATALOG_DENSE_MAX_BITS"):
    monkeypatch.delenv(k, raising=False)
  for k, v in env.items():
    monkeypatch.setenv(k, v)
  inp = inputs(seed, n, roles, m)
  key = (seed, n, roles, m)
  if key not in _models:
    _models[key] = brute.galen(inp["PInput"], inp["QInput"], inp["RInput"], inp["CInput"],
                               inp["UInput"], inp["SInput"])
  want, deltas = _models[key]
  print(f"galen seed={seed} n={n}: OutP {len(want['OutP'])}, OutQ {len(want['OutQ'])}, "
        f"{len(deltas['OutP'])} iterations", flush=True)
  lib = programs.load("galen")
  assert lib.init() == 0
  try:
    for rel, rows in inp.items():
      lib.load_rows(rel, rows)
    assert lib.run(0) == 0
    for rel, k in (("OutP", 2), ("OutQ", 3)):
      got = lib.export(rel, k)
      exp = np.array(sorted(want[rel]), dtype=np.uint32).reshape(-1, k)
      assert got.shape == exp.shape, (rel, got.shape, exp.shape)
      np.testing.assert_array_equal(got, exp, err_msg=rel)
      assert lib.delta_sizes(rel) == deltas[rel], rel
  finally:
    lib.shutdown()
