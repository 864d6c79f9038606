"""GPU parity of the type-lookup program (constants in heads, body atoms and
negations, a wildcard negation, a repeated head variable, a filter, three
recursive relations) against the brute-force model on random class
hierarchies: every derived relation equal as a set."""

import numpy as np
import pytest

import oracle.brute as brute
from srdatalog_hip import programs

pytestmark = pytest.mark.gpu

RELS = {"SubtypeOf": 2, "Implemented": 3, "Lookup": 3, "MainMethod": 1, "Initialized": 1}


def inputs(seed, n_cls, n_meth):
  rng = np.random.default_rng(seed)
  cls = np.arange(10, 10 + n_cls)
  # a forest plus a few extra superclass edges (some classes have two)
  sup = [(int(c), int(rng.integers(10, c))) for c in cls[1:] if rng.random() < 0.9]
  sup += [(int(rng.choice(cls)), int(rng.choice(cls))) for _ in range(n_cls // 20)]
  meth = np.arange(1000, 1000 + n_meth)
  decl = [(int(m), int(rng.choice(cls))) for m in meth]
  name = [(int(m), int(rng.integers(0, 12))) for m in meth]
  mod = sorted({(int(rng.choice([1, 2, 4])), int(m)) for m in meth for _ in range(2)
                if rng.random() < 0.4})
  # the initialiser id 3 as a static main method of a main class
  decl.append((3, int(cls[0])))
  name.append((3, 7))
  mod.append((2, 3))
  main_cls = [(int(c),) for c in rng.choice(cls, size=max(1, n_cls // 50), replace=False)]
  arr = lambda rows, k: np.array(sorted(set(rows)), dtype=np.uint32).reshape(-1, k)
  return {"DirectSuperclass": arr(sup, 2), "ClassType": arr([(int(c),) for c in cls], 1),
          "MethodDecl": arr(decl, 2), "MethodName": arr(name, 2), "MethodModifier": arr(mod, 2),
          "MainClass": arr(main_cls, 1)}


# the default engine (device-resident recursive steps where they qualify,
# dense-bitmap NEW), the host-driven loop, and the host loop writing rows
ENGINES = [({}, "default"), ({"SRDATALOG_DEVLOOP": "0"}, "host-loop"),
           ({"SRDATALOG_DEVLOOP": "0", "SRDATALOG_DENSE_MAX_BITS": "0"}, "host-loop-rows")]


@pytest.mark.parametrize("env,label", ENGINES, ids=[e[1] for e in ENGINES])
@pytest.mark.parametrize("seed,n_cls,n_meth", [(0, 40, 200), (1, 300, 3000), (2, 1500, 20000)])
def test_typelookup_matches_model(seed, n_cls, n_meth, env, label, monkeypatch):
  for k in ("SRDATALOG_DEVLOOP", "SRDATALOG_DENSE_MAX_BITS"):
    monkeypatch.delenv(k, raising=False)
  for k, v in env.items():
    monkeypatch.setenv(k, v)
  inp = inputs(seed, n_cls, n_meth)
  want = brute.typelookup(inp["DirectSuperclass"], inp["ClassType"], inp["MethodDecl"],
                          inp["MethodName"], inp["MethodModifier"], inp["MainClass"])
  print(f"typelookup seed={seed} classes={n_cls} methods={n_meth}: "
        + ", ".join(f"{k} {len(v)}" for k, v in want.items()), flush=True)
  lib = programs.load("typelookup")
  assert lib.init() == 0
  try:
    for rel, rows in inp.items():
      lib.load_rows(rel, rows)
    assert lib.run(0) == 0
    for rel, k in RELS.items():
      got = lib.export(rel, k)
      exp = np.array(sorted(want[rel]), dtype=np.uint32).reshape(-1, k)
      assert got.shape == exp.shape, (rel, got.shape, exp.shape)
      np.testing.assert_array_equal(got, exp, err_msg=rel)
  finally:
    lib.shutdown()
