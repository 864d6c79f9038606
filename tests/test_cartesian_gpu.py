"""GPU parity of the cartesian walks (flattened per-tuple decode and the
register-tiled walk of large products) on same-generation over forests of
depth 2, whose fixpoint has a closed form: Sg holds every ordered pair
(x, y), x != y, of nodes at the same depth under the same root.  SGRec's
last level is the cartesian product of the children of two siblings, so a
root of fanout f gives products of f * f tuples: fanout 48 and above takes
the tiled walk (products of >= 1,024 tuples with a side of >= 48 rows),
small fanouts the flattened one, and a mixed forest runs both in one launch.
Iteration 0 derives the cousins (f^3 (f - 1) per root), iteration 1 nothing."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def forest(fans):
  """Depth-2 forest, one root per entry of `fans` with that many children
  per node; ids in breadth-first order per root.  Returns (edges, groups):
  groups lists, per root, its depth-1 and depth-2 node id arrays."""
  edges, groups = [], []
  nxt = len(fans)
  for r, f in enumerate(fans):
    d1 = np.arange(nxt, nxt + f, dtype=np.int64)
    nxt += f
    edges.append(np.stack([np.full(f, r), d1], 1))
    d2 = np.arange(nxt, nxt + f * f, dtype=np.int64)
    nxt += f * f
    edges.append(np.stack([np.repeat(d1, f), d2], 1))
    groups.append((d1, d2))
  return np.concatenate(edges).astype(np.uint32), groups


def expected(groups):
  """The closed form, already in (x, y) order: ids grow root by root and,
  within a root, from depth 1 to depth 2."""
  parts = []
  for d1, d2 in groups:
    for g in (d1, d2):
      x = np.repeat(g, g.size)
      y = np.tile(g, g.size)
      keep = x != y
      parts.append(np.stack([x[keep], y[keep]], 1).astype(np.uint32))
  return np.concatenate(parts)


@pytest.mark.parametrize("fans", [(48, 48), (60, 5, 8, 1, 49), (3, 7, 2), (100,)],
                         ids=["tiled", "mixed", "flat", "tiled-wide"])
def test_sg_forest_closed_form(fans):
  from srdatalog_hip import programs
  edges, groups = forest(fans)
  rng = np.random.default_rng(len(fans))
  edges = edges[rng.permutation(edges.shape[0])]
  want = expected(groups)
  cousins = sum(f ** 3 * (f - 1) for f in fans)
  print(f"sg forest fans={fans}: |Sg|={want.shape[0]}", flush=True)
  lib = programs.load("sg")
  assert lib.init() == 0
  try:
    lib.load_rows("Arc", edges)
    assert lib.run(0) == 0
    got = lib.export("Sg", 2)
    assert got.shape == want.shape
    np.testing.assert_array_equal(got, want)
    assert lib.delta_sizes("Sg") == [cousins, 0]
  finally:
    lib.shutdown()
