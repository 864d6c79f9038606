"""Host logic of batched execution: packing, splitting, delta trimming and
admission (no GPU).  The packed-run semantics is checked against the oracle:
the fixpoint of the packed database, split by the instance table, must equal
every instance's own fixpoint, and the packed per-iteration delta sizes must
be the sums of the instances' own sequences."""

import numpy as np

import oracle
from srdatalog_hip import batch
from srdatalog_hip import datasets as ds


def _instances(seed, count=6):
  rng = np.random.default_rng(seed)
  out = []
  for i in range(count):
    n = int(rng.integers(20, 300))
    if i % 2 == 0:
      out.append(batch.Instance("tc", n, ds.tc_edges(seed * 100 + i, n, int(n * 1.6))))
    else:
      out.append(batch.Instance("sg", n, ds.sg_forest_edges(np.random.default_rng(seed + i), n)))
  return out


def test_pack_split_roundtrip():
  insts = [i for i in _instances(1) if i.kind == "tc"]
  rows, bounds = batch.pack(insts)
  assert bounds[0] == 0 and bounds[-1] == sum(i.n for i in insts)
  order = np.lexsort((rows[:, 1], rows[:, 0]))
  blocks = batch.split(rows[order], bounds)
  for inst, blk in zip(insts, blocks):
    want = inst.edges[np.lexsort((inst.edges[:, 1], inst.edges[:, 0]))]
    np.testing.assert_array_equal(blk, want)


def test_trim_deltas():
  assert batch.trim_deltas(np.array([5, 3, 0, 0, 0])) == [5, 3, 0]
  assert batch.trim_deltas(np.array([0, 0, 0])) == [0]
  assert batch.trim_deltas(np.array([4, 0])) == [4, 0]


def test_packed_fixpoint_is_disjoint_union():
  """What the batched path relies on, checked with the oracle on the CPU."""
  for kind, oname, rin, rout in (("tc", "tc", "ArcInput", "Path"), ("sg", "sg", "Arc", "Sg")):
    insts = [i for i in _instances(2, 8) if i.kind == kind]
    rows, bounds = batch.pack(insts)
    packed = oracle.run(oname, {rin: rows})
    blocks = batch.split(packed.relations[rout], bounds)
    alone = [oracle.run(oname, {rin: i.edges}) for i in insts]
    for blk, o in zip(blocks, alone):
      np.testing.assert_array_equal(blk, o.relations[rout])
    # packed deltas = per-iteration sums of the instances' own sequences
    width = len(packed.deltas[rout])
    mat = np.zeros((width, len(insts)), np.int64)
    for j, o in enumerate(alone):
      d = o.deltas[rout]
      assert len(d) <= width
      mat[:len(d), j] = d
    assert mat.sum(1).tolist() == packed.deltas[rout]
    for j, o in enumerate(alone):
      assert batch.trim_deltas(mat[:, j]) == o.deltas[rout]


def test_admission_caps_and_budget():
  insts = _instances(3, 6)
  for inst in insts:
    inst.est = 0.0
  ok, rej = batch.admit(insts, cap_tuples=1e12, budget_bytes=1e15)
  assert ok == list(range(6)) and rej == []
  big = max(range(6), key=lambda i: insts[i].est)
  ok, rej = batch.admit(insts, cap_tuples=insts[big].est - 1, budget_bytes=1e15)
  assert big not in ok and any(i == big for i, _ in rej)
  ok, rej = batch.admit(insts, cap_tuples=1e12, budget_bytes=1.0)
  assert ok == [] and len(rej) == 6


def test_estimates_track_true_sizes():
  insts = _instances(4, 8)
  rng = np.random.default_rng(0)
  for inst in insts:
    est = batch.estimate_output(inst, rng, samples=inst.n)   # exhaustive BFS: exact for TC
    oname, rin, rout = ("tc", "ArcInput", "Path") if inst.kind == "tc" else ("sg", "Arc", "Sg")
    true = oracle.run(oname, {rin: inst.edges}).relations[rout].shape[0]
    if inst.kind == "tc":
      assert est == true
    else:
      assert true <= 4 * est + 10


def test_packed_deltas_with_instances_converged_at_iteration_0():
  """Instances whose iteration-0 DELTA is empty record [0] alone; the
  packed run keeps iterating for the others."""
  e0 = np.zeros((0, 2), np.uint32)
  insts = [batch.Instance("tc", 4, e0),
           batch.Instance("tc", 3, np.array([[0, 1]], np.uint32)),
           batch.Instance("tc", 6, ds.tc_edges(3, 6, 12))]
  rows, bounds = batch.pack(insts)
  packed = oracle.run("tc", {"ArcInput": rows})
  width = len(packed.deltas["Path"])
  for j, inst in enumerate(insts):
    alone = oracle.run("tc", {"ArcInput": inst.edges}).deltas["Path"]
    col = np.zeros(width, np.int64)
    lo, hi = int(bounds[j]), int(bounds[j + 1])
    # per-iteration counts of instance j rebuilt from the alone sequence
    col[:len(alone)] = alone
    assert batch.trim_deltas(col) == alone
    assert alone[-1] == 0 and 0 not in alone[:-1]
    assert hi - lo == inst.n
