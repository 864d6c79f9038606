"""Batched execution on the GPU: every instance's result and per-iteration
delta sizes from the packed run equal its plain (one instance per session)
run and the oracle, bit-exact."""

import numpy as np
import pytest

import oracle
from srdatalog_hip import batch
from srdatalog_hip import datasets as ds

pytestmark = pytest.mark.gpu


def _batch(seed, kind, count, n_lo=30, n_hi=3000):
  rng = np.random.default_rng(seed)
  out = []
  for i in range(count):
    n = int(np.rint(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi)))))
    if kind == "tc":
      deg = rng.uniform(1.5, 4.0) if n < 400 else rng.uniform(0.6, 1.0)
      m = min(int(round(n * deg)), n * (n - 1))
      out.append(batch.Instance("tc", n, ds.tc_edges(seed * 1000 + i, n, m)))
    else:
      out.append(batch.Instance("sg", n, ds.sg_forest_edges(np.random.default_rng(seed + 7 * i), n)))
  return out


@pytest.mark.parametrize("kind,seed", [("tc", 0), ("tc", 1), ("sg", 0), ("sg", 1)])
def test_batched_equals_plain_and_oracle(kind, seed):
  insts = _batch(seed, kind, 12)
  got = batch.run_batched(kind, insts)
  plain = batch.run_plain(kind, insts)
  prog, rin, rout, _ = batch.KINDS[kind]
  for inst, b, p in zip(insts, got, plain):
    o = oracle.run(prog, {rin: inst.edges})
    np.testing.assert_array_equal(b.rows, o.relations[rout])
    np.testing.assert_array_equal(p.rows, o.relations[rout])
    assert b.deltas == p.deltas == o.deltas[rout], (inst.n, b.deltas, p.deltas)


def test_batch_with_empty_and_tiny_instances():
  e = np.zeros((0, 2), np.uint32)
  insts = [batch.Instance("tc", 5, e),
           batch.Instance("tc", 3, np.array([[0, 1], [1, 2]], np.uint32)),
           batch.Instance("tc", 4, np.array([[0, 1], [1, 0], [2, 3]], np.uint32)),
           batch.Instance("tc", 1, e)]
  got = batch.run_batched("tc", insts)
  plain = batch.run_plain("tc", insts)
  for inst, b, p in zip(insts, got, plain):
    np.testing.assert_array_equal(b.rows, p.rows)
    assert b.deltas == p.deltas
    if inst.edges.shape[0]:
      o = oracle.run("tc", {"ArcInput": inst.edges})
      np.testing.assert_array_equal(b.rows, o.relations["Path"])


def test_autotune_picks_a_candidate():
  insts = _batch(3, "tc", 6)
  g, times = batch.autotune_grain("tc", insts, candidates=(256, 4096))
  assert g in (256, 4096) and set(times) == {256, 4096}
  got = batch.run_batched("tc", insts, grain=g)
  for inst, b in zip(insts, got):
    o = oracle.run("tc", {"ArcInput": inst.edges})
    np.testing.assert_array_equal(b.rows, o.relations["Path"])
