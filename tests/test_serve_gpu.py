"""Serving on the GPU: a mixed request stream served with continuous
batching (three kinds overlapping, reach requests sharing one fixpoint over
the resident graphs) gives every request the same result as running it
alone through the plain path, and the oracle's."""

import copy

import numpy as np
import pytest

import oracle
from srdatalog_hip import serve

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def graphs():
  return serve.make_graphs(11, scales=(9, 10, 11, 12), edge_factor=8)


def _oracle_result(r, g):
  if r.kind == "reach":
    o = oracle.run("reach", {"Edge": g.edges, "Src": np.array([[r.src]], np.uint32)})
    return o.relations["Reach"][:, 0]
  if r.kind == "tri":
    return oracle.run("triangle_count", {"KnowsInput": r.sub}).count if r.sub.shape[0] else 0
  return oracle.run("tc", {"ArcInput": r.sub}).relations["Path"]


@pytest.mark.parametrize("iterative", [True, False], ids=["tc-iterative", "tc-batched"])
def test_served_equals_alone_and_oracle(graphs, iterative):
  reqs = serve.make_requests(4, graphs, rate=300.0, seconds=0.4, tri_log2=(6, 10),
                             tc_range=(50, 400))
  assert len(reqs) > 40
  plain = copy.deepcopy(reqs)
  srv = serve.Server(graphs, max_batch=64, iterative=iterative)
  try:
    srv.run(reqs)
  finally:
    srv.close()
  assert sum(srv.batches.values()) < len(reqs)      # requests did share batches
  serve.run_plain(plain, graphs)
  for a, b in zip(reqs, plain):
    assert not a.rejected
    assert a.t_done >= a.t and b.t_done >= b.t
    assert serve.same_result(a, b), (a.rid, a.kind)
    want = _oracle_result(a, graphs[a.graph])
    if a.kind == "tri":
      assert a.result == want
    else:
      np.testing.assert_array_equal(a.result, want)


def test_served_rejects_over_cap(graphs):
  reqs = serve.make_requests(5, graphs, rate=200.0, seconds=0.2, tri_log2=(6, 9),
                             tc_range=(50, 200))
  srv = serve.Server(graphs, cap_tuples=100.0, max_batch=16)
  try:
    srv.run(reqs)
  finally:
    srv.close()
  for r in reqs:
    if r.est > 100.0:
      assert r.rejected and r.t_done < 0
    else:
      assert not r.rejected and r.t_done >= r.t


def test_iterative_tc_joins_running_fixpoint(graphs):
  """TC requests absorbed into the running fixpoint between its iterations
  -- the second and third after its first iteration, the last two into an
  engine whose instances had all finished -- each end with the oracle's
  closure of their own subgraph; finished ones are read out and erased."""
  reqs = [r for r in serve.make_requests(9, graphs, rate=400.0, seconds=1.0, tc_range=(80, 600))
          if r.kind == "tc"][:5]
  assert len(reqs) == 5
  serve.prepare(reqs, graphs)
  runner = serve.IterativeTCRunner()
  try:
    runner.admit(reqs[:1])
    runner.step()
    runner.admit(reqs[1:3])
    assert len(runner.active) >= 2                   # joined a fixpoint still running
    finished = []
    while runner.active:
      finished += runner.step()
    runner.admit(reqs[3:])
    while runner.active:
      finished += runner.step()
  finally:
    runner.close()
  assert sorted(r.rid for r in finished) == sorted(r.rid for r in reqs)
  for r in reqs:
    np.testing.assert_array_equal(r.result, _oracle_result(r, graphs[r.graph]))
