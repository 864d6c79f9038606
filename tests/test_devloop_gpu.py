"""GPU parity of the device-resident fixpoint iteration (recursive steps
whose iterations are sized on the device and replayed from a HIP graph)
against the oracle and against the host-driven loop: final tuple sets and
every per-iteration delta size, bit-exact.

SRDATALOG_DEVLOOP=0 forces the host-driven loop, SRDATALOG_DEVLOOP_GRAPH=0
enqueues the device-sized iteration without capturing it,
SRDATALOG_DEVLOOP_CAPS=N starts every capacity at N (tiny values make the
loop abandon and re-run iterations while it grows its buffers), and
SRDATALOG_DEVLOOP_BATCH sets how many iterations are launched between two
reads of the control block, and SRDATALOG_DEVLOOP_MAX_BYTES and
SRDATALOG_DEVLOOP_MAX_ROWS bound the loop's buffers and its FULL / NEW rows
(beyond them, the loop hands its state to the host loop)."""

import numpy as np
import pytest

import oracle
from srdatalog_hip import datasets as ds

pytestmark = pytest.mark.gpu

ENV_KEYS = ("SRDATALOG_DEVLOOP", "SRDATALOG_DEVLOOP_GRAPH", "SRDATALOG_DEVLOOP_CAPS",
            "SRDATALOG_DEVLOOP_BATCH", "SRDATALOG_DEVLOOP_MAX_BYTES", "SRDATALOG_DEVLOOP_MAX_ROWS",
            "SRDATALOG_JOIN_MODE", "SRDATALOG_GRAIN")

VARIANTS = [({}, "default"),
            ({"SRDATALOG_DEVLOOP_GRAPH": "0"}, "eager"),
            ({"SRDATALOG_DEVLOOP_CAPS": "64"}, "tiny-caps"),
            ({"SRDATALOG_DEVLOOP_CAPS": "200", "SRDATALOG_DEVLOOP_GRAPH": "0"}, "tiny-caps-eager"),
            ({"SRDATALOG_DEVLOOP_BATCH": "1"}, "batch-1"),
            ({"SRDATALOG_DEVLOOP_BATCH": "7", "SRDATALOG_GRAIN": "64"}, "batch-7-grain-64"),
            ({"SRDATALOG_JOIN_MODE": "twopass"}, "twopass"),
            ({"SRDATALOG_DEVLOOP_CAPS": "64", "SRDATALOG_DEVLOOP_MAX_BYTES": "3000000"},
             "budget-handback"),
            ({"SRDATALOG_DEVLOOP_CAPS": "64", "SRDATALOG_DEVLOOP_MAX_ROWS": "40000"},
             "rows-handback"),
            ({"SRDATALOG_DEVLOOP": "0"}, "host-loop")]


def cases():
  return [
    ("tc", "tc", {"ArcInput": ds.tc_edges(0, 1000, 5000)}, "Path", 2),
    ("tc", "tc", {"ArcInput": ds.tc_edges(5, 3000, 6000)}, "Path", 2),
    ("sg", "sg", {"Arc": ds.sg_forest_edges(np.random.default_rng(3), 4000)}, "Sg", 2),
    ("reach", "reach", {"Edge": ds.tc_edges(1, 5000, 12000),
                        "Src": np.array([[17]], dtype=np.uint32)}, "Reach", 1),
    ("tc", "tc", {"ArcInput": np.array([[1, 2], [1, 2], [2, 3], [3, 3], [3, 1], [5, 6]],
                                       dtype=np.uint32)}, "Path", 2),
  ]


def run(prog, inputs, max_iters=0):
  from srdatalog_hip import programs
  lib = programs.load(prog)
  assert lib.init() == 0
  for rel, rows in inputs.items():
    lib.load_rows(rel, rows)
  assert lib.run(max_iters) == 0
  return lib


@pytest.mark.parametrize("env,label", VARIANTS, ids=[v[1] for v in VARIANTS])
def test_devloop_parity(env, label, monkeypatch):
  for k in ENV_KEYS:
    monkeypatch.delenv(k, raising=False)
  for k, v in env.items():
    monkeypatch.setenv(k, v)
  for prog, oname, inputs, rel, arity in cases():
    print(f"devloop {label}: {prog} {[x.shape for x in inputs.values()]}", flush=True)
    orc = oracle.run(oname, inputs)
    lib = run(prog, inputs)
    try:
      got = lib.export(rel, arity)
      want = orc.relations[rel]
      assert got.shape == want.shape, (prog, label, got.shape, want.shape)
      np.testing.assert_array_equal(got, want)
      assert lib.delta_sizes(rel) == orc.deltas[rel], (prog, label)
      assert lib.size(rel) == want.shape[0]
    finally:
      lib.shutdown()


@pytest.mark.parametrize("max_iters", [1, 3])
def test_devloop_max_iters(max_iters, monkeypatch):
  for k in ENV_KEYS:
    monkeypatch.delenv(k, raising=False)
  e = ds.tc_edges(0, 1000, 5000)
  orc = oracle.run("tc", {"ArcInput": e}, max_iters=max_iters)
  lib = run("tc", {"ArcInput": e}, max_iters=max_iters)
  try:
    np.testing.assert_array_equal(lib.export("Path", 2), orc.relations["Path"])
    assert lib.delta_sizes("Path") == orc.deltas["Path"]
  finally:
    lib.shutdown()


def test_devloop_repeat_runs(monkeypatch):
  """Runs of one session reuse the resident inputs; every run starts over."""
  for k in ENV_KEYS:
    monkeypatch.delenv(k, raising=False)
  e = ds.tc_edges(2, 1000, 5000)
  orc = oracle.run("tc", {"ArcInput": e})
  lib = run("tc", {"ArcInput": e})
  try:
    for _ in range(3):
      np.testing.assert_array_equal(lib.export("Path", 2), orc.relations["Path"])
      assert lib.delta_sizes("Path") == orc.deltas["Path"]
      assert lib.run(0) == 0
  finally:
    lib.shutdown()
