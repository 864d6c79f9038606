"""End-to-end parity of compiled programs (through the srdatalog_* C ABI)
against the CPU oracle: final sorted tuple sets AND per-iteration delta
sizes, bit-exact."""

import numpy as np
import pytest

import oracle
from srdatalog_hip import datasets as ds

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def progs():
  from srdatalog_hip import programs
  return programs


def run_prog(progs, name, inputs, max_iters=0):
  lib = progs.load(name)
  assert lib.init() == 0
  for rel, rows in inputs.items():
    lib.load_rows(rel, rows)
  rc = lib.run(max_iters)
  assert rc == 0, f"srdatalog_run returned {rc}"
  return lib


def check_rel(lib, orc, rel, arity):
  got = lib.export(rel, arity)
  want = orc.relations[rel]
  assert got.shape == want.shape, f"{rel}: {got.shape} vs {want.shape}"
  np.testing.assert_array_equal(got, want)
  assert lib.size(rel) == want.shape[0]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_tc_config(progs, seed):
  e = ds.tc_edges(seed, 1000, 5000)
  orc = oracle.run("tc", {"ArcInput": e})
  lib = run_prog(progs, "tc", {"ArcInput": e})
  try:
    check_rel(lib, orc, "Path", 2)
    check_rel(lib, orc, "Edge", 2)
    assert lib.delta_sizes("Path") == orc.deltas["Path"]
  finally:
    lib.shutdown()


@pytest.mark.parametrize("n,m,seed", [(10, 15, 3), (60, 150, 4), (400, 600, 5), (3000, 6000, 6)])
def test_tc_small(progs, n, m, seed):
  e = ds.tc_edges(seed, n, m)
  orc = oracle.run("tc", {"ArcInput": e})
  lib = run_prog(progs, "tc", {"ArcInput": e})
  try:
    check_rel(lib, orc, "Path", 2)
    assert lib.delta_sizes("Path") == orc.deltas["Path"]
  finally:
    lib.shutdown()


def test_tc_duplicates_and_loops(progs):
  # duplicate input rows and self-loops are legal input
  e = np.array([[1, 2], [1, 2], [2, 3], [3, 3], [3, 1], [5, 6]], dtype=np.uint32)
  orc = oracle.run("tc", {"ArcInput": e})
  lib = run_prog(progs, "tc", {"ArcInput": e})
  try:
    check_rel(lib, orc, "Path", 2)
    assert lib.delta_sizes("Path") == orc.deltas["Path"]
  finally:
    lib.shutdown()


def test_tc_max_iters(progs):
  e = ds.tc_edges(0, 1000, 5000)
  orc = oracle.run("tc", {"ArcInput": e}, max_iters=3)
  lib = run_prog(progs, "tc", {"ArcInput": e}, max_iters=3)
  try:
    check_rel(lib, orc, "Path", 2)
    assert lib.delta_sizes("Path") == orc.deltas["Path"]
  finally:
    lib.shutdown()


def test_tc_empty(progs):
  e = np.zeros((0, 2), dtype=np.uint32)
  lib = run_prog(progs, "tc", {"ArcInput": e})
  try:
    assert lib.size("Path") == 0
    assert lib.export("Path", 2).shape == (0, 2)
  finally:
    lib.shutdown()


def test_tc_large_ids(progs):
  rng = np.random.default_rng(9)
  ids = rng.choice(np.arange(2**32 - 5000, 2**32, dtype=np.uint64), size=300, replace=False)
  src = rng.choice(ids, 500)
  dst = rng.choice(ids, 500)
  e = np.stack([src, dst], 1).astype(np.uint32)
  orc = oracle.run("tc", {"ArcInput": e})
  lib = run_prog(progs, "tc", {"ArcInput": e})
  try:
    check_rel(lib, orc, "Path", 2)
    assert lib.delta_sizes("Path") == orc.deltas["Path"]
  finally:
    lib.shutdown()


@pytest.mark.parametrize("n,seed", [(50, 0), (2000, 1), (20000, 2)])
def test_sg(progs, n, seed):
  e = ds.sg_forest_edges(np.random.default_rng(seed), n)
  orc = oracle.run("sg", {"Arc": e})
  lib = run_prog(progs, "sg", {"Arc": e})
  try:
    check_rel(lib, orc, "Sg", 2)
    assert lib.delta_sizes("Sg") == orc.deltas["Sg"]
  finally:
    lib.shutdown()


@pytest.mark.parametrize("scale_div,seed", [(1000, 0), (100, 1), (100, 2), (30, 0)])
def test_andersen(progs, scale_div, seed):
  f = scale_div
  a = ds.andersen_inputs(seed, n_vars=200_000 // f, n_objs=50_000 // f,
                         sizes=(100_000 // f, 400_000 // f, 100_000 // f, 100_000 // f))
  orc = oracle.run("andersen", a)
  lib = run_prog(progs, "andersen", a)
  try:
    check_rel(lib, orc, "PointsTo", 2)
    assert lib.delta_sizes("PointsTo") == orc.deltas["PointsTo"]
  finally:
    lib.shutdown()


@pytest.mark.parametrize("scale,m,seed", [(8, 2000, 0), (12, 60_000, 1), (16, 1_000_000, 2)])
def test_triangle_listing(progs, scale, m, seed):
  e = ds.triangle_edges(seed, scale, m)
  orc = oracle.run("triangle", {"KnowsInput": e})
  lib = run_prog(progs, "triangle", {"KnowsInput": e})
  try:
    check_rel(lib, orc, "Triangle", 3)
    check_rel(lib, orc, "Knows", 2)
    assert lib.delta_sizes("Triangle") == orc.deltas["Triangle"]
  finally:
    lib.shutdown()


@pytest.mark.parametrize("scale,m,seed", [(12, 60_000, 1), (16, 1_000_000, 2)])
def test_triangle_count(progs, scale, m, seed):
  e = ds.triangle_edges(seed, scale, m)
  orc = oracle.run("triangle_count", {"KnowsInput": e})
  lib = run_prog(progs, "triangle_count", {"KnowsInput": e})
  try:
    assert lib.size("Triangle") == orc.count
  finally:
    lib.shutdown()


@pytest.mark.parametrize("seed", [0, 1])
def test_reach(progs, seed):
  e = ds.tc_edges(seed, 5000, 12000)
  src = np.array([[seed * 17]], dtype=np.uint32)
  orc = oracle.run("reach", {"Edge": e, "Src": src})
  lib = run_prog(progs, "reach", {"Edge": e, "Src": src})
  try:
    check_rel(lib, orc, "Reach", 1)
    assert lib.delta_sizes("Reach") == orc.deltas["Reach"]
  finally:
    lib.shutdown()


def test_repeat_runs_identical(progs):
  e = ds.tc_edges(1, 1000, 5000)
  lib = run_prog(progs, "tc", {"ArcInput": e})
  try:
    first = lib.export("Path", 2)
    d1 = lib.delta_sizes("Path")
    assert lib.run(0) == 0
    np.testing.assert_array_equal(lib.export("Path", 2), first)
    assert lib.delta_sizes("Path") == d1
  finally:
    lib.shutdown()


def test_csv_load_all(progs, tmp_path):
  e = ds.tc_edges(2, 300, 700)
  ds.write_csv(str(tmp_path / "Arc.csv"), e, delimiter="\t")
  orc = oracle.run("tc", {"ArcInput": e})
  lib = progs.load("tc")
  assert lib.init() == 0
  try:
    assert lib.load_all(str(tmp_path)) == 0
    assert lib.run(0) == 0
    check_rel(lib, orc, "Path", 2)
  finally:
    lib.shutdown()


# ------------------------------------------------------------------ modes ----
# The engine picks the direct / fused / two-pass join mode per rule; each is
# forced here (SRDATALOG_JOIN_MODE is read when the engine is created) and
# must give the same relation and deltas as the oracle.  SRDATALOG_STAGE_CAP
# shrinks the fused staging buffer so the overflow -> write-pass fallback runs;
# SRDATALOG_NO_SEGMENTS keeps every FULL as one index (no segmented FULL);
# SRDATALOG_NEW_CHUNK_ROWS writes join outputs in small deduplicated chunks;
# SRDATALOG_WEIGH_MIN=1 cuts cost-balanced slices for every rule with two or
# more nested levels (SRDATALOG_WEIGH_TARGET: cost per slice), and
# SRDATALOG_NO_WEIGH=1 keeps uniform slices everywhere; SRDATALOG_BITMAP_BYTES
# bounds the heavy-key bitmaps (0: none; a small budget keeps only the
# longest keys' bitmaps, so bitmap probes and searches mix in one wave);
# Rules whose destination's key space is small take their NEW as a bitmap
# the joins mark (default; SRDATALOG_DENSE_MAX_BITS=0: never), so every
# other mode turns that off to keep its own row-writing path under test;
# dense-host runs the default with the host-driven loop (TC and SG mark
# there too), dense-groups-in-order with the parallel group's members
# marking one after the other, dense-chunked-marks with every marking
# traversal split into launches of 3 slices.
# SRDATALOG_WINDOW_MIN=1 narrows the driving range of every nested item to
# the other sources' value window (0: of none);
# SRDATALOG_ROOT_POS=0 looks nested-level keys up by searching the root
# values instead of through the dense root map.  Every variant but the
# default ones runs recursive steps in the host-driven loop
# (SRDATALOG_DEVLOOP=0); devloop-small-caps runs them device-resident from
# capacities of 500 rows, growing as it goes.  The members of a parallel
# group (Andersen's five rules) run concurrently on their own streams;
# group-in-order runs them one after the other, group-chunked-overflow
# sends some members down the sequential chunked and overflow paths.
MODES = [({"SRDATALOG_JOIN_MODE": "twopass"}, "twopass"),
         ({"SRDATALOG_JOIN_MODE": "fused"}, "fused"),
         ({"SRDATALOG_JOIN_MODE": "fused", "SRDATALOG_STAGE_CAP": "3000"}, "fused-overflow"),
         ({"SRDATALOG_NO_SEGMENTS": "1"}, "one-segment-full"),
         ({"SRDATALOG_NEW_CHUNK_ROWS": "20000", "SRDATALOG_GRAIN": "256"}, "chunked-new"),
         ({"SRDATALOG_NEW_CHUNK_ROWS": "20000", "SRDATALOG_JOIN_MODE": "twopass"},
          "chunked-new-twopass"),
         ({"SRDATALOG_WEIGH_MIN": "1", "SRDATALOG_WEIGH_TARGET": "1"}, "weighted-1"),
         ({"SRDATALOG_WEIGH_MIN": "1", "SRDATALOG_WEIGH_TARGET": "500",
           "SRDATALOG_JOIN_MODE": "twopass"}, "weighted-twopass"),
         ({"SRDATALOG_WEIGH_MIN": "1", "SRDATALOG_NEW_CHUNK_ROWS": "20000"}, "weighted-chunked"),
         ({"SRDATALOG_NO_WEIGH": "1"}, "uniform-slices"),
         ({"SRDATALOG_BITMAP_BYTES": "0"}, "no-bitmaps"),
         ({"SRDATALOG_BITMAP_BYTES": "20000"}, "few-bitmaps"),
         ({"SRDATALOG_ROOT_POS": "0"}, "root-search"),
         ({"SRDATALOG_WINDOW_MIN": "1"}, "window-all"),
         ({"SRDATALOG_WINDOW_MIN": "0"}, "no-window"),
         ({"SRDATALOG_DEVLOOP": "1", "SRDATALOG_DEVLOOP_CAPS": "500"}, "devloop-small-caps"),
         ({"SRDATALOG_GROUP_STREAMS": "0"}, "group-in-order"),
         ({"SRDATALOG_NEW_CHUNK_ROWS": "3000", "SRDATALOG_JOIN_MODE": "fused",
           "SRDATALOG_STAGE_CAP": "5000"}, "group-chunked-overflow"),
         ({"SRDATALOG_DEVLOOP": "0"}, "dense-host"),
         ({"SRDATALOG_DEVLOOP": "0", "SRDATALOG_GROUP_STREAMS": "0"}, "dense-groups-in-order"),
         ({"SRDATALOG_DEVLOOP": "0", "SRDATALOG_MARK_CHUNK_SLICES": "3"}, "dense-chunked-marks"),
         ({}, "default")]


@pytest.mark.parametrize("env,label", MODES, ids=[m[1] for m in MODES])
def test_modes_tc_sg_triangle_andersen(progs, env, label, monkeypatch):
  for k in ("SRDATALOG_JOIN_MODE", "SRDATALOG_STAGE_CAP", "SRDATALOG_NO_SEGMENTS",
            "SRDATALOG_NEW_CHUNK_ROWS", "SRDATALOG_GRAIN", "SRDATALOG_WEIGH_MIN",
            "SRDATALOG_WEIGH_TARGET", "SRDATALOG_NO_WEIGH", "SRDATALOG_BITMAP_BYTES",
            "SRDATALOG_ROOT_POS", "SRDATALOG_WINDOW_MIN", "SRDATALOG_DEVLOOP", "SRDATALOG_DEVLOOP_CAPS",
            "SRDATALOG_GROUP_STREAMS", "SRDATALOG_DENSE_MAX_BITS", "SRDATALOG_MARK_CHUNK_SLICES"):
    monkeypatch.delenv(k, raising=False)
  if not label.startswith(("dense", "default")):
    monkeypatch.setenv("SRDATALOG_DENSE_MAX_BITS", "0")
  # the join modes of recursive rules are those of the host-driven loop (the
  # device-resident one has its own variants)
  if label not in ("default",) and "SRDATALOG_DEVLOOP" not in env:
    monkeypatch.setenv("SRDATALOG_DEVLOOP", "0")
  for k, v in env.items():
    monkeypatch.setenv(k, v)
  cases = [
    ("tc", {"ArcInput": ds.tc_edges(7, 800, 3000)}, "Path", 2, "tc"),
    ("sg", {"Arc": ds.sg_forest_edges(np.random.default_rng(3), 4000)}, "Sg", 2, "sg"),
    ("triangle", {"KnowsInput": ds.triangle_edges(4, 12, 80_000)}, "Triangle", 3, "triangle"),
    ("andersen", ds.andersen_inputs(5, n_vars=2000, n_objs=500, sizes=(1000, 4000, 1000, 1000)),
     "PointsTo", 2, "andersen"),
  ]
  for prog, inputs, rel, arity, oname in cases:
    orc = oracle.run(oname, inputs)
    lib = run_prog(progs, prog, inputs)
    try:
      check_rel(lib, orc, rel, arity)
      assert lib.delta_sizes(rel) == orc.deltas[rel], (prog, label)
    finally:
      lib.shutdown()


# Key-major bitmaps for over a thousand keys (ranges of 16+ rows on an R-MAT
# scale-14 graph), so bitmap slots run past the few keys of the smaller
# graphs: every key long enough, and a budget that doubles the threshold
# once (bitmap tests and searches in one wave).
BITMAP_MANY = [({"SRDATALOG_BITMAP_MIN": "16"}, "keys-16"),
               ({"SRDATALOG_BITMAP_MIN": "16", "SRDATALOG_BITMAP_BYTES": "4000000"}, "keys-16-budget")]


@pytest.mark.parametrize("env,label", BITMAP_MANY, ids=[m[1] for m in BITMAP_MANY])
def test_triangle_many_bitmaps(progs, env, label, monkeypatch):
  for k in ("SRDATALOG_BITMAP_MIN", "SRDATALOG_BITMAP_BYTES"):
    monkeypatch.delenv(k, raising=False)
  for k, v in env.items():
    monkeypatch.setenv(k, v)
  inputs = {"KnowsInput": ds.triangle_edges(6, 14, 300_000)}
  orc = oracle.run("triangle", inputs)
  lib = run_prog(progs, "triangle", inputs)
  try:
    check_rel(lib, orc, "Triangle", 3)
  finally:
    lib.shutdown()


def test_triangle_wide_value_bitmaps(progs):
  """Bitmaps over column-1 values up to 2.5 M: a bitmap no longer fits the
  LDS of the block that builds it (values below 1,179,648 do), so the keys'
  bits are set by global atomics instead.  The scale-12 graph relabelled
  x -> 600 x keeps its edges oriented and its 11 keys of 512+ rows."""
  e = ds.triangle_edges(4, 12, 80_000).astype(np.uint64) * 600
  inputs = {"KnowsInput": e.astype(np.uint32)}
  assert int(inputs["KnowsInput"].max()) > 1_179_648
  orc = oracle.run("triangle", inputs)
  lib = run_prog(progs, "triangle", inputs)
  try:
    check_rel(lib, orc, "Triangle", 3)
  finally:
    lib.shutdown()


# --------------------------------------------------------------- negation ---
def _neg2hop_inputs(seed, n, m, tags, mi):
  rng = np.random.default_rng(seed)
  k = ds.tc_edges(seed, n, m)
  hi = np.unique(np.stack([rng.integers(0, n, mi), rng.integers(0, tags, mi)], 1)
                 .astype(np.uint32), axis=0)
  return {"KnowsInput": k, "HasInterestInput": hi}


@pytest.mark.parametrize("n,m,tags,mi,seed", [(50, 200, 8, 100, 0), (2000, 20000, 50, 4000, 1),
                                              (20000, 150000, 200, 30000, 2)])
def test_neg2hop_listing_and_count(progs, n, m, tags, mi, seed):
  inp = _neg2hop_inputs(seed, n, m, tags, mi)
  orc = oracle.run("neg2hop", inp)
  lib = run_prog(progs, "neg2hop", inp)
  try:
    check_rel(lib, orc, "Path", 4)
    check_rel(lib, orc, "Knows", 2)
  finally:
    lib.shutdown()
  lib = run_prog(progs, "neg2hop_count", inp)
  try:
    assert lib.size("Path") == orc.count
  finally:
    lib.shutdown()


def _q7_inputs(seed, n_msg, n_tag, n_person, sizes):
  rng = np.random.default_rng(seed)

  def pairs(n, a, b):
    return np.unique(np.stack([rng.integers(0, a, n), rng.integers(0, b, n)], 1)
                     .astype(np.uint32), axis=0)
  return {"HasTag": pairs(sizes[0], n_msg, n_tag), "HasCreator": pairs(sizes[1], n_msg, n_person),
          "ReplyOf": pairs(sizes[2], 10 * n_msg, n_msg), "Likes": pairs(sizes[3], n_person, n_msg)}


@pytest.mark.parametrize("n_msg,sizes,seed", [(40, (80, 50, 60, 70), 0),
                                              (5000, (20000, 6000, 8000, 9000), 1),
                                              (200000, (600000, 200000, 150000, 150000), 2)])
def test_q7_counts(progs, n_msg, sizes, seed):
  inp = _q7_inputs(seed, n_msg, 20, 3000, sizes)
  orc = oracle.run("q7", inp)
  lib = run_prog(progs, "q7", inp)
  try:
    for case, want in orc.rule_counts.items():
      assert lib.size(case) == want, case
  finally:
    lib.shutdown()


def _q3_inputs(seed, n, m, cities, countries, extra_loc):
  rng = np.random.default_rng(seed)
  k = ds.tc_edges(seed, n, m)
  loc = np.stack([np.arange(n), rng.integers(0, cities, n)], 1)
  more = np.stack([rng.integers(0, n, extra_loc), rng.integers(0, cities, extra_loc)], 1)
  loc = np.unique(np.concatenate([loc, more]).astype(np.uint32), axis=0)
  part = np.stack([np.arange(cities), rng.integers(0, countries, cities)], 1)
  part = np.unique(part.astype(np.uint32), axis=0)
  return {"KnowsInput": k, "IsLocatedInInput": loc, "IsPartOfInput": part}


@pytest.mark.parametrize("n,m,cities,countries,extra,seed", [(40, 260, 10, 3, 6, 0),
                                                             (3000, 60000, 200, 8, 300, 1),
                                                             (50000, 800000, 2000, 20, 5000, 2)])
def test_q3_six_level_join(progs, n, m, cities, countries, extra, seed):
  inp = _q3_inputs(seed, n, m, cities, countries, extra)
  orc = oracle.run("q3", inp)
  lib = run_prog(progs, "q3", inp)
  try:
    check_rel(lib, orc, "Triangle", 3)
  finally:
    lib.shutdown()
