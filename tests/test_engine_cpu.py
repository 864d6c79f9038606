"""The runtime's MIR executor (csrc/srd_engine.hip) compiled for the host
against CPU primitives (tests/cpu_sim), driving the transitive-closure plan:
checks the engine's host logic -- MIR parsing, FULL/DELTA/NEW bookkeeping,
iteration-0 redirect, the convergence test and max_iters -- against the
oracle without a GPU."""

import ctypes
import os

import numpy as np
import pytest

import oracle
from conftest import build_cpu_sim
from srdatalog_hip import codegen, compiler, mir, programs
from srdatalog_hip import datasets as ds

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpu_sim")
LIB = os.path.join(HERE, "libengine_cpu.so")
_u32p = ctypes.POINTER(ctypes.c_uint32)


@pytest.fixture(scope="module")
def eng():
  lib = ctypes.CDLL(build_cpu_sim("libengine_cpu.so"))
  lib.cpu_tc_engine.restype = ctypes.c_void_p
  lib.cpu_tc_engine.argtypes = [ctypes.c_char_p]
  lib.srd_engine_ok.argtypes = [ctypes.c_void_p]
  lib.srd_engine_load_rows.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _u32p, ctypes.c_uint64,
                                       ctypes.c_uint32]
  lib.srd_engine_run.argtypes = [ctypes.c_void_p, ctypes.c_ulonglong]
  lib.srd_engine_size.restype = ctypes.c_ulonglong
  lib.srd_engine_size.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
  lib.srd_engine_export.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _u32p, ctypes.c_uint64,
                                    ctypes.POINTER(ctypes.c_uint64)]
  lib.srd_engine_delta_sizes.argtypes = [ctypes.c_void_p, ctypes.c_char_p,
                                         ctypes.POINTER(ctypes.c_ulonglong), ctypes.c_ulonglong]
  lib.srd_engine_destroy.argtypes = [ctypes.c_void_p]
  lib.srd_engine_set_quiet.argtypes = [ctypes.c_void_p, ctypes.c_int]
  return lib


@pytest.fixture(autouse=True)
def host_loop_by_default(monkeypatch):
  """The tests below drive the host-driven loop unless they ask for the
  device-resident one; the host-only build never captures graphs."""
  monkeypatch.setenv("SRDATALOG_DEVLOOP", "0")
  monkeypatch.setenv("SRDATALOG_DEVLOOP_GRAPH", "0")
  for k in ("SRDATALOG_DEVLOOP_CAPS", "SRDATALOG_DEVLOOP_BATCH", "SRDATALOG_DEVLOOP_MAX_BYTES",
            "SRDATALOG_DEVLOOP_MAX_ROWS"):
    monkeypatch.delenv(k, raising=False)


def tc_text():
  prog = programs.tc_program()
  return codegen.program_text(compiler.relation_decls(prog), compiler.compile_to_mir(prog))


def run_tc(lib, edges, max_iters=0):
  e = lib.cpu_tc_engine(tc_text().encode())
  assert lib.srd_engine_ok(e) == 1
  lib.srd_engine_set_quiet(e, 1)
  edges = np.ascontiguousarray(edges, dtype=np.uint32)
  lib.srd_engine_load_rows(e, b"ArcInput", edges.ctypes.data_as(_u32p), edges.shape[0], 2)
  assert lib.srd_engine_run(e, max_iters) == 0
  n = ctypes.c_uint64()
  lib.srd_engine_export(e, b"Path", None, 0, ctypes.byref(n))
  out = np.empty((max(n.value, 1), 2), np.uint32)
  lib.srd_engine_export(e, b"Path", out.ctypes.data_as(_u32p), n.value, ctypes.byref(n))
  cnt = lib.srd_engine_delta_sizes(e, b"Path", None, 0)
  buf = (ctypes.c_ulonglong * max(cnt, 1))()
  lib.srd_engine_delta_sizes(e, b"Path", buf, cnt)
  size = lib.srd_engine_size(e, b"Path")
  lib.srd_engine_destroy(e)
  return out[:n.value], [int(buf[i]) for i in range(cnt)], size


@pytest.mark.parametrize("mode", ["direct", "fused", "twopass"])
@pytest.mark.parametrize("n,m,seed", [(12, 20, 0), (60, 150, 1), (300, 700, 2)])
def test_engine_tc_matches_oracle(eng, n, m, seed, mode, monkeypatch):
  # the engine reads SRDATALOG_JOIN_MODE when it is created
  if mode != "direct":
    monkeypatch.setenv("SRDATALOG_JOIN_MODE", mode)
  else:
    monkeypatch.delenv("SRDATALOG_JOIN_MODE", raising=False)
  e = ds.tc_edges(seed, n, m)
  want = oracle.run("tc", {"ArcInput": e})
  path, deltas, size = run_tc(eng, e)
  np.testing.assert_array_equal(path, want.relations["Path"])
  assert deltas == want.deltas["Path"]
  assert size == want.size("Path")


def test_engine_tc_max_iters(eng):
  e = ds.tc_edges(3, 80, 200)
  want = oracle.run("tc", {"ArcInput": e}, max_iters=2)
  path, deltas, _ = run_tc(eng, e, max_iters=2)
  np.testing.assert_array_equal(path, want.relations["Path"])
  assert deltas == want.deltas["Path"]


def test_engine_tc_empty(eng):
  path, deltas, size = run_tc(eng, np.zeros((0, 2), np.uint32))
  assert path.shape[0] == 0 and size == 0


def test_engine_rejects_bad_plan(eng):
  e = eng.cpu_tc_engine(b"(relations (relation :name A :arity 2)) (program (step :recursive false "
                        b"(fixpoint-plan (bogus-op :schema A))))")
  assert eng.srd_engine_ok(e) == 0
  eng.srd_engine_destroy(e)


@pytest.mark.parametrize("segments", [True, False])
@pytest.mark.parametrize("n,extra,seed", [(300, 0, 0), (400, 40, 1), (150, 300, 2)])
def test_engine_segmented_full_long_tail(eng, n, extra, seed, segments, monkeypatch):
  """A chain 0->1->...->n-1 plus random edges: many iterations with small
  deltas, so FULL accumulates many segments and merge cascades."""
  monkeypatch.delenv("SRDATALOG_JOIN_MODE", raising=False)
  if segments:
    monkeypatch.delenv("SRDATALOG_NO_SEGMENTS", raising=False)
  else:
    monkeypatch.setenv("SRDATALOG_NO_SEGMENTS", "1")
  chain = np.stack([np.arange(n - 1), np.arange(1, n)], 1).astype(np.uint32)
  rng = np.random.default_rng(seed)
  rnd = rng.integers(0, n, size=(extra, 2)).astype(np.uint32)
  e = np.concatenate([chain, rnd])
  want = oracle.run("tc", {"ArcInput": e})
  path, deltas, size = run_tc(eng, e)
  np.testing.assert_array_equal(path, want.relations["Path"])
  assert deltas == want.deltas["Path"]
  assert size == want.relations["Path"].shape[0]


@pytest.mark.parametrize("chunk", [1, 37, 500])
@pytest.mark.parametrize("mode", ["direct", "twopass"])
def test_engine_chunked_new(eng, chunk, mode, monkeypatch):
  """Join outputs written in slice chunks, each deduplicated into a NEW
  part, give the same fixpoint and deltas as materialising them at once."""
  monkeypatch.setenv("SRDATALOG_NEW_CHUNK_ROWS", str(chunk))
  monkeypatch.setenv("SRDATALOG_GRAIN", "16")
  if mode == "twopass":
    monkeypatch.setenv("SRDATALOG_JOIN_MODE", "twopass")
  else:
    monkeypatch.delenv("SRDATALOG_JOIN_MODE", raising=False)
  e = ds.tc_edges(4, 120, 400)
  want = oracle.run("tc", {"ArcInput": e})
  path, deltas, size = run_tc(eng, e)
  np.testing.assert_array_equal(path, want.relations["Path"])
  assert deltas == want.deltas["Path"]


@pytest.mark.parametrize("target", [1, 50, 100000])
@pytest.mark.parametrize("mode", ["fused", "twopass", "chunked"])
def test_engine_weighted_slices(eng, target, mode, monkeypatch):
  """Cost-balanced slices (weigh launch, prefix of item costs, slice_bounds)
  give the same fixpoint and deltas as uniform slices; the CPU pipeline's
  costs are an uneven hash of the item, so slices of one item, empty slices
  and long runs of cheap items all occur."""
  monkeypatch.setenv("SRDATALOG_WEIGH_MIN", "1")
  monkeypatch.setenv("SRDATALOG_WEIGH_TARGET", str(target))
  monkeypatch.setenv("SRDATALOG_GRAIN", "16")
  monkeypatch.setenv("SRDATALOG_JOIN_MODE", "fused" if mode == "fused" else "twopass")
  if mode == "chunked":
    monkeypatch.setenv("SRDATALOG_NEW_CHUNK_ROWS", "300")
  e = ds.tc_edges(5, 150, 500)
  want = oracle.run("tc", {"ArcInput": e})
  path, deltas, size = run_tc(eng, e)
  np.testing.assert_array_equal(path, want.relations["Path"])
  assert deltas == want.deltas["Path"]


# ------------------------------------------------ device-resident loop ----
# The device-sized iteration (capacities, control block, bind / plan / out /
# check steps, growth after an overflow) driven by the host-only build: the
# steps and primitives run on the CPU in stream order, as the kernels would.
@pytest.mark.parametrize("caps,batch", [(0, 4), (8, 1), (33, 3), (200, 16), (64, 2)])
@pytest.mark.parametrize("mode", ["direct", "twopass"])
@pytest.mark.parametrize("n,m,seed", [(12, 20, 0), (60, 150, 1), (300, 700, 2)])
def test_devloop_tc_matches_oracle(eng, n, m, seed, mode, caps, batch, monkeypatch):
  monkeypatch.setenv("SRDATALOG_DEVLOOP", "1")
  monkeypatch.setenv("SRDATALOG_DEVLOOP_BATCH", str(batch))
  if caps:
    monkeypatch.setenv("SRDATALOG_DEVLOOP_CAPS", str(caps))
  if mode == "twopass":
    monkeypatch.setenv("SRDATALOG_JOIN_MODE", "twopass")
  else:
    monkeypatch.delenv("SRDATALOG_JOIN_MODE", raising=False)
  e = ds.tc_edges(seed, n, m)
  want = oracle.run("tc", {"ArcInput": e})
  path, deltas, size = run_tc(eng, e)
  np.testing.assert_array_equal(path, want.relations["Path"])
  assert deltas == want.deltas["Path"]
  assert size == want.size("Path")


@pytest.mark.parametrize("max_iters", [1, 2, 5])
@pytest.mark.parametrize("caps", [0, 16])
def test_devloop_tc_max_iters(eng, max_iters, caps, monkeypatch):
  monkeypatch.setenv("SRDATALOG_DEVLOOP", "1")
  if caps:
    monkeypatch.setenv("SRDATALOG_DEVLOOP_CAPS", str(caps))
  e = ds.tc_edges(3, 80, 200)
  want = oracle.run("tc", {"ArcInput": e}, max_iters=max_iters)
  path, deltas, _ = run_tc(eng, e, max_iters=max_iters)
  np.testing.assert_array_equal(path, want.relations["Path"])
  assert deltas == want.deltas["Path"]


def test_devloop_tc_empty_and_long_tail(eng, monkeypatch):
  monkeypatch.setenv("SRDATALOG_DEVLOOP", "1")
  monkeypatch.setenv("SRDATALOG_DEVLOOP_CAPS", "50")
  path, deltas, size = run_tc(eng, np.zeros((0, 2), np.uint32))
  assert path.shape[0] == 0 and size == 0
  n = 300   # a chain: 298 iterations, more than one iteration record of 50
  e = np.stack([np.arange(n - 1), np.arange(1, n)], 1).astype(np.uint32)
  want = oracle.run("tc", {"ArcInput": e})
  path, deltas, size = run_tc(eng, e)
  np.testing.assert_array_equal(path, want.relations["Path"])
  assert deltas == want.deltas["Path"]


# ------------------------------------------------------ parallel groups ----
def tc_group_text(copies: int) -> str:
  """The TC plan with its recursive rule run `copies` times as the members
  of one parallel group: every member writes the same rows into Path's NEW,
  which deduplicates them, so the fixpoint and deltas are TC's."""
  prog = programs.tc_program()
  mp = compiler.compile_to_mir(prog)
  for st in mp.steps:
    if st.recursive:
      ops = st.body.ops
      ops[0] = mir.ParallelGroup([ops[0]] * copies)
  return codegen.program_text(compiler.relation_decls(prog), mp)


def run_tc_text(lib, text, edges):
  e = lib.cpu_tc_engine(text.encode())
  assert lib.srd_engine_ok(e) == 1
  lib.srd_engine_set_quiet(e, 1)
  edges = np.ascontiguousarray(edges, dtype=np.uint32)
  lib.srd_engine_load_rows(e, b"ArcInput", edges.ctypes.data_as(_u32p), edges.shape[0], 2)
  assert lib.srd_engine_run(e, 0) == 0
  n = ctypes.c_uint64()
  lib.srd_engine_export(e, b"Path", None, 0, ctypes.byref(n))
  out = np.empty((max(n.value, 1), 2), np.uint32)
  lib.srd_engine_export(e, b"Path", out.ctypes.data_as(_u32p), n.value, ctypes.byref(n))
  cnt = lib.srd_engine_delta_sizes(e, b"Path", None, 0)
  buf = (ctypes.c_ulonglong * max(cnt, 1))()
  lib.srd_engine_delta_sizes(e, b"Path", buf, cnt)
  lib.srd_engine_destroy(e)
  return out[:n.value], [int(buf[i]) for i in range(cnt)]


@pytest.mark.parametrize("copies", [2, 3])
@pytest.mark.parametrize("env", [{}, {"SRDATALOG_GROUP_STREAMS": "0"},
                                 {"SRDATALOG_JOIN_MODE": "twopass"},
                                 {"SRDATALOG_JOIN_MODE": "fused"},
                                 {"SRDATALOG_JOIN_MODE": "fused", "SRDATALOG_STAGE_CAP": "7"},
                                 {"SRDATALOG_NEW_CHUNK_ROWS": "13", "SRDATALOG_GRAIN": "4"}],
                         ids=["streams", "in-order", "twopass", "fused", "fused-overflow", "chunked"])
def test_parallel_group_members(eng, copies, env, monkeypatch):
  for k in ("SRDATALOG_JOIN_MODE", "SRDATALOG_STAGE_CAP", "SRDATALOG_NEW_CHUNK_ROWS",
            "SRDATALOG_GRAIN", "SRDATALOG_GROUP_STREAMS"):
    monkeypatch.delenv(k, raising=False)
  for k, v in env.items():
    monkeypatch.setenv(k, v)
  e = ds.tc_edges(6, 200, 500)
  want = oracle.run("tc", {"ArcInput": e})
  path, deltas = run_tc_text(eng, tc_group_text(copies), e)
  np.testing.assert_array_equal(path, want.relations["Path"])
  assert deltas == want.deltas["Path"]


@pytest.mark.parametrize("budget", [1000, 60000, 250000, 400000])
def test_devloop_hands_back_to_host_loop(eng, budget, monkeypatch):
  """A device-resident loop whose buffers would exceed its memory budget
  hands its state to the host loop (at the start or mid-run, after the
  iterations it completed); the fixpoint and deltas are unchanged."""
  monkeypatch.setenv("SRDATALOG_DEVLOOP", "1")
  monkeypatch.setenv("SRDATALOG_DEVLOOP_CAPS", "16")
  monkeypatch.setenv("SRDATALOG_DEVLOOP_MAX_BYTES", str(budget))
  e = ds.tc_edges(2, 300, 700)
  want = oracle.run("tc", {"ArcInput": e})
  path, deltas, size = run_tc(eng, e)
  np.testing.assert_array_equal(path, want.relations["Path"])
  assert deltas == want.deltas["Path"]
  assert size == want.size("Path")


@pytest.mark.parametrize("rows", [10, 3000, 20000, 60000])
def test_devloop_row_limit_hands_back(eng, rows, monkeypatch):
  """A step whose FULL starts above SRDATALOG_DEVLOOP_MAX_ROWS runs on the
  host loop; one whose NEW or FULL outgrows it mid-run continues there from
  the last completed iteration; the fixpoint and deltas are unchanged."""
  monkeypatch.setenv("SRDATALOG_DEVLOOP", "1")
  monkeypatch.setenv("SRDATALOG_DEVLOOP_CAPS", "16")
  monkeypatch.setenv("SRDATALOG_DEVLOOP_MAX_ROWS", str(rows))
  e = ds.tc_edges(2, 300, 700)
  want = oracle.run("tc", {"ArcInput": e})
  path, deltas, size = run_tc(eng, e)
  np.testing.assert_array_equal(path, want.relations["Path"])
  assert deltas == want.deltas["Path"]
  assert size == want.size("Path")


# ---- resumable runs (iteration-granular serving) --------------------------
class Resumable:
  """srd_engine_run_begin / run_iterations / absorb / range_counts /
  export_ranges / erase_ranges over the host-only TC engine."""

  def __init__(self, lib):
    self.lib = lib
    ull_p = ctypes.POINTER(ctypes.c_ulonglong)
    vp = ctypes.c_void_p
    lib.srd_engine_run_begin.argtypes = [vp]
    lib.srd_engine_run_iterations.argtypes = [vp, ctypes.c_ulonglong, ull_p, ctypes.POINTER(ctypes.c_int)]
    lib.srd_engine_absorb.argtypes = [vp, vp]
    lib.srd_engine_range_counts.argtypes = [vp, ctypes.c_char_p, ctypes.c_int, _u32p, ctypes.c_uint32,
                                            ull_p]
    lib.srd_engine_export_ranges.argtypes = [vp, ctypes.c_char_p, _u32p, ctypes.c_uint32, ctypes.c_int,
                                             _u32p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64),
                                             ctypes.POINTER(ctypes.c_uint64)]
    lib.srd_engine_erase_ranges.argtypes = [vp, _u32p, ctypes.c_uint32]

  def begin(self, edges):
    e = self.lib.cpu_tc_engine(tc_text().encode())
    assert self.lib.srd_engine_ok(e) == 1
    self.lib.srd_engine_set_quiet(e, 1)
    edges = np.ascontiguousarray(edges, dtype=np.uint32).reshape(-1, 2)
    self.lib.srd_engine_load_rows(e, b"ArcInput", edges.ctypes.data_as(_u32p), edges.shape[0], 2)
    assert self.lib.srd_engine_run_begin(e) == 0
    return e

  def step(self, e, k=1):
    done, conv = ctypes.c_ulonglong(), ctypes.c_int()
    assert self.lib.srd_engine_run_iterations(e, k, ctypes.byref(done), ctypes.byref(conv)) == 0
    return done.value, bool(conv.value)

  def counts(self, e, ranges):
    lohi = np.ascontiguousarray(np.asarray(ranges, np.uint32).reshape(-1))
    out = (ctypes.c_ulonglong * max(len(ranges), 1))()
    assert self.lib.srd_engine_range_counts(e, b"Path", 1, lohi.ctypes.data_as(_u32p), len(ranges),
                                            out) == 0
    return [int(out[i]) for i in range(len(ranges))]

  def export(self, e, ranges, rebase=0):
    """Rows of each range, one array per range (one export call)."""
    lohi = np.ascontiguousarray(np.asarray(ranges, np.uint32).reshape(-1))
    m = len(ranges)
    n = ctypes.c_uint64()
    cuts = np.zeros(m + 1, np.uint64)
    u64p = ctypes.POINTER(ctypes.c_uint64)
    assert self.lib.srd_engine_export_ranges(e, b"Path", lohi.ctypes.data_as(_u32p), m, rebase,
                                             None, 0, ctypes.byref(n), cuts.ctypes.data_as(u64p)) == 0
    out = np.empty((max(n.value, 1), 2), np.uint32)
    assert self.lib.srd_engine_export_ranges(e, b"Path", lohi.ctypes.data_as(_u32p), m, rebase,
                                             out.ctypes.data_as(_u32p), n.value, ctypes.byref(n),
                                             cuts.ctypes.data_as(u64p)) == 0
    return [out[int(cuts[g]):int(cuts[g + 1])] for g in range(m)]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_resumable_instances_match_alone(eng, seed):
  """Instances absorbed into a running TC fixpoint at different iterations
  (one before the first iteration, the others later, one of them empty)
  each end with the tuples and per-iteration |DELTA| of a run of that
  instance alone; finished instances are read out and erased while the
  others keep running."""
  R = Resumable(eng)
  rng = np.random.default_rng(seed)
  sizes = [(40, 90), (25, 60), (0, 0), (70, 150), (12, 30)]
  inst = []
  base = 5
  for n, m in sizes:
    ed = ds.tc_edges(int(rng.integers(1 << 20)), max(n, 1), m) if m else np.zeros((0, 2), np.uint32)
    inst.append((base, base + max(n, 1), ed))
    base += max(n, 1) + int(rng.integers(0, 7))          # gaps between ranges
  arrive = [0, 0, 2, 3, 5]                                # iteration each instance is absorbed at
  main = R.begin(np.zeros((0, 2), np.uint32))
  active, deltas, done = {}, {}, {}
  pending = []

  def erase(rr):
    lohi = np.array(sorted(rr), np.uint32).reshape(-1)
    assert eng.srd_engine_erase_ranges(main, lohi.ctypes.data_as(_u32p), len(rr)) == 0

  it = 0
  while len(done) < len(inst):
    for j, a in enumerate(arrive):
      if a == it and j not in active and j not in done:
        lo, hi, ed = inst[j]
        st = R.begin(ed + np.uint32(lo))
        assert eng.srd_engine_absorb(main, st) == 0
        eng.srd_engine_destroy(st)
        active[j] = True
        deltas[j] = []
    if active:
      R.step(main, 1)
      order = sorted(active)
      cnt = R.counts(main, [(inst[j][0], inst[j][1]) for j in order])
      fin = []
      for j, c in zip(order, cnt):
        deltas[j].append(c)
        if c == 0:
          fin.append(j)
      if fin:
        # seed 2 lets the engine rebase the rows, the others rebase here
        rows = R.export(main, [(inst[j][0], inst[j][1]) for j in fin], rebase=int(seed == 2))
        for j, rj in zip(fin, rows):
          done[j] = rj if seed == 2 else rj - np.uint32(inst[j][0])
          pending.append((inst[j][0], inst[j][1]))
          del active[j]
      # seed 1 defers erasing until two finished instances wait, so that
      # finished rows sit beside running ones and several ranges go at once
      if pending and (len(pending) >= 2 or seed != 1):
        erase(pending)
        pending = []
    it += 1
    assert it < 200
  if pending:
    erase(pending)
  for j, (lo, hi, ed) in enumerate(inst):
    want = oracle.run("tc", {"ArcInput": ed})
    np.testing.assert_array_equal(done[j], want.relations["Path"])
    assert deltas[j] == want.deltas["Path"], j
  # everything erased: the engine holds no Path rows any more
  assert R.export(main, [(0, 1 << 31)])[0].shape[0] == 0
  eng.srd_engine_destroy(main)


def test_resumable_rejects_misuse(eng):
  R = Resumable(eng)
  e = eng.cpu_tc_engine(tc_text().encode())
  eng.srd_engine_set_quiet(e, 1)
  done, conv = ctypes.c_ulonglong(), ctypes.c_int()
  assert eng.srd_engine_run_iterations(e, 1, ctypes.byref(done), ctypes.byref(conv)) != 0
  main = R.begin(np.array([[1, 2]], np.uint32))
  assert eng.srd_engine_absorb(main, main) != 0
  R.step(main, 1)
  late = R.begin(np.array([[7, 8]], np.uint32))
  R.step(late, 1)
  assert eng.srd_engine_absorb(main, late) != 0           # absorbed engine past iteration 0
  bad = (ctypes.c_uint32 * 4)(5, 9, 3, 4)
  out = (ctypes.c_ulonglong * 2)()
  assert eng.srd_engine_range_counts(main, b"Path", 1, ctypes.cast(bad, _u32p), 2, out) != 0
  for x in (e, main, late):
    eng.srd_engine_destroy(x)
