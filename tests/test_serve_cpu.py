"""Host logic of the serving loop (no GPU): request stream, subgraph
extraction, estimates, and that the multi-query reachability program's
semantics (query id carried unchanged) splits into single-source results."""

import numpy as np

import oracle
from srdatalog_hip import datasets as ds
from srdatalog_hip import serve


def _graphs():
  return serve.make_graphs(5, scales=(8, 9, 10), edge_factor=8)


def test_graphs_packed_disjoint():
  gs = _graphs()
  assert [g.base for g in gs] == [0, 256, 768]
  for g in gs:
    assert g.edges.dtype == np.uint32 and (g.edges[:, 0] != g.edges[:, 1]).all()
    keys = g.edges[:, 0].astype(np.uint64) << np.uint64(32) | g.edges[:, 1]
    assert (np.diff(keys) > 0).all()
    assert g.ptr[-1] == g.edges.shape[0]
  pe = serve.packed_edges(gs)
  assert pe.shape[0] == sum(g.edges.shape[0] for g in gs)


def test_request_stream_rate_and_mix():
  gs = _graphs()
  reqs = serve.make_requests(1, gs, rate=400.0, seconds=5.0, tri_log2=(4, 7), tc_range=(20, 100))
  assert 1700 < len(reqs) < 2300
  ts = np.array([r.t for r in reqs])
  assert (np.diff(ts) > 0).all() and ts[-1] < 5.0
  frac = np.mean([r.kind == "reach" for r in reqs])
  assert 0.55 < frac < 0.65


def test_extract_matches_brute_force():
  gs = _graphs()
  reqs = serve.make_requests(2, gs, rate=200.0, seconds=0.5, tri_log2=(4, 8), tc_range=(30, 200))
  for r in reqs:
    if r.kind == "reach":
      continue
    g = gs[r.graph]
    sub = serve.extract(r, g)
    e = g.edges.astype(np.int64)
    if r.kind == "tri":
      m = (e[:, 0] >= r.lo) & (e[:, 0] < r.lo + r.size) & (e[:, 1] >= r.lo) & (e[:, 1] < r.lo + r.size)
      want = ds.orient_dedup((e[m] - r.lo).astype(np.uint32)) if m.any() else np.zeros((0, 2), np.uint32)
    else:
      idx = {int(v): i for i, v in enumerate(r.verts)}
      rows = sorted((idx[int(a)], idx[int(b)]) for a, b in e if int(a) in idx and int(b) in idx)
      want = np.array(rows, np.uint32).reshape(-1, 2)
    np.testing.assert_array_equal(sub, want)


def test_multi_query_reach_splits_into_single_source():
  """ReachQ(q, .) over the packed graphs equals Reach from q's source in its
  own graph (checked with the oracle's single-source reach)."""
  gs = _graphs()
  rng = np.random.default_rng(0)
  pe = serve.packed_edges(gs)
  for q in range(6):
    gi = int(rng.integers(0, len(gs)))
    g = gs[gi]
    s = int(rng.integers(0, g.n))
    alone = oracle.run("reach", {"Edge": g.edges, "Src": np.array([[s]], np.uint32)})
    packed = oracle.run("reach", {"Edge": pe, "Src": np.array([[g.base + s]], np.uint32)})
    np.testing.assert_array_equal(packed.relations["Reach"][:, 0] - g.base,
                                  alone.relations["Reach"][:, 0])


def test_prepare_estimates():
  gs = _graphs()
  reqs = serve.make_requests(3, gs, rate=100.0, seconds=0.3, tri_log2=(4, 8), tc_range=(30, 200))
  serve.prepare(reqs, gs)
  for r in reqs:
    assert r.prepared and r.est >= 0
    if r.kind == "reach":
      assert r.est == gs[r.graph].n


# ---- scheduling of the server, with stand-in runners (no GPU) --------------
class _FakePacked:
  def __init__(self, kind, log):
    self.kind, self.log = kind, log

  def run(self, reqs):
    self.log.append((self.kind, [r.rid for r in reqs]))
    for r in reqs:
      r.result = 1 if self.kind == "tri" else np.full((3, 2), r.rid, np.uint32)

  def close(self):
    pass


class _FakeIterative:
  """Requests finish after a number of steps set by their estimate; records
  the estimated tuples held at every admission."""

  def __init__(self, log):
    self.active, self.log, self.held_seen = [], log, []

  def admit(self, reqs):
    self.active += [[r, 1 + int(r.est) % 3] for r in reqs]
    self.held_seen.append(sum(r.est for r, _ in self.active))
    self.log.append(("tc", [r.rid for r in reqs]))

  def step(self):
    done = []
    for a in self.active:
      a[1] -= 1
      if a[1] == 0:
        a[0].result = np.full((2, 2), a[0].rid, np.uint32)
        done.append(a[0])
    self.active = [a for a in self.active if a[1] > 0]
    return done

  def close(self):
    pass


def _fake_server(monkeypatch, iterative, **kw):
  log = []
  it = _FakeIterative(log)
  monkeypatch.setattr(serve, "ReachRunner", lambda graphs: _FakePacked("reach", log))
  monkeypatch.setattr(serve, "PackedRunner", lambda kind: _FakePacked(kind, log))
  monkeypatch.setattr(serve, "IterativeTCRunner", lambda: it)
  monkeypatch.setattr(serve, "prepare", _fake_prepare)
  srv = serve.Server(_graphs(), iterative=iterative, **kw)
  return srv, log, it


def _fake_prepare(reqs, graphs, seed=0):
  for r in reqs:
    r.est = float(10 + r.rid % 7)
    r.prepared = True


def _stream(n=60, dt=0.002):
  reqs = []
  for i in range(n):
    r = serve.Request(i, i * dt, ("reach", "tri", "tc")[i % 3], 0)
    reqs.append(r)
  return reqs


def test_server_iterative_admission_respects_tuple_cap(monkeypatch):
  """TC requests join the running engine while its held estimate plus theirs
  stays within the batch tuple cap; every request completes once."""
  srv, log, it = _fake_server(monkeypatch, True, batch_tuples=40.0)
  reqs = _stream()
  srv.run(reqs)
  assert all(r.t_done >= r.t for r in reqs)
  assert max(it.held_seen) <= 40.0
  tc_ids = sorted(i for k, ids in log if k == "tc" for i in ids)
  assert tc_ids == sorted(r.rid for r in reqs if r.kind == "tc")


def test_server_wall_cap_and_keep_budget(monkeypatch):
  """No batch starts after the wall cap (later requests stay unserved) and
  results beyond the keep budget are delivered but not retained."""
  srv, log, _ = _fake_server(monkeypatch, True, keep_bytes=100.0)
  reqs = _stream(n=90, dt=0.01)
  seen = []
  srv.progress = lambda kind, t, done: seen.append(done)
  srv.run(reqs, wall_cap=0.3)
  served = [r for r in reqs if r.t_done >= 0]
  assert served and len(served) < len(reqs)
  assert all(r.t <= 0.3 + 0.05 for r in served)
  dropped = [r for r in served if r.result_dropped]
  assert dropped and all(r.result is None for r in dropped)
  assert srv.kept_bytes <= 100.0
  assert seen and seen[-1] == len(served)
