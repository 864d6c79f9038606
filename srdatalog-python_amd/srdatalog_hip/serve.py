"""Serving: a request loop over resident base graphs with continuous batching.

Base graphs (R-MAT, deduplicated, self-loops dropped) are uploaded to HBM
once, packed into disjoint id ranges of one Edge relation of the multi-query
reachability program (``reach_multi``), whose Edge index then stays resident
across requests.  Three request kinds:

* ``reach``  single-source reachability from a vertex of one graph;
* ``tri``    triangle count of the subgraph induced by a vertex range
             [lo, lo+size) of one graph (edges oriented low->high);
* ``tc``     transitive closure of the subgraph induced by a random vertex
             set of one graph (vertices relabelled 0..k-1 in id order).

Continuous batching: one worker per kind runs on its own program library
(own engine, own HIP stream), so the three kinds' launches overlap on the
device.  Whenever a worker is free it takes every request of its kind that
has arrived (up to the batch limit and the memory budget) and runs them as
one batch: reach requests become rows (q, source) of SrcQ over the resident
graphs (one fixpoint for all of them); tri and tc requests are packed into
disjoint id ranges (srdatalog_hip.batch).  Requests that arrive meanwhile
wait for the next batch of their kind; each result is split out of the
batch's output and is, by construction, the request's result alone.

TC requests can be admitted at iteration granularity instead
(``iterative=True``): one resumable fixpoint keeps running (include/srdatalog.h,
engine handles), and between two of its iterations the requests that
arrived are packed into fresh id ranges, their base step is run on a
second engine and absorbed into the running one, and the requests whose
rows produced an empty delta are read out together and later erased.  A request thus
starts at the next iteration boundary rather than after the running batch,
and leaves as soon as its own fixpoint is reached; its rows never meet
another request's (disjoint ranges), so its result and per-iteration deltas
are those of a run of it alone.  On the 200 req/s stream, which TC
requests overload, both admissions serve about the same (116.3 against
118.2 queries/s, 899 against 1,062 TC requests in the 90 s window), so
batch admission stays the default.

Admission: every request's output is estimated (reach: the graph's vertex
count; tri: the listing bound m^1.5; tc: sampled BFS reach) and a request
is admitted into a batch only while the summed estimate of in-flight
requests fits the HBM budget; a request whose estimate alone exceeds the
per-request cap is rejected with the reason.

The plain baseline executes the same stream serially, one shim session per
request, uploading that request's inputs each time (the reference's model,
examples/run_benchmark.py:195-238).
"""

from __future__ import annotations

import dataclasses
import threading
import time

import numpy as np

from . import batch, datasets, programs, runtime

BYTES_PER_TUPLE = batch.BYTES_PER_TUPLE
KIND_NAMES = ("reach", "tri", "tc")


# ---------------------------------------------------------------- graphs ----
@dataclasses.dataclass
class Graph:
  scale: int
  n: int
  edges: np.ndarray        # (m, 2) u32, distinct, no self-loops, sorted
  ptr: np.ndarray          # CSR row pointers (n + 1), int64
  base: int = 0            # offset of this graph in the packed id space
  label: np.ndarray | None = None   # scratch vertex -> local id table (tc extraction)


def make_graph(seed: int, scale: int, edge_factor: int = 16) -> Graph:
  raw = datasets.rmat_edges(seed, scale, edge_factor << scale)
  u = raw[:, 0].astype(np.uint64)
  v = raw[:, 1].astype(np.uint64)
  keep = u != v
  keys = np.unique((u[keep] << np.uint64(32)) | v[keep])
  e = np.stack([(keys >> np.uint64(32)).astype(np.uint32),
                (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)], axis=1)
  n = 1 << scale
  ptr = np.zeros(n + 1, np.int64)
  np.cumsum(np.bincount(e[:, 0], minlength=n), out=ptr[1:])
  return Graph(scale, n, e, ptr)


def make_graphs(seed: int, scales=(16, 18, 20, 22), edge_factor: int = 16) -> list[Graph]:
  gs = [make_graph(seed + i, s, edge_factor) for i, s in enumerate(scales)]
  base = 0
  for g in gs:
    g.base = base
    base += g.n
  if base >= 2**32:
    raise ValueError("packed graph ids exceed u32")
  return gs


def packed_edges(graphs: list[Graph]) -> np.ndarray:
  return np.concatenate([g.edges + np.uint32(g.base) for g in graphs])


# -------------------------------------------------------------- requests ----
@dataclasses.dataclass
class Request:
  rid: int
  t: float                 # arrival, seconds from stream start
  kind: str
  graph: int
  src: int = 0             # reach
  lo: int = 0              # tri: vertex range [lo, lo + size)
  size: int = 0
  verts: np.ndarray | None = None   # tc: sorted vertex set
  est: float = 0.0
  sub: np.ndarray | None = None     # extracted subgraph (tri, tc)
  result: object = None
  result_dropped: bool = False      # delivered, then not retained (keep budget)
  t_done: float = -1.0
  rejected: str = ""
  prepared: bool = False


def make_requests(seed: int, graphs: list[Graph], rate: float = 200.0, seconds: float = 60.0,
                  mix=(0.6, 0.2, 0.2), tri_log2=(12, 16), tc_range=(1000, 50_000)) -> list[Request]:
  """Poisson arrivals at `rate` per second for `seconds`; kinds drawn by
  `mix`; the graph uniformly."""
  rng = np.random.default_rng(seed)
  out = []
  t = 0.0
  while True:
    t += rng.exponential(1.0 / rate)
    if t >= seconds:
      break
    kind = KIND_NAMES[int(rng.choice(3, p=mix))]
    gi = int(rng.integers(0, len(graphs)))
    g = graphs[gi]
    r = Request(len(out), t, kind, gi)
    if kind == "reach":
      r.src = int(rng.integers(0, g.n))
    elif kind == "tri":
      r.size = min(g.n, int(round(2.0 ** rng.uniform(*tri_log2))))
      r.lo = int(rng.integers(0, g.n - r.size + 1))
    else:
      k = min(g.n, int(rng.integers(tc_range[0], tc_range[1] + 1)))
      r.verts = np.sort(rng.choice(g.n, size=k, replace=False)).astype(np.uint32)
    out.append(r)
  return out


def _out_edges(g: Graph, rows: np.ndarray) -> np.ndarray:
  """All edges whose source is in `rows` (ascending vertex ids)."""
  b, e = g.ptr[rows], g.ptr[rows + 1]
  cnt = e - b
  tot = int(cnt.sum())
  if tot == 0:
    return np.zeros((0, 2), np.uint32)
  starts = np.repeat(b - np.concatenate([[0], np.cumsum(cnt)[:-1]]), cnt)
  return g.edges[starts + np.arange(tot)]


def extract(r: Request, g: Graph) -> np.ndarray:
  """The request's input subgraph (tri: oriented, deduplicated; tc: relabelled)."""
  if r.kind == "tri":
    lo, hi = np.uint32(r.lo), np.uint32(r.lo + r.size)
    e = g.edges[g.ptr[r.lo]:g.ptr[r.lo + r.size]]
    e = e[(e[:, 1] >= lo) & (e[:, 1] < hi)] - lo
    return datasets.orient_dedup(e) if e.shape[0] else np.zeros((0, 2), np.uint32)
  if r.kind == "tc":
    # label table: vertex -> position in r.verts, or -1 (reset after use so
    # the table stays all -1 between requests; one table per graph, used by
    # one thread at a time)
    if g.label is None:
      g.label = np.full(g.n, -1, np.int32)
    lab = g.label
    lab[r.verts] = np.arange(r.verts.size, dtype=np.int32)
    try:
      e = _out_edges(g, r.verts.astype(np.int64))
      dst = lab[e[:, 1]]
      keep = dst >= 0
      src = lab[e[keep, 0]]
      return np.stack([src.astype(np.uint32), dst[keep].astype(np.uint32)], axis=1)
    finally:
      lab[r.verts] = -1
  raise ValueError(r.kind)


def estimate(r: Request, g: Graph, rng: np.random.Generator) -> float:
  if r.kind == "reach":
    return float(g.n)
  if r.kind == "tri":
    m = float(r.sub.shape[0])
    return m ** 1.5
  return batch.estimate_output(batch.Instance("tc", int(r.verts.size), r.sub), rng, samples=4)


def prepare(reqs: list[Request], graphs: list[Graph], seed: int = 0) -> None:
  """Host-side work done at admission: subgraph extraction and the estimate."""
  rng = np.random.default_rng(seed)
  for r in reqs:
    if r.prepared:
      continue
    g = graphs[r.graph]
    if r.kind != "reach" and r.sub is None:
      r.sub = extract(r, g)
    r.est = estimate(r, g, rng)
    r.prepared = True


# ---------------------------------------------------------- batch runners ---
class ReachRunner:
  """Multi-query reachability over the resident packed graphs."""

  def __init__(self, graphs: list[Graph]):
    self.graphs = graphs
    self.lib = programs.load("reach_multi")
    if self.lib.init() != 0:
      raise RuntimeError("srdatalog_init failed (reach_multi)")
    self.lib.load_rows("Edge", packed_edges(graphs))
    self.lib.load_rows("SrcQ", np.zeros((0, 2), np.uint32))
    self.lib.prepare()                      # Edge resident in HBM from here on

  def run(self, reqs: list[Request]) -> None:
    src = np.array([[q, self.graphs[r.graph].base + r.src] for q, r in enumerate(reqs)],
                   dtype=np.uint32).reshape(-1, 2)
    self.lib.load_rows("SrcQ", src)
    if self.lib.run(0) != 0:
      raise RuntimeError("srdatalog_run failed (reach batch)")
    rows = self.lib.export("ReachQ", 2)
    cuts = np.searchsorted(rows[:, 0], np.arange(len(reqs) + 1, dtype=np.uint32))
    for q, r in enumerate(reqs):
      r.result = rows[cuts[q]:cuts[q + 1], 1] - np.uint32(self.graphs[r.graph].base)

  def close(self) -> None:
    self.lib.shutdown()


class PackedRunner:
  """tri / tc requests packed into disjoint id ranges of one database."""

  def __init__(self, kind: str):
    self.kind = kind
    name, self.rin, self.rout, self.arity = (("triangle", "KnowsInput", "Triangle", 3)
                                             if kind == "tri" else batch.KINDS["tc"])
    self.lib = programs.load(name)
    if self.lib.init() != 0:
      raise RuntimeError(f"srdatalog_init failed ({name})")

  def run(self, reqs: list[Request]) -> None:
    insts = [batch.Instance(self.kind, r.size if self.kind == "tri" else int(r.verts.size), r.sub)
             for r in reqs]
    rows, bounds = batch.pack(insts)
    self.lib.load_rows(self.rin, rows)
    self.lib.set_instances(bounds)
    if self.lib.run(0) != 0:
      raise RuntimeError(f"srdatalog_run failed ({self.kind} batch)")
    out, cuts = self.lib.export_instances(self.rout, self.arity, len(reqs))
    for i, r in enumerate(reqs):
      blk = out[cuts[i]:cuts[i + 1]]
      r.result = int(blk.shape[0]) if self.kind == "tri" else blk

  def close(self) -> None:
    self.lib.shutdown()


class IterativeTCRunner:
  """TC requests admitted between iterations of one running fixpoint."""

  # fresh id ranges are taken from here on; the running engine starts over
  # (when idle) before they could reach the u32 limit
  ID_LIMIT = 1 << 30

  def __init__(self):
    self.prog = programs.load("tc")
    self.main = runtime.Engine(self.prog)
    self.stage = runtime.Engine(self.prog)
    self.active: list[tuple[int, int, Request]] = []   # (lo, hi, request), ascending lo
    self._restart()

  # finished requests' rows stay in the engine until nothing runs or this
  # many have accumulated (their memory): one compaction pass over FULL then
  # removes them all.  Meanwhile they cost an iteration nothing measurable
  # (60 requests: 0.50 s of iterations with them kept, while compacting
  # every few iterations doubled the total)
  ERASE_ROWS = 1 << 28

  def _restart(self) -> None:
    self.main.load_rows("ArcInput", np.zeros((0, 2), np.uint32))
    self.main.run_begin()
    self.next_id = 0
    self.dead: list[tuple[int, int]] = []
    self.dead_rows = 0

  def admit(self, reqs: list[Request]) -> None:
    if not self.active and self.next_id > self.ID_LIMIT:
      self._restart()
    parts = []
    for r in reqs:
      k = int(r.verts.size)
      lo, hi = self.next_id, self.next_id + max(k, 1)
      self.next_id = hi
      self.active.append((lo, hi, r))
      if r.sub.shape[0]:
        parts.append(r.sub.astype(np.uint32) + np.uint32(lo))
    rows = np.concatenate(parts) if parts else np.zeros((0, 2), np.uint32)
    self.stage.load_rows("ArcInput", rows)
    self.stage.run_begin()
    self.main.absorb(self.stage)

  def step(self) -> list[Request]:
    """One iteration; returns the requests that finished with it."""
    if not self.active:
      return []
    self.main.run_iterations(1)
    cnt = self.main.range_counts("Path", [(lo, hi) for lo, hi, _ in self.active], delta=True)
    done, keep, fin = [], [], []
    for (lo, hi, r), c in zip(self.active, cnt):
      (fin if c == 0 else keep).append((lo, hi, r))
    if fin:
      # one gather, sort and copy for every request finished this iteration
      rows, cuts = self.main.export_ranges("Path", 2, [(lo, hi) for lo, hi, _ in fin], rebase=True)
      for g, (lo, hi, r) in enumerate(fin):
        r.result = rows[cuts[g]:cuts[g + 1]]
        self.dead.append((lo, hi))
        self.dead_rows += int(r.result.shape[0]) + int(r.sub.shape[0])
        done.append(r)
    self.active = keep
    if self.dead and (not self.active or self.dead_rows >= self.ERASE_ROWS):
      self.main.erase_ranges(sorted(self.dead))
      self.dead, self.dead_rows = [], 0
    return done

  def close(self) -> None:
    self.main.close()
    self.stage.close()


# ----------------------------------------------------------------- server ---
class Server:
  """Continuous batching over one request stream (arrival times relative to
  the call of run())."""

  def __init__(self, graphs: list[Graph], budget_bytes: float = 200e9, cap_tuples: float = 5e7,
               max_batch: int = 256, batch_tuples: float = 2e8, progress=None,
               keep_bytes: float = 48e9, iterative: bool = False):
    self.graphs = graphs
    # progress(kind, seconds since the stream started, requests completed
    # so far): called by a worker after each batch it completes
    self.progress = progress
    self.completed = 0
    # Host bytes of results retained on the requests after delivery; a
    # batch whose results would exceed it is delivered and not retained
    # (its requests keep result None and result_dropped True).
    self.keep_bytes = keep_bytes
    self.kept_bytes = 0
    self.budget = budget_bytes
    self.cap = cap_tuples
    self.max_batch = max_batch
    # summed estimate of one batch: a batch's join output (NEW, before
    # deduplication, several times the estimate) must stay far below the
    # 2^32 rows an index holds
    self.batch_tuples = batch_tuples
    # iterative: TC requests join the running fixpoint between iterations
    self.iterative = iterative
    self.runners = {"reach": ReachRunner(graphs), "tri": PackedRunner("tri"),
                    "tc": IterativeTCRunner() if iterative else PackedRunner("tc")}
    self.lock = threading.Lock()
    self.in_flight = 0.0
    self.errors: list[str] = []
    self.batches = {k: 0 for k in KIND_NAMES}

  def close(self) -> None:
    for r in self.runners.values():
      r.close()

  def _take(self, queue: list[Request], pos: int, now: float,
            held: float = 0.0) -> tuple[list[Request], int]:
    """Admit arrived, prepared requests in arrival order while the in-flight
    estimate fits the budget.  `held`: estimated tuples of requests already
    running in the engine the new ones join (iterative TC), counted against
    the per-batch tuple cap with them."""
    take = []
    batch_est = held
    with self.lock:
      while pos < len(queue) and queue[pos].t <= now and len(take) < self.max_batch:
        r = queue[pos]
        if not r.prepared:
          break
        if r.est > self.cap:
          r.rejected = f"estimated {r.est:.3g} output tuples > cap {self.cap:.3g}"
          pos += 1
          continue
        need = r.est * BYTES_PER_TUPLE
        if self.in_flight > 0 and self.in_flight + need > self.budget:
          break                      # wait for memory to free up
        if (take or held > 0) and batch_est + r.est > self.batch_tuples:
          break                      # the rest goes into the next batch
        self.in_flight += need
        batch_est += r.est
        take.append(r)
        pos += 1
    return take, pos

  def _finish(self, kind: str, done: list[Request], t0: float) -> None:
    """Completion of `done` at the current time: memory, counts, progress."""
    t = time.perf_counter() - t0
    nbytes = sum(r.result.nbytes for r in done if isinstance(r.result, np.ndarray))
    with self.lock:
      keep = self.kept_bytes + nbytes <= self.keep_bytes
      if keep:
        self.kept_bytes += nbytes
      for r in done:
        r.t_done = t
        self.in_flight -= r.est * BYTES_PER_TUPLE
        if not keep and isinstance(r.result, np.ndarray):
          r.result = None
          r.result_dropped = True
      self.completed += len(done)
      completed = self.completed
    if self.progress is not None:
      self.progress(kind, t, completed)

  def _iter_worker(self, kind: str, queue: list[Request], t0: float,
                   wall_cap: float | None) -> None:
    """Iteration-granular admission: between two iterations of the running
    fixpoint, absorb the requests that arrived (budget and in-flight tuple
    cap permitting) and release the ones that finished."""
    pos = 0
    runner = self.runners[kind]
    try:
      while pos < len(queue) or runner.active:
        now = time.perf_counter() - t0
        open_ = wall_cap is None or now <= wall_cap
        if open_ and pos < len(queue) and queue[pos].t <= now:
          j = pos
          while j < len(queue) and queue[j].t <= now and j - pos < self.max_batch:
            prepare([queue[j]], self.graphs)
            j += 1
          take, pos = self._take(queue, pos, now, held=sum(r.est for _, _, r in runner.active))
          if take:
            runner.admit(take)
            with self.lock:
              self.batches[kind] += 1
        if runner.active:
          done = runner.step()
          if done:
            self._finish(kind, done, t0)
        elif not open_:
          break                      # the rest stays unserved (t_done < 0)
        elif pos < len(queue):
          time.sleep(min(max(queue[pos].t - now, 0.0), 0.002))
    except Exception as ex:  # surfaced by run()
      self.errors.append(f"{kind}: {ex!r}")

  def _worker(self, kind: str, queue: list[Request], t0: float, wall_cap: float | None) -> None:
    if kind == "tc" and self.iterative:
      return self._iter_worker(kind, queue, t0, wall_cap)
    pos = 0
    runner = self.runners[kind]
    try:
      while pos < len(queue):
        now = time.perf_counter() - t0
        if wall_cap is not None and now > wall_cap:
          break                      # the rest stays unserved (t_done < 0)
        if queue[pos].t > now:
          time.sleep(min(queue[pos].t - now, 0.002))
          continue
        j = pos
        while j < len(queue) and queue[j].t <= now and j - pos < self.max_batch:
          prepare([queue[j]], self.graphs)
          j += 1
        take, pos = self._take(queue, pos, now)
        if not take:
          time.sleep(0.0005)
          continue
        runner.run(take)
        with self.lock:
          self.batches[kind] += 1
        self._finish(kind, take, t0)
    except Exception as ex:  # surfaced by run()
      self.errors.append(f"{kind}: {ex!r}")

  def run(self, reqs: list[Request], wall_cap: float | None = None) -> float:
    """Serve the stream; returns the wall seconds until the last completion.
    With wall_cap, no batch starts after wall_cap seconds: requests not
    started by then stay unserved (t_done < 0)."""
    queues = {k: [r for r in reqs if r.kind == k] for k in KIND_NAMES}
    t0 = time.perf_counter()
    ths = [threading.Thread(target=self._worker, args=(k, q, t0, wall_cap), daemon=True)
           for k, q in queues.items()]
    for th in ths:
      th.start()
    for th in ths:
      th.join()
    if self.errors:
      raise RuntimeError("; ".join(self.errors))
    return time.perf_counter() - t0


# ------------------------------------------------------------ plain serial --
def run_plain(reqs: list[Request], graphs: list[Graph], wall_cap: float | None = None,
              on_result=None) -> float:
  """Serve the stream one request at a time, one shim session each; stops
  starting new requests after wall_cap seconds.  Returns the wall seconds.
  on_result(r), if given, is called with each request as it completes."""
  t0 = time.perf_counter()
  for r in reqs:
    now = time.perf_counter() - t0
    if wall_cap is not None and now > wall_cap:
      break
    if r.t > now:
      time.sleep(r.t - now)
    g = graphs[r.graph]
    if r.kind == "reach":
      lib = programs.load("reach")
      inputs = {"Edge": g.edges, "Src": np.array([[r.src]], np.uint32)}
      out_rel, arity = "Reach", 1
    else:
      if r.sub is None:
        r.sub = extract(r, g)
      name, rin, out_rel, arity = (("triangle", "KnowsInput", "Triangle", 3) if r.kind == "tri"
                                   else batch.KINDS["tc"])
      lib = programs.load(name)
      inputs = {rin: r.sub}
    if lib.init() != 0:
      raise RuntimeError("srdatalog_init failed")
    try:
      for rel, rows in inputs.items():
        lib.load_rows(rel, rows)
      if lib.run(0) != 0:
        raise RuntimeError(f"srdatalog_run failed (plain {r.kind})")
      if r.kind == "tri":
        r.result = int(lib.size(out_rel))
      else:
        res = lib.export(out_rel, arity)
        r.result = res[:, 0].copy() if r.kind == "reach" else res
    finally:
      lib.shutdown()
    r.t_done = time.perf_counter() - t0
    if on_result is not None:
      on_result(r)
  return time.perf_counter() - t0


def same_result(a: Request, b: Request) -> bool:
  if a.kind == "tri":
    return a.result == b.result
  return (isinstance(a.result, np.ndarray) and isinstance(b.result, np.ndarray)
          and a.result.shape == b.result.shape and np.array_equal(a.result, b.result))


def latency_stats(reqs: list[Request]) -> dict:
  lat = np.array([r.t_done - r.t for r in reqs if r.t_done >= 0])
  if lat.size == 0:
    return {"completed": 0}
  return {"completed": int(lat.size), "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
          "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
          "mean_ms": round(float(lat.mean()) * 1e3, 3)}
