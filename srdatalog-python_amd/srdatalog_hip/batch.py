"""Batched execution of many independent instances of one compiled program.

The plain path runs each instance through the program's shim on its own
(init -> load -> run -> export -> shutdown).  The batched path packs every
admitted instance of one program into a single database: instance i's
vertex ids are shifted into the disjoint range [base_i, base_i + n_i), so
no join can combine tuples of two instances and the fixpoint of the packed
database is the disjoint union of the per-instance fixpoints -- iteration
for iteration, because the semi-naive loop is Jacobi: the packed DELTA of
iteration k is the union of every instance's DELTA of iteration k (an
instance that has converged contributes nothing).  One launch per operator
then serves all instances; the persistent work-stealing join balances the
uneven instance sizes inside each launch.

The instance table (the ascending range bounds) is handed to the engine
(``srdatalog_set_instances``); after every DELTA computation the engine counts
each instance's DELTA rows on the device, which gives every instance's own
per-iteration delta sizes without a host round trip.

Admission: an instance's fixpoint size is estimated first (sampled BFS reach
for transitive closure, squared generation widths for same-generation);
instances whose estimate exceeds the per-instance cap, or that would push
the packed batch past the HBM budget, are rejected with the reason.

Reference counterpart: none -- the reference runs one instance per process
(examples/run_benchmark.py:195-238); the per-instance results are defined by
that single-instance semantics.
"""

from __future__ import annotations

import dataclasses
import time

import numpy as np

from . import programs

# kind -> (program, input relation, output relation, arity of the output)
KINDS = {
  "tc": ("tc", "ArcInput", "Path", 2),
  "sg": ("sg", "Arc", "Sg", 2),
}

# Bytes of HBM per output tuple of a binary relation while the fixpoint runs:
# FULL (8) + NEW/DELTA (<= 8) + sort keys and their ping-pong copy (16) +
# the merged FULL being built (8), rounded up.
BYTES_PER_TUPLE = 48


@dataclasses.dataclass
class Instance:
  kind: str
  n: int
  edges: np.ndarray          # (m, 2) u32, vertex ids in [0, n)
  est: float = 0.0           # estimated output tuples


@dataclasses.dataclass
class Result:
  rows: np.ndarray           # (t, arity) u32, sorted, instance-local ids
  deltas: list[int]          # per-iteration |DELTA| of the output relation
  session_s: float = 0.0     # run_plain: wall time of the instance's session (init to shutdown)


# ------------------------------------------------------------- estimates ----
def estimate_output(inst: Instance, rng: np.random.Generator, samples: int = 24) -> float:
  """Estimated size of the instance's output relation at fixpoint."""
  import scipy.sparse as sp
  import scipy.sparse.csgraph as csg
  n, e = inst.n, inst.edges
  if n == 0 or e.shape[0] == 0:
    return 0.0
  a = sp.csr_matrix((np.ones(e.shape[0], np.int8), (e[:, 0], e[:, 1])), shape=(n, n))
  if inst.kind == "tc":
    # |Path| = sum over x of |reach+(x)| (x itself counts when it lies on a
    # cycle: its strong component has more than one vertex or a self-loop)
    _, comp = csg.connected_components(a, directed=True, connection="strong")
    csize = np.bincount(comp)
    on_cycle = (csize[comp] > 1) | (a.diagonal() != 0)
    src = rng.choice(n, size=min(n, samples), replace=False)
    reach = [csg.breadth_first_order(a, int(s), directed=True,
                                     return_predecessors=False).size - 1 + int(on_cycle[s])
             for s in src]
    return float(n) * float(np.mean(reach))
  if inst.kind == "sg":
    # Pairs at equal depth below the source components of the condensation
    # (a forest's roots; a root that received a cross edge sits on a cycle
    # and stays a source component).  A cycle makes generations repeat
    # around it, so everything reachable from a cycle is counted as one
    # generation (pessimistic: admission rather overestimates).
    ncomp, comp = csg.connected_components(a, directed=True, connection="strong")
    csize = np.bincount(comp, minlength=ncomp)
    cross = comp[e[:, 0]] != comp[e[:, 1]]
    has_in = np.zeros(ncomp, bool)
    has_in[comp[e[cross, 1]]] = True
    depth = np.full(n, -1, np.int64)
    frontier = np.flatnonzero(~has_in[comp])
    depth[frontier] = 0
    d = 0
    while frontier.size:
      nxt = np.unique(a[frontier].indices)
      nxt = nxt[depth[nxt] < 0]
      d += 1
      depth[nxt] = d
      frontier = nxt
    cnt = np.bincount(depth[depth >= 0]).astype(np.float64)
    est = float((cnt * cnt).sum())
    cyclic = np.flatnonzero((csize[comp] > 1) | (a.diagonal() != 0))
    if cyclic.size:
      seen = np.zeros(n, bool)
      seen[cyclic] = True
      frontier = cyclic
      while frontier.size:
        nxt = np.unique(a[frontier].indices)
        nxt = nxt[~seen[nxt]]
        seen[nxt] = True
        frontier = nxt
      r = float(seen.sum())
      est += r * r
    return est
  raise ValueError(f"unknown instance kind {inst.kind!r}")


def admit(instances: list[Instance], cap_tuples: float = 1e8,
          budget_bytes: float = 200e9, seed: int = 0) -> tuple[list[int], list[tuple[int, str]]]:
  """Indices of admitted instances (in order) and (index, reason) of the rest."""
  rng = np.random.default_rng(seed)
  ok, rejected = [], []
  used = 0.0
  for i, inst in enumerate(instances):
    if not inst.est:
      inst.est = estimate_output(inst, rng)
    need = inst.est * BYTES_PER_TUPLE
    if inst.est > cap_tuples:
      rejected.append((i, f"estimated {inst.est:.3g} output tuples > cap {cap_tuples:.3g}"))
    elif used + need > budget_bytes:
      rejected.append((i, f"batch HBM budget {budget_bytes:.3g} B exhausted"))
    else:
      ok.append(i)
      used += need
  return ok, rejected


# --------------------------------------------------------------- packing ----
def pack(instances: list[Instance]) -> tuple[np.ndarray, np.ndarray]:
  """Packed edge rows (ids shifted by each instance's base) and the instance
  table: bounds[i] = base_i, bounds[-1] = total vertex count."""
  sizes = np.array([inst.n for inst in instances], dtype=np.int64)
  bounds = np.zeros(len(instances) + 1, dtype=np.int64)
  np.cumsum(sizes, out=bounds[1:])
  if bounds[-1] >= 2**32:
    raise ValueError("packed vertex ids exceed u32")
  parts = [inst.edges.astype(np.uint32) + np.uint32(bounds[i])
           for i, inst in enumerate(instances) if inst.edges.shape[0]]
  rows = np.concatenate(parts) if parts else np.zeros((0, 2), np.uint32)
  return rows, bounds.astype(np.uint32)


def split(rows: np.ndarray, bounds: np.ndarray) -> list[np.ndarray]:
  """Per-instance row blocks of a packed, lexicographically sorted export,
  with instance-local ids."""
  cuts = np.searchsorted(rows[:, 0], bounds) if rows.shape[0] else np.zeros(bounds.size, np.int64)
  out = []
  for i in range(bounds.size - 1):
    blk = rows[cuts[i]:cuts[i + 1]]
    out.append(blk - np.uint32(bounds[i]))
  return out


def trim_deltas(col: np.ndarray) -> list[int]:
  """An instance's own delta sequence from its column of the packed
  per-iteration counts.  Run alone, its loop checks at the top of iteration
  k > 0 whether iteration k-1 produced any DELTA
  (codegen/jit/orchestrator_jit.py:659-695), so its sequence ends with its
  first empty DELTA, iteration 0 included."""
  out = []
  for v in col.tolist():
    out.append(int(v))
    if v == 0:
      break
  return out


# ---------------------------------------------------------------- runners ---
def run_plain(kind: str, instances: list[Instance], grain: int = 0) -> list[Result]:
  """One shim session per instance (the reference's execution model)."""
  prog, rin, rout, arity = KINDS[kind]
  lib = programs.load(prog)
  out = []
  for inst in instances:
    t0 = time.perf_counter()
    if lib.init() != 0:
      raise RuntimeError("srdatalog_init failed")
    try:
      if grain:
        lib.set_grain(grain)
      lib.load_rows(rin, inst.edges)
      if lib.run(0) != 0:
        raise RuntimeError(f"srdatalog_run failed for a {kind} instance (n={inst.n})")
      res = Result(lib.export(rout, arity), lib.delta_sizes(rout))
    finally:
      lib.shutdown()
    res.session_s = time.perf_counter() - t0
    out.append(res)
  return out


def run_batched(kind: str, instances: list[Instance], grain: int = 0) -> list[Result]:
  """All instances in one packed database, one launch per operator."""
  prog, rin, rout, arity = KINDS[kind]
  if not instances:
    return []
  rows, bounds = pack(instances)
  lib = programs.load(prog)
  if lib.init() != 0:
    raise RuntimeError("srdatalog_init failed")
  try:
    if grain:
      lib.set_grain(grain)
    lib.load_rows(rin, rows)
    lib.set_instances(bounds)
    if lib.run(0) != 0:
      raise RuntimeError(f"srdatalog_run failed for the packed {kind} batch")
    # rows come back rebased to instance-local ids with per-instance offsets,
    # so each instance's result is a view of one buffer
    rows_out, cuts = lib.export_instances(rout, arity, len(instances))
    per_iter = lib.instance_delta_sizes(rout, len(instances))
  finally:
    lib.shutdown()
  return [Result(rows_out[cuts[i]:cuts[i + 1]], trim_deltas(per_iter[:, i]))
          for i in range(len(instances))]


def run_batched_kinds(groups: dict[str, list[Instance]], grains: dict[str, int] | None = None
                      ) -> dict[str, list[Result]]:
  """run_batched for every program kind at once, one host thread and one
  engine (with its own stream) per kind, so that one kind's export over
  PCIe overlaps the other's kernels.  The foreign calls release the GIL."""
  import threading
  out: dict[str, list[Result]] = {}
  errs: list[BaseException] = []

  def one(kind: str, lst: list[Instance]) -> None:
    try:
      out[kind] = run_batched(kind, lst, grain=(grains or {}).get(kind, 0))
    except BaseException as ex:  # re-raised below
      errs.append(ex)

  ths = [threading.Thread(target=one, args=(k, v)) for k, v in groups.items() if v]
  for th in ths:
    th.start()
  for th in ths:
    th.join()
  if errs:
    raise errs[0]
  return {k: out.get(k, []) for k in groups}


def autotune_grain(kind: str, instances: list[Instance],
                   candidates=(256, 1024, 4096, 16384)) -> tuple[int, dict]:
  """Pick the join work-slice size with the lowest batched wall time."""
  times = {}
  for g in candidates:
    t = time.perf_counter()
    run_batched(kind, instances, grain=g)
    times[g] = time.perf_counter() - t
  best = min(times, key=times.get)
  return best, times
