"""Benchmark: triangle listing fixpoint on an R-MAT scale-20 graph, one MI355X.

Workload (the headline configuration): Tri(x,y,z) <= Edge(x,y), Edge(y,z),
Edge(x,z) over an R-MAT graph of scale 20 with 16,777,216 generated edges,
(a,b,c,d) = (0.57,0.19,0.19,0.05), deduplicated, self-loops dropped and
oriented low->high (so every triangle has x<y<z), u32 columns.  A step is
one full srdatalog_run of the compiled program with the input edges already
resident in HBM: build the Knows index (sort+dedup), run the leapfrog join
(count + materialise), and the NEW->DELTA->FULL maintenance of Triangle.

Prints ONE JSON line: metric = derived tuples/s (|Triangle| / fixpoint s),
plus ms_per_step (fixpoint wall time), the roofline of the dominant kernel
(from HIP events on the runtime's stream during the timed steps) and a CPU
baseline (the C++ oracle on a bounded sample of root keys, on this host).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "srdatalog-python_amd")):
  if _p not in sys.path:
    sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0


def parse():
  p = argparse.ArgumentParser()
  p.add_argument("--gpus", type=int, default=1)
  p.add_argument("--steps", type=int, default=5)
  p.add_argument("--warmup", type=int, default=1)
  p.add_argument("--scale", type=int, default=20)
  p.add_argument("--edges", type=int, default=1 << 24)
  p.add_argument("--seed", type=int, default=0)
  p.add_argument("--cpu-stride-all", type=int, default=16,
                 help="CPU baseline on all threads: root keys a with a %% S == 0")
  p.add_argument("--cpu-stride-one", type=int, default=256,
                 help="CPU baseline on one thread: root keys a with a %% S == 0")
  p.add_argument("--cpu-reps", type=int, default=3)
  p.add_argument("--no-cpu-baseline", action="store_true")
  p.add_argument("--no-count-only", action="store_true")
  p.add_argument("--no-batch", action="store_true",
                 help="triangle workload: skip the uneven-batch sub-object")
  p.add_argument("--no-serve", action="store_true",
                 help="triangle workload: skip the serving sub-object")
  p.add_argument("--no-parity", action="store_true")
  p.add_argument("--no-kernel-breakdown", action="store_true",
                 help="tc/andersen: skip the extra fixpoint run whose every launch is timed")
  p.add_argument("--andersen-div", type=int, default=10,
                 help="andersen workload: run 1/div of the full Andersen input sizes")
  p.add_argument("--workload", choices=("triangle", "batch", "serve", "tc", "andersen"),
                 default="triangle",
                 help="triangle: the headline triangle-listing fixpoint; batch: the uneven "
                      "512-instance TC/SG batch, batched vs plain per-instance; serve: the "
                      "Poisson request stream, continuous batching vs plain serial")
  p.add_argument("--serve-rate", type=float, default=200.0)
  p.add_argument("--serve-seconds", type=float, default=60.0)
  p.add_argument("--serve-cap", type=float, default=float(1 << 31),
                 help="serve: per-request cap on the estimated output tuples (default 2^31: "
                      "an index holds at most 2^32-1 rows, so a request estimated beyond half "
                      "of that is refused; memory admission is the HBM budget)")
  p.add_argument("--serve-budget", type=float, default=200e9)
  p.add_argument("--serve-tc-iterative", action="store_true",
                 help="serve: TC requests join the running fixpoint between its iterations "
                      "instead of waiting for the running batch of their kind")
  p.add_argument("--serve-wall-cap", type=float, default=90.0,
                 help="serve: no batch starts after this many seconds of the served run; "
                      "requests not started by then are reported as unserved")
  p.add_argument("--serve-plain-cap", type=float, default=30.0,
                 help="serve: wall seconds after which the plain serial baseline starts "
                      "no new request")
  p.add_argument("--serve-scales", type=str, default="16,18,20,22")
  p.add_argument("--batch-instances", type=int, default=512)
  p.add_argument("--batch-cap", type=float, default=1e8,
                 help="batch: per-instance cap on the estimated output tuples")
  p.add_argument("--batch-budget", type=float, default=200e9,
                 help="batch: HBM budget (bytes) of the packed batch")
  p.add_argument("--prim-rows", type=int, default=1 << 26,
                 help="rows of the sort/set-difference/merge benchmarks (0 = skip)")
  p.add_argument("--calib-copy-bytes", type=int, default=0,
                 help="also run the plain streaming copy kernel over this many bytes "
                      "(profiler counter calibration)")
  return p.parse_args()


_T0 = time.time()


def log(msg: str) -> None:
  print(f"[bench {time.time() - _T0:7.1f}s] {msg}", file=sys.stderr, flush=True)


def primitives(native, n_sort: int, reps: int) -> dict:
  """Index-maintenance primitives at full scale on device-generated u32
  pairs (k=2, w=8 B/tuple): sort+dedup of n_sort unordered rows with 40-bit
  keys (5 radix passes), set-difference NEW(n) - FULL(2n), merge of FULL(2n)
  and DELTA(n/2).  Algorithmic bytes per call:
    sort+dedup      2*n*w*passes + n*w + u*w + n*4 + r*4 (r <= 2^20)
    set-difference  (|N|+|F|)*w + |D|*w
    merge           2*(|F|+|D|)*w
  """
  w, passes = 8, 5
  out = {}
  n = n_sort
  ms, u = native.prim_bench("sort_dedup", n, 0, 2, reps)
  alg = 2 * n * w * passes + n * w + u * w + n * 4 + (1 << 20) * 4
  out["sort_dedup"] = {"rows": n, "distinct": u, "ms": round(ms, 4), "alg_bytes": alg}
  na, nb = n, 2 * n
  ms, d = native.prim_bench("set_difference", na, nb, 2, reps)
  out["set_difference"] = {"new": na, "full": nb, "delta": d, "ms": round(ms, 4),
                           "alg_bytes": (na + nb) * w + d * w}
  na, nb = 2 * n, n // 2
  ms, m = native.prim_bench("merge", na, nb, 2, reps)
  out["merge"] = {"full": na, "delta": nb, "ms": round(ms, 4), "alg_bytes": 2 * m * w}
  for v in out.values():
    gbs = v["alg_bytes"] / (v["ms"] * 1e-3) / 1e9
    v["GB/s"] = round(gbs, 1)
    v["frac"] = round(gbs / HBM_PEAK_GBS, 4)
  # per-kernel split of one more call of each, from HIP events
  native.timing_reset()
  native.timing_enable(True)
  for op, (a, b) in (("sort_dedup", (n, 0)), ("set_difference", (n, 2 * n)),
                     ("merge", (2 * n, n // 2))):
    native.prim_bench(op, a, b, 2, 1)
  native.timing_enable(False)
  ks = native.timing_report()
  native.timing_reset()
  for k in ks:
    k["GB/s"] = round(k["bytes"] / (k["ms"] * 1e-3) / 1e9, 1) if k["ms"] > 0 else None
  out["kernels"] = sorted(ks, key=lambda k: -k["ms"])
  return out


def bench_batch(args) -> dict:
  """The uneven batch: 512 instances, even ones transitive closure over
  uniform random edges, odd ones same-generation over a random forest;
  n_i log-uniform in [500, 200000], mean out-degree U[1.5, 4].  Instances
  whose estimated fixpoint exceeds the per-instance cap are rejected (and
  listed).  The admitted ones run (a) batched: one packed database per
  program, one launch sequence, the two programs' batches concurrently
  (one host thread and stream each); (b) plain: one shim session per
  instance, one after the other.
  Both are timed end to end from host arrays to per-instance host results;
  every instance's tuples and per-iteration delta sizes must be identical."""
  import numpy as np

  from srdatalog_hip import batch, datasets

  t0 = time.time()
  raw = datasets.batch_instances(args.seed, n_instances=args.batch_instances)
  insts = [batch.Instance(d["kind"], d["n"], d["edges"]) for d in raw]
  gen_s = time.time() - t0
  t0 = time.time()
  ok, rejected = batch.admit(insts, cap_tuples=args.batch_cap, budget_bytes=args.batch_budget,
                             seed=args.seed)
  admit_s = time.time() - t0
  log(f"batch: {len(insts)} instances generated in {gen_s:.1f}s, admitted {len(ok)} "
      f"(rejected {len(rejected)}) in {admit_s:.1f}s")
  kinds = {k: [insts[i] for i in ok if insts[i].kind == k] for k in ("tc", "sg")}

  grains = {}
  for k, lst in kinds.items():
    sample = lst[::8] or lst
    g, times = batch.autotune_grain(k, sample)
    grains[k] = g
    log(f"batch: autotuned grain for {k}: {g} ({', '.join(f'{a}: {b * 1e3:.0f} ms' for a, b in times.items())})")

  batch.run_batched_kinds(kinds, grains)              # warm-up (pool growth)
  from srdatalog_hip import native
  p0 = native.pool_stats()
  t = time.perf_counter()
  got_b = batch.run_batched_kinds(kinds, grains)
  t_b = time.perf_counter() - t
  log(f"batch: batched {len(ok)} instances in {t_b:.2f}s")
  pm = native.pool_stats()
  t = time.perf_counter()
  got_p = {k: batch.run_plain(k, lst, grain=grains[k]) for k, lst in kinds.items()}
  t_p = time.perf_counter() - t
  log(f"batch: plain {len(ok)} instances in {t_p:.2f}s")
  p1 = native.pool_stats()
  pool_delta = {}
  for label, a, b in (("batched", p0, pm), ("plain", pm, p1)):
    pool_delta[label] = {k: b[k] - a[k] for k in ("driver_allocs", "reserved", "trims",
                                                  "cross_stream")}
    log(f"batch: device pool during the timed {label} run: " +
        ", ".join(f"{k} +{v}" for k, v in pool_delta[label].items()) +
        f"; reserved {b['reserved'] / 1e9:.1f} GB, cached {b['cached'] / 1e9:.1f} GB")
  log(f"batch: pool peak {p1['peak'] / 1e9:.1f} GB (reserved {p0['reserved'] / 1e9:.1f} GB, "
      f"cached {p0['cached'] / 1e9:.1f} GB before the timed runs)")

  # where the time goes: kernel time (HIP events) against wall time, for one
  # more batched run (its kinds one after the other) and for plain runs of a
  # 16-instance sample
  diag = {}
  pick = {k: list(range(0, len(v), max(1, len(v) // 8)))[:8] for k, v in kinds.items()}
  sample = {k: [kinds[k][j] for j in pick[k]] for k in kinds}
  late = {}
  for label, runner, sets in (("batched", batch.run_batched, kinds),
                              ("plain_sample16", batch.run_plain, sample)):
    native.timing_reset()
    native.timing_enable(True)
    t = time.perf_counter()
    for k, lst in sets.items():
      res = runner(k, lst, grain=grains[k])
      if label == "plain_sample16":
        late[k] = [r.session_s for r in res]
    wall = time.perf_counter() - t
    native.timing_enable(False)
    ks = native.timing_report()
    native.timing_reset()
    diag[label] = {"wall_s": round(wall, 4), "instances": sum(len(v) for v in sets.values()),
                   "kernel_ms": round(sum(k_["ms"] for k_ in ks), 2),
                   "launches": int(sum(k_["launches"] for k_ in ks)),
                   "top": [{"name": k_["name"], "ms": round(k_["ms"], 2), "launches": k_["launches"]}
                           for k_ in sorted(ks, key=lambda z: -z["ms"])[:6]]}
    log(f"batch: {label}: wall {wall:.2f}s, kernels {diag[label]['kernel_ms']:.0f} ms in "
        f"{diag[label]['launches']} launches")

  # the sampled instances' plain sessions in the timed plain run and again
  # after every session of the process (kernel timing on): equal times mean
  # the plain baseline does not slow down with the number of sessions
  early = {k: [got_p[k][j].session_s for j in pick[k]] for k in kinds}
  ratios = [b / a for k in kinds for a, b in zip(early[k], late[k]) if a > 0]
  diag["plain_sessions"] = {
    "early_s": {k: [round(x, 4) for x in v] for k, v in early.items()},
    "late_s": {k: [round(x, 4) for x in v] for k, v in late.items()},
    "median_late_over_early": round(float(np.median(ratios)), 3) if ratios else None}
  log(f"batch: plain sessions of the sample, late/early median "
      f"{diag['plain_sessions']['median_late_over_early']}")

  mismatch = 0
  tuples = 0
  for k in kinds:
    for b, p in zip(got_b[k], got_p[k]):
      tuples += b.rows.shape[0]
      if b.deltas != p.deltas or b.rows.shape != p.rows.shape or not np.array_equal(b.rows, p.rows):
        mismatch += 1
  n = len(ok)
  return {
    "metric": "instances/s (uneven TC/SG batch, batched vs plain per-instance)",
    "value": round(n / t_b, 3),
    "unit": "instances/s",
    "plain_value": round(n / t_p, 3),
    "speedup_vs_plain": round(t_p / t_b, 3),
    "batched_s": round(t_b, 4),
    "plain_s": round(t_p, 4),
    "admitted": n,
    "admitted_by_kind": {k: len(v) for k, v in kinds.items()},
    "rejected": len(rejected),
    "rejected_examples": [{"instance": i, "kind": insts[i].kind, "n": insts[i].n, "reason": r}
                          for i, r in rejected[:8]],
    "derived_tuples": int(tuples),
    "outputs_identical": mismatch == 0,
    "mismatched_instances": mismatch,
    "grain": grains,
    "admission_s": round(admit_s, 2),
    "pool_during_timed_runs": pool_delta,
    "time_split": diag,
  }


def bench_serve(args) -> dict:
  """Poisson arrivals against resident R-MAT graphs (edge factor 16), mix
  60% reach / 20% triangle count on a vertex range / 20% TC on an induced
  subgraph; served with continuous batching, then the same stream served
  serially by the plain path (bounded by --serve-plain-cap)."""
  import copy

  from srdatalog_hip import serve

  t0 = time.time()
  scales = tuple(int(x) for x in args.serve_scales.split(","))
  graphs = serve.make_graphs(args.seed, scales=scales)
  reqs = serve.make_requests(args.seed, graphs, rate=args.serve_rate, seconds=args.serve_seconds)
  log(f"serve: {len(graphs)} graphs ({', '.join(str(g.edges.shape[0]) for g in graphs)} edges) "
      f"and {len(reqs)} requests in {time.time() - t0:.1f}s")
  plain_reqs = copy.deepcopy(reqs)

  import threading
  last = [time.time()]
  plock = threading.Lock()

  def progress(kind, t, completed):
    with plock:
      if time.time() - last[0] < 15.0:
        return
      last[0] = time.time()
    log(f"serve: {completed} requests completed at t={t:.1f}s (last batch: {kind})")

  srv = serve.Server(graphs, budget_bytes=args.serve_budget, cap_tuples=args.serve_cap,
                     progress=progress, iterative=args.serve_tc_iterative)
  try:
    wall_s = srv.run(reqs, wall_cap=args.serve_wall_cap)
  finally:
    srv.close()
  served = serve.latency_stats(reqs)
  log(f"serve: served {served.get('completed')} in {wall_s:.1f}s, batches {srv.batches}")

  keep = [p for p, r in zip(plain_reqs, reqs) if not r.rejected]
  by_id = {r.rid: r for r in reqs}
  cmp = {"compared": 0, "mismatched": 0}

  def check(p):
    # compare with the served result while it is retained, then drop the
    # plain one: host memory holds one plain result at a time
    r = by_id[p.rid]
    if not r.result_dropped:
      cmp["compared"] += 1
      cmp["mismatched"] += not serve.same_result(r, p)
    if p.kind != "tri":
      p.result = None

  wall_p = serve.run_plain(keep, graphs, wall_cap=args.serve_plain_cap, on_result=check)
  plain = serve.latency_stats(keep)
  log(f"serve: plain served {plain.get('completed')} in {wall_p:.1f}s")
  compared, mismatched = cmp["compared"], cmp["mismatched"]
  dropped = sum(r.result_dropped for r in reqs)
  kinds = {}
  for k in serve.KIND_NAMES:
    kinds[k] = serve.latency_stats([r for r in reqs if r.kind == k])
  rejected = [r for r in reqs if r.rejected]
  return {
    "metric": "queries/s served (continuous batching) with p50/p99 latency",
    "value": round(served["completed"] / wall_s, 3),
    "unit": "queries/s",
    "p50_ms": served.get("p50_ms"), "p99_ms": served.get("p99_ms"),
    "plain_value": round(plain["completed"] / wall_p, 3) if plain.get("completed") else 0.0,
    "plain_p50_ms": plain.get("p50_ms"), "plain_p99_ms": plain.get("p99_ms"),
    "plain_completed": plain.get("completed", 0),
    "plain_unstarted": len(keep) - plain.get("completed", 0),
    "speedup_vs_plain": (round((served["completed"] / wall_s) / (plain["completed"] / wall_p), 3)
                         if plain.get("completed") else None),
    "requests": len(reqs), "completed": served["completed"], "rejected": len(rejected),
    "unserved": sum(1 for r in reqs if not r.rejected and r.t_done < 0),
    "served_wall_cap_s": args.serve_wall_cap,
    "rejected_examples": [{"rid": r.rid, "kind": r.kind, "graph_scale": graphs[r.graph].scale,
                           "reason": r.rejected} for r in rejected[:5]],
    "by_kind": kinds, "batches": srv.batches,
    "tc_admission": "iteration" if args.serve_tc_iterative else "batch",
    "results_compared": compared, "results_identical": mismatched == 0,
    "results_not_retained": dropped, "retained_result_gb": round(srv.kept_bytes / 1e9, 2),
    "served_wall_s": round(wall_s, 3), "plain_wall_s": round(wall_p, 3),
  }


def bench_fixpoint(args) -> dict:
  """Fixpoint wall time of a recursive program with resident inputs:
  tc = transitive closure of 5,000 distinct uniform pairs over 1,000
  vertices (self-loops dropped); andersen = the Andersen points-to program
  on 1/div of its full input sizes (200,000 variables and 50,000 objects at
  div 1; AddressOf 100k, Assign 400k, Load 100k, Store 100k pairs, Zipf 1.1
  endpoints).  Parity: every per-iteration delta and the final relation
  against the C++ oracle when it finishes in seconds (tc; andersen at
  div >= 30)."""
  import numpy as np

  from srdatalog_hip import datasets, programs
  if args.workload == "tc":
    inputs = {"ArcInput": datasets.tc_edges(args.seed, 1000, 5000)}
    prog, rel, arity, label = "tc", "Path", 2, "transitive closure, 5000 edges over 1000 vertices"
  else:
    d = args.andersen_div
    inputs = datasets.andersen_inputs(args.seed, n_vars=200_000 // d, n_objs=50_000 // d,
                                      sizes=(100_000 // d, 400_000 // d, 100_000 // d,
                                             100_000 // d))
    prog, rel, arity = "andersen", "PointsTo", 2
    label = f"Andersen points-to at 1/{d} of the full input sizes"
  lib = programs.load(prog)
  times = []
  for i in range(args.warmup + args.steps):
    if lib.init() != 0:
      raise RuntimeError("srdatalog_init failed")
    for r, rows in inputs.items():
      lib.load_rows(r, rows)
    lib.prepare()
    t = time.perf_counter()
    if lib.run(0) != 0:
      raise RuntimeError("srdatalog_run failed")
    dt = time.perf_counter() - t
    if i >= args.warmup:
      times.append(dt)
    if i == args.warmup + args.steps - 1:
      got = lib.export(rel, arity)
      deltas = lib.delta_sizes(rel)
    lib.shutdown()
  sec = float(np.median(times))
  # where the time goes: one more run with every launch timed (HIP events)
  from srdatalog_hip import native
  kern = None
  if not args.no_kernel_breakdown:
    kern = fixpoint_breakdown(lib, inputs, prog, native)
  parity = None
  digest = pair_hash(got) if arity == 2 else None
  if not args.no_parity and (args.workload == "tc" or args.andersen_div >= 30):
    import oracle
    want = oracle.run(prog, inputs)
    parity = bool(np.array_equal(got, want.relations[rel]) and deltas == want.deltas[rel])
  elif not args.no_parity and args.workload == "andersen":
    # sizes the oracle does not finish in seconds: the committed fixture of
    # the sparse-matrix restatement, when it holds this workload
    fx = os.path.join(ROOT, "tests", "golden", "andersen_scaled.json")
    case = next((c for c in json.load(open(fx)) if c["div"] == args.andersen_div and
                 c["seed"] == args.seed), None) if os.path.exists(fx) else None
    if case is not None:
      parity = bool(deltas == case["deltas"] and got.shape[0] == case["size"] and
                    str(digest) == case["hash"])
      log(f"parity vs the sparse restatement's fixture (1/{args.andersen_div}): {parity}")
  return {"metric": "fixpoint s + derived tuples/s", "value": round(got.shape[0] / sec, 1),
          "unit": "derived tuples/s", "fixpoint_s": round(sec, 4), "steps": args.steps,
          "warmup": args.warmup, "ms_per_step": round(sec * 1e3, 3), "n_gpus": 1,
          "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
          "data": "synthetic (seeded PCG64), generated in-process",
          "kernels": kern,
          # every timed session of the process, in order: later sessions are
          # not slower than the first
          "session_s": [round(x, 4) for x in times],
          "config": {"workload": label, "tuples": int(got.shape[0]), "iterations": len(deltas),
                     "deltas": deltas, "tuple_hash": None if digest is None else str(digest),
                     "parity_vs_oracle": parity}}


def pair_hash(rows) -> int:
  """Order-independent hash of a set of (a, b) u32 pairs: the sum mod 2^64
  of a mixed 64-bit key per pair (golden-ratio multiply, xor-shift 31), in
  chunks so a large result needs little extra host memory."""
  import numpy as np
  total = 0
  for lo in range(0, rows.shape[0], 1 << 26):
    part = rows[lo:lo + (1 << 26)]
    key = (part[:, 0].astype(np.uint64) << np.uint64(32)) | part[:, 1].astype(np.uint64)
    h = key * np.uint64(0x9E3779B97F4A7C15)
    h ^= h >> np.uint64(31)
    total = (total + int(h.sum(dtype=np.uint64))) % (1 << 64)
  return total


def fixpoint_breakdown(lib, inputs, prog, native) -> dict:
  """One more fixpoint with every launch timed (HIP events): kernel time
  and the heaviest kernels, against the run's wall time."""
  if lib.init() != 0:
    raise RuntimeError("srdatalog_init failed")
  for r, rows in inputs.items():
    lib.load_rows(r, rows)
  lib.prepare()
  native.timing_reset()
  native.timing_enable(True)
  t = time.perf_counter()
  lib.run(0)
  wall = time.perf_counter() - t
  native.timing_enable(False)
  ks = native.timing_report()
  native.timing_reset()
  lib.shutdown()
  kern = {"wall_s": round(wall, 4), "kernel_ms": round(sum(k["ms"] for k in ks), 2),
          "launches": int(sum(k["launches"] for k in ks)),
          "top": [{"name": k["name"], "ms": round(k["ms"], 2), "launches": k["launches"]}
                  for k in sorted(ks, key=lambda z: -z["ms"])[:8]]}
  log(f"{prog}: timed run wall {wall:.3f}s, kernels {kern['kernel_ms']:.1f} ms in "
      f"{kern['launches']} launches (device-resident iterations replayed from graphs are "
      "not timed per launch)")
  return kern


def main() -> int:
  args = parse()
  if args.gpus != 1:
    print(f"bench.py: this build runs on one GPU only (got --gpus {args.gpus}); "
          "multi-GPU sharding is designed but not part of this build", file=sys.stderr)
    return 2
  os.environ.setdefault("SRDATALOG_QUIET", "1")
  if args.workload == "batch":
    res = bench_batch(args)
    res.update({"n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": "u32", "data": "synthetic (seeded PCG64), generated in-process",
                "config": {"workload": f"uneven batch of {args.batch_instances} TC/SG instances, "
                                       "n log-uniform [500, 200000], mean out-degree U[1.5, 4]",
                           "cap_tuples": args.batch_cap, "budget_bytes": args.batch_budget}})
    print(json.dumps(res))
    return 0 if res["outputs_identical"] else 1
  if args.workload in ("tc", "andersen"):
    res = bench_fixpoint(args)
    print(json.dumps(res))
    return 0 if res["config"]["parity_vs_oracle"] is not False else 1
  if args.workload == "serve":
    res = bench_serve(args)
    res.update({"n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": "u32", "data": "synthetic R-MAT + Poisson stream (seeded), in-process",
                "config": {"workload": f"serving: Poisson {args.serve_rate} req/s for "
                                       f"{args.serve_seconds} s, R-MAT scales {args.serve_scales}, "
                                       "edge factor 16; 60% reach / 20% tri-count / 20% TC",
                           "budget_bytes": args.serve_budget, "cap_tuples": args.serve_cap}})
    print(json.dumps(res))
    return 0 if res["results_identical"] else 1

  out = bench_triangle(args)
  ok = out["config"]["parity_vs_oracle"] is not False
  if not args.no_batch:
    b = bench_batch(args)
    out["batch"] = b
    ok = ok and b["outputs_identical"]
  if not args.no_serve:
    sv = bench_serve(args)
    out["serve"] = sv
    ok = ok and sv["results_identical"]
  print(json.dumps(out))
  return 0 if ok else 1


def cpu_model() -> dict:
  model = ""
  try:
    for line in open("/proc/cpuinfo"):
      if line.startswith("model name"):
        model = line.split(":", 1)[1].strip()
        break
  except OSError:
    pass
  try:
    nproc = len(os.sched_getaffinity(0))
  except AttributeError:
    nproc = os.cpu_count() or 1
  return {"model": model, "nproc": nproc}


def cpu_baseline_triangle(edges, args) -> dict:
  """The oracle's triangle listing (join + sort/dedup of the listed rows) on
  root-key samples of the same graph: keys a with a % stride == 0 (an
  unbiased sample of the R-MAT key distribution), 1 thread and all threads,
  median of --cpu-reps runs each."""
  import numpy as np

  import oracle
  cm = cpu_model()
  res = {}
  for label, threads, stride in (("all_threads", cm["nproc"], args.cpu_stride_all),
                                 ("one_thread", 1, args.cpu_stride_one)):
    oracle.set_key_stride(stride)
    try:
      secs, n = [], 0
      for _ in range(args.cpu_reps):
        o = oracle.run("triangle", {"KnowsInput": edges}, threads=threads, relations=("Triangle",))
        secs.append(o.seconds)
        n = o.relations["Triangle"].shape[0]
        del o
    finally:
      oracle.set_key_stride(1)
    med = float(np.median(secs))
    res[label] = {"threads": threads, "key_stride": stride, "triangles": int(n),
                  "seconds_median": round(med, 4), "seconds": [round(x, 4) for x in secs],
                  "tuples_per_s": round(n / med, 1)}
    log(f"cpu baseline {label}: {n} triangles (keys a % {stride} == 0) in {med:.3f}s median "
        f"on {threads} threads")
  a = res["all_threads"]
  return {"value": a["tuples_per_s"], "unit": "derived tuples/s", "cores": a["threads"],
          "kind": "port",
          "sample": f"triangle listing (join + sort/dedup) of the same graph for root keys a with "
                    f"a % {a['key_stride']} == 0 ({a['triangles']} triangles), median of "
                    f"{args.cpu_reps}; one-thread figure on a % {args.cpu_stride_one} == 0",
          "one_thread": res["one_thread"], "all_threads": a,
          "cpu": cm}


def bench_triangle(args) -> dict:
  import numpy as np

  import oracle
  from srdatalog_hip import datasets, native, programs

  t0 = time.time()
  edges = datasets.triangle_edges(args.seed, args.scale, args.edges)
  # rows in a seeded random order: the Knows index build then really sorts
  edges = edges[np.random.default_rng(args.seed + 1).permutation(edges.shape[0])]
  gen_s = time.time() - t0
  log(f"generated {edges.shape[0]} edges in {gen_s:.1f}s")

  if args.calib_copy_bytes:
    ms = native.calib_copy(args.calib_copy_bytes, reps=1)
    log(f"calib copy {args.calib_copy_bytes} B: {ms:.3f} ms "
        f"({2 * args.calib_copy_bytes / ms / 1e6:.1f} GB/s)")

  lib = programs.load("triangle")
  if lib.init() != 0:
    raise RuntimeError("srdatalog_init failed")
  lib.load_rows("KnowsInput", edges)
  lib.prepare()                              # inputs resident in HBM before timing

  for i in range(args.warmup):
    if lib.run(0) != 0:
      raise RuntimeError("srdatalog_run failed")
    log(f"warmup {i} done")

  native.timing_reset()
  native.timing_enable(True)
  step_s = []
  for _ in range(args.steps):
    t = time.perf_counter()
    rc = lib.run(0)                          # synchronises the stream before returning
    step_s.append(time.perf_counter() - t)
    if rc != 0:
      raise RuntimeError("srdatalog_run failed")
    log(f"step {len(step_s) - 1}: {step_s[-1] * 1e3:.2f} ms")
  native.timing_enable(False)
  kernels = native.timing_report()

  n_tri = lib.size("Triangle")
  n_knows = lib.size("Knows")
  mean_s = float(np.mean(step_s))
  value = n_tri / mean_s
  pool = native.pool_stats()

  # Dominant kernel and its algorithmic bytes per launch.
  dom = max(kernels, key=lambda k: k["ms"]) if kernels else None
  roofline = None
  if dom is not None:
    avg_ms = dom["ms"] / max(dom["launches"], 1)
    if dom["name"].startswith("join."):
      # streaming floor: every source handle's rows once (3 Knows handles x
      # 8 B/row) + the output rows once (|Triangle| x 12 B) for every mode
      # that materialises them (fused, write, direct; not count)
      writes = not dom["name"].startswith(("join.count", "join.root", "join.weigh"))
      alg = 3 * n_knows * 8 + (n_tri * 12 if writes else 0)
    else:
      alg = dom["bytes"] / max(dom["launches"], 1)
    achieved = alg / (avg_ms * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
      traffic = json.load(open(tfile)).get(dom["name"])
    roofline = {"bound": "hbm", "kernel": dom["name"], "achieved": round(achieved, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": traffic, "avg_ms": round(avg_ms, 4), "alg_bytes": alg}

  # join probes: one more, untimed run counting the searches, root-map
  # lookups and bitmap tests (an extra probe-mode launch per pipeline)
  lib.count_probes(True)
  if lib.run(0) != 0:
    raise RuntimeError("srdatalog_run failed (probe run)")
  probes = lib.join_probes()
  lib.count_probes(False)
  join_ms = sum(k["ms"] for k in kernels if k["name"].startswith("join.fused")
                or k["name"].startswith("join.write") or k["name"].startswith("join.count")
                or k["name"].startswith("join.direct")) / max(args.steps, 1)
  jp = {"per_run": int(probes),
        "per_s_join_kernel": round(probes / (join_ms * 1e-3), 1) if join_ms > 0 else None,
        "per_s_fixpoint": round(probes / mean_s, 1),
        "join_kernel_ms": round(join_ms, 4),
        "definition": "lower/upper-bound searches, root-map lookups and heavy-key bitmap tests "
                      "of the join's root and nested levels, counted by the probe mode"}
  log(f"join probes: {probes} per run ({jp['per_s_join_kernel']} /s of join kernel)")

  prims = None
  if args.prim_rows:
    prims = primitives(native, args.prim_rows, reps=max(args.steps, 1))
    log("primitives: " + ", ".join(f"{k} {v['GB/s']} GB/s" for k, v in prims.items()
                                    if k != "kernels"))

  parity = None
  digest = None
  if not args.no_parity:
    want, orc_s = oracle.triangle_digest(edges)
    rows = lib.export("Triangle", 3)
    got = oracle.digest_rows(rows)
    sorted_ok = bool(rows.shape[0] < 2 or np.all(np.any(np.diff(rows[:1 << 20].astype(np.int64),
                                                                 axis=0) != 0, axis=1)))
    del rows
    parity = bool(got == want and sorted_ok and n_tri == want[0])
    digest = {"device": [str(x) for x in got], "oracle": [str(x) for x in want],
              "oracle_enumeration_s": round(orc_s, 3)}
    log(f"parity: device digest {got} vs oracle {want}: {parity}")
  lib.shutdown()

  count_only = None
  if not args.no_count_only:
    clib = programs.load("triangle_count")
    if clib.init() != 0:
      raise RuntimeError("srdatalog_init failed (triangle_count)")
    clib.load_rows("KnowsInput", edges)
    clib.prepare()
    for _ in range(args.warmup):
      clib.run(0)
    cs = []
    for _ in range(args.steps):
      t = time.perf_counter()
      if clib.run(0) != 0:
        raise RuntimeError("srdatalog_run failed (triangle_count)")
      cs.append(time.perf_counter() - t)
    cnt = clib.size("Triangle")
    clib.shutdown()
    cm = float(np.mean(cs))
    count_only = {"ms_per_step": round(cm * 1e3, 3), "ms_median": round(float(np.median(cs)) * 1e3, 3),
                  "triangles": int(cnt), "triangles_per_s": round(cnt / cm, 1),
                  "matches_listing": cnt == n_tri}
    log(f"count-only: {cnt} triangles in {cm * 1e3:.2f} ms mean")

  cpu = None
  if not args.no_cpu_baseline:
    cpu = cpu_baseline_triangle(edges, args)

  med = float(np.median(step_s))
  return {
    "metric": "derived tuples/s (triangle listing fixpoint)",
    "value": round(value, 1),
    "unit": "derived tuples/s",
    "n_gpus": 1,
    "steps": args.steps,
    "warmup": args.warmup,
    "ms_per_step": round(mean_s * 1e3, 3),
    "higher_is_better": True,
    "scaling": "weak",
    "vs_baseline": None,
    "dtype": "u32",
    "data": "synthetic R-MAT (seeded PCG64), generated in-process",
    "config": {"workload": f"triangle listing, R-MAT scale {args.scale}, "
                           f"{args.edges} generated edges, (0.57,0.19,0.19,0.05), "
                           "dedup + no self-loops + oriented low->high",
               "edges": int(edges.shape[0]), "triangles": int(n_tri),
               "fixpoint_s": round(mean_s, 5), "fixpoint_s_median": round(med, 5),
               "step_ms_max": round(max(step_s) * 1e3, 3),
               "parity_vs_oracle": parity, "parity_digest": digest,
               "gen_s": round(gen_s, 2)},
    "roofline": roofline,
    "join_probes": jp,
    "count_only": count_only,
    "cpu_baseline": cpu,
    "pool": pool,
    "primitives": prims,
    "kernels": sorted(kernels, key=lambda k: -k["ms"])[:12],
  }


if __name__ == "__main__":
  raise SystemExit(main())
