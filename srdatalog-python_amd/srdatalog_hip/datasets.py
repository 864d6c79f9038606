"""Seeded synthetic inputs for the five benchmark workloads.

Every generator uses numpy's PCG64 (``np.random.default_rng(seed)``) so the
GPU engine, the CPU oracle and the CPU baseline all consume the identical
tuples.  All columns are ``uint32``; relations are returned as ``(n, arity)``
C-contiguous arrays of distinct rows in *draw order* (not sorted), which is
what a CSV written by a user would look like.

Workloads (sizes as specified for the benchmark):

* ``tc_edges``        -- uniform random directed pairs over ``n`` vertices,
                         self-loops dropped, distinct, exactly ``m`` pairs.
* ``rmat_edges``      -- R-MAT (a,b,c,d) with ``scale`` quadrant draws per edge.
* ``triangle_edges``  -- R-MAT, self-loops dropped, oriented (min,max), distinct.
* ``andersen_inputs`` -- AddressOf / Assign / Load / Store with Zipf(1.1)
                         endpoints in disjoint variable / object id ranges.
* ``batch_instances`` -- the uneven 512-instance TC / same-generation batch.
"""

from __future__ import annotations

import numpy as np

U32 = np.uint32


def _distinct_in_order(keys: np.ndarray) -> np.ndarray:
  """Indices of the first occurrence of every distinct key, in draw order."""
  _, first = np.unique(keys, return_index=True)
  first.sort()
  return first


def _pairs_to_rows(src: np.ndarray, dst: np.ndarray) -> np.ndarray:
  out = np.empty((src.shape[0], 2), dtype=U32)
  out[:, 0] = src
  out[:, 1] = dst
  return out


def tc_edges(seed: int = 0, n_vertices: int = 1000, n_edges: int = 5000) -> np.ndarray:
  """Uniform random distinct directed pairs, no self-loops.

  Pairs are drawn in batches, self-loops dropped, and the first ``n_edges``
  distinct pairs (in draw order) are kept.
  """
  if n_edges > n_vertices * (n_vertices - 1):
    raise ValueError("more edges requested than distinct non-loop pairs exist")
  rng = np.random.default_rng(seed)
  keys = np.empty(0, dtype=np.uint64)
  while True:
    batch = max(2 * (n_edges - keys.shape[0]), 1024)
    u = rng.integers(0, n_vertices, size=batch, dtype=np.uint64)
    v = rng.integers(0, n_vertices, size=batch, dtype=np.uint64)
    keep = u != v
    keys = np.concatenate([keys, (u[keep] << np.uint64(32)) | v[keep]])
    first = _distinct_in_order(keys)
    if first.shape[0] >= n_edges:
      keys = keys[first[:n_edges]]
      break
    keys = keys[first]
  return _pairs_to_rows((keys >> np.uint64(32)).astype(U32), (keys & np.uint64(0xFFFFFFFF)).astype(U32))


def rmat_edges(
  seed: int,
  scale: int,
  n_edges: int,
  abcd: tuple[float, float, float, float] = (0.57, 0.19, 0.19, 0.05),
  chunk: int = 1 << 22,
) -> np.ndarray:
  """Raw R-MAT edge list (with duplicates and self-loops), ``(n_edges, 2)`` u32.

  Each edge takes ``scale`` independent quadrant draws; quadrant b sets the
  destination bit, c the source bit, d both.  No noise is applied.
  """
  a, b, c, d = abcd
  if abs(a + b + c + d - 1.0) > 1e-9:
    raise ValueError("R-MAT probabilities must sum to 1")
  rng = np.random.default_rng(seed)
  out = np.empty((n_edges, 2), dtype=U32)
  t_ab = a + b
  t_abc = a + b + c
  for lo in range(0, n_edges, chunk):
    n = min(chunk, n_edges - lo)
    src = np.zeros(n, dtype=U32)
    dst = np.zeros(n, dtype=U32)
    for level in range(scale):
      r = rng.random(n)
      bit = U32(1 << (scale - 1 - level))
      src_bit = r >= t_ab                       # quadrants c, d
      dst_bit = ((r >= a) & (r < t_ab)) | (r >= t_abc)   # quadrants b, d
      src |= np.where(src_bit, bit, U32(0)).astype(U32)
      dst |= np.where(dst_bit, bit, U32(0)).astype(U32)
    out[lo:lo + n, 0] = src
    out[lo:lo + n, 1] = dst
  return out


def orient_dedup(edges: np.ndarray) -> np.ndarray:
  """Drop self-loops, orient every edge low->high id, keep distinct rows (sorted)."""
  u = edges[:, 0].astype(np.uint64)
  v = edges[:, 1].astype(np.uint64)
  keep = u != v
  lo = np.minimum(u[keep], v[keep])
  hi = np.maximum(u[keep], v[keep])
  keys = np.unique((lo << np.uint64(32)) | hi)
  return _pairs_to_rows((keys >> np.uint64(32)).astype(U32), (keys & np.uint64(0xFFFFFFFF)).astype(U32))


def triangle_edges(seed: int = 0, scale: int = 20, n_edges: int = 1 << 24) -> np.ndarray:
  """R-MAT graph for triangle listing: deduplicated, loop-free, oriented low->high."""
  return orient_dedup(rmat_edges(seed, scale, n_edges))


def _zipf_ranks(rng: np.random.Generator, s: float, n_max: int, size: int) -> np.ndarray:
  """Zipf(s) ranks in [1, n_max], truncated by rejection."""
  out = np.empty(0, dtype=np.int64)
  while out.shape[0] < size:
    draw = rng.zipf(s, size=2 * (size - out.shape[0]) + 64)
    out = np.concatenate([out, draw[draw <= n_max]])
  return out[:size]


def _zipf_pairs(rng, s, n_left, map_left, n_right, map_right, count):
  keys = np.empty(0, dtype=np.uint64)
  while True:
    need = count - keys.shape[0]
    batch = max(2 * need, 4096)
    left = map_left[_zipf_ranks(rng, s, n_left, batch) - 1].astype(np.uint64)
    right = map_right[_zipf_ranks(rng, s, n_right, batch) - 1].astype(np.uint64)
    keys = np.concatenate([keys, (left << np.uint64(32)) | right])
    first = _distinct_in_order(keys)
    if first.shape[0] >= count:
      keys = keys[first[:count]]
      break
    keys = keys[first]
  return _pairs_to_rows((keys >> np.uint64(32)).astype(U32), (keys & np.uint64(0xFFFFFFFF)).astype(U32))


def andersen_inputs(
  seed: int = 0,
  n_vars: int = 200_000,
  n_objs: int = 50_000,
  sizes: tuple[int, int, int, int] = (100_000, 400_000, 100_000, 100_000),
  s: float = 1.1,
) -> dict[str, np.ndarray]:
  """Andersen points-to inputs with Zipf(s) endpoints.

  Variables are ids ``[0, n_vars)``, objects ``[n_vars, n_vars + n_objs)``.
  Endpoint ranks are drawn Zipf(s), truncated to the range by rejection and
  mapped through a seeded permutation of the range, so popular ids are
  scattered.  Each relation holds exactly the requested number of distinct
  pairs.
  """
  rng = np.random.default_rng(seed)
  var_map = rng.permutation(n_vars).astype(U32)
  obj_map = (rng.permutation(n_objs) + n_vars).astype(U32)
  n_addr, n_assign, n_load, n_store = sizes
  return {
    "AddressOf": _zipf_pairs(rng, s, n_vars, var_map, n_objs, obj_map, n_addr),
    "Assign": _zipf_pairs(rng, s, n_vars, var_map, n_vars, var_map, n_assign),
    "Load": _zipf_pairs(rng, s, n_vars, var_map, n_vars, var_map, n_load),
    "Store": _zipf_pairs(rng, s, n_vars, var_map, n_vars, var_map, n_store),
  }


def sg_forest_edges(rng: np.random.Generator, n: int, mean_children: float = 3.0,
                    cross_frac: float = 0.02) -> np.ndarray:
  """Galton-Watson forest (Geometric children, given mean) plus random cross edges.

  Trees grow breadth-first with consecutive node ids: the children of the
  frontier nodes ``[lo, hi)`` get the next ids in frontier order.  A new root
  starts whenever the forest dies out, until ``n`` nodes exist.  Tree edges
  point parent -> child.  ``round(cross_frac * n_tree_edges)`` extra uniform
  non-loop edges are appended and duplicates dropped (draw order kept).
  """
  p = 1.0 / (mean_children + 1.0)       # Geometric on {0,1,2,...} with this mean
  srcs, dsts = [], []
  made = 1
  lo, hi = 0, 1                         # frontier: nodes [lo, hi) await children
  while made < n:
    if lo == hi:                        # extinct: start a new root
      lo, hi = made, made + 1
      made += 1
      continue
    kids = rng.geometric(p, size=hi - lo) - 1
    total = min(int(kids.sum()), n - made)
    if total > 0:
      srcs.append(np.repeat(np.arange(lo, hi, dtype=np.int64), kids)[:total])
      dsts.append(np.arange(made, made + total, dtype=np.int64))
    lo, hi = hi, made + total
    made += total
  src = np.concatenate(srcs) if srcs else np.empty(0, dtype=np.int64)
  dst = np.concatenate(dsts) if dsts else np.empty(0, dtype=np.int64)
  m_cross = int(round(cross_frac * src.shape[0]))
  if m_cross > 0:
    cu = rng.integers(0, n, size=2 * m_cross + 16)
    cv = rng.integers(0, n, size=2 * m_cross + 16)
    keep = cu != cv
    src = np.concatenate([src, cu[keep][:m_cross]])
    dst = np.concatenate([dst, cv[keep][:m_cross]])
  keys = (src.astype(np.uint64) << np.uint64(32)) | dst.astype(np.uint64)
  keys = keys[_distinct_in_order(keys)]
  return _pairs_to_rows((keys >> np.uint64(32)).astype(U32), (keys & np.uint64(0xFFFFFFFF)).astype(U32))


def batch_instances(seed: int = 0, n_instances: int = 512, n_lo: int = 500,
                    n_hi: int = 200_000, deg_lo: float = 1.5, deg_hi: float = 4.0):
  """The uneven TC / same-generation batch.

  Instance ``i`` has ``n_i = round(exp(U[ln n_lo, ln n_hi]))`` vertices and
  mean out-degree ``d_i ~ U[deg_lo, deg_hi]``.  Even ``i`` are transitive
  closure over uniform random edges (``round(n_i * d_i)`` distinct non-loop
  pairs); odd ``i`` are same-generation over a Galton-Watson forest with
  mean 3 children plus 2 % cross edges.  Returns a list of dicts
  ``{"kind", "n", "deg", "edges"}``.
  """
  rng = np.random.default_rng(seed)
  ns = np.rint(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi), size=n_instances))).astype(np.int64)
  degs = rng.uniform(deg_lo, deg_hi, size=n_instances)
  out = []
  for i in range(n_instances):
    sub_seed = int(rng.integers(0, 2**63 - 1))
    n = int(ns[i])
    if i % 2 == 0:
      m = int(round(n * degs[i]))
      out.append({"kind": "tc", "n": n, "deg": float(degs[i]),
                  "edges": tc_edges(sub_seed, n, m)})
    else:
      sub = np.random.default_rng(sub_seed)
      out.append({"kind": "sg", "n": n, "deg": float(degs[i]),
                  "edges": sg_forest_edges(sub, n)})
  return out


def write_csv(path: str, rows: np.ndarray, delimiter: str = "\t") -> None:
  """Write rows as integer CSV (tab-delimited by default), one tuple per line."""
  np.savetxt(path, rows, fmt="%d", delimiter=delimiter)
