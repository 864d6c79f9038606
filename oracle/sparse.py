"""Andersen points-to fixpoint as boolean sparse-matrix products (scipy).
Test infrastructure only: the checker for scales the C++ oracle cannot
reach in seconds, never the thing measured or shipped.

Same semantics as oracle.cpp / brute.py (Jacobi semi-naive over the four
rules of the Andersen program, src/srdatalog examples' andersen shape with
the rules as stated for the Andersen workload):

  PointsTo(v,o) <= AddressOf(v,o)
  PointsTo(v,o) <= Assign(v,w), PointsTo(w,o)
  PointsTo(x,o) <= Load(x,y), PointsTo(y,a), PointsTo(a,o)
  PointsTo(a,o) <= Store(x,y), PointsTo(x,a), PointsTo(y,o)

With P the 0/1 matrix of PointsTo (row = first column), one application of
the rules is T(P) = Assign.P + Load.P.P + P^T.Store.P (boolean products);
F_0 = AddressOf, Delta_i = T(F_i) - F_i, F_{i+1} = F_i + Delta_i, stopping
after the first empty Delta -- the delta sequence of the reference's loop
(iteration 0 reads FULL for DELTA; see the oracle.cpp header for the
reference citations).  Only the sets matter here, so the products need not
split into the semi-naive rule versions: every tuple of T(F_i) - F_i is
derived by some version that reads a Delta_{i-1} tuple, and conversely.
Pinned by agreement with the C++ oracle (tests/test_oracle_sparse.py).
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def tuple_hash(rows: np.ndarray) -> int:
  """Order-independent hash of a set of (a, b) u32 pairs: the sum mod 2^64
  of a mixed 64-bit key per pair (golden-ratio multiply, xor-shift 31)."""
  rows = np.asarray(rows)
  if rows.size == 0:
    return 0
  key = (rows[:, 0].astype(np.uint64) << np.uint64(32)) | rows[:, 1].astype(np.uint64)
  h = key * np.uint64(0x9E3779B97F4A7C15)
  h ^= h >> np.uint64(31)
  return int(h.sum(dtype=np.uint64))


def _mat(rows: np.ndarray, n: int) -> sp.csr_matrix:
  rows = np.asarray(rows, dtype=np.int64).reshape(-1, 2)
  m = sp.csr_matrix((np.ones(len(rows), np.int8), (rows[:, 0], rows[:, 1])), shape=(n, n))
  m.sum_duplicates()
  return _bool(m)


def _bool(m) -> sp.csr_matrix:
  m = sp.csr_matrix(m)
  m.data = np.ones_like(m.data, dtype=np.int8)
  m.eliminate_zeros()
  return m


def _prod(a: sp.csr_matrix, b: sp.csr_matrix) -> sp.csr_matrix:
  # int64 accumulation so that path counts cannot wrap to zero
  return _bool(a.astype(np.int64) @ b.astype(np.int64))


def andersen(addr, assign, load, store):
  """Returns (PointsTo rows sorted lexicographically, per-iteration |Delta|)."""
  all_rows = [np.asarray(x).reshape(-1, 2) for x in (addr, assign, load, store)]
  n = int(max((int(r.max()) for r in all_rows if r.size), default=0)) + 1
  A, L, S = _mat(assign, n), _mat(load, n), _mat(store, n)
  F = _mat(addr, n)
  deltas = []
  while True:
    T = _prod(A, F) + _prod(_prod(L, F), F) + _prod(_prod(F.T.tocsr(), S), F)
    T = _bool(T)
    new = _bool(T - T.multiply(F))
    deltas.append(int(new.nnz))
    if new.nnz == 0:
      break
    F = _bool(F + new)
  c = F.tocoo()
  rows = np.stack([c.row, c.col], 1).astype(np.uint32)
  rows = rows[np.lexsort((rows[:, 1], rows[:, 0]))]
  return rows, deltas
