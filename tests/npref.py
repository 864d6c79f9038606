"""numpy set semantics used as the expected values of primitive parity tests.

These are definitions (lexicographic order, set difference, disjoint union),
independent of the device code paths they check.
"""

import numpy as np


def sort_unique(rows: np.ndarray) -> np.ndarray:
  rows = np.asarray(rows, dtype=np.uint32)
  if rows.ndim == 1:
    rows = rows.reshape(-1, 1)
  if rows.shape[0] == 0:
    return rows.copy()
  idx = np.lexsort(rows.T[::-1])
  s = rows[idx]
  keep = np.ones(s.shape[0], dtype=bool)
  keep[1:] = np.any(s[1:] != s[:-1], axis=1)
  return s[keep]


def roots(sorted_unique: np.ndarray):
  c0 = sorted_unique[:, 0]
  if c0.shape[0] == 0:
    return c0[:0], np.zeros(0, dtype=np.uint32)
  starts = np.flatnonzero(np.concatenate([[True], c0[1:] != c0[:-1]]))
  offs = np.concatenate([starts, [c0.shape[0]]]).astype(np.uint32)
  return c0[starts], offs


def _void(rows):
  rows = np.ascontiguousarray(np.asarray(rows, dtype=">u4"))
  return rows.view(np.dtype((np.void, rows.dtype.itemsize * rows.shape[1]))).ravel()


def set_difference(a, b):
  a = np.asarray(a, dtype=np.uint32)
  if a.shape[0] == 0 or np.asarray(b).shape[0] == 0:
    return a.copy()
  mask = ~np.isin(_void(a), _void(b))
  return a[mask]


def union_sorted(a, b):
  return sort_unique(np.concatenate([np.asarray(a, np.uint32), np.asarray(b, np.uint32)]))
