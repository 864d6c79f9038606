"""The merge-path merge / set-difference tiles (mp_partition_kernel and
mp_tile_kernel in csrc/srd_prims.hip), modelled tile by tile and thread by
thread with small tiles so that tile boundaries fall everywhere, and
checked against numpy's set operations: ties broken B-first, the staged
predecessor B[j0-1], per-thread sub-splits and the keep rule."""

import numpy as np
import pytest


def split(a, b, d):
  lo, hi = max(0, d - len(b)), min(d, len(a))
  while lo < hi:
    i = (lo + hi + 1) >> 1
    if a[i - 1] < b[d - i]:
      lo = i
    else:
      hi = i - 1
  return lo


def model(a, b, mode, tile=16, items=4):
  na, nb = len(a), len(b)
  tiles = (na + nb + tile - 1) // tile
  sp = [split(a, b, min(t * tile, na + nb)) for t in range(tiles + 1)]
  out = []
  for t in range(tiles):
    d0, d1 = t * tile, min(t * tile + tile, na + nb)
    i0, i1 = sp[t], sp[t + 1]
    j0, j1 = d0 - i0, d1 - i1
    an, bn = i1 - i0, j1 - j0
    pre = 1 if (mode == 0 and j0 > 0) else 0
    st = list(a[i0:i1]) + ([b[j0 - 1]] if pre else []) + list(b[j0:j1])
    bb = an + pre
    tn = an + bn
    tile_rows = []
    for tid in range(tile // items):
      p0 = min(tid * items, tn)
      lo, hi = max(0, p0 - bn), min(p0, an)
      while lo < hi:
        i = (lo + hi + 1) >> 1
        if st[i - 1] < st[bb + (p0 - i)]:
          lo = i
        else:
          hi = i - 1
      ai, bi = lo, p0 - lo
      for it in range(items):
        if p0 + it >= tn:
          break
        take_a = bi >= bn or (ai < an and st[ai] < st[bb + bi])
        if take_a:
          x = st[ai]
          if mode == 1:
            tile_rows.append(x)
          else:
            has_prev = bi > 0 or pre
            if not (has_prev and st[bb + bi - 1] == x):
              tile_rows.append(x)
          ai += 1
        else:
          if mode == 1:
            tile_rows.append(st[bb + bi])
          bi += 1
    out += tile_rows
  return np.array(out, dtype=np.int64)


@pytest.mark.parametrize("seed", range(60))
def test_merge_path_set_difference_and_merge(seed):
  rng = np.random.default_rng(seed)
  u = int(rng.integers(5, 400))
  a = np.unique(rng.integers(0, u, int(rng.integers(0, 120))))
  b = np.unique(rng.integers(0, u, int(rng.integers(0, 120))))
  np.testing.assert_array_equal(model(a, b, 0), np.setdiff1d(a, b))
  d = np.setdiff1d(b, a)   # disjoint from a
  np.testing.assert_array_equal(model(a, d, 1), np.union1d(a, d))


def test_all_equal_and_nested():
  a = np.arange(0, 200, 2)
  np.testing.assert_array_equal(model(a, a, 0), np.zeros(0))
  b = np.arange(0, 200, 6)
  np.testing.assert_array_equal(model(a, b, 0), np.setdiff1d(a, b))
  np.testing.assert_array_equal(model(b, a, 0), np.zeros(0))
