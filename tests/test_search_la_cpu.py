"""The two-steps-per-round-trip binary search (srd::bound_la in
csrc/srd_device.h), modelled line by line and checked against bisect on
random sorted arrays, ranges and probe values (present, absent, out of
range, duplicates)."""

import bisect

import numpy as np


def bound_la(a, lo, hi, v, upper):
  less = (lambda x: x <= v) if upper else (lambda x: x < v)
  n = hi - lo
  while n > 1:
    half = n >> 1
    mid = lo + half
    nl, nr = half, n - half - 1
    ml, mr = lo + (nl >> 1), mid + 1 + (nr >> 1)
    x, xl = a[mid], a[ml]
    xr = a[mr] if nr else 0
    if less(x):
      lo, n = mid + 1, nr
      if n > 0:
        h2 = n >> 1
        if less(xr):
          lo, n = mr + 1, n - h2 - 1
        else:
          n = h2
    else:
      n = nl
      h2 = n >> 1
      if less(xl):
        lo, n = ml + 1, n - h2 - 1
      else:
        n = h2
  if n == 1 and less(a[lo]):
    lo += 1
  return lo


def test_bound_la_matches_bisect():
  rng = np.random.default_rng(0)
  for _ in range(6000):
    a = sorted(rng.integers(0, 300, int(rng.integers(1, 400))).tolist())
    lo = int(rng.integers(0, len(a) + 1))
    hi = int(rng.integers(lo, len(a) + 1))
    v = int(rng.integers(-3, 55))
    assert bound_la(a, lo, hi, v, False) == bisect.bisect_left(a, v, lo, hi)
    assert bound_la(a, lo, hi, v, True) == bisect.bisect_right(a, v, lo, hi)
