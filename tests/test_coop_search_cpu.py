"""The wave-cooperative lower bound used by the generated join kernels
(srdj::seg_lower_bound in csrc/srd_join.h), modelled lane by lane and
checked against numpy.searchsorted on random run layouts: runs of equal
segment id with non-decreasing values, inactive lanes in between, singleton
runs, duplicates and values outside the range."""

import numpy as np


def _lower(a, lo, n, v, upper=False):
  while n > 0:
    half = n >> 1
    mid = lo + half
    x = a[mid]
    if (x <= v) if upper else (x < v):
      lo = mid + 1
      n -= half + 1
    else:
      n = half
  return lo


def seg_lower_bound_model(col, b, e, v, on, seg):
  """64-lane model of the device function (b, e, v, on, seg: per lane)."""
  W = 64
  first = [on[l] and not (l > 0 and on[l - 1] and seg[l - 1] == seg[l]) for l in range(W)]
  last = [on[l] and not (l < W - 1 and on[l + 1] and seg[l + 1] == seg[l]) for l in range(W)]
  lo = [0] * W
  for l in range(W):
    n = (e[l] - b[l]) if (first[l] or last[l]) else 0
    lo[l] = _lower(col, b[l], n, v[l], upper=last[l] and not first[l])
  out = [0] * W
  for l in range(W):
    if not on[l] or first[l]:
      out[l] = lo[l]
      continue
    src_f = max(i for i in range(l + 1) if first[i])
    src_l = min(i for i in range(l, W) if last[i])
    w_lo, w_hi = lo[src_f], lo[src_l]
    out[l] = _lower(col, w_lo, w_hi - w_lo, v[l])
  return out


def test_seg_lower_bound_model_random():
  rng = np.random.default_rng(0)
  for trial in range(400):
    col = np.sort(rng.integers(0, 200, size=int(rng.integers(1, 300))))
    n = col.size
    # lanes: a few runs, each over its own [b, e) with sorted values
    b, e, v, on, seg = [0] * 64, [0] * 64, [0] * 64, [False] * 64, [0] * 64
    lane = 0
    sid = 0
    while lane < 64:
      length = int(rng.integers(1, 20))
      lo_r = int(rng.integers(0, n + 1))
      hi_r = int(rng.integers(lo_r, n + 1))
      vals = np.sort(rng.integers(-5, 210, size=length))
      for j in range(length):
        if lane >= 64:
          break
        b[lane], e[lane], seg[lane] = lo_r, hi_r, sid
        v[lane] = int(vals[j])
        on[lane] = bool(rng.random() < 0.85)
        lane += 1
      sid += 1
    got = seg_lower_bound_model(col, b, e, v, on, seg)
    for l in range(64):
      if not on[l]:
        continue
      want = b[l] + int(np.searchsorted(col[b[l]:e[l]], v[l], side="left"))
      assert got[l] == want, (trial, l)


def test_one_run_over_all_lanes():
  col = np.arange(0, 100, 2)
  b = [10] * 64
  e = [40] * 64
  v = list(range(15, 15 + 64))
  on = [True] * 64
  seg = [7] * 64
  got = seg_lower_bound_model(col, b, e, v, on, seg)
  for l in range(64):
    assert got[l] == 10 + int(np.searchsorted(col[10:40], v[l]))
