"""The device binary searches of csrc/srd_device.h -- srd::lower_bound /
upper_bound and the two-steps-per-round-trip srd::lower_bound_la /
upper_bound_la -- compiled for the host from the same source (g++ through
tests/cpu_sim's stand-in HIP header) and checked against bisect on random
sorted arrays, sub-ranges and probe values (present, absent, below and above
the range, long runs of duplicates)."""

import bisect
import ctypes
import os

import numpy as np
import pytest

from conftest import build_cpu_sim

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpu_sim")
_u32p = ctypes.POINTER(ctypes.c_uint32)


@pytest.fixture(scope="module")
def lib():
  lib = ctypes.CDLL(build_cpu_sim("libsearch_host.so"))
  for f in (lib.host_bound_la, lib.host_bound):
    f.restype = ctypes.c_uint32
    f.argtypes = [_u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]
  return lib


@pytest.mark.parametrize("fn", ["host_bound_la", "host_bound"])
def test_bounds_match_bisect(lib, fn):
  f = getattr(lib, fn)
  rng = np.random.default_rng(0)
  for _ in range(20000):
    n = int(rng.integers(1, 400))
    top = int(rng.choice([3, 50, 300, 1 << 30]))
    a = np.sort(rng.integers(0, top, n)).astype(np.uint32)
    lo = int(rng.integers(0, n + 1))
    hi = int(rng.integers(lo, n + 1))
    v = int(rng.integers(0, top + 3))
    al = a.tolist()
    p = a.ctypes.data_as(_u32p)
    assert f(p, lo, hi, v, 0) == bisect.bisect_left(al, v, lo, hi), (fn, al, lo, hi, v)
    assert f(p, lo, hi, v, 1) == bisect.bisect_right(al, v, lo, hi), (fn, al, lo, hi, v)


def test_bounds_every_length_and_probe(lib):
  """Every range length up to 70 and every probe around every element."""
  a = (np.arange(70, dtype=np.uint32) // 2) * 3 + 1   # pairs of duplicates
  al = a.tolist()
  p = a.ctypes.data_as(_u32p)
  for lo in range(0, 71, 7):
    for hi in range(lo, 71):
      for v in range(0, 110):
        assert lib.host_bound_la(p, lo, hi, v, 0) == bisect.bisect_left(al, v, lo, hi)
        assert lib.host_bound_la(p, lo, hi, v, 1) == bisect.bisect_right(al, v, lo, hi)


def test_find_slot_and_gallop_match_bisect(lib):
  """srd::find_slot (bisection) and srd::find_slot_from (gallop from a slot
  known to be at or below the answer) -- the join kernels' item -> key
  lookup over the scanned per-key work -- against bisect: the largest k in
  [0, n) with start[k] <= x, keys with no work (equal starts) included."""
  f = lib.host_find_slot
  f.restype = ctypes.c_uint32
  f.argtypes = [ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint32, ctypes.c_uint32,
                ctypes.c_uint64, ctypes.c_int]
  rng = np.random.default_rng(1)
  for _ in range(4000):
    n = int(rng.integers(1, 300))
    work = rng.integers(0, int(rng.choice([2, 5, 100])), n).astype(np.uint64)
    work[0] = max(int(work[0]), 1)
    start = np.zeros(n + 1, dtype=np.uint64)
    start[1:] = np.cumsum(work)
    p = start.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
    total = int(start[n])
    for _ in range(8):
      x = int(rng.integers(0, total))
      want = bisect.bisect_right(start[:n].tolist(), x) - 1
      assert f(p, 0, n, x, 0) == want
      lo = int(rng.integers(0, want + 1))   # any slot whose start is <= x
      assert f(p, lo, n, x, 1) == want, (start.tolist(), lo, x)
