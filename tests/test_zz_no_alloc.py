"""GPU parity of srd_prim_sort_dedup_noalloc (the sort + dedup entry point
that allocates nothing) against srd_prim_sort_dedup, through the C ABI.

The inputs are copied to device buffers the test allocates with the HIP
runtime; the entry point runs on the null stream and the outputs (sorted
distinct rows, distinct column-0 values, their first-row offsets and the
counts) must equal srd_prim_sort_dedup's bit for bit.  Two calls in a row,
the second with fewer rows, check that the shared static scratch carries
nothing from one call into the next."""

import ctypes

import numpy as np
import pytest

from srdatalog_hip import native

pytestmark = pytest.mark.gpu

MAX_ROWS = 1 << 20


class Hip:
  """Just enough of the HIP runtime for device buffers (test plumbing)."""

  def __init__(self):
    self.lib = ctypes.CDLL("libamdhip64.so", mode=ctypes.RTLD_GLOBAL)
    self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    self.lib.hipFree.argtypes = [ctypes.c_void_p]
    self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    self.bufs = []

  def check(self, rc, what):
    assert rc == 0, f"{what} returned {rc}"

  def alloc(self, nbytes):
    p = ctypes.c_void_p()
    self.check(self.lib.hipMalloc(ctypes.byref(p), max(nbytes, 256)), "hipMalloc")
    self.bufs.append(p)
    return p

  def upload(self, a):
    a = np.ascontiguousarray(a)
    p = self.alloc(a.nbytes)
    if a.nbytes:
      self.check(self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
    return p

  def download(self, p, dtype, count):
    out = np.empty(count, dtype=dtype)
    if count:
      self.check(self.lib.hipMemcpy(out.ctypes.data, p, out.nbytes, 2), "hipMemcpy D2H")
    return out

  def sync(self):
    self.check(self.lib.hipDeviceSynchronize(), "hipDeviceSynchronize")

  def free_all(self):
    for p in self.bufs:
      self.lib.hipFree(p)
    self.bufs = []


@pytest.fixture(scope="module")
def hip():
  lib = native.load_runtime()
  assert lib.srd_prim_init(0) == 0
  h = Hip()
  yield h
  h.free_all()


def noalloc_sort(hip, rows):
  """Run srd_prim_sort_dedup_noalloc on (n, k) rows; returns host copies."""
  lib = native.load_runtime()
  n, k = rows.shape
  cols = hip.upload(np.ascontiguousarray(rows.T))        # column c at c * n
  out = hip.alloc(4 * n * k)
  rv = hip.alloc(4 * n)
  ro = hip.alloc(4 * (n + 1))
  cnt = hip.alloc(8 * 3)
  print(f"noalloc sort: n={n} k={k}", flush=True)
  rc = lib.srd_prim_sort_dedup_noalloc(cols, n, n, k, out, n, rv, ro, cnt, None)
  assert rc == 0, lib.srd_last_error().decode()
  hip.sync()
  u, r, err = (int(x) for x in hip.download(cnt, np.uint64, 3))
  assert err == 0, f"error bits {err}"
  got = hip.download(out, np.uint32, n * k).reshape(k, n)[:, :u].T
  return got, hip.download(rv, np.uint32, r), hip.download(ro, np.uint32, r + 1)


def reference_sort(rows):
  want, rv, ro = native.sort_dedup(rows)
  return want, rv, ro


def draw(rng, n, k, hi):
  return rng.integers(0, hi, size=(n, k), dtype=np.uint64).astype(np.uint32)


@pytest.mark.parametrize("k,hi", [(1, 5000), (2, 1 << 12), (2, 1 << 32), (3, 1 << 20), (4, 300)])
def test_noalloc_matches_sort_dedup(hip, k, hi):
  rng = np.random.default_rng(100 + k)
  for n in (300_000, 47_000):           # two calls in a row, the second smaller
    rows = draw(rng, n, k, hi)
    got, rv, ro = noalloc_sort(hip, rows)
    want, wrv, wro = reference_sort(rows)
    assert got.shape == want.shape
    assert np.array_equal(got, want)
    assert np.array_equal(rv, wrv)
    assert np.array_equal(ro, wro)
  hip.free_all()


def test_noalloc_edge_sizes(hip):
  rng = np.random.default_rng(7)
  cases = [draw(rng, MAX_ROWS, 2, 1 << 21),                       # the maximum size
           np.zeros((5, 2), dtype=np.uint32),                     # all equal
           np.sort(draw(rng, 10_000, 2, 100), axis=0),            # presorted with duplicates
           np.full((3, 3), 0xFFFFFFFF, dtype=np.uint32),          # max u32
           draw(rng, 1, 2, 10)]
  for rows in cases:
    got, rv, ro = noalloc_sort(hip, rows)
    want, wrv, wro = reference_sort(rows)
    assert np.array_equal(got, want)
    assert np.array_equal(rv, wrv)
    assert np.array_equal(ro, wro)
    hip.free_all()


def test_noalloc_rejects_above_maximum(hip):
  lib = native.load_runtime()
  # rejected on the host before any device access: the pointers are never read
  p = hip.alloc(256)
  rc = lib.srd_prim_sort_dedup_noalloc(p, MAX_ROWS + 1, MAX_ROWS + 1, 2, p, MAX_ROWS + 1, p, p, p,
                                       None)
  assert rc != 0
  assert b"exceed the maximum" in lib.srd_last_error()
  hip.free_all()
