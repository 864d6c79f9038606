"""Parity of the device set primitives (through the C ABI of
include/srdatalog_prims.h) against numpy set definitions."""

import numpy as np
import pytest

import npref

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nat():
  from srdatalog_hip import native
  native.load_runtime()
  return native


def gen(kind, n, k, seed):
  rng = np.random.default_rng(seed)
  if kind == "small":
    return rng.integers(0, 50, size=(n, k), dtype=np.uint32)
  if kind == "full":
    return rng.integers(0, 2**32, size=(n, k), dtype=np.uint64).astype(np.uint32)
  if kind == "equal":
    return np.full((n, k), 7, dtype=np.uint32)
  if kind == "zeros":
    return np.zeros((n, k), dtype=np.uint32)
  if kind == "max":
    a = rng.integers(0, 2, size=(n, k), dtype=np.uint32)
    return np.where(a == 1, np.uint32(0xFFFFFFFF), np.uint32(0xFFFFFFFE)).astype(np.uint32)
  if kind == "sorted":
    return npref.sort_unique(rng.integers(0, 1000, size=(n, k), dtype=np.uint32))
  if kind == "reversed":
    return npref.sort_unique(rng.integers(0, 1000, size=(n, k), dtype=np.uint32))[::-1].copy()
  if kind == "skew":
    z = np.minimum(rng.zipf(1.3, size=(n, k)), 1 << 20).astype(np.uint32)
    return z
  raise ValueError(kind)


@pytest.mark.parametrize("k", [1, 2, 3, 4])
@pytest.mark.parametrize("n", [0, 1, 63, 4095, 4096, 4097, 100_000])
@pytest.mark.parametrize("kind", ["small", "full", "equal", "max", "reversed", "skew"])
def test_sort_dedup(nat, kind, n, k):
  rows = gen(kind, n, k, n * 31 + k)
  got, rv, ro = nat.sort_dedup(rows)
  want = npref.sort_unique(rows)
  np.testing.assert_array_equal(got, want)
  wv, wo = npref.roots(want)
  np.testing.assert_array_equal(rv, wv)
  if want.shape[0]:
    np.testing.assert_array_equal(ro, wo)


@pytest.mark.parametrize("k,n", [(2, 3_000_000), (3, 1_000_000), (2, 17_000_000)])
def test_sort_dedup_large(nat, k, n):
  rows = gen("skew", n, k, 5)
  got, rv, ro = nat.sort_dedup(rows)
  want = npref.sort_unique(rows)
  np.testing.assert_array_equal(got, want)
  wv, wo = npref.roots(want)
  np.testing.assert_array_equal(rv, wv)
  np.testing.assert_array_equal(ro, wo)


@pytest.mark.parametrize("variant", ["gather", "lookback", "twopass"])
@pytest.mark.parametrize("k", [1, 2, 3])
@pytest.mark.parametrize("na,nb", [(0, 10), (10, 0), (1000, 1000), (50, 200_000),
                                   (200_000, 50), (300_000, 300_000), (3_000_000, 2_000_000)])
def test_set_difference(nat, k, na, nb, variant, monkeypatch):
  # SRDATALOG_SETDIFF picks the dense variant; it is read on every call
  monkeypatch.setenv("SRDATALOG_SETDIFF", variant)
  rng = np.random.default_rng(na + nb + k)
  hi = 300 if k > 1 else 400_000
  a = npref.sort_unique(rng.integers(0, hi, size=(na, k), dtype=np.uint32))
  b = npref.sort_unique(rng.integers(0, hi, size=(nb, k), dtype=np.uint32))
  got = nat.set_difference(a, b)
  np.testing.assert_array_equal(got, npref.set_difference(a, b))


@pytest.mark.parametrize("k", [1, 2, 3])
@pytest.mark.parametrize("na,nb", [(0, 10), (10, 0), (1000, 1000), (50, 200_000),
                                   (200_000, 50), (300_000, 300_000)])
def test_merge_disjoint(nat, k, na, nb):
  rng = np.random.default_rng(na * 3 + nb + k)
  hi = 300 if k > 1 else 1_000_000
  pool = npref.sort_unique(rng.integers(0, hi, size=(na + nb + 100, k), dtype=np.uint32))
  perm = rng.permutation(pool.shape[0])
  a = npref.sort_unique(pool[perm[:min(na, pool.shape[0])]])
  b = npref.set_difference(npref.sort_unique(pool[perm[min(na, pool.shape[0]):][:nb]]), a)
  got = nat.merge(a, b)
  np.testing.assert_array_equal(got, npref.union_sorted(a, b))


@pytest.mark.parametrize("n", [0, 1, 4096, 4097, 5_000_000])
def test_exclusive_scan(nat, n):
  x = np.random.default_rng(n).integers(0, 1000, size=n, dtype=np.uint32)
  got = nat.exclusive_scan(x)
  want = np.concatenate([[0], np.cumsum(x, dtype=np.uint64)])
  np.testing.assert_array_equal(got, want)


# ------------------------------------------------------ device benchmark ----
def _splitmix(x):
  x = (x + np.uint64(0x9E3779B97F4A7C15)).astype(np.uint64)
  x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)).astype(np.uint64)
  x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)).astype(np.uint64)
  return x ^ (x >> np.uint64(31))


@pytest.mark.parametrize("k", [1, 2, 3])
def test_prim_bench_output_sizes(k):
  from srdatalog_hip import native
  n = 300_000
  # sort+dedup of hashed keys: distinct count from the same hash on the host
  bits = 32 if k == 1 else 20 * (k - 1) + 20
  with np.errstate(over="ignore"):
    keys = _splitmix(np.arange(n, dtype=np.uint64)) % np.uint64(1 << bits)
  _, got = native.prim_bench("sort_dedup", n, 0, k, reps=2)
  assert got == np.unique(keys).size
  # NEW = {2i}, FULL = {3j}: NEW - FULL drops the multiples of 6 both reach
  na, nb = 200_000, 150_000
  m = min(2 * (na - 1), 3 * (nb - 1))
  _, got = native.prim_bench("set_difference", na, nb, k, reps=2)
  assert got == na - (m // 6 + 1)
  _, got = native.prim_bench("merge", na, nb, k, reps=2)
  assert got == na + nb
