"""The runtime's device memory pool (csrc/srd_pool.cpp) compiled for the host
over a simulated hipMalloc with a capacity (tests/cpu_sim/pool_sim.cpp):
live blocks never overlap and lie inside driver blocks, freed memory serves
later requests of other sizes (ranges split and merge), a range freed on a
stream serves another stream only after that stream is retired, or at once
behind an event wait when both are registered library streams, and a
failed hipMalloc hands back only cached memory (least recently freed first)
before the request is retried.  Each test loads its own copy of the library
so the pool starts empty."""

import ctypes
import os
import shutil

import numpy as np
import pytest

from conftest import build_cpu_sim

GiB = 1 << 30
MiB = 1 << 20
STATS = ("in_use", "peak", "budget", "allocs", "driver_allocs", "reserved", "cached", "trims",
         "cross_stream")


class Pool:
  def __init__(self, path):
    self.lib = ctypes.CDLL(path)
    u64 = ctypes.c_uint64
    self.lib.sim_pool_alloc.restype = u64
    self.lib.sim_pool_alloc.argtypes = [u64, u64]
    self.lib.sim_pool_free.argtypes = [u64, u64, u64]
    self.lib.sim_pool_retire.argtypes = [u64]
    self.lib.sim_pool_register.argtypes = [u64]
    self.lib.sim_event_log.restype = u64
    self.lib.sim_event_log.argtypes = [ctypes.POINTER(u64), u64]
    self.lib.sim_pool_budget.argtypes = [u64]
    self.lib.sim_pool_stats.argtypes = [ctypes.POINTER(u64)]
    self.lib.sim_device_capacity.argtypes = [u64]
    self.lib.sim_device_stats.argtypes = [ctypes.POINTER(u64)]
    self.lib.sim_device_contains.restype = ctypes.c_int
    self.lib.sim_device_contains.argtypes = [u64, u64]
    self.live = {}  # pointer -> bytes

  def alloc(self, n, s=1):
    p = self.lib.sim_pool_alloc(n, s)
    if p:
      assert self.lib.sim_device_contains(p, n) == 1
      for q, m in self.live.items():   # no overlap with any live block
        assert p + n <= q or q + m <= p, (hex(p), n, hex(q), m)
      self.live[p] = n
    return p

  def free(self, p, s=1):
    self.lib.sim_pool_free(p, self.live.pop(p), s)

  def retire(self, s):
    self.lib.sim_pool_retire(s)

  def register(self, s):
    self.lib.sim_pool_register(s)

  def events(self):
    """(op, stream) of every event record (op 0) and stream wait (op 1)."""
    n = self.lib.sim_event_log(None, 0)
    out = (ctypes.c_uint64 * max(n, 1))()
    self.lib.sim_event_log(out, n)
    return [(int(x) >> 32, int(x) & 0xffffffff) for x in out[:n]]

  def stats(self):
    out = (ctypes.c_uint64 * len(STATS))()
    self.lib.sim_pool_stats(out)
    return dict(zip(STATS, out))

  def device(self):
    out = (ctypes.c_uint64 * 5)()
    self.lib.sim_device_stats(out)
    return dict(zip(("used", "mallocs", "frees", "syncs", "blocks"), out))


@pytest.fixture
def pool(tmp_path):
  src = build_cpu_sim("libpool_sim.so")
  dst = os.path.join(tmp_path, "libpool_sim_copy.so")
  shutil.copy(src, dst)
  return Pool(dst)


@pytest.mark.parametrize("registered", [False, True], ids=["own-stream", "cross-stream"])
def test_random_workload_no_overlap(pool, registered):
  """Random sizes (small classes and arena ranges) on three streams, frees in
  random order with retirements: blocks never overlap, in_use is exact, and
  the driver memory stays within a small factor of the peak in use; with
  the streams registered, ranges also move between streams (every such
  take is a record on the freeing stream followed by a wait)."""
  if registered:
    for s in (1, 2, 3):
      pool.register(s)
  rng = np.random.default_rng(0)
  for step in range(3000):
    if pool.live and rng.random() < 0.45:
      p = list(pool.live)[int(rng.integers(len(pool.live)))]
      pool.free(p, s=int(rng.integers(1, 4)))
    else:
      n = int(rng.choice([rng.integers(1, 4096), rng.integers(1, 2 * MiB), rng.integers(MiB, 300 * MiB)]))
      assert pool.alloc(n, s=int(rng.integers(1, 4)))
    if step % 500 == 499:
      pool.retire(int(rng.integers(1, 4)))
    assert pool.stats()["in_use"] == sum(pool.live.values())
  st, dev = pool.stats(), pool.device()
  assert dev["used"] <= max(4 * st["peak"], 2 * GiB)
  ev = pool.events()
  assert len(ev) == 2 * st["cross_stream"]
  assert (st["cross_stream"] > 0) == registered
  for (op_r, s_r), (op_w, s_w) in zip(ev[::2], ev[1::2]):
    assert op_r == 0 and op_w == 1 and s_r != s_w


def test_freed_ranges_serve_other_sizes(pool):
  """Memory freed by one run serves a second run of different sizes without
  new driver allocations once the stream is retired (split and merge)."""
  a = [pool.alloc(n * MiB) for n in (100, 300, 50, 700, 200)]
  for p in a:
    pool.free(p)
  pool.retire(1)
  before = pool.device()["mallocs"]
  b = [pool.alloc(n * MiB, s=2) for n in (650, 350, 250, 90)]   # 1,340 MiB <= 1,350 freed
  assert all(b)
  assert pool.device()["mallocs"] == before
  for p in b:
    pool.free(p, s=2)
  pool.retire(2)
  # everything merged back: each 1 GiB arena serves one request of its size
  dev = pool.device()
  assert dev["blocks"] == 2
  assert pool.alloc(1024 * MiB, s=3) and pool.alloc(1024 * MiB, s=3)
  assert pool.device()["mallocs"] == dev["mallocs"]


def test_stream_tag_until_retired(pool):
  """A range freed on stream 1 serves stream 1 at once but stream 2 only after
  stream 1 is retired."""
  p = pool.alloc(512 * MiB, s=1)
  pool.free(p, s=1)
  q = pool.alloc(512 * MiB, s=2)
  assert q != p                        # not the range stream 1 may still use
  r = pool.alloc(512 * MiB, s=1)
  assert r == p                        # same stream: reused in stream order
  pool.free(r, s=1)
  pool.retire(1)
  assert pool.alloc(512 * MiB, s=2) == p


def test_cross_stream_reuse_of_registered_streams(pool):
  """A range freed on a registered stream serves another stream that finds
  no range of its own, after an event recorded on the freeing stream and a
  wait on the requesting one (in that order); no hipMalloc is made.  The
  rest of a split range stays with the freeing stream, and a range freed on
  an unregistered stream is never taken across."""
  pool.register(1)
  pool.register(2)
  p = pool.alloc(1024 * MiB, s=1)
  pool.free(p, s=1)
  mallocs = pool.device()["mallocs"]
  q = pool.alloc(256 * MiB, s=2)
  assert q == p
  assert pool.events() == [(0, 1), (1, 2)]
  assert pool.stats()["cross_stream"] == 1
  r = pool.alloc(768 * MiB, s=1)       # the rest, still stream 1's: no wait
  assert r == p + 256 * MiB
  assert pool.events() == [(0, 1), (1, 2)]
  assert pool.device()["mallocs"] == mallocs
  x = pool.alloc(512 * MiB, s=3)       # stream 3 is not registered
  pool.free(x, s=3)
  y = pool.alloc(512 * MiB, s=2)
  assert y != x
  assert pool.stats()["cross_stream"] == 1


def test_oom_hands_back_cached_memory(pool):
  """At capacity, a request that no free range fits makes the pool hand back
  wholly free arenas (least recently freed first) and retry; memory still
  in use is never handed back."""
  pool.lib.sim_device_capacity(6 * GiB + 4 * MiB)
  keep = pool.alloc(2 * GiB, s=1)              # arena 1, stays live
  old = pool.alloc(2 * GiB, s=1)               # arena 2, freed first
  young = pool.alloc(1 * GiB + 2 * MiB, s=1)   # arena 3, freed later
  pool.free(old, s=1)
  pool.free(young, s=1)
  assert pool.stats()["trims"] == 0
  big = pool.alloc(3 * GiB, s=2)               # 5 GiB + 2 MiB reserved: needs arena 2 back
  assert big
  st, dev = pool.stats(), pool.device()
  assert st["trims"] == 1
  assert dev["frees"] == 1                     # only the least recently freed arena
  assert dev["used"] <= 6 * GiB + 4 * MiB
  assert pool.lib.sim_device_contains(keep, 2 * GiB) == 1
  # the younger cached arena is still there and serves its size again
  before = dev["mallocs"]
  assert pool.alloc(1 * GiB, s=1)
  assert pool.device()["mallocs"] == before


def test_oom_when_nothing_cached_fails_cleanly(pool):
  pool.lib.sim_device_capacity(2 * GiB)
  a = pool.alloc(1536 * MiB)
  assert a
  assert pool.alloc(1 * GiB) == 0              # the pool raises; nothing handed out
  assert pool.stats()["in_use"] == 1536 * MiB


def test_budget(pool):
  pool.lib.sim_pool_budget(100 * MiB)
  assert pool.alloc(60 * MiB)
  assert pool.alloc(60 * MiB) == 0             # beyond the budget: an error, no allocation
  assert pool.stats()["in_use"] == 60 * MiB
