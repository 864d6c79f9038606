"""Build and load the native MI355X runtime (``libsrdatalog_hip.so``).

The runtime is compiled in-tree with ``hipcc --offload-arch=gfx950`` into
``srdatalog_hip/_lib``.  Nothing here falls back to a CPU path: when the
library is missing or fails to load, callers get an exception.
"""

from __future__ import annotations

import ctypes
import os
import shutil
import subprocess

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT_DIR = os.path.dirname(PKG_DIR)                   # srdatalog-python_amd/
CSRC_DIR = os.path.join(ROOT_DIR, "csrc")
REPO_DIR = os.path.dirname(ROOT_DIR)
INCLUDE_DIR = os.path.join(REPO_DIR, "include")
LIB_DIR = os.path.join(PKG_DIR, "_lib")
RUNTIME_LIB = os.path.join(LIB_DIR, "libsrdatalog_hip.so")

OFFLOAD_ARCH = "gfx950"
RUNTIME_SOURCES = ["srd_prims.hip", "srd_engine.hip", "srd_prims_capi.cpp", "srd_noalloc.hip",
                   "srd_pool.cpp"]


def hipcc() -> str:
  for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
    if cand and os.path.exists(cand):
      return cand
  raise RuntimeError("hipcc not found (ROCm is required to build the MI355X runtime)")


def hip_flags() -> list[str]:
  """Flags for every HIP translation unit of the runtime and generated programs."""
  return [
    f"--offload-arch={OFFLOAD_ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-I", CSRC_DIR, "-I", INCLUDE_DIR,
  ]


def _compile(src: str, obj: str) -> None:
  cmd = [hipcc(), *hip_flags(), "-x", "hip", "-c", src, "-o", obj]
  r = subprocess.run(cmd, capture_output=True, text=True)
  if r.returncode != 0:
    raise RuntimeError(f"hipcc failed on {os.path.basename(src)}:\n{r.stderr[-6000:]}")


def build_runtime(jobs: int = 4) -> str:
  """Compile every runtime source and link libsrdatalog_hip.so (always rebuilds)."""
  from concurrent.futures import ThreadPoolExecutor

  os.makedirs(LIB_DIR, exist_ok=True)
  objs = []
  srcs = [s for s in RUNTIME_SOURCES if os.path.exists(os.path.join(CSRC_DIR, s))]
  with ThreadPoolExecutor(max_workers=jobs) as ex:
    futs = []
    for s in srcs:
      obj = os.path.join(LIB_DIR, s.rsplit(".", 1)[0] + ".o")
      objs.append(obj)
      futs.append(ex.submit(_compile, os.path.join(CSRC_DIR, s), obj))
    for f in futs:
      f.result()
  tmp = RUNTIME_LIB + ".tmp"
  cmd = [hipcc(), f"--offload-arch={OFFLOAD_ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
  r = subprocess.run(cmd, capture_output=True, text=True)
  if r.returncode != 0:
    raise RuntimeError(f"link of libsrdatalog_hip.so failed:\n{r.stderr[-4000:]}")
  os.replace(tmp, RUNTIME_LIB)
  for o in objs:
    os.remove(o)
  return RUNTIME_LIB


_runtime = None


def load_runtime() -> ctypes.CDLL:
  """dlopen the runtime (RTLD_GLOBAL so generated programs resolve against it)."""
  global _runtime
  if _runtime is None:
    if not os.path.exists(RUNTIME_LIB):
      raise RuntimeError(f"{RUNTIME_LIB} is not built; run __graft_entry__.build() first")
    _runtime = ctypes.CDLL(RUNTIME_LIB, mode=ctypes.RTLD_GLOBAL)
    _bind_prims(_runtime)
  return _runtime


_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)


def _bind_prims(lib) -> None:
  lib.srd_last_error.restype = ctypes.c_char_p
  lib.srd_prim_init.argtypes = [ctypes.c_int]
  lib.srd_prim_sort_dedup.argtypes = [_u32p, ctypes.c_uint64, ctypes.c_uint32, _u32p, _u64p,
                                      _u32p, _u32p, _u64p]
  lib.srd_prim_set_difference.argtypes = [_u32p, ctypes.c_uint64, _u32p, ctypes.c_uint64,
                                          ctypes.c_uint32, _u32p, _u64p]
  lib.srd_prim_merge.argtypes = lib.srd_prim_set_difference.argtypes
  lib.srd_prim_exclusive_scan.argtypes = [_u32p, ctypes.c_uint64, _u64p]
  lib.srd_prim_time_sort_dedup.argtypes = [_u32p, ctypes.c_uint64, ctypes.c_uint32,
                                           ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]
  lib.srd_calib_copy.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
  lib.srd_prim_bench.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32,
                                 ctypes.c_uint32, ctypes.POINTER(ctypes.c_double), _u64p]
  lib.srd_pool_stats.argtypes = [_u64p, ctypes.c_uint32]
  _vp = ctypes.c_void_p
  lib.srd_prim_sort_dedup_noalloc.argtypes = [_vp, ctypes.c_uint64, ctypes.c_uint64,
                                              ctypes.c_uint32, _vp, ctypes.c_uint64, _vp, _vp,
                                              _vp, _vp]
  lib.srd_timing_enable.argtypes = [ctypes.c_int]
  lib.srd_timing_report.restype = ctypes.c_uint64
  lib.srd_timing_report.argtypes = [ctypes.c_char_p, ctypes.c_uint64]


POOL_FIELDS = ("in_use", "peak", "budget", "allocs", "driver_allocs", "reserved", "cached",
               "trims", "cross_stream")


def pool_stats() -> dict:
  """Counters of the runtime's device memory pool (bytes and counts)."""
  lib = load_runtime()
  buf = (ctypes.c_uint64 * len(POOL_FIELDS))()
  _check(lib, lib.srd_pool_stats(buf, len(POOL_FIELDS)), "pool_stats")
  return {k: int(buf[i]) for i, k in enumerate(POOL_FIELDS)}


def calib_copy(nbytes: int, reps: int = 5) -> float:
  """ms per launch of the plain streaming copy kernel over nbytes."""
  lib = load_runtime()
  ms = ctypes.c_double()
  _check(lib, lib.srd_calib_copy(nbytes, reps, ctypes.byref(ms)), "calib_copy")
  return ms.value


PRIM_OPS = {"sort_dedup": 0, "set_difference": 1, "merge": 2}


def prim_bench(op: str, na: int, nb: int = 0, k: int = 2, reps: int = 5) -> tuple[float, int]:
  """(ms per call, output rows) of one primitive on device-generated inputs
  (see srd_prim_bench in include/srdatalog_prims.h)."""
  lib = load_runtime()
  ms = ctypes.c_double()
  n = ctypes.c_uint64()
  _check(lib, lib.srd_prim_bench(PRIM_OPS[op], na, nb, k, reps, ctypes.byref(ms), ctypes.byref(n)),
         "prim_bench")
  return ms.value, n.value


def timing_enable(on: bool = True) -> None:
  load_runtime().srd_timing_enable(1 if on else 0)


def timing_reset() -> None:
  load_runtime().srd_timing_reset()


def timing_report() -> list[dict]:
  """Per-kernel-name launches / summed ms / algorithmic bytes since the last reset."""
  import json
  lib = load_runtime()
  n = lib.srd_timing_report(None, 0)
  buf = ctypes.create_string_buffer(n + 1)
  lib.srd_timing_report(buf, n + 1)
  return json.loads(buf.value.decode())


def _ptr32(a: np.ndarray):
  return a.ctypes.data_as(_u32p)


def _check(lib, rc: int, what: str) -> None:
  if rc != 0:
    raise RuntimeError(f"{what} failed: {lib.srd_last_error().decode()}")


def _rows(a: np.ndarray) -> np.ndarray:
  a = np.ascontiguousarray(a, dtype=np.uint32)
  return a.reshape(-1, 1) if a.ndim == 1 else a


def sort_dedup(rows: np.ndarray):
  """Device sort+dedup of (n, k) u32 rows -> (sorted unique rows, root_vals, root_offs)."""
  lib = load_runtime()
  rows = _rows(rows)
  n, k = rows.shape
  out = np.empty((max(n, 1), k), dtype=np.uint32)
  rv = np.empty(max(n, 1), dtype=np.uint32)
  ro = np.empty(max(n, 1) + 1, dtype=np.uint32)
  on, orr = ctypes.c_uint64(), ctypes.c_uint64()
  _check(lib, lib.srd_prim_sort_dedup(_ptr32(rows), n, k, _ptr32(out), ctypes.byref(on),
                                      _ptr32(rv), _ptr32(ro), ctypes.byref(orr)), "sort_dedup")
  r = orr.value
  return out[:on.value], rv[:r], ro[:r + 1] if on.value else ro[:0]


def set_difference(a: np.ndarray, b: np.ndarray) -> np.ndarray:
  lib = load_runtime()
  a, b = _rows(a), _rows(b)
  k = a.shape[1]
  out = np.empty((max(a.shape[0], 1), k), dtype=np.uint32)
  on = ctypes.c_uint64()
  _check(lib, lib.srd_prim_set_difference(_ptr32(a), a.shape[0], _ptr32(b), b.shape[0], k,
                                          _ptr32(out), ctypes.byref(on)), "set_difference")
  return out[:on.value]


def merge(a: np.ndarray, b: np.ndarray) -> np.ndarray:
  lib = load_runtime()
  a, b = _rows(a), _rows(b)
  k = a.shape[1]
  out = np.empty((max(a.shape[0] + b.shape[0], 1), k), dtype=np.uint32)
  on = ctypes.c_uint64()
  _check(lib, lib.srd_prim_merge(_ptr32(a), a.shape[0], _ptr32(b), b.shape[0], k, _ptr32(out),
                                 ctypes.byref(on)), "merge")
  return out[:on.value]


def exclusive_scan(x: np.ndarray) -> np.ndarray:
  lib = load_runtime()
  x = np.ascontiguousarray(x, dtype=np.uint32)
  out = np.empty(x.shape[0] + 1, dtype=np.uint64)
  _check(lib, lib.srd_prim_exclusive_scan(_ptr32(x), x.shape[0],
                                          out.ctypes.data_as(_u64p)), "exclusive_scan")
  return out


def time_sort_dedup(rows: np.ndarray, reps: int = 10) -> float:
  lib = load_runtime()
  rows = _rows(rows)
  ms = ctypes.c_double()
  _check(lib, lib.srd_prim_time_sort_dedup(_ptr32(rows), rows.shape[0], rows.shape[1], reps,
                                           ctypes.byref(ms)), "time_sort_dedup")
  return ms.value
