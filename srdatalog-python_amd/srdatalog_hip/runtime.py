"""ctypes binding of a compiled program (include/srdatalog.h).

``ProgramLib`` binds exactly the reference's six shim symbols with the
reference's argument/return types (examples/run_benchmark.py:195-203) plus
the additions this backend exports.  Loading requires the HIP runtime; there
is no CPU fallback.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import native

_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)


class SrdatalogError(RuntimeError):
  pass


def _bind_handles(lib) -> None:
  vp, c_int, u32 = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32
  ull_p = ctypes.POINTER(ctypes.c_ulonglong)
  lib.srdatalog_engine_new.restype = vp
  lib.srdatalog_engine_new.argtypes = []
  lib.srdatalog_engine_free.argtypes = [vp]
  lib.srdatalog_engine_free.restype = None
  lib.srdatalog_engine_set_quiet.argtypes = [vp, c_int]
  lib.srdatalog_engine_set_quiet.restype = None
  lib.srdatalog_engine_load_rows.argtypes = [vp, ctypes.c_char_p, _u32p, ctypes.c_uint64, u32]
  lib.srdatalog_engine_run_begin.argtypes = [vp]
  lib.srdatalog_engine_run_iterations.argtypes = [vp, ctypes.c_ulonglong, ull_p,
                                                  ctypes.POINTER(c_int)]
  lib.srdatalog_engine_absorb.argtypes = [vp, vp]
  lib.srdatalog_engine_range_counts.argtypes = [vp, ctypes.c_char_p, c_int, _u32p, u32, ull_p]
  lib.srdatalog_engine_export_ranges.argtypes = [vp, ctypes.c_char_p, _u32p, u32, c_int, _u32p,
                                                 ctypes.c_uint64, _u64p, _u64p]
  lib.srdatalog_engine_erase_ranges.argtypes = [vp, _u32p, u32]
  for f in ("load_rows", "run_begin", "run_iterations", "absorb", "range_counts", "export_ranges",
            "erase_ranges"):
    getattr(lib, f"srdatalog_engine_{f}").restype = c_int


class ProgramLib:
  def __init__(self, so_path: str):
    native.load_runtime()
    lib = ctypes.CDLL(so_path, mode=ctypes.RTLD_GLOBAL)
    lib.srdatalog_init.restype = ctypes.c_int
    lib.srdatalog_load_csv.restype = ctypes.c_int
    lib.srdatalog_load_csv.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    lib.srdatalog_load_all.restype = ctypes.c_int
    lib.srdatalog_load_all.argtypes = [ctypes.c_char_p]
    lib.srdatalog_run.restype = ctypes.c_int
    lib.srdatalog_run.argtypes = [ctypes.c_ulonglong]
    lib.srdatalog_size.restype = ctypes.c_ulonglong
    lib.srdatalog_size.argtypes = [ctypes.c_char_p]
    lib.srdatalog_shutdown.restype = ctypes.c_int
    lib.srdatalog_export.restype = ctypes.c_int
    lib.srdatalog_export.argtypes = [ctypes.c_char_p, _u32p, ctypes.c_uint64, _u64p]
    lib.srdatalog_delta_sizes.restype = ctypes.c_int
    lib.srdatalog_delta_sizes.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_ulonglong),
                                          ctypes.c_ulonglong]
    lib.srdatalog_num_iterations.restype = ctypes.c_ulonglong
    lib.srdatalog_num_iterations.argtypes = [ctypes.c_char_p]
    lib.srdatalog_load_rows.restype = ctypes.c_int
    lib.srdatalog_load_rows.argtypes = [ctypes.c_char_p, _u32p, ctypes.c_uint64, ctypes.c_uint32]
    lib.srdatalog_prepare.restype = ctypes.c_int
    lib.srdatalog_run_seconds.restype = ctypes.c_double
    lib.srdatalog_join_tuples.restype = ctypes.c_ulonglong
    lib.srdatalog_count_probes.argtypes = [ctypes.c_int]
    lib.srdatalog_count_probes.restype = None
    lib.srdatalog_join_probes.restype = ctypes.c_ulonglong
    lib.srdatalog_set_grain.argtypes = [ctypes.c_uint]
    lib.srdatalog_set_grain.restype = None
    lib.srdatalog_set_instances.argtypes = [_u32p, ctypes.c_uint64]
    lib.srdatalog_set_instances.restype = ctypes.c_int
    lib.srdatalog_instance_delta_sizes.argtypes = [ctypes.c_char_p,
                                                   ctypes.POINTER(ctypes.c_ulonglong),
                                                   ctypes.c_ulonglong]
    lib.srdatalog_instance_delta_sizes.restype = ctypes.c_longlong
    lib.srdatalog_export_instances.argtypes = [ctypes.c_char_p, _u32p, ctypes.c_uint64, _u64p,
                                               _u64p]
    lib.srdatalog_export_instances.restype = ctypes.c_int
    if hasattr(lib, "srdatalog_engine_new"):
      _bind_handles(lib)
    self.lib = lib
    self.path = so_path

  # -- the reference's six ---------------------------------------------------
  def init(self) -> int:
    return self.lib.srdatalog_init()

  def load_csv(self, rel: str, path: str) -> int:
    return self.lib.srdatalog_load_csv(rel.encode(), path.encode())

  def load_all(self, data_dir: str) -> int:
    return self.lib.srdatalog_load_all(data_dir.encode())

  def run(self, max_iters: int = 0) -> int:
    return self.lib.srdatalog_run(max_iters)

  def size(self, rel: str) -> int:
    return int(self.lib.srdatalog_size(rel.encode()))

  def shutdown(self) -> int:
    return self.lib.srdatalog_shutdown()

  # -- additions ---------------------------------------------------------------
  def load_rows(self, rel: str, rows: np.ndarray) -> None:
    rows = np.ascontiguousarray(rows, dtype=np.uint32)
    if rows.ndim == 1:
      rows = rows.reshape(-1, 1)
    rc = self.lib.srdatalog_load_rows(rel.encode(), rows.ctypes.data_as(_u32p), rows.shape[0],
                                      rows.shape[1])
    if rc != 0:
      raise SrdatalogError(f"srdatalog_load_rows({rel}) returned {rc}")

  def prepare(self) -> None:
    rc = self.lib.srdatalog_prepare()
    if rc != 0:
      raise SrdatalogError(f"srdatalog_prepare returned {rc}")

  def export(self, rel: str, arity: int) -> np.ndarray:
    n = ctypes.c_uint64()
    rc = self.lib.srdatalog_export(rel.encode(), None, 0, ctypes.byref(n))
    if rc != 0:
      raise SrdatalogError(f"srdatalog_export({rel}) returned {rc}")
    out = np.empty((max(n.value, 1), arity), dtype=np.uint32)
    rc = self.lib.srdatalog_export(rel.encode(), out.ctypes.data_as(_u32p), n.value, ctypes.byref(n))
    if rc != 0:
      raise SrdatalogError(f"srdatalog_export({rel}) returned {rc}")
    return out[:n.value]

  def delta_sizes(self, rel: str) -> list[int]:
    cnt = self.lib.srdatalog_delta_sizes(rel.encode(), None, 0)
    if cnt < 0:
      raise SrdatalogError(f"unknown relation {rel}")
    buf = (ctypes.c_ulonglong * max(cnt, 1))()
    self.lib.srdatalog_delta_sizes(rel.encode(), buf, cnt)
    return [int(buf[i]) for i in range(cnt)]

  def run_seconds(self) -> float:
    return float(self.lib.srdatalog_run_seconds())

  def join_tuples(self) -> int:
    return int(self.lib.srdatalog_join_tuples())

  def count_probes(self, on: bool) -> None:
    """While on, every run also counts the joins' searches (an extra launch
    per pipeline; leave it off in timed runs)."""
    self.lib.srdatalog_count_probes(1 if on else 0)

  def join_probes(self) -> int:
    """Searches, root-map lookups and bitmap tests of the last probe run."""
    return int(self.lib.srdatalog_join_probes())

  def set_grain(self, grain: int) -> None:
    self.lib.srdatalog_set_grain(grain)

  def set_instances(self, bounds: np.ndarray) -> None:
    """Instance table of a packed batch (ascending column-0 bounds)."""
    b = np.ascontiguousarray(bounds, dtype=np.uint32)
    rc = self.lib.srdatalog_set_instances(b.ctypes.data_as(_u32p), b.size)
    if rc != 0:
      raise SrdatalogError(f"srdatalog_set_instances returned {rc}")

  def instance_delta_sizes(self, rel: str, n_inst: int) -> np.ndarray:
    """(iterations, n_inst) per-instance |DELTA| of rel."""
    it = self.lib.srdatalog_instance_delta_sizes(rel.encode(), None, 0)
    if it < 0:
      raise SrdatalogError(f"srdatalog_instance_delta_sizes({rel}) failed")
    out = np.zeros((max(it, 1), max(n_inst, 1)), dtype=np.uint64)
    if it:
      rc = self.lib.srdatalog_instance_delta_sizes(
          rel.encode(), out.ctypes.data_as(ctypes.POINTER(ctypes.c_ulonglong)), it)
      if rc != it:
        raise SrdatalogError(f"srdatalog_instance_delta_sizes({rel}) returned {rc}")
    return out[:it, :n_inst].astype(np.int64)

  def export_instances(self, rel: str, arity: int, n_inst: int) -> tuple[np.ndarray, np.ndarray]:
    """Rows of rel with instance-local ids, and the n_inst+1 row offsets of
    the instances (instance g owns rows cuts[g]:cuts[g+1])."""
    n = ctypes.c_uint64()
    cuts = np.zeros(n_inst + 1, dtype=np.uint64)
    rc = self.lib.srdatalog_export_instances(rel.encode(), None, 0, ctypes.byref(n), None)
    if rc != 0:
      raise SrdatalogError(f"srdatalog_export_instances({rel}) returned {rc}")
    out = np.empty((max(n.value, 1), arity), dtype=np.uint32)
    rc = self.lib.srdatalog_export_instances(rel.encode(), out.ctypes.data_as(_u32p), n.value,
                                             ctypes.byref(n), cuts.ctypes.data_as(_u64p))
    if rc != 0:
      raise SrdatalogError(f"srdatalog_export_instances({rel}) returned {rc}")
    if n.value == 0:
      cuts[:] = 0
    return out[:n.value], cuts.astype(np.int64)


class Engine:
  """Another engine of a loaded program (srdatalog_engine_new), with the
  resumable-run calls of iteration-granular serving (include/srdatalog.h):
  run_begin, run_iterations, absorb, range_counts, export_ranges,
  erase_ranges.  Not thread-safe; one thread drives an engine."""

  def __init__(self, prog: ProgramLib, quiet: bool = True):
    if not hasattr(prog.lib, "srdatalog_engine_new"):
      raise SrdatalogError(f"{prog.path} exports no engine handles; rebuild it")
    self.lib = prog.lib
    self.h = self.lib.srdatalog_engine_new()
    if not self.h:
      raise SrdatalogError("srdatalog_engine_new failed")
    if quiet:
      self.lib.srdatalog_engine_set_quiet(self.h, 1)

  def close(self) -> None:
    if self.h:
      self.lib.srdatalog_engine_free(self.h)
      self.h = None

  def _ok(self, rc: int, what: str) -> None:
    if rc != 0:
      raise SrdatalogError(f"srdatalog_engine_{what} returned {rc}")

  def load_rows(self, rel: str, rows: np.ndarray) -> None:
    rows = np.ascontiguousarray(rows, dtype=np.uint32)
    if rows.ndim == 1:
      rows = rows.reshape(-1, 1)
    self._ok(self.lib.srdatalog_engine_load_rows(self.h, rel.encode(), rows.ctypes.data_as(_u32p),
                                                 rows.shape[0], rows.shape[1]), "load_rows")

  def run_begin(self) -> None:
    self._ok(self.lib.srdatalog_engine_run_begin(self.h), "run_begin")

  def run_iterations(self, k: int = 1) -> tuple[int, bool]:
    done, conv = ctypes.c_ulonglong(), ctypes.c_int()
    self._ok(self.lib.srdatalog_engine_run_iterations(self.h, k, ctypes.byref(done),
                                                      ctypes.byref(conv)), "run_iterations")
    return int(done.value), bool(conv.value)

  def absorb(self, src: "Engine") -> None:
    self._ok(self.lib.srdatalog_engine_absorb(self.h, src.h), "absorb")

  def range_counts(self, rel: str, ranges, delta: bool = True) -> np.ndarray:
    lohi = np.ascontiguousarray(np.asarray(ranges, dtype=np.uint32).reshape(-1))
    m = lohi.size // 2
    out = np.zeros(max(m, 1), dtype=np.uint64)
    self._ok(self.lib.srdatalog_engine_range_counts(
        self.h, rel.encode(), 1 if delta else 0, lohi.ctypes.data_as(_u32p), m,
        out.ctypes.data_as(ctypes.POINTER(ctypes.c_ulonglong))), "range_counts")
    return out[:m].astype(np.int64)

  def export_ranges(self, rel: str, arity: int, ranges,
                    rebase: bool = False) -> tuple[np.ndarray, np.ndarray]:
    """FULL rows of rel in each range (ascending, disjoint), in relation
    column order, and the m+1 row cuts: range g owns rows cuts[g]:cuts[g+1]
    (with rebase, its values minus the range's first value)."""
    lohi = np.ascontiguousarray(np.asarray(ranges, dtype=np.uint32).reshape(-1))
    m = lohi.size // 2
    n = ctypes.c_uint64()
    cuts = np.zeros(m + 1, dtype=np.uint64)
    self._ok(self.lib.srdatalog_engine_export_ranges(
        self.h, rel.encode(), lohi.ctypes.data_as(_u32p), m, int(rebase), None, 0,
        ctypes.byref(n), cuts.ctypes.data_as(_u64p)), "export_ranges")
    out = np.empty((max(n.value, 1), arity), dtype=np.uint32)
    if n.value:
      self._ok(self.lib.srdatalog_engine_export_ranges(
          self.h, rel.encode(), lohi.ctypes.data_as(_u32p), m, int(rebase),
          out.ctypes.data_as(_u32p), n.value, ctypes.byref(n), cuts.ctypes.data_as(_u64p)),
          "export_ranges")
    return out[:n.value], cuts.astype(np.int64)

  def erase_ranges(self, ranges) -> None:
    lohi = np.ascontiguousarray(np.asarray(ranges, dtype=np.uint32).reshape(-1))
    self._ok(self.lib.srdatalog_engine_erase_ranges(self.h, lohi.ctypes.data_as(_u32p),
                                                    lohi.size // 2), "erase_ranges")
