"""ctypes front of the C++ oracle (``liborc.so``).  Test infrastructure only."""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")
_lib = None


class _Input(ctypes.Structure):
  _fields_ = [
    ("name", ctypes.c_char_p),
    ("rows", ctypes.POINTER(ctypes.c_uint32)),
    ("n", ctypes.c_uint64),
    ("arity", ctypes.c_uint32),
  ]


def build() -> str:
  """Compile liborc.so from oracle.cpp (make)."""
  subprocess.run(["make", "-s", "-C", _HERE], check=True)
  return _LIB_PATH


def _load():
  global _lib
  if _lib is None:
    if not os.path.exists(_LIB_PATH):
      build()
    lib = ctypes.CDLL(_LIB_PATH)
    lib.orc_run.restype = ctypes.c_void_p
    lib.orc_run.argtypes = [ctypes.c_char_p, ctypes.POINTER(_Input), ctypes.c_int,
                            ctypes.c_uint64, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32]
    lib.orc_error.restype = ctypes.c_char_p
    lib.orc_error.argtypes = [ctypes.c_void_p]
    lib.orc_words.restype = ctypes.c_uint64
    lib.orc_words.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    lib.orc_export.restype = ctypes.c_int
    lib.orc_export.argtypes = [ctypes.c_void_p, ctypes.c_char_p,
                               ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint64]
    lib.orc_num_deltas.restype = ctypes.c_uint64
    lib.orc_num_deltas.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    lib.orc_deltas.restype = ctypes.c_int
    lib.orc_deltas.argtypes = [ctypes.c_void_p, ctypes.c_char_p,
                               ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint64]
    lib.orc_rule_count.restype = ctypes.c_uint64
    lib.orc_rule_count.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    for f in ("orc_iterations", "orc_count"):
      getattr(lib, f).restype = ctypes.c_uint64
      getattr(lib, f).argtypes = [ctypes.c_void_p]
    lib.orc_seconds.restype = ctypes.c_double
    lib.orc_seconds.argtypes = [ctypes.c_void_p]
    lib.orc_free.restype = None
    lib.orc_free.argtypes = [ctypes.c_void_p]
    lib.orc_set_key_stride.restype = None
    lib.orc_set_key_stride.argtypes = [ctypes.c_uint32]
    lib.orc_digest_rows.restype = None
    lib.orc_digest_rows.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint64,
                                    ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]
    lib.orc_triangle_digest.restype = None
    lib.orc_triangle_digest.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint64,
                                        ctypes.c_int, ctypes.POINTER(ctypes.c_uint64),
                                        ctypes.POINTER(ctypes.c_double)]
    _lib = lib
  return _lib


def digest_rows(rows: np.ndarray, threads: int = 0) -> tuple[int, int, int]:
  """Order-independent digest (count, sum1, sum2) of distinct tuples."""
  lib = _load()
  rows = np.ascontiguousarray(rows, dtype=np.uint32)
  if rows.ndim == 1:
    rows = rows.reshape(-1, 1)
  out = (ctypes.c_uint64 * 3)()
  lib.orc_digest_rows(rows.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), rows.shape[0],
                      rows.shape[1], threads, out)
  return int(out[0]), int(out[1]), int(out[2])


def triangle_digest(knows: np.ndarray, threads: int = 0) -> tuple[tuple[int, int, int], float]:
  """Digest of the oriented triangle listing over the KnowsInput pairs, and
  the enumeration's wall seconds."""
  lib = _load()
  knows = np.ascontiguousarray(knows, dtype=np.uint32)
  out = (ctypes.c_uint64 * 3)()
  sec = ctypes.c_double()
  lib.orc_triangle_digest(knows.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), knows.shape[0],
                          threads, out, ctypes.byref(sec))
  return (int(out[0]), int(out[1]), int(out[2])), float(sec.value)


def set_key_stride(stride: int) -> None:
  """Triangle programs: sample the root keys a with a % stride == 0."""
  _load().orc_set_key_stride(stride)


ARITY = {"Triangle": 3, "Reach": 1}
# relations whose arity depends on the program
PROGRAM_ARITY = {("neg2hop", "Path"): 4, ("neg2hop_count", "Path"): 4}


class OracleResult:
  """Sorted relation contents, per-iteration deltas and timing of one run."""

  def __init__(self, relations, deltas, iterations, count, seconds, rule_counts=None):
    self.relations = relations      # name -> (n, arity) uint32, lexicographically sorted
    self.deltas = deltas            # name -> list[int], |DELTA| per iteration
    self.iterations = iterations
    self.count = count
    self.seconds = seconds
    self.rule_counts = rule_counts or {}   # count-only head relation -> count

  def size(self, rel: str) -> int:
    return int(self.relations[rel].shape[0])


def run(program: str, inputs: dict[str, np.ndarray], max_iters: int = 0, threads: int = 0,
        key_range: tuple[int, int] | None = None,
        relations: tuple[str, ...] = ("Edge", "Path", "Sg", "PointsTo", "Knows", "Triangle",
                                      "Reach", "HasInterest")) -> OracleResult:
  """Run ``program`` ("tc", "sg", "andersen", "triangle", "triangle_count",
  "reach", "neg2hop", "neg2hop_count", "q3", "q7") on ``inputs`` (relation name ->
  (n, arity) uint32 rows)."""
  lib = _load()
  keep = []
  arr = (_Input * max(len(inputs), 1))()
  for i, (name, rows) in enumerate(inputs.items()):
    rows = np.ascontiguousarray(rows, dtype=np.uint32)
    if rows.ndim == 1:
      rows = rows.reshape(-1, 1)
    keep.append(rows)
    arr[i].name = name.encode()
    arr[i].rows = rows.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    arr[i].n = rows.shape[0]
    arr[i].arity = rows.shape[1]
  lo, hi = key_range if key_range else (0, 0)
  h = lib.orc_run(program.encode(), arr, len(inputs), max_iters, threads, lo, hi)
  if not h:
    raise ValueError(f"unknown oracle program {program!r}")
  try:
    err = lib.orc_error(h)
    if err:
      raise ValueError(err.decode())
    rels, deltas = {}, {}
    for name in relations:
      nw = lib.orc_words(h, name.encode())
      arity = PROGRAM_ARITY.get((program, name), ARITY.get(name, 2))
      buf = np.empty(nw, dtype=np.uint32)
      if lib.orc_export(h, name.encode(), buf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                        nw) == 0:
        rels[name] = buf.reshape(-1, arity)
      nd = lib.orc_num_deltas(h, name.encode())
      if nd:
        d = np.empty(nd, dtype=np.uint64)
        lib.orc_deltas(h, name.encode(), d.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), nd)
        deltas[name] = [int(x) for x in d]
    rule_counts = {name: int(lib.orc_rule_count(h, name.encode()))
                   for name in ("Case1", "Case2", "Case3", "Case4")}
    return OracleResult(rels, deltas, int(lib.orc_iterations(h)), int(lib.orc_count(h)),
                        float(lib.orc_seconds(h)), rule_counts)
  finally:
    lib.orc_free(h)
