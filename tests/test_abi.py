"""CPU checks of the C ABI: every symbol declared in include/*.h is exported
by the library that implements it (no device calls are made)."""

import ctypes
import os
import re

import pytest

from srdatalog_hip import native, programs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
  text = open(os.path.join(ROOT, "include", header)).read()
  text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
  return sorted(set(re.findall(r"\b(srd_\w+|srdatalog_\w+)\s*\(", text)))


def exported(path):
  import subprocess
  out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                       check=True).stdout
  return {line.split()[-1] for line in out.splitlines() if line.strip()}


@pytest.fixture(scope="module")
def built():
  if not os.path.exists(native.RUNTIME_LIB) or not os.path.exists(programs.so_path("tc")):
    pytest.skip("libraries not built (run __graft_entry__.build())")


def test_prims_header_exported(built):
  syms = exported(native.RUNTIME_LIB)
  missing = [s for s in declared("srdatalog_prims.h") if s not in syms]
  assert not missing, missing


@pytest.mark.parametrize("name", sorted(programs.PROGRAMS))
def test_program_header_exported(built, name):
  syms = exported(programs.so_path(name))
  decl = declared("srdatalog.h")
  assert {"srdatalog_init", "srdatalog_load_csv", "srdatalog_load_all", "srdatalog_run",
          "srdatalog_size", "srdatalog_shutdown"} <= set(decl)
  missing = [s for s in decl if s not in syms]
  assert not missing, missing


def test_runtime_loads_without_gpu(built):
  lib = ctypes.CDLL(native.RUNTIME_LIB, mode=ctypes.RTLD_GLOBAL)
  assert hasattr(lib, "srd_prim_sort_dedup")
  assert hasattr(lib, "srd_engine_create")
