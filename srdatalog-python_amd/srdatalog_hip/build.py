"""Project builder: program -> HIP source tree -> shared library.

Mirrors the reference's two-call flow (src/srdatalog/build.py:42
``build_project`` and src/srdatalog/codegen/jit/compiler.py:300
``compile_jit_project``): ``build_project`` writes the generated HIP
translation unit into a cache directory, ``compile_jit_project`` compiles it
for gfx950 with hipcc and links it against the in-tree runtime
(``libsrdatalog_hip.so``).  The resulting library exports the reference's
six ``srdatalog_*`` symbols (include/srdatalog.h).
"""

from __future__ import annotations

import hashlib
import os
import subprocess
from dataclasses import dataclass, field

from . import mir as M
from . import native
from .codegen import RelDecl, generate_program_source


@dataclass
class BuildResult:
  artifact: str = ""
  returncode: int = 0
  stderr: str = ""
  compile_results: list = field(default_factory=list)

  def ok(self) -> bool:
    return self.returncode == 0 and bool(self.artifact)


def default_cache_base() -> str:
  return os.environ.get("SRDATALOG_HIP_CACHE",
                        os.path.join(os.path.expanduser("~"), ".cache", "srdatalog_hip"))


def build_project_from_mir(project_name: str, decls: list[RelDecl], prog: M.Program,
                           cache_base: str | None = None) -> dict:
  """Write the HIP source of a MIR program; returns {"dir", "main", "name"}."""
  src = generate_program_source(project_name, decls, prog)
  digest = hashlib.sha256(src.encode()).hexdigest()[:8]
  base = cache_base or default_cache_base()
  d = os.path.join(base, f"{project_name}_{digest}")
  os.makedirs(d, exist_ok=True)
  main = os.path.join(d, f"{project_name}.hip")
  with open(main, "w") as f:
    f.write(src)
  return {"dir": d, "main": main, "name": project_name}


def build_project(program, project_name: str, *, cache_base: str | None = None) -> dict:
  """Compile a DSL Program (srdatalog_hip.dsl) through HIR/MIR and write its HIP tree."""
  from .compiler import compile_to_mir, relation_decls
  return build_project_from_mir(project_name, relation_decls(program), compile_to_mir(program),
                                cache_base=cache_base)


def compile_jit_project(project: dict, out_path: str | None = None) -> BuildResult:
  """hipcc-compile a project written by build_project into a shared library."""
  main = project["main"]
  so = out_path or os.path.join(project["dir"], f"lib{project['name']}.so")
  if not os.path.exists(native.RUNTIME_LIB):
    native.build_runtime()
  lib_dir = os.path.dirname(native.RUNTIME_LIB)
  tmp = so + ".tmp"
  rpaths = [f"-Wl,-rpath,{lib_dir}"]
  rel = os.path.relpath(lib_dir, os.path.dirname(os.path.abspath(so)))
  rpaths.insert(0, "-Wl,-rpath,$ORIGIN/" + rel)
  cmd = [native.hipcc(), *native.hip_flags(), "-shared", "-x", "hip", main, "-o", tmp,
         f"-L{lib_dir}", "-lsrdatalog_hip", *rpaths]
  r = subprocess.run(cmd, capture_output=True, text=True)
  res = BuildResult(returncode=r.returncode, stderr=r.stderr, compile_results=[r])
  if r.returncode == 0:
    os.replace(tmp, so)
    res.artifact = so
  return res
