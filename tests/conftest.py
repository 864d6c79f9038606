import fcntl
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "srdatalog-python_amd")
for p in (ROOT, PKG):
  if p not in sys.path:
    sys.path.insert(0, p)


def pytest_configure(config):
  config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


CPU_SIM = os.path.join(ROOT, "tests", "cpu_sim")


def build_cpu_sim(target: str) -> str:
  """make one library of tests/cpu_sim (rebuilt when a source changed),
  serialised across test processes by a file lock; returns its path."""
  with open(os.path.join(CPU_SIM, ".build.lock"), "w") as lk:
    fcntl.flock(lk, fcntl.LOCK_EX)
    subprocess.run(["make", "-s", "-C", CPU_SIM, target], check=True)
  return os.path.join(CPU_SIM, target)
