"""CPU oracle for the semi-naive fixpoint path.

TEST INFRASTRUCTURE ONLY.  ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import it as the checker / CPU
baseline; nothing in the ``srdatalog_hip`` package may import it.

* ``oracle.run``   -- ctypes front of ``liborc.so`` (oracle.cpp): the C++
                      restatement of the reference semi-naive loop.
* ``oracle.brute`` -- pure-Python naive Jacobi evaluation for tiny inputs,
                      used to pin the C++ restatement.
"""

from .oracle import OracleResult, build, run  # noqa: F401
