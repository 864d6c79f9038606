"""CPU oracle for the semi-naive fixpoint path.

TEST INFRASTRUCTURE ONLY.  ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import it as the checker / CPU
baseline; nothing in the ``srdatalog_hip`` package may import it.

* ``oracle.run``   -- ctypes front of ``liborc.so`` (oracle.cpp): the C++
                      restatement of the reference semi-naive loop.
* ``oracle.brute`` -- pure-Python naive Jacobi evaluation for tiny inputs,
                      used to pin the C++ restatement.
* ``oracle.digest_rows`` / ``oracle.triangle_digest`` -- order-independent
                      tuple-set digests, for parity at sizes whose tuple sets
                      are too large to compare row by row.
"""

from .oracle import (OracleResult, build, digest_rows, run, set_key_stride,  # noqa: F401
                     triangle_digest)
