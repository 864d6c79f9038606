"""Writes andersen_scaled.json: the Andersen fixpoint (per-iteration |Delta|,
|PointsTo| and the order-independent tuple hash of oracle/sparse.py) of the
seeded Andersen workload at 1/10 of its full size (20,000 variables, 5,000
objects; AddressOf 10k, Assign 40k, Load 10k, Store 10k pairs, Zipf 1.1
endpoints, seed 0), computed by the sparse-matrix restatement, which
tests/test_oracle_sparse.py pins against the C++ oracle.  Takes a few
minutes on one core:  python tests/golden/gen_andersen_scaled.py"""

import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "srdatalog-python_amd"))

from oracle import sparse  # noqa: E402
from srdatalog_hip import datasets as ds  # noqa: E402

CASES = [{"seed": 0, "div": 10}]


def main():
  out = []
  for c in CASES:
    d = c["div"]
    a = ds.andersen_inputs(c["seed"], n_vars=200_000 // d, n_objs=50_000 // d,
                           sizes=(100_000 // d, 400_000 // d, 100_000 // d, 100_000 // d))
    t = time.time()
    rows, deltas = sparse.andersen(a["AddressOf"], a["Assign"], a["Load"], a["Store"])
    out.append({**c, "deltas": deltas, "size": int(rows.shape[0]),
                "hash": str(sparse.tuple_hash(rows)), "seconds": round(time.time() - t, 1)})
    print(out[-1], flush=True)
  with open(os.path.join(HERE, "andersen_scaled.json"), "w") as f:
    json.dump(out, f, indent=1)


if __name__ == "__main__":
  main()
