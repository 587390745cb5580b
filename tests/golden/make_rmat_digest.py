"""Writes tests/golden/rmat16.json: size and SHA-256 of the oracle's RMAT
scale-16 CSR (edge factor 8, a/b/c = .57/.19/.19, key 42), the config-1 graph.
The GPU generator must reproduce the same bytes (tests/test_gpu_parity.py)."""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import numpy as np  # noqa: E402
import oracle as O  # noqa: E402

off, adj = O.rmat(16, 8)
out = {"V": len(off) - 1, "E": int(len(adj)), "max_degree": int(np.diff(off).max()),
       "sha256": hashlib.sha256(off.tobytes() + adj.tobytes()).hexdigest()}
json.dump(out, open(os.path.join(HERE, "rmat16.json"), "w"), indent=1)
print(out)
