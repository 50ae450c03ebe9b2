"""Regression fixture for the product's R-MAT generator (no reference generator
exists): sha256 of the CSR (row offsets then column ids, little-endian uint32)
and the edge count, for scales 10 and 12 at edge factor 16, (a,b,c) =
(0.57, 0.19, 0.19).  Run: python tests/golden/make_rmat_fixture.py
"""
import ctypes as C
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "tpds22-artifact_amd"))
import hsync  # noqa: E402


def main():
    lib = hsync.load()
    out = {}
    for scale, ef, seed in ((10, 16, 1), (12, 16, 7)):
        m = C.c_uint64(0)
        lib.hs_rmat_csr_host(scale, ef, 0.57, 0.19, 0.19, seed, None, None, C.byref(m))
        row = np.empty((1 << scale) + 1, np.uint32)
        col = np.empty(m.value, np.uint32)
        lib.hs_rmat_csr_host(scale, ef, 0.57, 0.19, 0.19, seed, row.ctypes.data, col.ctypes.data,
                             C.byref(m))
        out[f"{scale}:{ef}:{seed}"] = {"sha256": hashlib.sha256(row.tobytes() + col.tobytes()).hexdigest(),
                                       "nedges": int(m.value)}
    json.dump(out, open(os.path.join(HERE, "rmat.json"), "w"), indent=1)


if __name__ == "__main__":
    main()
