"""Writes tests/golden/golden.npz: oracle outputs (long double, rounded to
double) for small seeded problems, so GPU tests can check against frozen
values without rebuilding the oracle.

Inputs are re-derivable: uniform [-1, 1) from SplitMix64 with the seed
listed in the 'meta' entry (same generator as oracle.uniform and the device
fill).  Run from the repo root: python3 tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

C2C_SIZES = [1, 2, 3, 4, 5, 7, 8, 11, 12, 13, 16, 17, 31, 32, 60, 64, 97, 100, 128, 1000, 1001, 4093]
R2C_SIZES = [1, 2, 3, 8, 15, 16, 100, 1000]
MULTI = [(8, 12), (4, 6, 5)]
R2R_SIZES = [2, 3, 8, 15, 16, 64, 100]          # every kind 0..10 (fftw_r2r_kind)
R2R_MULTI = [((8, 6), (5, 9)), ((4, 5, 6), (3, 2, 10))]


def main():
    out = {}
    meta = {"c2c": [], "r2c": [], "multi": [], "r2r": [], "r2r_multi": []}
    for n in C2C_SIZES:
        seed = 100 + n
        x = oracle.uniform_complex(n, seed)
        out["c2c_fwd_%d" % n] = oracle.fftn(x, -1)
        out["c2c_bwd_%d" % n] = oracle.fftn(x, +1)
        meta["c2c"].append([n, seed])
    for n in R2C_SIZES:
        seed = 200 + n
        x = oracle.uniform(n, seed)
        X = oracle.rfftn(x)
        out["r2c_%d" % n] = X
        out["c2r_%d" % n] = oracle.irfftn(X, n)
        meta["r2c"].append([n, seed])
    for shape in MULTI:
        seed = 300 + len(shape)
        x = oracle.uniform_complex(int(np.prod(shape)), seed).reshape(shape)
        out["multi_%s" % "x".join(map(str, shape))] = oracle.fftn(x, -1)
        meta["multi"].append([list(shape), seed])
    for n in R2R_SIZES:
        seed = 400 + n
        x = oracle.uniform(n, seed)
        for kind in range(11):
            out["r2r_%d_%d" % (kind, n)] = oracle.r2rn(x, [kind])
        meta["r2r"].append([n, seed])
    for shape, kinds in R2R_MULTI:
        seed = 500 + len(shape)
        x = oracle.uniform(int(np.prod(shape)), seed).reshape(shape)
        out["r2r_multi_%s" % "x".join(map(str, shape))] = oracle.r2rn(x, list(kinds))
        meta["r2r_multi"].append([list(shape), list(kinds), seed])
    out["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.npz"), **out)


if __name__ == "__main__":
    main()
