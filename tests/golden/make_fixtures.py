"""Writes the seeded pivot fixtures used by the GPU parity tests.

Inputs come from lair_amd.inputs (SplitMix64 + Marsaglia polar, basic IEEE
operations only, so the bits are identical on every host); pivots come from the
CPU oracle's restatement of getrf_row_major (src/lapack/getrf.rs:46-120).  Each
fixture records the generator, seed, shape and SHA-256 prefix of the matrix so
a test can prove it regenerated the same input.

The oracle's per-row updates are independent within a step, so its
multi-threaded run is bit-identical to the single-threaded one.

Usage: python tests/golden/make_fixtures.py [--threads 8]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from lair_amd import inputs  # noqa: E402
from oracle import oracle as O  # noqa: E402

CASES = [
    # name, seed, n, nrhs
    ("c1_n512", 512, 512, 1),
    ("c2_n8192", 8192, 8192, 16),
]


def make(name, seed, n, nrhs, threads):
    t0 = time.time()
    a = inputs.gaussian_matrix(seed, n, n)
    b = inputs.gaussian_matrix(seed + 1, n, nrhs)
    da, db = inputs.digest(a), inputs.digest(b)
    lu = a.copy()
    piv, sing = O.getrf(lu, nthreads=threads)
    x = O.getrs(lu, piv, b)
    r = O.residual(a, x, b, nthreads=threads)
    na = np.abs(a).sum(1).max()
    berr = float((np.abs(r).max(0) / (na * np.abs(x).max(0))).max())
    out = {
        "name": name,
        "generator": "lair_amd.inputs.gaussian_matrix (SplitMix64 + Marsaglia polar)",
        "seed_a": seed, "seed_b": seed + 1, "n": n, "nrhs": nrhs,
        "digest_a": da, "digest_b": db,
        "singular": sing,
        "pivots": [int(p) for p in piv],
        "oracle_backward_error": berr,
        "oracle": "oracle/lair_oracle.c lo_dgetrf_row_major + lo_dgetrs",
    }
    path = os.path.join(HERE, f"{name}.json")
    with open(path, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(f"{name}: n={n} digest={da} berr={berr:.2e} swaps={int((piv != np.arange(n)).sum())} "
          f"{time.time() - t0:.1f}s -> {path}", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--only", default=None)
    args = ap.parse_args()
    for name, seed, n, nrhs in CASES:
        if args.only and args.only != name:
            continue
        make(name, seed, n, nrhs, args.threads)


if __name__ == "__main__":
    main()
