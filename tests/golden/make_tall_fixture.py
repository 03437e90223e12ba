"""Writes tests/golden/tall_8192x2048.npz: the CPU oracle's solve of the
8192 x 2048 geometric-spectrum matrix (sigma_k = 10^(-10k/2047), Haar factors,
seed 20260515; jsvd.gen.tall_geometric) at eps = 1e-15, max_sweeps = 80.

Stored (small: the matrix itself is regenerated from its seed):
  * sigma_sorted: the oracle's singular values, sorted descending;
  * sigma_order: the oracle's sigma in its own final column order;
  * sweeps: the oracle's sweep count (the verifying sweep included);
  * input_checksum: sum, sum of squares and the first eight entries of A, so a
    test can tell a different input (another numpy/LAPACK) from a solver
    difference;
  * seconds, threads: how long the oracle took here.
The oracle runs round-parallel over all host cores (OpenMP over the disjoint
pairs of a round); that is bitwise equal to its serial solve.  Run from the
repository root:
    python tests/golden/make_tall_fixture.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, os.path.join(REPO, "jacobi-svd-cuda-opt_amd"))
import oracle  # noqa: E402
from jsvd import gen  # noqa: E402

EPS, MAX_SWEEPS = 1e-15, 80


def checksum(a: np.ndarray) -> np.ndarray:
    flat = a.ravel()
    return np.concatenate([[flat.sum(), np.dot(flat, flat)], flat[:8]])


def main():
    a, sig = gen.tall_geometric()
    threads = os.cpu_count() or 1
    t0 = time.time()
    _, s, _, sw = oracle.solve(a, EPS, MAX_SWEEPS, threads=threads)
    dt = time.time() - t0
    err = float(np.max(np.abs(np.sort(s)[::-1] - sig)) / sig[0])
    print(f"oracle 8192x2048: {sw} sweeps in {dt:.1f} s on {threads} threads; "
          f"sigma vs constructed spectrum {err:.2e}", flush=True)
    np.savez(os.path.join(HERE, "tall_8192x2048.npz"), sigma_sorted=np.sort(s)[::-1].copy(), sigma_order=s,
             sweeps=np.array(sw, dtype=np.int32), input_checksum=checksum(a),
             seconds=np.array(dt), threads=np.array(threads, dtype=np.int32),
             eps=np.array(EPS), max_sweeps=np.array(MAX_SWEEPS, dtype=np.int32))


if __name__ == "__main__":
    main()
