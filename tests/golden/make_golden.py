"""Writes tests/golden/golden_small.npz.

Inputs are small hand-picked and seeded matrices.  Expected values:
  * sigma_mp_<k>: singular values from mpmath at 50 digits (descending),
    an independent ground truth;
  * u/s/v/sweeps_<k>: the CPU oracle's full result at eps=1e-15,
    max_sweeps=30 (a regression fixture for the oracle and the GPU path).
The reference ships no golden vectors for this path, so these fixtures are
this repository's own.  Run from the repository root:
    python tests/golden/make_golden.py
"""
import os
import sys

import mpmath
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))
import oracle  # noqa: E402

EPS, MAX_SWEEPS = 1e-15, 30


def cases():
    rng = np.random.default_rng(20260515)
    out = [
        np.array([[3.0, 1.0], [1.0, 2.0]]),
        np.array([[1.0, 2.0], [3.0, 4.0], [5.0, 6.0]]),
        np.array([[4.0, 1.0, -2.0], [1.0, 2.0, 0.0], [-2.0, 0.0, 3.0], [1.0, 1.0, 1.0]]),
        np.array([[1.0 / (i + j + 1) for j in range(5)] for i in range(5)]),
    ]
    z = rng.standard_normal((6, 4))
    z[:, 2] = 0.0
    out.append(z)
    out.append(rng.standard_normal((8, 6)))
    g = rng.standard_normal((12, 7)) * (10.0 ** -np.arange(7))
    out.append(g)
    out.append(rng.standard_normal((9, 9)))
    return out


def main():
    mpmath.mp.dps = 50
    data = {}
    for k, a in enumerate(cases()):
        s_mp = mpmath.svd_r(mpmath.matrix(a.tolist()), compute_uv=False)
        s_mp = np.array(sorted([float(x) for x in s_mp], reverse=True))
        u, s, v, sw = oracle.solve(a, EPS, MAX_SWEEPS)
        data[f"a_{k}"] = a
        data[f"sigma_mp_{k}"] = s_mp
        data[f"u_{k}"] = u
        data[f"s_{k}"] = s
        data[f"v_{k}"] = v
        data[f"sweeps_{k}"] = np.array(sw, dtype=np.int32)
    data["count"] = np.array(len(cases()), dtype=np.int32)
    np.savez(os.path.join(HERE, "golden_small.npz"), **data)
    print("wrote", len(cases()), "cases")


if __name__ == "__main__":
    main()
