"""Seeded synthetic workloads of the benchmark configurations.

Base seed 20260515 follows the reference generator's default
(psrc/generator/common.py:167); matrix k of a batch uses seed base + k
(common.py:459-460).  numpy's PCG64 stream is the generator, so the same
inputs come out on every host.  Haar-random factors are the Q of a Gaussian
QR with columns sign-fixed by sign(diag R).
"""
from __future__ import annotations

import numpy as np

BASE_SEED = 20260515


def haar(rng: np.random.Generator, m: int, n: int) -> np.ndarray:
    g = rng.standard_normal((m, n))
    q, r = np.linalg.qr(g)
    d = np.sign(np.diag(r))
    d[d == 0] = 1.0
    return q * d


def constructed(rng: np.random.Generator, m: int, n: int, sigma: np.ndarray) -> np.ndarray:
    """A = Q1 diag(sigma) Q2^T with Haar Q1 (m x n) and Q2 (n x n)."""
    q1 = haar(rng, m, n)
    q2 = haar(rng, n, n)
    return np.ascontiguousarray((q1 * sigma) @ q2.T)


def gaussian_square(n: int = 256, seed: int = BASE_SEED) -> np.ndarray:
    """Single n x n matrix with iid N(0,1) entries (256 x 256 FP64 workload)."""
    return np.random.default_rng(seed).standard_normal((n, n))


def uneven_batch(count: int = 4096, seed: int = BASE_SEED, n_lo: int = 8, n_hi: int = 64, decades: float = 6.0):
    """count matrices, n ~ U{n_lo..n_hi}, m ~ U{n..2n}, sigma log-spaced 1 .. 10^-decades.

    Returns (mats, sigmas): the row-major matrices and their constructed
    singular values (descending)."""
    mats, sigmas = [], []
    for k in range(count):
        rng = np.random.default_rng(seed + k)
        n = int(rng.integers(n_lo, n_hi + 1))
        m = int(rng.integers(n, 2 * n + 1))
        sig = 10.0 ** (-decades * np.arange(n) / max(n - 1, 1))
        mats.append(constructed(rng, m, n, sig))
        sigmas.append(sig)
    return mats, sigmas


def conditioned_batch(count: int = 1024, m: int = 128, n: int = 128, seed: int = BASE_SEED,
                      log_kappa=(3.0, 8.0)):
    """count m x n matrices with kappa = 10^U(log_kappa) and sigma_k = kappa^(-k/(n-1))."""
    mats, sigmas = [], []
    for k in range(count):
        rng = np.random.default_rng(seed + k)
        kappa = 10.0 ** rng.uniform(*log_kappa)
        sig = kappa ** (-np.arange(n) / (n - 1))
        mats.append(constructed(rng, m, n, sig))
        sigmas.append(sig)
    return mats, sigmas


def gapped_batch(count: int = 512, m: int = 96, n: int = 64, seed: int = BASE_SEED, decades: float = 2.0):
    """count m x n matrices with sigma_k = 10^(-decades k/(n-1)) (relative gaps >= 1e-2)."""
    mats, sigmas = [], []
    for k in range(count):
        rng = np.random.default_rng(seed + k)
        sig = 10.0 ** (-decades * np.arange(n) / (n - 1))
        mats.append(constructed(rng, m, n, sig))
        sigmas.append(sig)
    return mats, sigmas


def tall_geometric(m: int = 8192, n: int = 2048, seed: int = BASE_SEED, decades: float = 10.0):
    """Single m x n matrix with sigma_k = 10^(-decades k/(n-1)) and Haar factors."""
    rng = np.random.default_rng(seed)
    sig = 10.0 ** (-decades * np.arange(n) / (n - 1))
    return constructed(rng, m, n, sig), sig


def pack(mats):
    """Concatenate row-major matrices into the packed layout; returns (flat, ms, ns)."""
    ms = np.array([x.shape[0] for x in mats], dtype=np.int64)
    ns = np.array([x.shape[1] for x in mats], dtype=np.int64)
    flat = np.concatenate([np.ascontiguousarray(x, dtype=np.float64).ravel() for x in mats])
    return flat, ms, ns
