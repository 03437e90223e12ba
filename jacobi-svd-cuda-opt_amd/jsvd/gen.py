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


# ---- stress mix (robustness cases) ---------------------------------------
# The same scenario kinds as the reference generator
# (psrc/generator/common.py:340-451): dense (uniform, normal, log-uniform, each
# with a small smooth sin term), diagonal-dominant, low-rank plus noise,
# ill-conditioned factor product plus noise, sparse and deterministic zero
# columns.  The random streams are numpy PCG64, not the reference's
# random.Random, so the values differ; the structure of every kind is the
# same.  Defaults follow the reference CLI: rank 4, condition 1e8,
# sparsity 0.5 (psrc/generator/common.py:195-210).

STRESS_KINDS = ("uniform", "normal", "log-uniform", "diagonal-dominant", "low-rank", "ill-conditioned",
                "sparse", "zero-columns")


def stress_matrix(kind: str, m: int, n: int, rng: np.random.Generator, rank: int = 4, condition: float = 1e8,
                  sparsity: float = 0.5, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    """One m x n matrix of the given stress kind."""
    if kind in ("uniform", "normal", "log-uniform"):
        if kind == "uniform":
            x = rng.uniform(lo, hi, (m, n))
        elif kind == "normal":
            x = rng.standard_normal((m, n))
        else:
            amin = max(min(abs(lo), abs(hi)), 1e-300)
            amax = max(abs(lo), abs(hi), amin)
            mag = np.exp(rng.uniform(np.log(amin), np.log(amax), (m, n)))
            x = np.where(rng.random((m, n)) < 0.5, mag, -mag)
        rows = np.arange(1, m + 1)[:, None]
        cols = np.arange(1, n + 1)[None, :]
        return x + 0.05 * np.sin(0.013 * rows * cols)
    if kind == "diagonal-dominant":
        scale = max(abs(lo), abs(hi), 1.0)
        x = 0.01 * rng.uniform(lo, hi, (m, n))
        k = min(m, n)
        x[np.arange(k), np.arange(k)] += scale + np.abs(rng.uniform(lo, hi, k))
        return x
    if kind in ("low-rank", "ill-conditioned"):
        r = min(rank, m, n)
        left = rng.standard_normal((m, r))
        right = rng.standard_normal((r, n))
        if kind == "low-rank":
            return left @ right + 1e-6 * rng.standard_normal((m, n))
        scales = condition ** (-np.arange(r) / max(r - 1, 1))
        return (left * scales) @ right + 1e-12 * rng.standard_normal((m, n))
    if kind == "sparse":
        x = rng.uniform(lo, hi, (m, n))
        x[rng.random((m, n)) < sparsity] = 0.0
        return x
    if kind == "zero-columns":
        x = rng.uniform(lo, hi, (m, n))
        stride = max(2, n // 4)
        x[:, np.arange(n) % stride == 0] = 0.0
        return x
    raise ValueError(f"unsupported stress kind: {kind}")


def stress_mix(count: int, shapes, seed: int = BASE_SEED, kinds=STRESS_KINDS, **kw):
    """count matrices cycling through kinds and shapes ((m, n) pairs, m >= n);
    matrix k uses seed + k.  Returns (mats, kinds_used)."""
    mats, used = [], []
    for k in range(count):
        rng = np.random.default_rng(seed + k)
        kind = kinds[k % len(kinds)]
        m, n = shapes[(k // len(kinds)) % len(shapes)]
        mats.append(np.ascontiguousarray(stress_matrix(kind, m, n, rng, **kw)))
        used.append(kind)
    return mats, used
