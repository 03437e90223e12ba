"""Error metrics of a computed SVD: sigma error against a reference solve,
orthogonality of U and V, and the reconstruction residual.  None of these
exist in the reference; they are the numbers every result is reported with.
"""
from __future__ import annotations

import numpy as np


def sigma_errors(sigma, sigma_ref):
    """(normwise, componentwise) error of sigma against sigma_ref, both sorted.

    normwise = max_k |s_k - r_k| / max_k r_k; componentwise = max_k |s_k - r_k| / r_k
    over the r_k > 0."""
    s = np.sort(np.asarray(sigma, dtype=np.float64))[::-1]
    r = np.sort(np.asarray(sigma_ref, dtype=np.float64))[::-1]
    if s.shape != r.shape:
        raise ValueError("sigma shapes differ")
    if r.size == 0:
        return 0.0, 0.0
    d = np.abs(s - r)
    top = r.max()
    norm = float(d.max() / top) if top > 0 else float(d.max())
    pos = r > 0
    comp = float((d[pos] / r[pos]).max()) if pos.any() else 0.0
    return norm, comp


def orthogonality(q, mask=None) -> float:
    """||Q^T Q - I||_F over the columns selected by mask (all by default)."""
    q = np.asarray(q, dtype=np.float64)
    if mask is not None:
        q = q[:, np.asarray(mask, dtype=bool)]
    k = q.shape[1]
    return float(np.linalg.norm(q.T @ q - np.eye(k)))


def residual(a, u, sigma, v) -> float:
    """||A - U diag(sigma) V^T||_F / ||A||_F."""
    a = np.asarray(a, dtype=np.float64)
    rec = (np.asarray(u) * np.asarray(sigma)) @ np.asarray(v).T
    na = np.linalg.norm(a)
    return float(np.linalg.norm(a - rec) / (na if na > 0 else 1.0))


def report(a, u, sigma, v, sigma_ref=None, epsilon: float = 1e-9) -> dict:
    out = {
        "orth_u": orthogonality(u, np.asarray(sigma) > epsilon),
        "orth_v": orthogonality(v),
        "residual": residual(a, u, sigma, v),
    }
    if sigma_ref is not None:
        out["sigma_err_norm"], out["sigma_err_comp"] = sigma_errors(sigma, sigma_ref)
    return out
