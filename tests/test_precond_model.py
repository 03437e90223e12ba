"""CPU model of the mixed path's QR preconditioner (the steps the kernels of
jacobi-svd-cuda-opt_amd/csrc/jsvd_kernels_precond.hip take, restated in numpy
loops) with the oracle's one-sided Jacobi on X = R^T.

No reference counterpart (the reference is plain FP64 Jacobi).  What is
pinned here, independently of the GPU:
  * the Householder QR (no pivoting; LAPACK dlarfg's beta = -sign(alpha)
    ||x||, tau = (beta - alpha) / beta) reproduces A = Q1 R to rounding and
    matches numpy's (LAPACK) R up to row signs;
  * U = Q1 Vx (reflectors applied in reverse) and V = P Ux turn the SVD of
    X = R^T into an SVD of A: ||A - U S V^T|| / ||A||, ||U^T U - I|| and
    ||V^T V - I|| at rounding level, sigma equal to the oracle's own solve of
    A, and the sweep count on X about half of that on A;
  * the completion of exactly zero singular values keeps V orthogonal.
A plausible slip (tau from the wrong sign, the permutation applied to rows
instead of the inverse, reflectors applied forward) fails these checks.
"""
import numpy as np
import pytest

import oracle


def qr_model(a):
    """Householder QR as the kernel computes it: no pivoting (perm = identity),
    tau and beta by LAPACK dlarfg's choice, the trailing columns' norms below
    the next row recomputed after every update."""
    m, n = a.shape
    A = a.astype(np.float64).copy()
    refl = np.zeros((m, n))
    tau = np.zeros(n)
    for k in range(n):
        alpha = A[k, k]
        x2 = float(np.dot(A[k + 1:, k], A[k + 1:, k]))
        if x2 == 0.0:
            t, beta, scale = 0.0, alpha, 0.0
        else:
            beta = -np.copysign(np.sqrt(alpha * alpha + x2), alpha)
            t, scale = (beta - alpha) / beta, 1.0 / (alpha - beta)
        v = np.zeros(m)
        v[k] = 1.0
        v[k + 1:] = A[k + 1:, k] * scale
        A[k, k] = beta
        A[k + 1:, k] = 0.0
        refl[:, k] = v
        tau[k] = t
        for j in range(k + 1, n):
            d = t * np.dot(v[k:], A[k:, j])
            A[k:, j] -= d * v[k:]
    R = np.triu(A[:n, :])
    return R, refl, tau, np.arange(n)


def finish_model(refl, tau, perm, ux, s, vx, eps):
    m, n = refl.shape
    Y = np.zeros((m, n))
    Y[:n, :] = vx
    for k in range(n - 1, -1, -1):
        v = refl[:, k]
        Y -= tau[k] * np.outer(v, v @ Y)
    U = np.where(s > eps, Y, 0.0)
    V = np.zeros((n, n))
    V[perm, :] = ux
    for j in np.where(s == 0.0)[0]:
        others = [c for c in range(n) if c != j]
        cov = np.sum(V[:, others] ** 2, axis=1)
        r = np.zeros(n)
        r[int(np.argmin(cov))] = 1.0
        for _ in range(2):
            for c in others:
                r -= (V[:, c] @ r) * V[:, c]
        V[:, j] = r / np.linalg.norm(r)
    return U, V


def solve_model(a, eps=1e-15, max_sweeps=60):
    R, refl, tau, perm = qr_model(a)
    X = R.T.copy()
    ux, s, vx, sweeps = oracle.solve(X, eps, max_sweeps)
    # The kernels normalise every nonzero column of X's final working matrix
    # (B / sigma); the oracle returns those columns only above eps (its U
    # rule), so columns with 0 < sigma <= eps go through the completion
    # below here -- the kernels' path for sigma = 0 exactly.
    s = np.where(s > eps, s, 0.0)
    U, V = finish_model(refl, tau, perm, ux, s, vx, eps)
    return U, s, V, sweeps, R, refl, tau, perm


@pytest.mark.parametrize("shape,kappa", [((40, 30), 1e4), ((64, 64), 1e7), ((96, 50), 1e2), ((17, 5), 1e10)])
def test_preconditioned_solve_is_an_svd_of_a(shape, kappa):
    rng = np.random.default_rng(int(kappa) % 97 + shape[0])
    m, n = shape
    sig = kappa ** (-np.arange(n) / max(n - 1, 1))
    q1, _ = np.linalg.qr(rng.standard_normal((m, n)))
    q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
    a = (q1 * sig) @ q2.T
    U, s, V, sw, R, refl, tau, perm = solve_model(a)
    # The QR itself: A = Q1 R, and the same R as LAPACK's up to row signs.
    Q1 = np.eye(m)[:, :n].copy()
    for k in range(n - 1, -1, -1):
        v = refl[:, k]
        Q1 -= tau[k] * np.outer(v, v @ Q1)
    assert np.linalg.norm(a[:, perm] - Q1 @ R) / np.linalg.norm(a) < 1e-14
    _, r_np = np.linalg.qr(a)
    np.testing.assert_allclose(np.abs(R), np.abs(r_np), rtol=1e-8, atol=1e-13 * abs(R[0, 0]))
    # The SVD of A it yields.
    _, so, _, swo = oracle.solve(a, 1e-15, 60)
    assert np.max(np.abs(np.sort(s) - np.sort(so))) <= 1e-13 * so.max()
    assert np.linalg.norm(a - (U * s) @ V.T) / np.linalg.norm(a) < 1e-13
    assert np.linalg.norm(U.T @ U - np.eye(n)) < 1e-12 * n
    assert np.linalg.norm(V.T @ V - np.eye(n)) < 1e-12 * n
    assert sw <= swo


def test_zero_singular_values_keep_v_orthogonal():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((20, 8))
    a[:, 2] = 0.0
    a[:, 5] = 0.0
    U, s, V, *_ = solve_model(a, eps=1e-13)
    # Without pivoting the zero columns of A become zero rows of R, so X =
    # R^T keeps them as rounding-level singular values (not exact zeros);
    # the U rule zeroes their U columns and V stays orthogonal.
    tiny = s <= 1e-14 * s.max()
    assert np.sum(tiny) == 2
    assert np.linalg.norm(V.T @ V - np.eye(8)) < 1e-13
    assert np.all(U[:, tiny] == 0.0)
    # An all-zero matrix: every sigma is exactly 0 and V is completed to an
    # orthogonal matrix from unit vectors.
    U0, s0, V0, *_ = solve_model(np.zeros((6, 4)), eps=1e-13)
    assert np.all(s0 == 0.0) and np.all(U0 == 0.0)
    assert np.linalg.norm(V0.T @ V0 - np.eye(4)) < 1e-15
    assert np.linalg.norm(a - (U * s) @ V.T) / np.linalg.norm(a) < 1e-14


def test_a_planted_slip_fails():
    # Applying the reflectors forward instead of in reverse (a plausible
    # mistake) breaks the reconstruction.
    rng = np.random.default_rng(4)
    a = rng.standard_normal((30, 12))
    R, refl, tau, perm = qr_model(a)
    ux, s, vx, _ = oracle.solve(R.T.copy(), 1e-15, 60)
    m, n = refl.shape
    Y = np.zeros((m, n))
    Y[:n] = vx
    for k in range(n):
        v = refl[:, k]
        Y -= tau[k] * np.outer(v, v @ Y)
    V = np.zeros((n, n))
    V[perm] = ux
    assert np.linalg.norm(a - (Y * s) @ V.T) / np.linalg.norm(a) > 1e-3
