"""Pins the CPU oracle (oracle/jsvd_oracle.c) before it is trusted as the checker.

Each check is independent of the oracle's own formulas: hand-executed
reference schedules, brute-force invariants, the 2x2 Gram identity, LAPACK
(numpy gesdd, scipy dgejsv), mpmath at 50 digits, and closed-form special
cases.  The reference ships no numeric test or golden vector for this path
(its only test is a CLI-parser test), so these are the pins.
"""
import itertools
import math
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "golden_small.npz")


# ---------------------------------------------------------------- schedule --
# Tables written out by hand-executing the reference's host schedule
# (src/domain/jacobi_schedule.cuh:16-57): slot 0 fixed, slots 1..E-1 rotated
# right by one after each round, pair (slot i, slot E-1-i), dummies dropped.
HAND_TABLES = {
    4: [[(0, 3), (1, 2)], [(0, 2), (3, 1)], [(0, 1), (2, 3)]],
    5: [[(1, 4), (2, 3)], [(0, 4), (1, 2)], [(0, 3), (4, 2)], [(0, 2), (3, 1)], [(0, 1), (3, 4)]],
    2: [[(0, 1)]],
    3: [[(1, 2)], [(0, 2)], [(0, 1)]],
}


@pytest.mark.parametrize("n", sorted(HAND_TABLES))
def test_schedule_matches_hand_table(n):
    assert oracle.schedule(n) == HAND_TABLES[n]


@pytest.mark.parametrize("n", list(range(2, 17)) + [31, 64, 65])
def test_schedule_round_robin_invariants(n):
    rounds = oracle.schedule(n)
    even = n + (n & 1)
    assert len(rounds) == even - 1
    seen = set()
    for pairs in rounds:
        cols = [c for p in pairs for c in p]
        assert len(cols) == len(set(cols)), "pairs in a round must be disjoint"
        assert all(0 <= c < n for c in cols)
        for p, q in pairs:
            key = (min(p, q), max(p, q))
            assert key not in seen, "each pair appears once per sweep"
            seen.add(key)
    assert seen == set(itertools.combinations(range(n), 2))


@pytest.mark.parametrize("n", [2, 3, 4, 5, 8, 9, 33, 64])
def test_schedule_closed_form_equals_table(n):
    # The device form (column at a slot position, jacobi_rotation_kernels.cuh:32-42)
    # must reproduce the host table, orientation included.
    even = n + (n & 1)
    rounds = oracle.schedule(n)
    for r, pairs in enumerate(rounds):
        closed = []
        for k in range(even // 2):
            p, q = oracle.column_at(k, r, n), oracle.column_at(even - 1 - k, r, n)
            if p >= 0 and q >= 0:
                closed.append((p, q))
        assert closed == pairs


# ---------------------------------------------------------------- rotation --
@pytest.mark.parametrize("pp,qq,pq", [(4.0, 1.0, 0.5), (1.0, 4.0, 0.5), (1.0, 4.0, -0.5), (2.0, 2.0, 1.0),
                                      (1e-8, 1.0, 1e-5), (3.0, 3.0, -2.9), (1e3, 1e-3, 0.5)])
def test_rotation_annihilates_gram_offdiagonal(pp, qq, pq):
    rot, c, s = oracle.rotation(pp, qq, pq, 1e-15)
    assert rot
    assert abs(c * c + s * s - 1.0) < 4e-16
    # Rotated Gram: a_p' = c a_p - s a_q, a_q' = s a_p + c a_q.
    pq_new = (c * c - s * s) * pq + c * s * (pp - qq)
    assert abs(pq_new) <= 1e-15 * max(pp, qq, abs(pq))
    # Inner (small-angle) rotation: |t| <= 1, i.e. |s| <= c.
    assert abs(s) <= c + 1e-16
    # Diagonal update identities pp' = pp - t pq, qq' = qq + t pq.
    t = s / c
    pp_new = c * c * pp - 2 * c * s * pq + s * s * qq
    assert math.isclose(pp_new, pp - t * pq, rel_tol=1e-12, abs_tol=1e-15 * max(pp, qq))


def test_rotation_threshold_semantics():
    # |pq| <= eps*sqrt(pp*qq): no rotation, identity returned.
    assert oracle.rotation(1.0, 1.0, 1e-10, 1e-9) == (False, 1.0, 0.0)
    assert oracle.rotation(1.0, 1.0, 2e-9, 1e-9)[0]
    # The absolute 1e-300 floor.
    assert oracle.rotation(0.0, 0.0, 1e-301, 1e-9) == (False, 1.0, 0.0)
    # Negative (rounded) norms are clamped to zero before the sqrt.
    assert oracle.rotation(-1e-30, 1.0, 1e-3, 1e-9)[0]


# ------------------------------------------------------------------- solve --
def _lapack_sigma(a):
    return np.linalg.svd(a, compute_uv=False)


@pytest.mark.parametrize("shape", [(2, 2), (3, 2), (10, 7), (33, 33), (64, 17), (96, 64)])
def test_solve_against_lapack(shape):
    rng = np.random.default_rng(1234 + shape[0] * 7 + shape[1])
    a = rng.standard_normal(shape)
    u, s, v, sw = oracle.solve(a, 1e-15, 30)
    ref = _lapack_sigma(a)
    assert np.max(np.abs(np.sort(s)[::-1] - ref)) <= 1e-13 * ref[0]
    n = shape[1]
    assert np.linalg.norm(u.T @ u - np.eye(n)) < 1e-12
    assert np.linalg.norm(v.T @ v - np.eye(n)) < 1e-12
    assert np.linalg.norm(a - (u * s) @ v.T) <= 1e-13 * np.linalg.norm(a)
    assert 1 <= sw <= 30


def test_solve_graded_against_dgejsv():
    # Column-graded matrix: one-sided Jacobi resolves the tiny singular values
    # to high relative accuracy, as LAPACK's preconditioned Jacobi does.
    from scipy.linalg import lapack

    rng = np.random.default_rng(7)
    a = rng.standard_normal((40, 12)) * (10.0 ** -np.arange(12))
    u, s, v, _ = oracle.solve(a, 1e-15, 30)
    out = lapack.dgejsv(a, joba=0, jobu=3, jobv=3)
    sva, work = out[0], out[3]
    ref = np.sort(sva * (work[1] / work[0]))[::-1]
    rel = np.abs(np.sort(s)[::-1] - ref) / ref
    assert rel.max() < 1e-12


def test_golden_against_mpmath():
    d = np.load(GOLDEN)
    for k in range(int(d["count"])):
        a, ref = d[f"a_{k}"], d[f"sigma_mp_{k}"]
        _, s, _, _ = oracle.solve(a, 1e-15, 30)
        got = np.sort(s)[::-1]
        assert np.max(np.abs(got - ref)) <= 4e-15 * ref[0], k


def test_golden_oracle_regression():
    d = np.load(GOLDEN)
    for k in range(int(d["count"])):
        u, s, v, sw = oracle.solve(d[f"a_{k}"], 1e-15, 30)
        assert sw == int(d[f"sweeps_{k}"])
        np.testing.assert_allclose(s, d[f"s_{k}"], rtol=0, atol=1e-14 * max(1.0, s.max()))
        np.testing.assert_allclose(u, d[f"u_{k}"], rtol=0, atol=1e-12)
        np.testing.assert_allclose(v, d[f"v_{k}"], rtol=0, atol=1e-12)


def test_two_by_two_order_and_sweeps():
    # 2x2 closed form: the inner rotation grows the longer column, so sigma
    # keeps the order of the initial column norms (not sorted), and the solve
    # takes exactly one rotating sweep plus the verifying sweep.
    a = np.array([[1.0, 2.0], [0.5, 3.0]])
    u, s, v, sw = oracle.solve(a, 1e-15, 30)
    ref = _lapack_sigma(a)
    assert s[0] < s[1]
    assert np.allclose(sorted(s), sorted(ref), rtol=1e-15, atol=0)
    assert sw == 2
    b = a[:, ::-1].copy()
    _, s2, _, _ = oracle.solve(b, 1e-15, 30)
    assert s2[0] > s2[1]


def test_orthogonal_columns_need_one_sweep():
    rng = np.random.default_rng(3)
    q, _ = np.linalg.qr(rng.standard_normal((20, 6)))
    scale = np.array([3.0, 0.5, 7.0, 1.0, 2.0, 0.25])
    a = q * scale
    u, s, v, sw = oracle.solve(a, 1e-9, 30)
    assert sw == 1
    np.testing.assert_allclose(s, scale, rtol=1e-14)
    np.testing.assert_array_equal(v, np.eye(6))


def test_diagonal_unsorted_sigma():
    a = np.diag([2.0, -5.0, 1.0, 3.0])
    u, s, v, sw = oracle.solve(a, 1e-9, 30)
    assert sw == 1
    np.testing.assert_array_equal(s, [2.0, 5.0, 1.0, 3.0])
    np.testing.assert_array_equal(u, np.diag([1.0, -1.0, 1.0, 1.0]))
    np.testing.assert_array_equal(v, np.eye(4))


def test_single_column():
    a = np.array([[3.0], [4.0]])
    u, s, v, sw = oracle.solve(a, 1e-9, 30)
    assert sw == 1 and s[0] == 5.0
    np.testing.assert_allclose(u[:, 0], [0.6, 0.8], rtol=1e-15)
    assert v[0, 0] == 1.0


def test_zero_and_tiny_columns_use_absolute_eps():
    rng = np.random.default_rng(5)
    a = rng.standard_normal((7, 4))
    a[:, 1] = 0.0
    u, s, v, _ = oracle.solve(a, 1e-9, 30)
    zero = np.where(s == 0.0)[0]
    assert len(zero) == 1
    assert np.all(u[:, zero[0]] == 0.0)
    # sigma below the absolute epsilon: U column zeroed although sigma > 0.
    b = np.diag([1.0, 5e-10])
    u, s, _, _ = oracle.solve(np.vstack([b, np.zeros((1, 2))]), 1e-9, 30)
    assert s[1] == 5e-10 and np.all(u[:, 1] == 0.0)


def test_nan_column_quirk():
    # The reference's fmax(NaN, 0) = 0 turns a NaN column into sigma 0 and a
    # zero U column; NaN dot products fail the threshold test, so the other
    # columns are never rotated against it.
    a = np.eye(4, 3)
    a[0, 1] = np.nan
    u, s, v, _ = oracle.solve(a, 1e-9, 5)
    assert s[1] == 0.0 and np.all(u[:, 1] == 0.0)
    assert s[0] == 1.0 and s[2] == 1.0


def test_nan_epsilon_accepted_like_the_reference():
    # src/domain/jacobi_svd.cu:211 rejects `epsilon <= 0.0`, which is false
    # for NaN: the solve runs, every threshold comparison with NaN is false,
    # so nothing rotates (one sweep, V = I), sigma = the column norms and the
    # absolute-epsilon U rule (sigma > NaN is false) gives U = 0.
    rng = np.random.default_rng(12)
    a = rng.standard_normal((9, 5))
    assert oracle.validate(9, 5, 45, float("nan"), 8) == (0, None)
    u, s, v, sw = oracle.solve(a, float("nan"), 8)
    assert sw == 1
    np.testing.assert_array_equal(v, np.eye(5))
    np.testing.assert_allclose(s, np.linalg.norm(a, axis=0), rtol=1e-15)
    assert np.all(u == 0.0)
    # -0.0 and negative values stay rejected.
    assert oracle.validate(9, 5, 45, -0.0, 8)[0] == 1


def test_max_sweeps_stops_silently():
    rng = np.random.default_rng(11)
    a = rng.standard_normal((30, 30))
    _, _, _, sw = oracle.solve(a, 1e-15, 1)
    assert sw == 1


def test_round_parallel_is_bitwise_identical():
    rng = np.random.default_rng(21)
    a = rng.standard_normal((70, 50))
    r1 = oracle.solve(a, 1e-15, 30, threads=1)
    r8 = oracle.solve(a, 1e-15, 30, threads=8)
    for x, y in zip(r1[:3], r8[:3]):
        np.testing.assert_array_equal(x, y)
    assert r1[3] == r8[3]


def test_batch_matches_single():
    rng = np.random.default_rng(22)
    mats = [rng.standard_normal((int(m), int(n))) for n, m in [(3, 5), (8, 8), (12, 20), (1, 4)]]
    out = oracle.solve_batch(mats, 1e-12, 30, threads=4)
    for a, (u, s, v, sw) in zip(mats, out):
        u1, s1, v1, sw1 = oracle.solve(a, 1e-12, 30)
        np.testing.assert_array_equal(s, s1)
        np.testing.assert_array_equal(u, u1)
        np.testing.assert_array_equal(v, v1)
        assert sw == sw1


@pytest.mark.parametrize("rows,cols,size,eps,ms,msg", [
    (0, 3, 0, 1e-9, 8, "Input matrix shape must be non-zero."),
    (2, 3, 6, 1e-9, 8, "One-sided Jacobi SVD requires rows >= columns in this implementation."),
    (3, 2, 5, 1e-9, 8, "Input buffer size mismatch."),
    (3, 2, 6, 0.0, 8, "epsilon must be positive."),
    (3, 2, 6, 1e-9, 0, "max_sweeps must be positive."),
])
def test_validation_messages(rows, cols, size, eps, ms, msg):
    st, text = oracle.validate(rows, cols, size, eps, ms)
    assert st == 1 and text == msg
