"""Mixed-precision path (two FP64 QRs, FP32 sweeps with FP64 rotation
parameters and convergence test, the FP64 polar factor, B = X Q and two
first-order simultaneous Jacobi steps on MFMA, FP64 refinement sweeps, the
back-transform) against the FP64 CPU oracle, through the C-ABI.

Bar (the FP64 path's own bar against the oracle): normwise sigma error
max_k |s_k - r_k| / max_k r_k <= 1e-12 on sigma compared sorted (the
column order of a different pair sequence is not the reference's), and
||U^T U - I||, ||V^T V - I||, ||A - U S V^T|| / ||A|| <= 1e-12 (n/64 + 1).
The componentwise error is reported, not bounded: for kappa up to 1e8 two
valid FP64 solves already differ by ~kappa u on the smallest sigma.
"""
import sys

import numpy as np
import pytest

import jsvd
import oracle
from jsvd import gen, metrics

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _check(a, got, ref_sigma, eps, tag):
    u, s, v, sw = got
    n = a.shape[1]
    tol = 1e-12 * (n / 64 + 1)
    norm, comp = metrics.sigma_errors(s, ref_sigma)
    assert norm <= 1e-12, f"{tag}: normwise sigma error {norm:.3e}"
    assert metrics.orthogonality(u, s > eps) <= tol, f"{tag}: U orthogonality"
    assert metrics.orthogonality(v) <= tol, f"{tag}: V orthogonality"
    assert metrics.residual(a, u, s, v) <= tol, f"{tag}: residual"
    return norm, comp


# From 65 to 128 columns (at most 128 rows) the QR preconditioner and the
# back-transform run blocked on MFMA (16-column panels): 128 x 65 ends in a
# one-column panel, 120 x 120 and 96 x 80 pad rows and columns to the tiles.
# Up to 64 columns, and 200 x 50 / 256 x 64, take the one-column-per-lane-group
# QR and the reflector-by-reflector back-transform.
@pytest.mark.parametrize("shape", [(40, 30), (64, 64), (100, 60), (128, 128), (128, 100), (128, 65), (120, 120),
                                   (96, 80), (200, 50), (256, 64)])
def test_mixed_single_shapes(shape):
    _note(f"mixed shape {shape}")
    rng = np.random.default_rng(3 + shape[0] + shape[1])
    sig = 10.0 ** (-5 * np.arange(shape[1]) / (shape[1] - 1))
    a = gen.constructed(rng, shape[0], shape[1], sig)
    ref = oracle.solve(a, 1e-15, 30)
    got = jsvd.svd(a, epsilon=1e-15, max_sweeps=30, precision=jsvd.PRECISION_MIXED)
    norm, comp = _check(a, got, ref[1], 1e-15, str(shape))
    _note(f"  normwise {norm:.2e} componentwise {comp:.2e} sweeps {got[3]} (fp64 oracle {ref[3]})")


def test_mixed_conditioned_batch_against_fp64_oracle():
    # 128 x 128 matrices with kappa = 10^U(3, 8), the mixed-precision workload
    # shape, on a device plan; every matrix against the FP64 oracle.
    import torch

    _note("mixed batch of 64 conditioned 128x128")
    mats, sig = gen.conditioned_batch(64)
    flat, ms, ns = gen.pack(mats)
    plan = jsvd.BatchPlan(ms, ns, epsilon=1e-15, max_sweeps=30, precision=jsvd.PRECISION_MIXED)
    dev = torch.device("cuda:0")
    a = torch.from_numpy(flat).to(dev)
    u = torch.empty(plan.a_elems, dtype=torch.float64, device=dev)
    s = torch.empty(plan.s_elems, dtype=torch.float64, device=dev)
    v = torch.empty(plan.v_elems, dtype=torch.float64, device=dev)
    sw = torch.zeros(len(ms), dtype=torch.int32, device=dev)
    plan.execute(a, u, s, v, sw)
    torch.cuda.synchronize()
    u, s, v, sw = (t.cpu().numpy() for t in (u, s, v, sw))
    refs = oracle.solve_batch(mats, 1e-15, 30)
    worst = 0.0
    for k in range(len(mats)):
        m, n = int(ms[k]), int(ns[k])
        g = (u[plan.a_off[k]:plan.a_off[k] + m * n].reshape(m, n), s[plan.s_off[k]:plan.s_off[k] + n],
             v[plan.v_off[k]:plan.v_off[k] + n * n].reshape(n, n), int(sw[k]))
        norm, _ = _check(mats[k], g, refs[k][1], 1e-15, f"mixed{k}")
        worst = max(worst, norm)
    _note(f"worst normwise sigma error vs FP64 oracle: {worst:.3e}")
    plan.close()


def test_mixed_uneven_batch():
    _note("mixed uneven batch")
    mats, sig = gen.uneven_batch(200, seed=99)
    got = jsvd.svd_batched(mats, epsilon=1e-14, max_sweeps=30, precision=jsvd.PRECISION_MIXED)
    refs = oracle.solve_batch(mats, 1e-14, 30)
    for k, (a, g, r) in enumerate(zip(mats, got, refs)):
        _check(a, g, r[1], 1e-14, f"uneven{k}")


def test_mixed_rank_deficient_keeps_v_orthogonal():
    # Exactly zero columns: the FP64 oracle gives exactly zero singular
    # values (zero columns never rotate); the QR-preconditioned mixed path
    # gives rounding-level ones (the zero columns of A become zero rows of R).
    # Either way U's columns there are zero (U rule) and V is orthogonal; an
    # all-zero matrix gives exact zeros and a completed orthogonal V.
    _note("mixed rank-deficient")
    rng = np.random.default_rng(21)
    mats = []
    for shape, zero in [((40, 30), [3, 17]), ((128, 128), [0, 64, 127]), ((90, 45), [44])]:
        a = rng.standard_normal(shape)
        a[:, zero] = 0.0
        mats.append(a)
    mats.append(np.zeros((50, 20)))
    got = jsvd.svd_batched(mats, epsilon=1e-14, max_sweeps=40, precision=jsvd.PRECISION_MIXED)
    refs = oracle.solve_batch(mats, 1e-14, 40)
    for a, g, r in zip(mats, got, refs):
        u, s, v, _ = g
        tiny = s <= 1e-14 * max(s.max(), 1e-300)
        assert np.sum(tiny) == np.sum(r[1] == 0.0)
        assert np.all(u[:, tiny] == 0.0)
        assert metrics.orthogonality(v) <= 1e-12 * (a.shape[1] / 64 + 1)
        if np.linalg.norm(a) > 0:
            _check(a, g, r[1], 1e-14, f"rank-deficient {a.shape}")
        else:
            assert np.all(s == 0.0) and np.all(u == 0.0)


def test_mixed_stress_kinds_and_clustered_spectra():
    # The stress mix's kinds, and spectra of repeated and of clustered
    # singular values (the first-order steps before the FP64 refinement leave
    # pairs of close singular values to the sweeps), through the blocked QR
    # (128 x 128, 100 x 70) and the one-column-per-lane-group QR (60 x 40),
    # against the oracle with the stress suite's bar.
    mats = []
    for k, (m, n) in enumerate([(128, 128), (100, 70), (60, 40)]):
        rng = np.random.default_rng(900 + k)
        for kind in gen.STRESS_KINDS:
            mats.append(np.ascontiguousarray(gen.stress_matrix(kind, m, n, rng)))
        mats.append(gen.constructed(rng, m, n, np.repeat([1.0, 0.5, 1e-3, 1e-6], (n + 3) // 4)[:n]))
        groups = 10.0 ** -np.arange((n + 7) // 8, dtype=np.float64)
        mats.append(gen.constructed(rng, m, n, np.repeat(groups, 8)[:n] * (1.0 + 1e-10 * rng.standard_normal(n))))
    _note(f"mixed stress: {len(mats)} matrices")
    got = jsvd.svd_batched(mats, epsilon=1e-15, max_sweeps=30, precision=jsvd.PRECISION_MIXED)
    refs = oracle.solve_batch(mats, 1e-15, 30)
    for k, (a, g, r) in enumerate(zip(mats, got, refs)):
        u, s, v, sw = g
        top = max(r[1].max(), 1e-300)
        tag = f"mixed stress #{k} {a.shape}"
        assert np.max(np.abs(np.sort(s) - np.sort(r[1]))) <= 1e-12 * top, tag
        assert metrics.residual(a, u, s, v) <= 1e-12, tag
        assert metrics.orthogonality(v) <= 1e-11, tag
        assert metrics.orthogonality(u, s > 1e-6 * top) <= 1e-9, tag
