"""GPU tests of the per-matrix stop reason and the device error metrics
(no reference counterpart: the reference stops silently at max_sweeps,
src/domain/jacobi_svd.cu:156-185, and reports no error norm).

Stop reason: a solve is JSVD_STOP_CONVERGED when its last sweep rotated no
pair and JSVD_STOP_MAX_SWEEPS when max_sweeps sweeps ran and the last one
still rotated -- checked on every solver path against the oracle, whose
sweep loop has the same semantics (converged <=> sweeps < max_sweeps, or
sweeps == max_sweeps with the last sweep clean).

Metrics: ||U^T U - D||_F (D = diag(sigma > eps)), ||V^T V - I||_F and
||A - U S V^T||_F / ||A||_F from the device kernel against numpy's
(jsvd.metrics) on the same U, S, V; they are computed in different orders,
so they agree to the rounding of the norms themselves: |gpu - numpy| <=
1e-13 + 1e-6 * numpy (the norms are ~1e-14..1e-12; a wrong index or term
moves them by orders of magnitude).
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


# One shape per solver path: LDS class kernels (0/1, 4, 5 and the
# rotation-log class 6), the generic LDS kernel (700 x 24: no class takes
# more than 256 rows), the grouped wave kernel (rows in registers: 300 x 150;
# streamed: 1200 x 64).  The one-workgroup-per-matrix forms get their own
# batch below.
PATH_SHAPES = [(12, 8), (100, 60), (200, 100), (128, 128), (700, 24), (300, 150), (1200, 64)]


@pytest.mark.parametrize("max_sweeps", [1, 2, 30])
def test_stop_reason_matches_oracle_on_every_path(max_sweeps):
    _note(f"stop reason, max_sweeps={max_sweeps}")
    rng = np.random.default_rng(40 + max_sweeps)
    mats = [rng.standard_normal(sh) for sh in PATH_SHAPES]
    mats += [np.diag([3.0, 1.0, 2.0]), rng.standard_normal((40, 1))]  # converge in one sweep
    got = jsvd.svd_batched(mats, epsilon=1e-13, max_sweeps=max_sweeps, with_stop=True)
    ref = oracle.solve_batch(mats, 1e-13, max_sweeps)
    for k, (a, g, r) in enumerate(zip(mats, got, ref)):
        sweeps, stop = g[3], g[4]
        random = k < len(PATH_SHAPES)
        assert stop in (jsvd.STOP_CONVERGED, jsvd.STOP_MAX_SWEEPS)
        if sweeps < max_sweeps:
            assert stop == jsvd.STOP_CONVERGED, a.shape
        if max_sweeps <= 2 and random:
            # A random matrix is nowhere near orthogonal after one or two sweeps.
            assert stop == jsvd.STOP_MAX_SWEEPS and sweeps == max_sweeps, a.shape
            assert r[3] == max_sweeps
        if max_sweeps == 30:
            assert stop == jsvd.STOP_CONVERGED and abs(sweeps - r[3]) <= 1, a.shape
    # Diagonal input and a single column: one clean sweep whatever the cap.
    assert got[-2][3:] == (1, jsvd.STOP_CONVERGED) and got[-1][3:] == (1, jsvd.STOP_CONVERGED)
    # One workgroup per matrix, rows in registers (100 x 160 x 150) and
    # streamed (100 x 1500 x 24).
    for shape in [(160, 150), (1500, 24)]:
        batch = [rng.standard_normal(shape) for _ in range(100)]
        plan = jsvd.BatchPlan([shape[0]] * 100, [shape[1]] * 100)
        assert plan.info["large_matrices"] == 100
        plan.close()
        out = jsvd.svd_batched(batch, epsilon=1e-13, max_sweeps=max_sweeps, with_stop=True)
        want = jsvd.STOP_CONVERGED if max_sweeps == 30 else jsvd.STOP_MAX_SWEEPS
        assert all(o[4] == want for o in out), shape


@pytest.mark.parametrize("bw", [16, 32])
def test_stop_reason_blocked_and_mixed(bw):
    _note(f"stop reason, blocked width {bw} and mixed")
    rng = np.random.default_rng(77)
    big = rng.standard_normal((400, 200))
    for cap, want in ((1, jsvd.STOP_MAX_SWEEPS), (40, jsvd.STOP_CONVERGED)):
        (_, _, _, sw, st), = jsvd.svd_batched([big], epsilon=1e-13, max_sweeps=cap, block_width=bw, with_stop=True)
        assert st == want and (sw == cap if cap == 1 else sw < cap), (cap, sw, st)
    small = [rng.standard_normal((64, 48)) for _ in range(3)]
    out = jsvd.svd_batched(small, epsilon=1e-13, max_sweeps=40, precision=jsvd.PRECISION_MIXED, with_stop=True)
    assert all(o[4] == jsvd.STOP_CONVERGED for o in out)
    # Pairs of singular values 1e-9 apart: the FP32 sweeps leave such a pair
    # at an FP32 cosine and the first-order steps leave it to the FP64
    # refinement, whose first sweep then rotates; capped at one refinement
    # sweep, the solve stops at the cap.  (On well-separated spectra the
    # first-order steps leave the refinement a first sweep without rotations,
    # so one sweep already converges.)
    pairs = np.repeat(10.0 ** (-np.arange(24, dtype=np.float64) / 4), 2) * (1.0 + 1e-9 * np.tile([0.0, 1.0], 24))
    close = [gen.constructed(rng, 64, 48, pairs) for _ in range(3)]
    out = jsvd.svd_batched(close, epsilon=1e-13, max_sweeps=40, precision=jsvd.PRECISION_MIXED, with_stop=True)
    assert all(o[4] == jsvd.STOP_CONVERGED for o in out)
    out = jsvd.svd_batched(close, epsilon=1e-13, max_sweeps=40, refine_sweeps=1, precision=jsvd.PRECISION_MIXED,
                           with_stop=True)
    assert all(o[4] == jsvd.STOP_MAX_SWEEPS for o in out)


def test_plan_stop_tensor_and_metrics_against_numpy():
    import torch

    _note("plan stop tensor and device metrics")
    rng = np.random.default_rng(5)
    shapes = [(12, 8), (100, 60), (200, 100), (128, 128), (300, 150), (70, 33), (1, 1), (5, 1)]
    mats = [rng.standard_normal(sh) for sh in shapes]
    mats[5][:, 4] = 0.0  # a zero column: sigma 0, zero U column (masked out of orth_u)
    mats.append(np.zeros((6, 4)))  # A = 0: residual reported as 0
    flat, ms, ns = gen.pack(mats)
    dev = torch.device("cuda:0")
    plan = jsvd.BatchPlan(ms, ns, epsilon=1e-13, max_sweeps=30)
    A = torch.from_numpy(flat).to(dev)
    U = torch.empty(plan.a_elems, dtype=torch.float64, device=dev)
    S = torch.empty(plan.s_elems, dtype=torch.float64, device=dev)
    V = torch.empty(plan.v_elems, dtype=torch.float64, device=dev)
    sw = torch.zeros(len(mats), dtype=torch.int32, device=dev)
    stop = torch.full((len(mats),), -1, dtype=torch.int32, device=dev)
    plan.execute(A, U, S, V, sw, stop=stop)
    met = plan.metrics(A, U, S, V).cpu().numpy()
    met2 = plan.metrics(A, U, S, V).cpu().numpy()  # second call reuses the tile list
    np.testing.assert_array_equal(met, met2)
    assert np.all(stop.cpu().numpy() == jsvd.STOP_CONVERGED)
    u, s, v = U.cpu().numpy(), S.cpu().numpy(), V.cpu().numpy()
    for k, a in enumerate(mats):
        m, n = a.shape
        uk = u[plan.a_off[k]:plan.a_off[k] + m * n].reshape(m, n)
        sk = s[plan.s_off[k]:plan.s_off[k] + n]
        vk = v[plan.v_off[k]:plan.v_off[k] + n * n].reshape(n, n)
        want = (metrics.orthogonality(uk, sk > 1e-13), metrics.orthogonality(vk),
                metrics.residual(a, uk, sk, vk) if np.linalg.norm(a) > 0 else 0.0)
        for g, w, name in zip(met[k], want, ("orth_u", "orth_v", "residual")):
            assert abs(g - w) <= 1e-13 + 1e-6 * w, (k, a.shape, name, g, w)
        assert met[k].max() < 1e-11, (k, met[k])
    plan.close()


def test_metrics_see_a_planted_error():
    # The metrics must move when U, S or V is wrong: scale one U column, one
    # sigma and one V column, and each norm jumps far above the rounding
    # level in the way numpy's does.
    import torch

    _note("device metrics on corrupted factors")
    rng = np.random.default_rng(8)
    a = rng.standard_normal((90, 40))
    dev = torch.device("cuda:0")
    plan = jsvd.BatchPlan([90], [40], epsilon=1e-13, max_sweeps=30)
    A = torch.from_numpy(a.ravel()).to(dev)
    U = torch.empty(90 * 40, dtype=torch.float64, device=dev)
    S = torch.empty(40, dtype=torch.float64, device=dev)
    V = torch.empty(40 * 40, dtype=torch.float64, device=dev)
    plan.execute(A, U, S, V)
    U.view(90, 40)[:, 3] *= 1.001
    S[7] *= 1.01
    V.view(40, 40)[:, 5] *= 1.001
    met = plan.metrics(A, U, S, V).cpu().numpy()[0]
    uk, sk, vk = U.view(90, 40).cpu().numpy(), S.cpu().numpy(), V.view(40, 40).cpu().numpy()
    want = (metrics.orthogonality(uk, sk > 1e-13), metrics.orthogonality(vk), metrics.residual(a, uk, sk, vk))
    for g, w in zip(met, want):
        assert w > 1e-4 and abs(g - w) <= 1e-9 * w, (g, w)
    plan.close()
