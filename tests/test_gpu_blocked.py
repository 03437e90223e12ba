"""GPU tests of the blocked Jacobi path (block_width 16 or 32): column blocks,
Gram blocks [A_I A_J]^T [A_I A_J] and block updates on FP64 MFMA.

The reference has no blocked variant (it rotates single column pairs,
src/infrastructure/jacobi_rotation_kernels.cuh:241-359), so results are
checked against the oracle's column-pair solve where that converges within
seconds, against LAPACK's singular values (numpy.linalg.svd), and at the
8192 x 2048 workload against the oracle's own solve of that matrix
(tests/golden/tall_8192x2048.npz: sorted sigma and the sweep count, written
by tests/golden/make_tall_fixture.py in about four minutes of host time) and
through the invariants that hold at any size: reconstruction
||A - U S V^T|| / ||A||, ||V^T V - I||, sigma against the constructed
spectrum, and the sweep semantics (a sweep that applies no rotation ends the
solve).
"""
import os
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


def _matrix(m, n, kappa, seed):
    rng = np.random.default_rng(seed)
    sig = kappa ** (-np.arange(n) / (n - 1))
    return gen.constructed(rng, m, n, sig), sig


# n = 150 and 70 are not multiples of 2 * block: zero columns pad them.
SMALL = [(300, 150, 1e6), (256, 96, 1e4), (130, 70, 1e2)]


@pytest.mark.parametrize("bw", [16, 32])
@pytest.mark.parametrize("m,n,kappa", SMALL)
def test_blocked_against_oracle_and_lapack(m, n, kappa, bw):
    _note(f"blocked {m}x{n} kappa={kappa:g} block_width={bw}")
    a, _ = _matrix(m, n, kappa, 1000 + m + n)
    eps = 1e-13
    ref = oracle.solve(a, eps, 30, threads=8)
    u, s, v, sw = jsvd.svd(a, block_width=bw, epsilon=eps, max_sweeps=30)
    lap = np.linalg.svd(a, compute_uv=False)
    assert metrics.sigma_errors(s, ref[1])[0] <= 1e-12
    assert metrics.sigma_errors(s, lap)[0] <= 1e-12
    assert metrics.residual(a, u, s, v) <= 1e-12
    assert metrics.orthogonality(v) <= 1e-11
    assert metrics.orthogonality(u, s > 1e-9 * s.max()) <= 1e-9
    # Converged within the cap, like the oracle.
    assert ref[3] < 30 and 1 < sw < 30


def test_blocked_plan_info_and_routing():
    a, _ = _matrix(300, 150, 1e3, 7)
    plan = jsvd.BatchPlan([300, 40], [150, 20], block_width=16)
    info = plan.info
    assert info["large_matrices"] == 1 and info["blocked_matrices"] == 1 and info["lds_matrices"] == 1
    plan.close()
    plan = jsvd.BatchPlan([300], [150])
    assert plan.info["blocked_matrices"] == 0
    plan.close()
    # n < 2 * block_width: the grid-barrier kernel keeps the matrix.
    plan = jsvd.BatchPlan([4000], [40], block_width=32)
    assert plan.info["large_matrices"] == 1 and plan.info["blocked_matrices"] == 0
    plan.close()


def test_blocked_orthogonal_columns_take_one_sweep():
    # Mutually orthogonal columns: no pair fails the test, so the first
    # (verifying) sweep ends the solve and sigma are the column norms.
    rng = np.random.default_rng(11)
    q = gen.haar(rng, 400, 128)
    d = np.linspace(3.0, 1.0, 128)
    a = q * d
    for bw in (16, 32):
        u, s, v, sw = jsvd.svd(a, block_width=bw, epsilon=1e-10, max_sweeps=30)
        assert sw == 1
        np.testing.assert_allclose(s, np.linalg.norm(a, axis=0), rtol=1e-14, atol=0)
        np.testing.assert_array_equal(v, np.eye(128))


def test_blocked_repeat_runs_are_bitwise_identical():
    a, _ = _matrix(520, 200, 1e5, 12)
    r1 = jsvd.svd(a, block_width=32, epsilon=1e-14, max_sweeps=30)
    r2 = jsvd.svd(a, block_width=32, epsilon=1e-14, max_sweeps=30)
    for x, y in zip(r1[:3], r2[:3]):
        np.testing.assert_array_equal(x, y)
    assert r1[3] == r2[3]


def test_blocked_batch_with_lds_matrices():
    # A blocked matrix next to LDS-resident ones in one plan.
    rng = np.random.default_rng(13)
    mats = [rng.standard_normal((40, 24)), _matrix(600, 160, 1e4, 14)[0], rng.standard_normal((9, 9))]
    out = jsvd.svd_batched(mats, block_width=16, epsilon=1e-13, max_sweeps=30)
    refs = oracle.solve_batch(mats, 1e-13, 30)
    for a, (u, s, v, sw), ref in zip(mats, out, refs):
        assert metrics.sigma_errors(s, ref[1])[0] <= 1e-12
        assert metrics.residual(a, u, s, v) <= 1e-12


TALL_FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "tall_8192x2048.npz")


def _tall_input():
    """The 8192 x 2048 geometric-spectrum matrix and the oracle's fixture for
    it (tests/golden/make_tall_fixture.py: sorted sigma and the sweep count of
    the CPU oracle at eps = 1e-15).  The matrix is regenerated from its seed;
    its checksum must match the one the fixture was written from."""
    a, sig = gen.tall_geometric()
    fx = np.load(TALL_FIXTURE)
    flat = a.ravel()
    chk = np.concatenate([[flat.sum(), np.dot(flat, flat)], flat[:8]])
    assert np.allclose(chk, fx["input_checksum"], rtol=1e-12, atol=1e-15), "input differs from the fixture's"
    return a, sig, fx


def _solve_tall(a, **cfg):
    import torch

    m, n = a.shape
    dev = torch.device("cuda:0")
    plan = jsvd.BatchPlan([m], [n], epsilon=1e-15, max_sweeps=80, **cfg)
    A = torch.from_numpy(a.ravel()).to(dev)
    U = torch.empty(m * n, dtype=torch.float64, device=dev)
    S = torch.empty(n, dtype=torch.float64, device=dev)
    V = torch.empty(n * n, dtype=torch.float64, device=dev)
    sw = torch.zeros(1, dtype=torch.int32, device=dev)
    plan.execute(A, U, S, V, sw)
    torch.cuda.synchronize()
    plan.close()
    Am, Um, Vm = A.view(m, n), U.view(m, n), V.view(n, n)
    res = (torch.linalg.norm(Am - (Um * S) @ Vm.T) / torch.linalg.norm(Am)).item()
    orth_v = torch.linalg.norm(Vm.T @ Vm - torch.eye(n, dtype=torch.float64, device=dev)).item()
    return S.cpu().numpy(), int(sw.item()), res, orth_v


def test_plain_tall_geometric_8192x2048_against_oracle_fixture():
    # The plain FP64 path (column pairs in the reference's order) on the
    # 8192 x 2048 matrix, run to convergence and compared with the oracle's
    # own solve of the same input: sigma sorted (the column order of a wide
    # matrix is chaotic under rounding) within 1e-12 of sigma_max, sweeps
    # within three of the oracle's: at eps = 1e-15 the cosine test of
    # 8192-row columns is decided below the rounding of their dot products
    # (about sqrt(m) u = 2e-14 relative), so the last sweeps' decisions are
    # rounding noise (measured: 51 sweeps against the oracle's 53, sigma
    # within 6.9e-14).
    a, sig, fx = _tall_input()
    _note("plain 8192x2048 geometric kappa=1e10 vs oracle fixture")
    s, sweeps, res, orth_v = _solve_tall(a)
    vs_oracle = metrics.sigma_errors(s, fx["sigma_sorted"])[0]
    _note(f"  sweeps {sweeps} (oracle {int(fx['sweeps'])}) residual {res:.2e} ||V^T V - I||_F {orth_v:.2e} "
          f"sigma vs oracle {vs_oracle:.2e}")
    assert abs(sweeps - int(fx["sweeps"])) <= 3
    assert vs_oracle <= 1e-12
    assert res <= 1e-11 and orth_v <= 1e-9


@pytest.mark.parametrize("bw", [32, 16])
def test_blocked_tall_geometric_8192x2048(bw):
    _note(f"blocked 8192x2048 geometric kappa=1e10 block_width={bw}")
    a, sig, fx = _tall_input()
    s, sweeps, res, orth_v = _solve_tall(a, block_width=bw)
    norm_err, comp_err = metrics.sigma_errors(s, sig)
    vs_oracle = metrics.sigma_errors(s, fx["sigma_sorted"])[0]
    _note(f"  sweeps {sweeps} (oracle, plain order: {int(fx['sweeps'])}) residual {res:.2e} "
          f"||V^T V - I||_F {orth_v:.2e} sigma normwise {norm_err:.2e} componentwise {comp_err:.2e} "
          f"vs oracle {vs_oracle:.2e}")
    assert sweeps < 80  # converged: the last sweep applied no rotation
    # The block-cyclic order visits the same pairs per sweep as the plain
    # order, so the sweep counts are alike (not equal: another order).
    assert abs(sweeps - int(fx["sweeps"])) <= 3
    assert vs_oracle <= 1e-12
    # Tolerances: about 3.5e3 block updates touch every column (54 sweeps x
    # 63 rounds at block width 32); rounding accumulates to ~1e-12 relative.
    assert res <= 1e-11
    assert norm_err <= 1e-11
    assert orth_v <= 1e-9
