"""GPU solves of matrices past the reference's 32-bit index range.

The reference indexes a matrix with 32-bit ints (src/domain/matrix_index.cuh:12-27),
so a matrix of 2^31 elements or more overflows there.  These tests run the
large-matrix path (column-major workspace, scaled transpose epilogue) on:
  * a (2^31 + 2) x 2 matrix, 2^32 + 4 elements (32 GiB): more rows than a
    32-bit int holds, so it takes the grid-barrier kernel, whose row
    indices are 64-bit.  It is built on the device with known singular
    values and vectors;
  * a (3 * 2^20 + 7) x 5 matrix on the wave kernels: 98305 row tiles of 32,
    more than the grid-y limit (65536) a two-dimensional transpose grid
    would need.
Expected values come from the construction (closed form) or from LAPACK.
Tolerances follow the summation length: a thread of the big solve adds 2^21
products per dot and a thread of the norm pass 2^23 squares, so results carry
up to ~1e-9 relative rounding; an index that wrapped at 2^31 or 2^32 would
give O(1) errors instead.  epsilon is raised above that rounding noise for
the same reason (the rotation test compares |pq| with eps * sqrt(pp * qq)).
The big solve needs about 110 GiB of device memory at its peak.
"""
import math
import sys

import numpy as np
import pytest

import jsvd

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _solve(a, m, n, eps=1e-15, sweeps=12):
    import torch

    plan = jsvd.BatchPlan([m], [n], epsilon=eps, max_sweeps=sweeps)
    assert plan.info["large_matrices"] == 1
    u = torch.empty_like(a)
    s = torch.empty(n, dtype=torch.float64, device=a.device)
    v = torch.empty(n * n, dtype=torch.float64, device=a.device)
    sw = torch.zeros(1, dtype=torch.int32, device=a.device)
    stop = torch.full((1,), -1, dtype=torch.int32, device=a.device)
    plan.execute(a, u, s, v, sw, stop=stop)
    torch.cuda.synchronize()
    plan.close()
    return u, s.cpu().numpy(), v.cpu().numpy().reshape(n, n), int(sw.item()), int(stop.item())


def test_two_column_matrix_with_more_rows_than_int32():
    import torch

    m, n = (1 << 31) + 2, 2
    assert m > 2**31 - 1 and m % 2 == 0
    _note(f"large index: {m} x {n} = {m * n} elements")
    s1, s2, th = 3.0, 0.25, 0.3
    c, sn = math.cos(th), math.sin(th)
    root = math.sqrt(m)
    # Orthonormal columns e = 1/sqrt(m) and f = (-1)^i / sqrt(m) (m even),
    # A = s1 e (c, sn) + s2 f (-sn, c): singular values s1, s2, right
    # singular vectors (c, sn) and (-sn, c).
    f = torch.ones(m, dtype=torch.float64, device="cuda:0")
    f[1::2] = -1.0
    a = torch.empty((m, n), dtype=torch.float64, device="cuda:0")
    a[:, 0] = f * (-s2 * sn / root) + s1 * c / root
    a[:, 1] = f * (s2 * c / root) + s1 * sn / root
    del f
    u, s, v, sweeps, stop = _solve(a.view(-1), m, n, eps=1e-9)
    _note(f"large index: sigma {s.tolist()} sweeps {sweeps} stop {stop}")
    assert stop == jsvd.STOP_CONVERGED and 1 <= sweeps <= 12
    order = np.argsort(-s)
    np.testing.assert_allclose(s[order], [s1, s2], rtol=1e-8)
    vs = v[:, order]
    for j, want in enumerate(([c, sn], [-sn, c])):
        sign = np.sign(vs[0, j] * want[0] + vs[1, j] * want[1])
        np.testing.assert_allclose(sign * vs[:, j], want, atol=1e-8)
    # U rows at both ends, the last one at element offset 2^32 + 2.
    # Column j of U is +-e or +-f; its sign is fixed by row 0.
    u2 = u.view(m, n)
    col_sign = None
    for row in (0, 1, m - 2, m - 1):
        got = u2[row].cpu().numpy()[order]
        fi = 1.0 if row % 2 == 0 else -1.0
        ref = np.array([1.0, fi]) / root
        if col_sign is None:
            col_sign = np.sign(got * ref)
        np.testing.assert_allclose(col_sign * got, ref, rtol=1e-7)
    # Column norms of U over all rows (device reduction).
    norms = torch.linalg.vector_norm(u2, dim=0).cpu().numpy()
    np.testing.assert_allclose(norms, [1.0, 1.0], rtol=1e-8)


def test_tall_matrix_beyond_65536_row_tiles():
    import torch

    m, n = 3 * (1 << 20) + 7, 5
    _note(f"large index: {m} x {n} tall")
    rng = np.random.default_rng(41)
    a_h = rng.standard_normal((m, n)) * np.array([1.0, 0.5, 1e-3, 2.0, 1e-6])
    a = torch.from_numpy(a_h.ravel()).to("cuda:0")
    u, s, v, sweeps, stop = _solve(a, m, n, eps=1e-12, sweeps=20)
    ref = np.linalg.svd(a_h, compute_uv=False)
    assert stop == jsvd.STOP_CONVERGED
    np.testing.assert_allclose(np.sort(s)[::-1], ref, rtol=1e-11, atol=1e-12 * ref[0])
    u_h = u.cpu().numpy().reshape(m, n)
    recon = (u_h * s) @ v.T
    assert np.linalg.norm(recon - a_h) <= 1e-12 * np.linalg.norm(a_h)
    np.testing.assert_allclose(u_h.T @ u_h, np.eye(n), atol=1e-11)
