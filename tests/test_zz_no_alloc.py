"""jsvd_solve_device_noalloc_f64 (the entry point that allocates nothing)
against jsvd_plan_create + jsvd_plan_execute_f64 of the same batch, through
the C-ABI, on two calls in a row (the second with a smaller batch), and
against the CPU oracle.

Both entry points pick the same shape class for every matrix, so the results
must agree bitwise; the oracle bar is the FP64 path's: sigma position by
position within 1e-12 * max(sigma), orthogonality and residual within
1e-12 * (n/64 + 1), sweeps within one.
"""
import sys

import numpy as np
import pytest

import jsvd
import oracle
from jsvd import gen, metrics

pytestmark = pytest.mark.gpu

EPS, MAX_SWEEPS = 1e-13, 30


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _batch(count, seed, n_hi):
    rng = np.random.default_rng(seed)
    ns = rng.integers(1, n_hi + 1, count)
    ms = np.array([rng.integers(n, min(2 * n, 160) + 1) for n in ns])
    mats = [rng.standard_normal((int(m), int(n))) for m, n in zip(ms, ns)]
    return mats, ms.astype(np.int64), ns.astype(np.int64)


def _device(mats):
    import torch

    flat, ms, ns = gen.pack(mats)
    _, _, _, (ae, ve, se) = jsvd.pack_offsets(ms, ns)
    dev = torch.device("cuda:0")
    a = torch.from_numpy(flat).to(dev)
    out = lambda k: torch.full((k,), np.nan, dtype=torch.float64, device=dev)
    return a, (out(ae), out(se), out(ve), torch.full((len(mats),), -1, dtype=torch.int32, device=dev))


def _unpack(mats, u, s, v, sw):
    ms = [x.shape[0] for x in mats]
    ns = [x.shape[1] for x in mats]
    a_off, v_off, s_off, _ = jsvd.pack_offsets(ms, ns)
    u, s, v, sw = u.cpu().numpy(), s.cpu().numpy(), v.cpu().numpy(), sw.cpu().numpy()
    return [(u[a_off[k]:a_off[k] + m * n].reshape(m, n), s[s_off[k]:s_off[k] + n], v[v_off[k]:v_off[k] + n * n]
             .reshape(n, n), int(sw[k])) for k, (m, n) in enumerate(zip(ms, ns))]


def _check_call(mats, ms, ns, tag):
    import torch

    a, (u0, s0, v0, w0) = _device(mats)
    plan = jsvd.BatchPlan(ms, ns, epsilon=EPS, max_sweeps=MAX_SWEEPS)
    plan.execute(a, u0, s0, v0, w0)
    _, (u1, s1, v1, w1) = _device(mats)
    jsvd.solve_device_noalloc(ms, ns, a, u1, s1, v1, w1, epsilon=EPS, max_sweeps=MAX_SWEEPS)
    torch.cuda.synchronize()
    for x, y in ((u0, u1), (s0, s1), (v0, v1), (w0, w1)):
        assert torch.equal(x, y), f"{tag}: no-alloc result differs from the plan's"
    got = _unpack(mats, u1, s1, v1, w1)
    ref = oracle.solve_batch(mats, EPS, MAX_SWEEPS)
    for k, (aa, g, r) in enumerate(zip(mats, got, ref)):
        n = aa.shape[1]
        tol = 1e-12 * (n / 64 + 1)
        assert np.max(np.abs(g[1] - r[1])) <= 1e-12 * max(r[1].max(), 1e-300), f"{tag} matrix {k} sigma"
        assert metrics.orthogonality(g[0], g[1] > EPS) <= tol, f"{tag} matrix {k} U"
        assert metrics.orthogonality(g[2]) <= tol, f"{tag} matrix {k} V"
        assert metrics.residual(aa, *g[:3]) <= tol, f"{tag} matrix {k} residual"
        assert abs(g[3] - r[3]) <= 1, f"{tag} matrix {k} sweeps {g[3]} vs {r[3]}"
    plan.close()


def test_noalloc_matches_plan_twice_second_smaller():
    # First call: 600 uneven matrices over every non-logging class (n up to
    # 128, m up to 160); second: 50 narrower ones, so the static work-list
    # still holds entries of the first call beyond the second's count.
    _note("no-alloc call 1: 600 matrices, n <= 128")
    mats, ms, ns = _batch(600, 31, 128)
    _check_call(mats, ms, ns, "call 1")
    _note("no-alloc call 2: 50 matrices, n <= 40")
    mats, ms, ns = _batch(50, 32, 40)
    _check_call(mats, ms, ns, "call 2")


def test_noalloc_rejects_what_it_cannot_take():
    import torch

    dev = torch.device("cuda:0")
    z = torch.zeros(1, dtype=torch.float64, device=dev)
    with pytest.raises(jsvd.JsvdInvalidArgument, match="at most 32768"):
        jsvd.solve_device_noalloc(np.full(32769, 1), np.full(32769, 1), torch.zeros(32769, dtype=torch.float64,
                                  device=dev), torch.zeros(32769, dtype=torch.float64, device=dev),
                                  torch.zeros(32769, dtype=torch.float64, device=dev),
                                  torch.zeros(32769, dtype=torch.float64, device=dev))
    big = torch.zeros(300 * 200, dtype=torch.float64, device=dev)
    with pytest.raises(jsvd.JsvdInvalidArgument, match="LDS-resident"):
        jsvd.solve_device_noalloc([300], [200], big, big.clone(), torch.zeros(200, dtype=torch.float64, device=dev),
                                  torch.zeros(200 * 200, dtype=torch.float64, device=dev))
    with pytest.raises(jsvd.JsvdInvalidArgument, match="FP64 only"):
        jsvd.solve_device_noalloc([4], [4], z.repeat(16), z.repeat(16), z.repeat(4), z.repeat(16),
                                  precision=jsvd.PRECISION_MIXED)
