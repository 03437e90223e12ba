"""Failure path of the software barriers of the large-matrix kernels.

The grouped wave kernel (flag barriers among a group's workgroups), its ring
form (mailbox waits between neighbouring workgroups, a flag barrier per
sweep) and the grid-barrier kernel assume that every workgroup of their grid
is resident.
Each barrier wait is bounded: past the bound the launch is marked faulted,
every workgroup leaves its loops, the matrices get stop reason
STOP_BARRIER_TIMEOUT and the synchronous entry points return JSVD_EDEVICE.

The condition is forced from the host: JSVD_BARRIER_TIMEOUT_MS = 0 makes the
first wait that is not already satisfied give up at once, so every kernel
still runs to its end.  After the variable is removed, the same plan must
solve correctly again (the barrier and fault words are reset per launch).
"""
import sys

import numpy as np
import pytest

import jsvd
from jsvd import _lib

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run_plan(a_h, eps=1e-13, sweeps=30):
    import torch

    m, n = a_h.shape
    plan = jsvd.BatchPlan([m], [n], epsilon=eps, max_sweeps=sweeps)
    a = torch.from_numpy(a_h.ravel()).to("cuda:0")
    u = torch.empty_like(a)
    s = torch.empty(n, dtype=torch.float64, device="cuda:0")
    v = torch.empty(n * n, dtype=torch.float64, device="cuda:0")
    sw = torch.zeros(1, dtype=torch.int32, device="cuda:0")
    stop = torch.full((1,), -1, dtype=torch.int32, device="cuda:0")
    return plan, (a, u, s, v, sw, stop)


# 256 x 200 alone runs the ring form with two LDS slot sets (a failed wait is
# agreed on one round later), 400 x 300 with one.
@pytest.mark.parametrize("shape", [(400, 300), (256, 200), (1100, 520), (2100, 2050)],
                         ids=["ring", "ring-two-sets", "grouped-wave", "grid-barrier"])
def test_barrier_timeout_is_reported_and_the_plan_recovers(shape, monkeypatch):
    import torch

    m, n = shape
    rng = np.random.default_rng(m + n)
    a_h = rng.standard_normal((m, n))
    plan, (a, u, s, v, sw, stop) = _run_plan(a_h, sweeps=30 if n < 1000 else 2)
    assert plan.info["large_matrices"] == 1
    _note(f"barrier timeout: {m} x {n}, forced")
    monkeypatch.setenv("JSVD_BARRIER_TIMEOUT_MS", "0")
    plan.execute(a, u, s, v, sw, stop=stop)
    torch.cuda.synchronize()
    assert int(stop.item()) == jsvd.STOP_BARRIER_TIMEOUT
    # The synchronous entry point turns it into JSVD_EDEVICE.
    with pytest.raises(_lib.JsvdError, match="barrier timed out") as err:
        jsvd.svd_batched([a_h], epsilon=1e-13, max_sweeps=2)
    assert err.value.status == _lib.JSVD_EDEVICE
    monkeypatch.delenv("JSVD_BARRIER_TIMEOUT_MS")
    _note(f"barrier timeout: {m} x {n}, normal run after the fault")
    plan.execute(a, u, s, v, sw, stop=stop)
    torch.cuda.synchronize()
    if n < 1000:
        assert int(stop.item()) == jsvd.STOP_CONVERGED
        ref = np.linalg.svd(a_h, compute_uv=False)
        got = np.sort(s.cpu().numpy())[::-1]
        assert np.max(np.abs(got - ref)) <= 1e-12 * ref[0]
    else:
        # Two sweeps of the grid-barrier kernel: both run, no fault.
        assert int(stop.item()) == jsvd.STOP_MAX_SWEEPS and int(sw.item()) == 2
        # Orthogonal rotations keep the Frobenius norm.
        ss = s.cpu().numpy()
        assert abs(np.sqrt(np.sum(ss * ss)) - np.linalg.norm(a_h)) <= 1e-12 * np.linalg.norm(a_h)
    plan.close()
