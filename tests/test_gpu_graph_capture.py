"""Plan executions captured in a HIP graph (torch.cuda.graph) and replayed.

A plan's execution enqueues only kernels and device-to-device work on the
caller's stream (and on the device's barrier stream, joined by events), and
the blocked path keeps its convergence decision on the device, so a whole
solve can be captured once and replayed.  Each replay is compared bitwise
with an eager execution of the same plan on the same input, with eager work
and synchronisations between the replays.  Shapes: an uneven batch on the
LDS work-list kernel plus a 700 x 300 matrix on the grouped wave kernel, and
a 1536 x 768 matrix on the blocked path (block width 32).
"""
import sys

import numpy as np
import pytest

import jsvd
from jsvd import gen

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _buffers(plan, flat, torch):
    dev = torch.device("cuda:0")
    a = torch.from_numpy(flat).to(dev)
    u = torch.zeros(plan.a_elems, dtype=torch.float64, device=dev)
    s = torch.zeros(plan.s_elems, dtype=torch.float64, device=dev)
    v = torch.zeros(plan.v_elems, dtype=torch.float64, device=dev)
    sw = torch.zeros(len(plan.ms), dtype=torch.int32, device=dev)
    stop = torch.full((len(plan.ms),), -1, dtype=torch.int32, device=dev)
    return a, u, s, v, sw, stop


def _snapshot(bufs):
    return [t.detach().cpu().numpy().copy() for t in bufs[1:]]


@pytest.mark.parametrize("case", ["worklist+wave", "blocked"])
def test_captured_execution_replays_bitwise(case):
    import torch

    rng = np.random.default_rng(12)
    if case == "blocked":
        mats = [rng.standard_normal((1536, 768))]
        kw = dict(epsilon=1e-13, max_sweeps=40, block_width=32)
    else:
        mats, _ = gen.uneven_batch(300, seed=3)
        mats = list(mats) + [rng.standard_normal((700, 300))]
        kw = dict(epsilon=1e-14, max_sweeps=30)
    flat, ms, ns = gen.pack(mats)
    plan = jsvd.BatchPlan(ms, ns, **kw)
    if case == "blocked":
        assert plan.info["blocked_matrices"] == 1
    else:
        assert plan.info["large_matrices"] == 1 and plan.info["lds_matrices"] == 300
    bufs = _buffers(plan, flat, torch)
    _note(f"graph capture {case}: eager")
    plan.execute(*bufs[:5], stop=bufs[5])
    torch.cuda.synchronize()
    eager = _snapshot(bufs)
    assert np.all(eager[-1] == jsvd.STOP_CONVERGED)
    for t in bufs[1:]:
        t.fill_(-7)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
            plan.execute(*bufs[:5], stop=bufs[5])
    torch.cuda.synchronize()
    for rep in range(3):
        _note(f"graph capture {case}: replay {rep}")
        for t in bufs[1:]:
            t.fill_(-3 - rep)          # eager work between the replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        got = _snapshot(bufs)
        for x, y in zip(got, eager):
            assert np.array_equal(x, y)
    plan.close()
