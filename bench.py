#!/usr/bin/env python3
"""Benchmark of the batched FP64 one-sided Jacobi SVD on one MI355X.

Workload (one step = one pass of the hot path over one batch): 4096
matrices, n ~ U{8..64}, m ~ U{n..2n}, A = Q1 diag(sigma) Q2^T with Haar
factors and sigma log-spaced 1 .. 1e-6, eps = 1e-15, at most 30 sweeps,
solved in one persistent work-list launch.  Inputs are resident in HBM
before the timed region; outputs stay on the device.

Prints one JSON line: matrices/s, the work-list kernel's FP64 throughput
against the MI355X FP64 peak (HIP events on the launch stream), and the
CPU oracle timed on a bounded sample of the same batch.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "jacobi-svd-cuda-opt_amd"))

FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix peaks are equal)
HBM_PEAK_GBS = 8000.0


def sweep_flops(m, n):
    """Algorithmic FP64 flops of one plain sweep: n(n-1)/2 pairs x (12m + 6n)."""
    m = np.asarray(m, dtype=np.float64)
    n = np.asarray(n, dtype=np.float64)
    return n * (n - 1) / 2 * (12 * m + 6 * n)


def compulsory_bytes(m, n):
    """Read A, write U, sigma, V once: 8 (2mn + n^2 + n) bytes per matrix."""
    m = np.asarray(m, dtype=np.float64)
    n = np.asarray(n, dtype=np.float64)
    return 8 * (2 * m * n + n * n + n)


def load_traffic(name):
    path = os.path.join(REPO, "profiles", name)
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f)
    except Exception:
        return None


def cpu_baseline(mats, eps, max_sweeps, budget_s=15.0):
    """CPU oracle (one matrix per OpenMP thread) on a bounded prefix sample."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle

    cores = os.cpu_count() or 1
    cores = min(cores, 64)
    sample = mats[:64]
    t0 = time.perf_counter()
    oracle.solve_batch(sample, eps, max_sweeps, threads=cores)
    dt = time.perf_counter() - t0
    per = dt / len(sample)
    count = int(min(len(mats), max(64, budget_s / max(per, 1e-9))))
    sample = mats[:count]
    t0 = time.perf_counter()
    oracle.solve_batch(sample, eps, max_sweeps, threads=cores)
    dt = time.perf_counter() - t0
    return {"value": len(sample) / dt, "unit": "matrices/s", "cores": cores, "kind": "port",
            "sample": f"first {len(sample)} matrices of the same 4096-matrix batch, {dt:.1f} s, "
                      f"one matrix per OpenMP thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "batches would shard by matrix across GPUs with no exchange, which is not built here.",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import jsvd
    from jsvd import gen, metrics

    eps, max_sweeps = 1e-15, 30
    mats, sig = gen.uneven_batch(args.count)
    flat, ms, ns = gen.pack(mats)
    plan = jsvd.BatchPlan(ms, ns, epsilon=eps, max_sweeps=max_sweeps)
    dev = torch.device("cuda:0")
    a = torch.from_numpy(flat).to(dev)
    u = torch.empty(plan.a_elems, dtype=torch.float64, device=dev)
    s = torch.empty(plan.s_elems, dtype=torch.float64, device=dev)
    v = torch.empty(plan.v_elems, dtype=torch.float64, device=dev)
    sw = torch.zeros(len(ms), dtype=torch.int32, device=dev)

    for _ in range(args.warmup):
        plan.execute(a, u, s, v, sw)
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        plan.execute(a, u, s, v, sw)
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = float(np.mean([b.elapsed_time(e) for b, e in ev]))

    # Sanity of the last step's results (outside the timed region).
    sweeps = sw.cpu().numpy()
    s_h = s.cpu().numpy()
    worst = 0.0
    for k in range(0, len(mats), 61):
        n = int(ns[k])
        worst = max(worst, metrics.sigma_errors(s_h[plan.s_off[k]:plan.s_off[k] + n], sig[k])[0])
    if not worst < 1e-12:
        print(f"bench.py: sigma error {worst:.3e} exceeds 1e-12", file=sys.stderr)
        sys.exit(1)

    flops = float(np.sum(sweep_flops(ms, ns) * sweeps))
    ms_per_step = wall / args.steps * 1e3
    achieved_tflops = flops / (kernel_ms * 1e-3) / 1e12
    traffic = load_traffic("lds_batch_f64_uneven4096_traffic.json")
    result = {
        "metric": "matrices/s (batched FP64 one-sided Jacobi, 4096 uneven matrices)",
        "value": len(ms) * args.steps / wall,
        "unit": "matrices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "batched_uneven_fp64", "matrices": int(len(ms)), "n": "U{8..64}",
                   "m": "U{n..2n}", "sigma": "logspace(0,-6,n)", "epsilon": eps, "max_sweeps": max_sweeps,
                   "mean_sweeps": float(sweeps.mean()), "max_sigma_err_normwise": worst},
        "roofline": {"bound": "mfma", "achieved": achieved_tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tflops / FP64_PEAK_TFLOPS,
                     "traffic": traffic.get("bytes_per_launch") if traffic else None,
                     "kernel": "lds_batch_f64", "kernel_ms": kernel_ms,
                     "flops_per_launch": flops,
                     "compulsory_bytes_per_launch": float(np.sum(compulsory_bytes(ms, ns)))},
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(mats, eps, max_sweeps)
    print(json.dumps(result), flush=True)
    plan.close()


if __name__ == "__main__":
    main()
