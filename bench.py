#!/usr/bin/env python3
"""Benchmark of the batched FP64 one-sided Jacobi SVD on one MI355X.

Workload (one step = one pass of the hot path over one batch): 4096
matrices, n ~ U{8..64}, m ~ U{n..2n}, A = Q1 diag(sigma) Q2^T with Haar
factors and sigma log-spaced 1 .. 1e-6, eps = 1e-15, at most 30 sweeps,
solved by one device work-list plan.  Before the timed region the shape
class of each column-count bucket is picked by timing whole plans of this
batch (jsvd_autotune_batch; --no-tune keeps the defaults).  Inputs are
resident in HBM before the timed region; outputs stay on the device.

Prints one JSON line: matrices/s, the work-list kernels' FP64 throughput
against the MI355X peak and their LDS stores against the measured LDS store
rate (HIP events on the launch stream), the
HBM traffic measured with rocprofv3 counters, the CPU oracle timed on a
bounded sample, the mixed-precision vs FP64 comparison on 1024 matrices
of 128 x 128 with kappa = 10^U(3,8) and on the uneven batch (both legs'
sigma against the FP64 oracle, per matrix), the blocked and plain solves of
one 8192 x 2048 matrix (sweeps, HBM GB/s and FP64 rate per sweep on the
algorithmic definitions, sigma against the oracle's solve of it), the plain
solve of one 256 x 256 matrix beside the oracle (one thread and
round-parallel) and the host-API call latency, and the adjoint kernel on
512 matrices of 96 x 64 (FP64 and FP32).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-mixed] [--no-blocked] [--no-single]
                    [--no-backward] [--no-tune]
    python bench.py --traffic-probe    # one 256 MiB copy + one plan execution,
                                        # for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
    python bench.py --blocked-probe    # one 256 MiB copy + two blocked sweeps of
                                        # 8192 x 2048 at widths 32 and 16, for --pmc
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

FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix peaks are equal)
HBM_PEAK_GBS = 8000.0
LDS_WRITE_TBS = 47.0      # measured: conflict-free 16-byte LDS stores, 76 B/clk/CU (loads: 228 B/clk/CU)
EPS, MAX_SWEEPS = 1e-15, 30


def sweep_flops(m, n):
    """Algorithmic FP64 flops of one plain sweep: n(n-1)/2 pairs x (12m + 6n)."""
    m = np.asarray(m, dtype=np.float64)
    n = np.asarray(n, dtype=np.float64)
    return n * (n - 1) / 2 * (12 * m + 6 * n)


def sweep_lds_write_bytes(m, n):
    """LDS bytes stored per sweep if every pair rotates: each pair group holds
    its top column in registers and stores only its bottom A and V columns,
    8 (m + n) bytes per pair: 4 n(n-1)(m+n)."""
    m = np.asarray(m, dtype=np.float64)
    n = np.asarray(n, dtype=np.float64)
    return 4 * n * (n - 1) * (m + n)


def compulsory_bytes(m, n):
    """Read A, write U, sigma, V once: 8 (2mn + n^2 + n) bytes per matrix."""
    m = np.asarray(m, dtype=np.float64)
    n = np.asarray(n, dtype=np.float64)
    return 8 * (2 * m * n + n * n + n)


def cpu_model():
    """Host CPU model name (for every CPU figure in the report)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores():
    """Host threads for the CPU legs: the affinity set, capped by
    OMP_NUM_THREADS when the environment sets it (and at 64)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        n = min(n, int(os.environ.get("OMP_NUM_THREADS", n)))
    except ValueError:
        pass
    return max(1, min(n, 64))


def load_json(name):
    path = os.path.join(REPO, "profiles", name)
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def device_buffers(plan, flat, torch):
    dev = torch.device("cuda:0")
    a = torch.from_numpy(flat).to(dev)
    u = torch.empty(plan.a_elems, dtype=torch.float64, device=dev)
    s = torch.empty(plan.s_elems, dtype=torch.float64, device=dev)
    v = torch.empty(plan.v_elems, dtype=torch.float64, device=dev)
    sw = torch.zeros(len(plan.ms), dtype=torch.int32, device=dev)
    return a, u, s, v, sw


def time_plan(plan, bufs, steps, warmup, torch):
    a, u, s, v, sw = bufs
    for _ in range(warmup):
        plan.execute(a, u, s, v, sw)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        ev[k][0].record(stream)
        plan.execute(a, u, s, v, sw)
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, float(np.mean([b.elapsed_time(e) for b, e in ev]))


def unpack(plan, s, k):
    n = int(plan.ns[k])
    return s[plan.s_off[k]:plan.s_off[k] + n]


def cpu_baseline(mats, budget_s=12.0):
    """CPU oracle, one matrix per OpenMP thread, on a bounded sample of the batch
    (whole batch repeated until about budget_s of CPU work)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle

    cores = host_cores()
    t0 = time.perf_counter()
    oracle.solve_batch(mats, EPS, MAX_SWEEPS, threads=cores)
    once = time.perf_counter() - t0
    reps = max(1, int(budget_s / max(once, 1e-6)))
    t0 = time.perf_counter()
    for _ in range(reps):
        oracle.solve_batch(mats, EPS, MAX_SWEEPS, threads=cores)
    dt = time.perf_counter() - t0
    return {"value": reps * len(mats) / dt, "unit": "matrices/s", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"the full {len(mats)}-matrix batch solved {reps} times ({dt:.1f} s) by the C oracle, "
                      f"one matrix per OpenMP thread on {cores} threads"}


def sigma_error_stats(metrics, got, ref):
    """Per-matrix sigma errors (sorted, normwise and componentwise) of got
    against ref: max and percentiles."""
    norm = np.array([metrics.sigma_errors(g, r)[0] for g, r in zip(got, ref)])
    comp = np.array([metrics.sigma_errors(g, r)[1] for g, r in zip(got, ref)])
    return {"normwise_max": float(norm.max()), "normwise_p50": float(np.percentile(norm, 50)),
            "normwise_p99": float(np.percentile(norm, 99)), "componentwise_max": float(comp.max()),
            "matrices_above_1e-12": int(np.sum(norm > 1e-12))}


def mixed_vs_fp64(torch, jsvd, gen, metrics, cpu, workload, steps=3):
    """FP64 plan vs mixed-precision plan on one batch: the 1024 x 128x128
    batch of kappa = 10^U(3,8) ("cfg4") or the 4096 uneven batch ("cfg2").
    With cpu, every matrix of both legs is compared with the FP64 oracle's
    solve of it (sigma sorted), and the oracle's rate is reported."""
    if workload == "cfg4":
        mats, _ = gen.conditioned_batch(1024)
        desc = "1024 x 128x128, kappa=10^U(3,8), sigma_k = kappa^(-k/127)"
    else:
        mats, _ = gen.uneven_batch(4096)
        desc = "4096 uneven, n ~ U{8..64}, m ~ U{n..2n}, sigma logspace(0,-6,n)"
    flat, ms, ns = gen.pack(mats)
    out = {"workload": desc, "epsilon": EPS, "max_sweeps": MAX_SWEEPS}
    sig = {}
    for name, prec in (("fp64", jsvd.PRECISION_FP64), ("mixed", jsvd.PRECISION_MIXED)):
        plan = jsvd.BatchPlan(ms, ns, epsilon=EPS, max_sweeps=MAX_SWEEPS, precision=prec)
        bufs = device_buffers(plan, flat, torch)
        wall, kernel_ms = time_plan(plan, bufs, steps, 1, torch)
        s_h = bufs[2].cpu().numpy()
        sweeps = bufs[4].cpu().numpy()
        sig[name] = [unpack(plan, s_h, k).copy() for k in range(len(mats))]
        out[name] = {"matrices_per_s": len(mats) * steps / wall, "ms_per_batch": wall / steps * 1e3,
                     "mean_sweeps": float(sweeps.mean())}
        plan.close()
    if cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle

        cores = host_cores()
        t0 = time.perf_counter()
        refs = oracle.solve_batch(mats, EPS, MAX_SWEEPS, threads=cores)
        dt = time.perf_counter() - t0
        ref_sig = [r[1] for r in refs]
        out["oracle_fp64"] = {"matrices_per_s": len(mats) / dt, "cores": cores, "cpu_model": cpu_model(),
                              "sample": f"the whole batch, one matrix per OpenMP thread"}
        out["fp64"]["sigma_err_vs_oracle"] = sigma_error_stats(metrics, sig["fp64"], ref_sig)
        out["mixed"]["sigma_err_vs_oracle"] = sigma_error_stats(metrics, sig["mixed"], ref_sig)
    out["mixed_vs_fp64_sigma_err_normwise_max"] = float(
        max(metrics.sigma_errors(sm, sf)[0] for sm, sf in zip(sig["mixed"], sig["fp64"])))
    out["speedup_mixed_over_fp64"] = out["mixed"]["matrices_per_s"] / out["fp64"]["matrices_per_s"]
    return out


def blocked_bytes(m, n, b):
    """Algorithmic HBM bytes of one blocked sweep (the survey's definition):
    one read and one write of the two b-column panels of A and of V per
    block-pair visit, 16 b Nb (Nb-1) (m + n)."""
    nb = -(-n // (2 * b)) * 2
    return 16.0 * b * nb * (nb - 1) * (m + n)


def blocked_flops(m, n, b):
    """Algorithmic FP64 flops of one blocked sweep (the survey's definition):
    4 b^2 Nb (Nb-1) (2m + n)."""
    nb = -(-n // (2 * b)) * 2
    return 4.0 * b * b * nb * (nb - 1) * (2 * m + n)


def blocked_traffic_model(m, n, b):
    """This build's own HBM byte model of a blocked sweep (kept beside the
    algorithmic figure): each visit also reads the panels of A once more for
    the Gram (16 b m), 8 b Nb (Nb-1) (3m + 2n)."""
    nb = -(-n // (2 * b)) * 2
    return 8.0 * b * nb * (nb - 1) * (3 * m + 2 * n)


def blocked_tall(torch, jsvd, gen, metrics, cpu):
    """Single 8192 x 2048 matrix, sigma_k = 10^(-10k/2047), Haar factors, blocked
    Jacobi at block widths 32 and 16 (eps 1e-15, up to 80 sweeps), and the
    plain column-pair path run to convergence; sigma is compared with the
    oracle's solve of the same matrix (tests/golden/tall_8192x2048.npz, written
    once by tests/golden/make_tall_fixture.py).  With cpu, the oracle's plain
    sweep is timed beside it (one sweep, all host cores)."""
    a, sig = gen.tall_geometric()
    m, n = a.shape
    out = {"workload": "8192x2048, sigma_k=10^(-10k/2047), Haar factors", "epsilon": EPS, "max_sweeps": 80}
    fixture = np.load(os.path.join(REPO, "tests", "golden", "tall_8192x2048.npz"))
    out["oracle_sweeps"] = int(fixture["sweeps"])
    if cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle

        cores = host_cores()
        t0 = time.perf_counter()
        oracle.solve(a, EPS, 1, threads=cores)
        out["oracle_plain_s_per_sweep"] = time.perf_counter() - t0
        out["oracle_cores"] = cores
        out["oracle_cpu_model"] = cpu_model()
    for bw in (32, 16):
        plan = jsvd.BatchPlan([m], [n], epsilon=EPS, max_sweeps=80, block_width=bw)
        bufs = device_buffers(plan, a.ravel(), torch)
        wall, kernel_ms = time_plan(plan, bufs, 1, 1, torch)
        A, U, S, V, sw = bufs
        sweeps = int(sw.item())
        Am, Um, Vm = A.view(m, n), U.view(m, n), V.view(n, n)
        res = (torch.linalg.norm(Am - (Um * S) @ Vm.T) / torch.linalg.norm(Am)).item()
        s_h = S.cpu().numpy()
        norm_err, comp_err = metrics.sigma_errors(s_h, sig)
        per_sweep_s = kernel_ms * 1e-3 / sweeps
        byt, fl = blocked_bytes(m, n, bw), blocked_flops(m, n, bw)
        out[f"block_width_{bw}"] = {
            "ms": kernel_ms, "sweeps": sweeps, "ms_per_sweep": per_sweep_s * 1e3,
            "bytes_per_sweep": byt, "hbm_GBs_per_sweep": byt / per_sweep_s / 1e9,
            "hbm_frac": byt / per_sweep_s / 1e9 / HBM_PEAK_GBS,
            "flops_per_sweep": fl, "fp64_TFs": fl / per_sweep_s / 1e12,
            "mfma_frac": fl / per_sweep_s / 1e12 / FP64_PEAK_TFLOPS,
            "traffic_model_bytes_per_sweep": blocked_traffic_model(m, n, bw),
            "traffic_model_GBs": blocked_traffic_model(m, n, bw) / per_sweep_s / 1e9,
            "residual": res, "sigma_err_normwise": norm_err, "sigma_err_componentwise": comp_err,
            "sigma_err_vs_oracle": metrics.sigma_errors(s_h, fixture["sigma_sorted"])[0]}
        plan.close()
    # Plain column-pair order on the GPU (grouped wave kernel, rows streamed)
    # run to convergence; algorithmic traffic 16 n (n-1) (m+n) bytes per sweep.
    plan = jsvd.BatchPlan([m], [n], epsilon=EPS, max_sweeps=80)
    bufs = device_buffers(plan, a.ravel(), torch)
    wall, kernel_ms = time_plan(plan, bufs, 1, 0, torch)
    sweeps = int(bufs[4].item())
    per_sweep_s = kernel_ms * 1e-3 / sweeps
    byt = 16.0 * n * (n - 1) * (m + n)
    out["plain_gpu"] = {"sweeps": sweeps, "ms": kernel_ms, "ms_per_sweep": per_sweep_s * 1e3,
                        "bytes_per_sweep": byt, "hbm_GBs_per_sweep": byt / per_sweep_s / 1e9,
                        "hbm_frac": byt / per_sweep_s / 1e9 / HBM_PEAK_GBS,
                        "sigma_err_vs_oracle": metrics.sigma_errors(bufs[2].cpu().numpy(),
                                                                    fixture["sigma_sorted"])[0]}
    plan.close()
    return out


def plain_single(torch, jsvd, gen, cpu):
    """The reference's CPU-runnable case: one 256 x 256 N(0,1) matrix, plain
    column-pair Jacobi (grouped wave kernel), with the oracle beside it."""
    a = gen.gaussian_square(256)
    plan = jsvd.BatchPlan([256], [256], epsilon=EPS, max_sweeps=MAX_SWEEPS)
    bufs = device_buffers(plan, a.ravel(), torch)
    wall, kernel_ms = time_plan(plan, bufs, 3, 1, torch)
    out = {"workload": "256x256 N(0,1), eps 1e-15, 30 sweeps", "gpu_ms": kernel_ms,
           "sweeps": int(bufs[4].item())}
    plan.close()
    # 48 such matrices (seeds base + k) in one plan: grouped launches of the
    # same kernel, several matrices side by side.
    mats = [gen.gaussian_square(256, seed=gen.BASE_SEED + k) for k in range(48)]
    flat, ms, ns = gen.pack(mats)
    plan = jsvd.BatchPlan(ms, ns, epsilon=EPS, max_sweeps=MAX_SWEEPS)
    bufs = device_buffers(plan, flat, torch)
    wall, batch_ms = time_plan(plan, bufs, 3, 1, torch)
    out["batch_48"] = {"gpu_ms": batch_ms, "matrices_per_s": 48 / (batch_ms * 1e-3),
                       "mean_sweeps": float(bufs[4].double().mean().item())}
    plan.close()
    # 64 tall 1024 x 64 N(0,1) matrices: one workgroup per matrix, rows
    # streamed (taller than the register kernels).
    tall = [np.random.default_rng(gen.BASE_SEED + 100 + k).standard_normal((1024, 64)) for k in range(64)]
    flat, ms, ns = gen.pack(tall)
    plan = jsvd.BatchPlan(ms, ns, epsilon=EPS, max_sweeps=MAX_SWEEPS)
    bufs = device_buffers(plan, flat, torch)
    wall, tall_ms = time_plan(plan, bufs, 3, 1, torch)
    out["tall_batch_64x1024x64"] = {"gpu_ms": tall_ms, "matrices_per_s": 64 / (tall_ms * 1e-3),
                                    "mean_sweeps": float(bufs[4].double().mean().item())}
    plan.close()
    # The host-array entry point (one_sided_jacobi_svd's C-ABI, jsvd_solve_f64):
    # copies in and out included, the first call building the plan (device
    # allocations, occupancy queries), later calls reusing the cached one.
    t0 = time.perf_counter()
    jsvd.svd(a, epsilon=EPS, max_sweeps=MAX_SWEEPS)
    first = time.perf_counter() - t0
    calls = 10
    t0 = time.perf_counter()
    for _ in range(calls):
        jsvd.svd(a, epsilon=EPS, max_sweeps=MAX_SWEEPS)
    out["host_api_ms_per_call"] = {"first_call": first * 1e3, "cached_plan": (time.perf_counter() - t0) / calls * 1e3,
                                   "device_solve": kernel_ms}
    if cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle

        t0 = time.perf_counter()
        ref = oracle.solve(a, EPS, MAX_SWEEPS, threads=1)
        out["oracle_ms_1_thread"] = (time.perf_counter() - t0) * 1e3
        out["oracle_sweeps"] = int(ref[3])
        cores = host_cores()
        # Round-parallel oracle: the disjoint pairs of a round over all host
        # threads, one barrier per round (bitwise equal to the serial solve).
        t0 = time.perf_counter()
        oracle.solve(a, EPS, MAX_SWEEPS, threads=cores)
        out["oracle_ms_round_parallel"] = (time.perf_counter() - t0) * 1e3
        out["oracle_cores"] = cores
        out["oracle_cpu_model"] = cpu_model()
        t0 = time.perf_counter()
        oracle.solve_batch(mats, EPS, MAX_SWEEPS, threads=cores)
        dt = time.perf_counter() - t0
        out["batch_48"]["oracle_matrices_per_s"] = 48 / dt
        out["batch_48"]["oracle_cores"] = cores
    return out


def backward_gapped(torch, jsvd, gen, steps=10):
    """512 matrices 96 x 64, sigma log-spaced 1 .. 1e-2: forward solve, then the
    adjoint kernel on seeded upstream gradients, FP64 and FP32.  The adjoint
    reads U, gU (m n), V, gV (n^2), sigma, gsigma and writes gA (m n): its HBM
    bytes are 8 (3mn + 2n^2 + 2n) per matrix in FP64 (half in FP32); its flops
    are 6 m n^2 + 2 n^3 (U^T gU, V^T gV, U N, W V^T)."""
    mats, _ = gen.gapped_batch(512)
    flat, ms, ns = gen.pack(mats)
    plan = jsvd.BatchPlan(ms, ns, epsilon=EPS, max_sweeps=MAX_SWEEPS)
    a, u, s, v, sw = device_buffers(plan, flat, torch)
    plan.execute(a, u, s, v, sw)
    torch.cuda.synchronize()
    g = torch.Generator(device="cpu").manual_seed(5)
    gu = torch.randn(plan.a_elems, generator=g, dtype=torch.float64).to(u.device)
    gs = torch.randn(plan.s_elems, generator=g, dtype=torch.float64).to(u.device)
    gv = torch.randn(plan.v_elems, generator=g, dtype=torch.float64).to(u.device)
    m, n = 96, 64
    out = {"workload": "512 x 96x64, sigma_k = 10^(-2k/63), gU, gsigma, gV ~ N(0,1)"}
    for name, dt in (("fp64", torch.float64), ("fp32", torch.float32)):
        tens = [t.to(dt) for t in (u, s, v, gu, gs, gv)]
        ga = torch.empty_like(tens[0])
        for _ in range(2):
            plan.backward(*tens, ga)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            plan.backward(*tens, ga)
        e1.record()
        torch.cuda.synchronize()
        ms_per = e0.elapsed_time(e1) / steps
        elem = 8 if dt == torch.float64 else 4
        byt = 512 * elem * (3 * m * n + 2 * n * n + 2 * n)
        fl = 512 * (6.0 * m * n * n + 2.0 * n ** 3)
        out[name] = {"ms": ms_per, "matrices_per_s": 512 / (ms_per * 1e-3),
                     "hbm_GBs": byt / (ms_per * 1e-3) / 1e9, "hbm_frac": byt / (ms_per * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "TFs": fl / (ms_per * 1e-3) / 1e12}
    plan.close()
    return out


def traffic_probe():
    """One 256 MiB device copy (the counter calibration) and one execution of
    the benchmark plan, for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs."""
    import torch

    import jsvd
    from jsvd import gen

    mats, _ = gen.uneven_batch(4096)
    flat, ms, ns = gen.pack(mats)
    plan = jsvd.BatchPlan(ms, ns, epsilon=EPS, max_sweeps=MAX_SWEEPS)
    bufs = device_buffers(plan, flat, torch)
    src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    plan.execute(*bufs)
    torch.cuda.synchronize()
    print("traffic probe done", flush=True)


def blocked_probe():
    """One 256 MiB device copy (the counter calibration) and two blocked
    sweeps of the 8192 x 2048 matrix at block widths 32 and 16, for
    rocprofv3 --pmc runs on gram_partial / gram_solve / block_update."""
    import torch

    import jsvd
    from jsvd import gen

    a, _ = gen.tall_geometric()
    m, n = a.shape
    src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    for bw in (32, 16):
        plan = jsvd.BatchPlan([m], [n], epsilon=EPS, max_sweeps=2, block_width=bw)
        bufs = device_buffers(plan, a.ravel(), torch)
        plan.execute(*bufs)
        torch.cuda.synchronize()
        plan.close()
    print("blocked probe done", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed vs FP64 comparison")
    ap.add_argument("--no-blocked", action="store_true", help="skip the 8192 x 2048 blocked solve")
    ap.add_argument("--no-single", action="store_true", help="skip the single 256 x 256 plain solve")
    ap.add_argument("--no-backward", action="store_true", help="skip the 512 x 96x64 adjoint")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default shape classes (skip the batch autotune before the timed region)")
    ap.add_argument("--traffic-probe", action="store_true")
    ap.add_argument("--blocked-probe", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "batches would shard by matrix across GPUs with no exchange, which is not built here.",
              file=sys.stderr)
        sys.exit(2)
    if args.traffic_probe:
        traffic_probe()
        return
    if args.blocked_probe:
        blocked_probe()
        return

    import torch

    import jsvd
    from jsvd import gen, metrics

    mats, sig = gen.uneven_batch(args.count)
    flat, ms, ns = gen.pack(mats)
    plan = jsvd.BatchPlan(ms, ns, epsilon=EPS, max_sweeps=MAX_SWEEPS)
    bufs = device_buffers(plan, flat, torch)
    tune = None
    if not args.no_tune:
        # Shape classes per column-count bucket from whole-plan timings of this
        # batch (outside the timed region); the timed plan is built afterwards.
        tune = jsvd.autotune_batch(ms, ns, *bufs[:4], reps=3, epsilon=EPS, max_sweeps=MAX_SWEEPS)
        plan.close()
        plan = jsvd.BatchPlan(ms, ns, epsilon=EPS, max_sweeps=MAX_SWEEPS)
    wall, kernel_ms = time_plan(plan, bufs, args.steps, args.warmup, torch)

    # Sanity of the last step's results (outside the timed region).
    sweeps = bufs[4].cpu().numpy()
    s_h = bufs[2].cpu().numpy()
    worst = max(metrics.sigma_errors(unpack(plan, s_h, k), sig[k])[0] for k in range(0, len(mats), 61))
    if not worst < 1e-12:
        print(f"bench.py: sigma error {worst:.3e} exceeds 1e-12", file=sys.stderr)
        sys.exit(1)

    flops = float(np.sum(sweep_flops(ms, ns) * sweeps))
    lds_write_bytes = float(np.sum(sweep_lds_write_bytes(ms, ns) * sweeps))
    achieved_tflops = flops / (kernel_ms * 1e-3) / 1e12
    traffic = load_json("lds_class_f64_uneven4096_traffic.json")
    info = plan.info
    result = {
        "metric": "matrices/s (batched FP64 one-sided Jacobi, 4096 uneven matrices)",
        "value": len(ms) * args.steps / wall,
        "unit": "matrices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "batched_uneven_fp64", "matrices": int(len(ms)), "n": "U{8..64}",
                   "m": "U{n..2n}", "sigma": "logspace(0,-6,n)", "epsilon": EPS, "max_sweeps": MAX_SWEEPS,
                   "mean_sweeps": float(sweeps.mean()), "max_sigma_err_normwise": worst,
                   "buckets": info["buckets"],
                   "shape_classes": ({"per_bucket": tune["classes"], "plans_timed": tune["samples"],
                                      "ms_default": tune["ms_initial"], "ms_tuned": tune["ms_tuned"]}
                                     if tune else "default")},
        # The work-list kernel issues FP64 VALU instructions (no MFMA); MI355X's
        # FP64 vector and matrix peaks are both 78.6 TFLOP/s.
        "roofline": {"bound": "fp64-valu", "achieved": achieved_tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tflops / FP64_PEAK_TFLOPS,
                     "traffic": traffic.get("bytes_per_launch") if traffic else None,
                     "kernel": "lds_class_f64 (all buckets of one plan execution)", "kernel_ms": kernel_ms,
                     "flops_per_launch": flops,
                     "lds_stores": {"achieved_TBs": lds_write_bytes / (kernel_ms * 1e-3) / 1e12,
                                    "peak_TBs": LDS_WRITE_TBS,
                                    "frac": lds_write_bytes / (kernel_ms * 1e-3) / 1e12 / LDS_WRITE_TBS,
                                    "bytes_per_launch_upper": lds_write_bytes},
                     "compulsory_bytes_per_launch": float(np.sum(compulsory_bytes(ms, ns)))},
    }
    plan.close()
    if not args.no_mixed:
        result["mixed_vs_fp64"] = mixed_vs_fp64(torch, jsvd, gen, metrics, not args.no_cpu, "cfg4")
        result["mixed_vs_fp64_uneven4096"] = mixed_vs_fp64(torch, jsvd, gen, metrics, not args.no_cpu, "cfg2")
    if not args.no_blocked:
        result["blocked_8192x2048"] = blocked_tall(torch, jsvd, gen, metrics, not args.no_cpu)
    if not args.no_backward:
        result["backward_512x96x64"] = backward_gapped(torch, jsvd, gen)
    if not args.no_single:
        result["plain_256x256"] = plain_single(torch, jsvd, gen, not args.no_cpu)
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(mats)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
