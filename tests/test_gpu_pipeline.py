"""GPU tests of the C++ API, the batched pipeline and the CLI end to end.

The pipeline writes U, Sigma (1 x n) and V for every input matrix in input
order; every output is checked against the CPU oracle (sigma position by
position at eps above the rounding floor, U and V to 1e-10 where the
singular values are well separated) or, for large streams, against the
reconstruction and orthogonality invariants.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from jsvd import gen
from jsvd import io as jio
from jsvd import metrics

pytestmark = pytest.mark.gpu

PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "jacobi-svd-cuda-opt_amd")
BIN = os.path.join(PKG, "bin")
CLI = os.path.join(BIN, "jacobi-svd-amd")


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run(args):
    _note("run: " + " ".join(str(a) for a in args))
    return subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=300)


def _triples(path, reader):
    out = reader(path)
    assert len(out) % 3 == 0
    return [(out[3 * k], out[3 * k + 1][0], out[3 * k + 2]) for k in range(len(out) // 3)]


def test_cxx_api_binary():
    r = _run([os.path.join(BIN, "jsvd-api-tests")])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "api_tests passed" in r.stdout


def test_dispatch_reader_binary():
    # MatDispatchReader / PinnedHostTaskBuffer with real pinned memory.
    r = _run([os.path.join(BIN, "jsvd-dispatch-tests")])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "dispatch_tests passed" in r.stdout


def test_cli_mat_device_byte_swap_matches_txt(tmp_path):
    # A *.mat stream is staged in network byte order and swapped on the
    # device; a *.txt stream (max_digits10, exact) is decoded on the host.
    # Both must hand the solver the same bits, so the outputs are identical.
    rng = np.random.default_rng(17)
    mats = [rng.standard_normal((int(m), int(n))) * 10.0 ** rng.uniform(-3, 3)
            for n in rng.integers(1, 40, size=300) for m in [n + rng.integers(0, 40)]]
    mats.append(rng.standard_normal((700, 300)))  # a large matrix next to the batched ones
    src_m, src_t = tmp_path / "in.mat", tmp_path / "in.txt"
    dst_m, dst_t = tmp_path / "out_m.mat", tmp_path / "out_t.mat"
    jio.write_mat(src_m, mats)
    jio.write_txt(src_t, mats)
    for src, dst in ((src_m, dst_m), (src_t, dst_t)):
        r = _run([CLI, src, dst, "-e", "1e-14", "-s", "40", "-q"])
        assert r.returncode == 0, r.stderr
    a, b = jio.read_mat(dst_m), jio.read_mat(dst_t)
    assert len(a) == len(b) == 3 * len(mats)
    for x, y in zip(a, b):
        assert x.shape == y.shape and np.array_equal(x, y)
    refs = oracle.solve_batch(mats[:40], 1e-14, 40)
    for k, ref in enumerate(refs):
        assert np.max(np.abs(a[3 * k + 1][0] - ref[1])) <= 1e-12 * ref[1].max()


def test_cli_mat_stream_two_batches(tmp_path):
    # 5000 uneven matrices: two device batches (4096 + 904) in flight.
    mats, sig = gen.uneven_batch(5000, seed=321)
    src, dst = tmp_path / "in.mat", tmp_path / "out.mat"
    jio.write_mat(src, mats)
    r = _run([CLI, src, dst, "-e", "1e-13", "-s", "30", "--json-report", "-q"])
    assert r.returncode == 0, r.stderr
    rep = json.loads(r.stdout)
    assert rep["testcase_count"] == 5000 and rep["emitted_matrix_count"] == 15000
    res = _triples(dst, jio.read_mat)
    assert len(res) == 5000
    total = 0
    subset = list(range(0, 5000, 50))
    refs = oracle.solve_batch([mats[k] for k in subset], 1e-13, 30)
    for k, ref in zip(subset, refs):
        u, s, v = res[k]
        assert u.shape == mats[k].shape and v.shape == (mats[k].shape[1],) * 2
        assert np.max(np.abs(s - ref[1])) <= 1e-12 * ref[1].max()
        total += 1
    for k in range(0, 5000, 7):
        u, s, v = res[k]
        assert metrics.residual(mats[k], u, s, v) < 1e-12
        assert metrics.sigma_errors(s, sig[k])[0] < 1e-12
    assert rep["total_sweeps"] > 5000


def test_cli_graph_replay_same_shapes(tmp_path):
    # One slot (--queue-capacity 1) and three batches of identical shapes: the
    # second and third batches replay the captured graph of the first.
    rng = np.random.default_rng(8)
    mats = [rng.standard_normal((24, 16)) for _ in range(3 * 4096)]
    src, dst = tmp_path / "same.mat", tmp_path / "same.out.mat"
    jio.write_mat(src, mats)
    r = _run([CLI, src, dst, "-e", "1e-13", "-s", "30", "-c", "1", "--json-report", "-q"])
    assert r.returncode == 0, r.stderr
    res = _triples(dst, jio.read_mat)
    refs = oracle.solve_batch(mats, 1e-13, 30)
    for k in range(len(mats)):
        u, s, v = res[k]
        assert np.max(np.abs(s - refs[k][1])) <= 1e-12 * refs[k][1].max(), k
    rep = json.loads(r.stdout)
    assert rep["total_sweeps"] == pytest.approx(sum(x[3] for x in refs), rel=0.02)


def test_cli_txt_mixed_and_large(tmp_path):
    # Text in and out, mixed precision, plus a matrix large enough for the
    # persistent kernel in the same stream.
    rng = np.random.default_rng(9)
    mats = [gen.constructed(rng, 60, 40, 10.0 ** (-4 * np.arange(40) / 39)), rng.standard_normal((300, 150)),
            rng.standard_normal((5, 5))]
    src, dst = tmp_path / "in.txt", tmp_path / "out.txt"
    jio.write_txt(src, mats)
    r = _run([CLI, src, dst, "-e", "1e-14", "-s", "30", "--precision", "mixed"])
    assert r.returncode == 0, r.stderr
    assert "Pipeline completed." in r.stdout and "testcases       : 3" in r.stdout
    res = _triples(dst, jio.read_txt)
    for a, (u, s, v) in zip(mats, res):
        ref = np.linalg.svd(a, compute_uv=False)
        assert metrics.sigma_errors(s, ref)[0] < 1e-12
        assert metrics.residual(a, u, s, v) < 1e-12


def test_cli_rejects_wide_matrix_in_stream(tmp_path):
    src, dst = tmp_path / "bad.mat", tmp_path / "bad.out.mat"
    jio.write_mat(src, [np.eye(3), np.ones((2, 3))])
    r = _run([CLI, src, dst])
    assert r.returncode == 1
    assert "Execution failed: One-sided Jacobi SVD requires rows >= columns" in r.stderr


def test_cli_report_errors_stop_reason_and_reference_sigma(tmp_path):
    # The JSON report carries per-testcase sweeps, stop reasons and (with
    # --report-errors / --reference-sigma) the error norms; with
    # --max-sweeps 1 every testcase is reported as stopped at max_sweeps.
    rng = np.random.default_rng(99)
    mats = [rng.standard_normal((int(m), int(n))) for n, m in [(8, 12), (30, 45), (64, 128), (150, 300)]]
    inp, out = tmp_path / "in.mat", tmp_path / "out.mat"
    jio.write_mat(inp, mats)
    refs = oracle.solve_batch(mats, 1e-12, 40)
    ref_path = tmp_path / "ref.mat"
    jio.write_mat(ref_path, [r[1].reshape(1, -1) for r in refs])
    r = _run([CLI, inp, out, "-e", "1e-12", "-s", "40", "--reference-sigma", ref_path, "--json-report", "-q"])
    assert r.returncode == 0, r.stderr
    rep = json.loads(r.stdout)
    assert rep["unconverged_count"] == 0 and rep["matrix_stop"] == ["converged"] * 4
    assert [abs(s - ref[3]) <= 1 for s, ref in zip(rep["matrix_sweeps"], refs)] == [True] * 4
    assert rep["total_sweeps"] == sum(rep["matrix_sweeps"])
    # The reported norms are those of the written factors (numpy on the
    # output file), and small: orthogonality within n * eps (each converged
    # pair is orthogonal to eps), residual near the rounding level.
    for a, (u, s, v), e in zip(mats, _triples(out, jio.read_mat), rep["matrix_errors"]):
        want = (metrics.orthogonality(u, s > 1e-12), metrics.orthogonality(v), metrics.residual(a, u, s, v))
        for name, w in zip(("orth_u", "orth_v", "residual"), want):
            assert abs(e[name] - w) <= 1e-13 + 1e-6 * w, (a.shape, name, e[name], w)
        assert e["orth_u"] < a.shape[1] * 1e-12 and e["orth_v"] < a.shape[1] * 1e-12 and e["residual"] < 1e-13
        assert e["sigma_err_norm"] < 1e-12
    r = _run([CLI, inp, out, "-y", "-s", "1", "--json-report", "-q"])
    assert r.returncode == 0, r.stderr
    rep = json.loads(r.stdout)
    assert rep["unconverged_count"] == 4 and rep["matrix_stop"] == ["max_sweeps"] * 4
    assert rep["matrix_sweeps"] == [1] * 4 and "matrix_errors" not in rep
    bad = tmp_path / "short.mat"
    jio.write_mat(bad, [refs[0][1].reshape(1, -1)])
    r = _run([CLI, inp, out, "-y", "--reference-sigma", bad, "--json-report", "-q"])
    assert r.returncode == 1 and "Reference sigma" in r.stderr
