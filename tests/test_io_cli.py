"""CPU tests of the matrix-stream codecs and of the CLI front end (no device
call: help, version, dry-run and argument errors stop before the pipeline).

The C++ checks live in tests/cli/*.cpp and are built by the package
Makefile into jacobi-svd-cuda-opt_amd/bin/; the first four CLI checks restate
the reference's own CLI parser tests.
"""
import os
import subprocess

import numpy as np
import pytest

from jsvd import io as jio

PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "jacobi-svd-cuda-opt_amd")
BIN = os.path.join(PKG, "bin")
CLI = os.path.join(BIN, "jacobi-svd-amd")


def _run(args, **kw):
    return subprocess.run(args, capture_output=True, text=True, timeout=120, **kw)


def test_cli_parser_binary():
    r = _run([os.path.join(BIN, "jsvd-cli-tests")])
    assert r.returncode == 0, r.stderr
    assert "cli_parser_tests passed" in r.stdout


def test_io_binary_round_trips_and_errors():
    r = _run([os.path.join(BIN, "jsvd-io-tests")])
    assert r.returncode == 0, r.stderr
    assert "io_tests passed" in r.stdout


def test_python_and_cpp_mat_codecs_agree(tmp_path):
    rng = np.random.default_rng(4)
    mats = [rng.standard_normal((3, 2)), np.array([[np.pi]]), rng.standard_normal((7, 7)) * 1e-200]
    path = tmp_path / "x.mat"
    jio.write_mat(path, mats)
    assert path.stat().st_size == sum(16 + 8 * a.size for a in mats)
    r = _run([os.path.join(BIN, "jsvd-io-tests"), str(path)])
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().split("\n")
    assert len(lines) == len(mats)
    for line, a in zip(lines, mats):
        tok = line.split()
        assert (int(tok[0]), int(tok[1])) == a.shape
        np.testing.assert_array_equal(np.array([float(t) for t in tok[2:]]), a.ravel())
    back = jio.read_mat(path)
    for x, y in zip(back, mats):
        np.testing.assert_array_equal(x, y)


def test_python_txt_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    mats = [rng.standard_normal((4, 3)), rng.standard_normal((2, 2))]
    path = tmp_path / "x.txt"
    jio.write_txt(path, mats)
    for x, y in zip(jio.read_txt(path), mats):
        np.testing.assert_array_equal(x, y)


def test_cli_help_version_and_errors(tmp_path):
    r = _run([CLI, "--help"])
    assert r.returncode == 0 and "--queue-capacity" in r.stdout and "<input-stem>.svd" in r.stdout
    r = _run([CLI, "--version"])
    assert r.returncode == 0 and r.stdout.strip() == "jacobi-svd-amd CLI v0.1.0"
    r = _run([CLI, "--max-sweeps", "0", "in.mat"])
    assert r.returncode == 2 and "Argument error: Option --max-sweeps must be positive." in r.stderr
    r = _run([CLI, str(tmp_path / "absent.mat")])
    assert r.returncode == 2 and "Input file does not exist" in r.stderr


def test_cli_dry_run_normalisation(tmp_path):
    src = tmp_path / "cases.txt"
    jio.write_txt(src, [np.eye(3, 2)])
    r = _run([CLI, "--dry-run", str(src), "-e", "1e-12", "--precision", "mixed", "--block-width", "16"])
    assert r.returncode == 0, r.stderr
    out = dict(line.split(":", 1) for line in r.stdout.strip().split("\n")[1:])
    assert out["output          "].strip() == str(tmp_path / "cases.svd.txt")
    assert out["output-format   "].strip() == "txt"
    assert out["epsilon         "].strip() == "1e-12"
    assert out["precision       "].strip() == "mixed"
    assert out["block-width     "].strip() == "16"
    (tmp_path / "cases.svd.txt").write_text("x")
    r = _run([CLI, "--dry-run", str(src)])
    assert r.returncode == 2 and "Use --force" in r.stderr
