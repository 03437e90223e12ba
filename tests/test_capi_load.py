"""CPU checks of the C-ABI library: it loads, exports every symbol that
include/jsvd_c.h declares and the reference's C++ API, reports the
reference defaults, and rejects invalid arguments with the reference's
messages before touching a device.  No compute call is made here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import jsvd
from jsvd import _lib

HEADER = _lib.HEADER_PATH


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(jsvd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = jsvd.load()
    names = declared_functions()
    assert len(names) >= 10
    for name in names:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"


def test_cxx_api_symbols_exported():
    out = subprocess.run(["nm", "-D", "--defined-only", "-C", _lib.LIB_PATH], check=True,
                         capture_output=True, text=True).stdout
    for sym in ["jacobi::svd::one_sided_jacobi_svd(", "jacobi::svd::auto_tune_layout_transpose_threshold(",
                "jacobi::svd::DeviceMatrix::DeviceMatrix(", "jacobi::svd::CudaError::CudaError(",
                "jacobi::svd::DeviceMatrix::copy_to_host() const"]:
        assert sym in out, sym


def test_config_defaults_match_reference():
    # include/jacobi/svd/domain/jacobi_svd_config.cuh:17-57 of the reference.
    c = jsvd.default_config()
    assert c.epsilon == 1e-9 and c.max_sweeps == 128 and c.threads_per_block == 256
    assert c.layout_mode == jsvd.LAYOUT_AUTO and c.layout_min_columns == 16 and c.layout_min_elements == 4096
    assert c.layout_auto_tune == 0 and c.layout_bench_reps == 2 and c.layout_bench_sweeps == 8
    assert c.precision == jsvd.PRECISION_FP64 and c.block_width == 0


@pytest.mark.parametrize("shape,over,msg", [
    ((0, 3), {}, "Input matrix shape must be non-zero."),
    ((2, 3), {}, "One-sided Jacobi SVD requires rows >= columns in this implementation."),
    ((3, 2), {"epsilon": 0.0}, "epsilon must be positive."),
    ((3, 2), {"epsilon": -1.0}, "epsilon must be positive."),
    ((3, 2), {"max_sweeps": 0}, "max_sweeps must be positive."),
    ((3, 2), {"layout_min_columns": 0}, "layout_transpose_min_columns must be positive."),
    ((3, 2), {"layout_min_elements": 0}, "layout_transpose_min_elements must be positive."),
    ((3, 2), {"layout_bench_reps": 0}, "layout_transpose_benchmark_repetitions must be positive."),
    ((3, 2), {"layout_bench_sweeps": 0}, "layout_transpose_benchmark_sweeps must be positive."),
])
def test_validation_before_device(shape, over, msg):
    with pytest.raises(jsvd.JsvdInvalidArgument) as e:
        jsvd.svd(np.zeros(shape), **over)
    assert str(e.value) == msg


def test_autotune_batch_rejects_bad_arguments_before_device():
    lib = jsvd.load()
    rep = _lib.JsvdBatchTuneReport()
    cfg = jsvd.default_config()
    m = (ctypes.c_int64 * 1)(8)
    n = (ctypes.c_int64 * 1)(4)
    # count 0, null buffers and reps out of range are refused without a device call.
    assert lib.jsvd_autotune_batch(0, m, n, ctypes.byref(cfg), 1, 1, 1, 1, 1, None, ctypes.byref(rep)) == 1
    assert lib.jsvd_autotune_batch(1, m, n, ctypes.byref(cfg), None, 1, 1, 1, 1, None, ctypes.byref(rep)) == 1
    assert lib.jsvd_autotune_batch(1, m, n, ctypes.byref(cfg), 1, 1, 1, 1, 0, None, ctypes.byref(rep)) == 1
    bad = (ctypes.c_int64 * 1)(2)
    assert lib.jsvd_autotune_batch(1, bad, n, ctypes.byref(cfg), 1, 1, 1, 1, 1, None, ctypes.byref(rep)) == 1
    assert b"rows >= columns" in lib.jsvd_last_error()


def test_tune_class_table_round_trip():
    assert jsvd.tune_class(3, 4) == -1
    assert jsvd.tune_class(3, -1) == 4
    with pytest.raises(jsvd.JsvdInvalidArgument):
        jsvd.tune_class(7, 0)
