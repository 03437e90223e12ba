"""ctypes binding of libjsvd_amd.so (the C-ABI declared in include/jsvd_c.h).

The library is built in-tree (``make -C jacobi-svd-cuda-opt_amd``).  Loading
fails loudly when it is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os
import threading

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "lib", "libjsvd_amd.so")
HEADER_PATH = os.path.join(REPO_DIR, "include", "jsvd_c.h")

JSVD_OK, JSVD_EINVAL, JSVD_EDEVICE, JSVD_ENOMEM = 0, 1, 2, 3
PRECISION_FP64, PRECISION_MIXED = 0, 1
STOP_CONVERGED, STOP_MAX_SWEEPS, STOP_BARRIER_TIMEOUT = 0, 1, 2
LAYOUT_AUTO, LAYOUT_ON, LAYOUT_OFF = 0, 1, 2


class JsvdConfig(ctypes.Structure):
    """Mirror of ``jsvd_config`` (include/jsvd_c.h)."""

    _fields_ = [
        ("epsilon", ctypes.c_double),
        ("max_sweeps", ctypes.c_int32),
        ("threads_per_block", ctypes.c_int32),
        ("layout_mode", ctypes.c_int32),
        ("layout_min_columns", ctypes.c_int32),
        ("layout_min_elements", ctypes.c_uint64),
        ("layout_auto_tune", ctypes.c_int32),
        ("layout_bench_reps", ctypes.c_int32),
        ("layout_bench_sweeps", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("block_width", ctypes.c_int32),
        ("refine_sweeps", ctypes.c_int32),
    ]


class JsvdPlanInfo(ctypes.Structure):
    _fields_ = [
        ("count", ctypes.c_int64),
        ("lds_matrices", ctypes.c_int64),
        ("large_matrices", ctypes.c_int64),
        ("a_elements", ctypes.c_int64),
        ("v_elements", ctypes.c_int64),
        ("sigma_elements", ctypes.c_int64),
        ("workgroup_size", ctypes.c_int32),
        ("grid_size", ctypes.c_int32),
        ("lds_bytes", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("buckets", ctypes.c_int32),
        ("blocked_matrices", ctypes.c_int32),
        ("ring_matrices", ctypes.c_int32),
        ("pad", ctypes.c_int32),
    ]


class JsvdAutotuneReport(ctypes.Structure):
    _fields_ = [
        ("executed", ctypes.c_int32),
        ("recommended_min_columns", ctypes.c_int32),
        ("recommended_min_elements", ctypes.c_uint64),
        ("estimated_best_speedup", ctypes.c_double),
        ("sample_count", ctypes.c_uint64),
    ]


TUNE_BUCKETS = 7


class JsvdBatchTuneReport(ctypes.Structure):
    _fields_ = [
        ("classes", ctypes.c_int32 * TUNE_BUCKETS),
        ("samples", ctypes.c_int32),
        ("ms_initial", ctypes.c_double),
        ("ms_tuned", ctypes.c_double),
    ]


class JsvdBlockedLayout(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int64) for name in ("lda", "nv", "blocks", "pairs", "chunk_rows", "chunks",
                                                    "a_elements", "v_elements", "gram_elements")]


class JsvdError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


class JsvdInvalidArgument(JsvdError, ValueError):
    pass


_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_vp = ctypes.c_void_p

# name -> (restype, argtypes).  Every symbol declared in include/jsvd_c.h.
SIGNATURES = {
    "jsvd_config_default": (None, [ctypes.POINTER(JsvdConfig)]),
    "jsvd_last_error": (ctypes.c_char_p, []),
    "jsvd_abi_version": (ctypes.c_int32, []),
    "jsvd_solve_f64": (ctypes.c_int, [_dp, ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(JsvdConfig), _dp, _dp,
                                      _dp, _i32p]),
    "jsvd_solve_batched_f64": (ctypes.c_int, [ctypes.c_int64, ctypes.POINTER(_dp), _u64p, _u64p,
                                              ctypes.POINTER(JsvdConfig), ctypes.POINTER(_dp),
                                              ctypes.POINTER(_dp), ctypes.POINTER(_dp), _i32p]),
    "jsvd_solve_batched_status_f64": (ctypes.c_int, [ctypes.c_int64, ctypes.POINTER(_dp), _u64p, _u64p,
                                                     ctypes.POINTER(JsvdConfig), ctypes.POINTER(_dp),
                                                     ctypes.POINTER(_dp), ctypes.POINTER(_dp), _i32p, _i32p]),
    "jsvd_plan_execute_status_f64": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "jsvd_packed_layout": (ctypes.c_int, [ctypes.c_int64, _i64p, _i64p, _i64p, _i64p, _i64p, _i64p]),
    "jsvd_blocked_layout_f64": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int32,
                                               ctypes.POINTER(JsvdBlockedLayout)]),
    "jsvd_plan_create": (ctypes.c_int, [ctypes.c_int64, _i64p, _i64p, ctypes.POINTER(JsvdConfig),
                                        ctypes.POINTER(_vp)]),
    "jsvd_plan_execute_f64": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "jsvd_plan_get_info": (ctypes.c_int, [_vp, ctypes.POINTER(JsvdPlanInfo)]),
    "jsvd_plan_metrics_f64": (ctypes.c_int, [_vp] * 7),
    "jsvd_plan_destroy": (None, [_vp]),
    "jsvd_solve_device_noalloc_f64": (ctypes.c_int, [ctypes.c_int64, _i64p, _i64p, ctypes.POINTER(JsvdConfig),
                                                     _vp, _vp, _vp, _vp, _vp, _vp]),
    "jsvd_plan_backward_f64": (ctypes.c_int, [_vp] * 9),
    "jsvd_plan_backward_f32": (ctypes.c_int, [_vp] * 9),
    "jsvd_autotune": (ctypes.c_int, [ctypes.POINTER(JsvdConfig), ctypes.POINTER(JsvdAutotuneReport)]),
    "jsvd_tune_class": (ctypes.c_int32, [ctypes.c_int32, ctypes.c_int32]),
    "jsvd_class_count": (ctypes.c_int32, []),
    "jsvd_autotune_batch": (ctypes.c_int, [ctypes.c_int64, _i64p, _i64p, ctypes.POINTER(JsvdConfig), _vp, _vp, _vp,
                                           _vp, ctypes.c_int32, _vp, ctypes.POINTER(JsvdBatchTuneReport)]),
}

_lib = None
_lock = threading.Lock()


def load(path: str | None = None):
    """Load libjsvd_amd.so and bind every C-ABI symbol; raises if it is absent."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        p = path or LIB_PATH
        # torch ships its own libamdhip64.so (soname libamdhip64.so.7).  Loading
        # torch first makes this library's libamdhip64.so.7 dependency resolve
        # to that copy, so one HIP runtime serves both and torch streams and
        # device pointers are valid here.  Two runtimes in one process cannot
        # share the device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(p):
            raise RuntimeError(
                f"libjsvd_amd.so not found at {p}: build it with `make -C jacobi-svd-cuda-opt_amd` "
                "(or __graft_entry__.build()); there is no CPU fallback")
        lib = ctypes.CDLL(p)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def check(status: int):
    if status == JSVD_OK:
        return
    msg = load().jsvd_last_error().decode(errors="replace")
    if status == JSVD_EINVAL:
        raise JsvdInvalidArgument(status, msg)
    raise JsvdError(status, msg)


def default_config(**overrides) -> JsvdConfig:
    cfg = JsvdConfig()
    load().jsvd_config_default(ctypes.byref(cfg))
    for k, v in overrides.items():
        if not hasattr(cfg, k):
            raise AttributeError(f"jsvd_config has no field {k}")
        setattr(cfg, k, v)
    return cfg
