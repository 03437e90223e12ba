"""MI355X one-sided Jacobi SVD: Python front end of libjsvd_amd.so.

Host-array entry points mirror ``jacobi::svd::one_sided_jacobi_svd``
(reference include/jacobi/svd/domain/jacobi_svd.cuh:25-28): row-major m x n
input with m >= n, results U (m x n), sigma (n, unsorted, final column
order), V (n x n) and the sweep count.  ``BatchPlan`` runs device-resident
packed batches on the caller's HIP stream.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import (JsvdConfig, JsvdError, JsvdInvalidArgument, LAYOUT_AUTO, LAYOUT_OFF, LAYOUT_ON,
                   PRECISION_FP64, PRECISION_MIXED, STOP_CONVERGED, STOP_MAX_SWEEPS, STOP_BARRIER_TIMEOUT, default_config, load)

__all__ = [
    "JsvdConfig", "JsvdError", "JsvdInvalidArgument", "PRECISION_FP64", "PRECISION_MIXED", "LAYOUT_AUTO",
    "STOP_CONVERGED", "STOP_MAX_SWEEPS", "STOP_BARRIER_TIMEOUT",
    "LAYOUT_ON", "LAYOUT_OFF", "default_config", "load", "svd", "svd_batched", "BatchPlan", "pack_offsets", "blocked_layout",
    "autotune_batch", "tune_class", "solve_device_noalloc",
]

_dp = ctypes.POINTER(ctypes.c_double)


def _as_cfg(cfg: JsvdConfig | None, overrides) -> JsvdConfig:
    if cfg is None:
        return default_config(**overrides)
    for k, v in overrides.items():
        setattr(cfg, k, v)
    return cfg


def svd(a, cfg: JsvdConfig | None = None, **overrides):
    """One solve from host memory.  Returns (U, sigma, V, sweeps)."""
    lib = load()
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.ndim != 2:
        raise ValueError("input must be a 2-D array")
    m, n = a.shape
    c = _as_cfg(cfg, overrides)
    u = np.zeros((m, n))
    s = np.zeros((n,))
    v = np.zeros((n, n))
    sw = ctypes.c_int32()
    _lib.check(lib.jsvd_solve_f64(a.ctypes.data_as(_dp), m, n, ctypes.byref(c), u.ctypes.data_as(_dp),
                                  s.ctypes.data_as(_dp), v.ctypes.data_as(_dp), ctypes.byref(sw)))
    return u, s, v, sw.value


def svd_batched(mats, cfg: JsvdConfig | None = None, with_stop: bool = False, **overrides):
    """Many independent solves from host memory in one work-list launch.
    Returns [(U, sigma, V, sweeps)], or with_stop=True [(U, sigma, V, sweeps,
    stop)] with stop = STOP_CONVERGED or STOP_MAX_SWEEPS."""
    lib = load()
    mats = [np.ascontiguousarray(x, dtype=np.float64) for x in mats]
    count = len(mats)
    c = _as_cfg(cfg, overrides)
    outs = [(np.zeros(x.shape), np.zeros(x.shape[1]), np.zeros((x.shape[1], x.shape[1]))) for x in mats]
    arr = lambda xs: (_dp * count)(*[x.ctypes.data_as(_dp) for x in xs])
    ms = (ctypes.c_uint64 * count)(*[x.shape[0] for x in mats])
    ns = (ctypes.c_uint64 * count)(*[x.shape[1] for x in mats])
    sw = (ctypes.c_int32 * count)()
    stop = (ctypes.c_int32 * count)()
    _lib.check(lib.jsvd_solve_batched_status_f64(count, arr(mats), ms, ns, ctypes.byref(c),
                                                 arr([o[0] for o in outs]), arr([o[1] for o in outs]),
                                                 arr([o[2] for o in outs]), sw, stop))
    if with_stop:
        return [(o[0], o[1], o[2], int(sw[k]), int(stop[k])) for k, o in enumerate(outs)]
    return [(o[0], o[1], o[2], int(sw[k])) for k, o in enumerate(outs)]


def pack_offsets(ms, ns):
    """Element offsets of the packed layout in include/jsvd_c.h: (a_off, v_off, s_off, totals).

    Computed by the library (``jsvd_packed_layout``, 64-bit and overflow
    checked, the same code the plans use); needs no device.
    """
    lib = load()
    ms = np.ascontiguousarray(ms, dtype=np.int64)
    ns = np.ascontiguousarray(ns, dtype=np.int64)
    count = len(ms)
    if len(ns) != count:
        raise ValueError("ms and ns must have the same length")
    a_off, v_off, s_off = (np.zeros(count, dtype=np.int64) for _ in range(3))
    totals = np.zeros(3, dtype=np.int64)
    p = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    _lib.check(lib.jsvd_packed_layout(count, p(ms), p(ns), p(a_off), p(v_off), p(s_off), p(totals)))
    return a_off, v_off, s_off, (int(totals[0]), int(totals[1]), int(totals[2]))


def blocked_layout(m, n, block_width):
    """Padded shape and workspace sizes of one matrix on the blocked path
    (``jsvd_blocked_layout_f64``); needs no device."""
    lib = load()
    out = _lib.JsvdBlockedLayout()
    _lib.check(lib.jsvd_blocked_layout_f64(int(m), int(n), int(block_width), ctypes.byref(out)))
    return {name: int(getattr(out, name)) for name, _ in out._fields_}


class BatchPlan:
    """Device work-list for a fixed batch of shapes (``jsvd_plan``).

    ``execute`` takes torch CUDA tensors (float64, packed layout) and runs on
    the current torch stream without allocating.
    """

    def __init__(self, ms, ns, cfg: JsvdConfig | None = None, **overrides):
        lib = load()
        self.ms = np.asarray(ms, dtype=np.int64)
        self.ns = np.asarray(ns, dtype=np.int64)
        self.cfg = _as_cfg(cfg, overrides)
        self._handle = ctypes.c_void_p()
        _lib.check(lib.jsvd_plan_create(len(self.ms), self.ms.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                        self.ns.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                        ctypes.byref(self.cfg), ctypes.byref(self._handle)))
        self.a_off, self.v_off, self.s_off, (self.a_elems, self.v_elems, self.s_elems) = pack_offsets(
            self.ms, self.ns)

    @property
    def info(self) -> dict:
        info = _lib.JsvdPlanInfo()
        _lib.check(load().jsvd_plan_get_info(self._handle, ctypes.byref(info)))
        return {k: getattr(info, k) for k, _ in info._fields_}

    def execute(self, a, u, s, v, sweeps=None, stream=None, stop=None):
        """Solve on device tensors; ``stop`` (optional int32 CUDA tensor, one
        slot per matrix) receives STOP_CONVERGED or STOP_MAX_SWEEPS."""
        import torch

        for t, nel in ((a, self.a_elems), (u, self.a_elems), (s, self.s_elems), (v, self.v_elems)):
            if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous() and t.numel() >= nel):
                raise ValueError("BatchPlan.execute needs contiguous float64 CUDA tensors of the packed sizes")
        if sweeps is not None and not (sweeps.is_cuda and sweeps.dtype == torch.int32
                                       and sweeps.numel() >= len(self.ms)):
            raise ValueError("sweeps must be an int32 CUDA tensor with one slot per matrix")
        if stop is not None and not (stop.is_cuda and stop.dtype == torch.int32 and stop.numel() >= len(self.ms)):
            raise ValueError("stop must be an int32 CUDA tensor with one slot per matrix")
        st = stream if stream is not None else torch.cuda.current_stream()
        ptr = lambda t: t.data_ptr() if t is not None else None
        _lib.check(load().jsvd_plan_execute_status_f64(self._handle, a.data_ptr(), u.data_ptr(), s.data_ptr(),
                                                       v.data_ptr(), ptr(sweeps), ptr(stop), st.cuda_stream))

    def metrics(self, a, u, s, v, out=None, stream=None):
        """Per-matrix error metrics on the device (``jsvd_plan_metrics_f64``):
        a float64 CUDA tensor of shape (count, 3) holding ||U^T U - D||_F
        (D = diag(sigma > epsilon)), ||V^T V - I||_F and the relative
        residual ||A - U S V^T||_F / ||A||_F of every matrix."""
        import torch

        for t, nel in ((a, self.a_elems), (u, self.a_elems), (s, self.s_elems), (v, self.v_elems)):
            if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous() and t.numel() >= nel):
                raise ValueError("BatchPlan.metrics needs contiguous float64 CUDA tensors of the packed sizes")
        if out is None:
            out = torch.empty((len(self.ms), 3), dtype=torch.float64, device=a.device)
        elif not (out.is_cuda and out.dtype == torch.float64 and out.is_contiguous() and out.numel() >= 3 * len(self.ms)):
            raise ValueError("out must be a contiguous float64 CUDA tensor with three slots per matrix")
        st = stream if stream is not None else torch.cuda.current_stream()
        _lib.check(load().jsvd_plan_metrics_f64(self._handle, a.data_ptr(), u.data_ptr(), s.data_ptr(), v.data_ptr(),
                                                out.data_ptr(), st.cuda_stream))
        return out

    def backward(self, u, s, v, gu, gs, gv, ga, stream=None):
        """SVD adjoint on device tensors of the packed layout (float64 or float32,
        all the same dtype); gu / gs / gv may be None (zero gradient)."""
        import torch

        dt = u.dtype
        if dt not in (torch.float64, torch.float32):
            raise ValueError("backward supports float64 and float32")
        if ga is None:
            raise ValueError("backward needs an output tensor ga")
        # The kernel reads and writes the packed sizes: a short tensor would
        # be overrun on the device, so sizes are checked here as in execute().
        sizes = ((u, self.a_elems), (s, self.s_elems), (v, self.v_elems), (gu, self.a_elems),
                 (gs, self.s_elems), (gv, self.v_elems), (ga, self.a_elems))
        for t, nel in sizes:
            if t is not None and not (t.is_cuda and t.dtype == dt and t.is_contiguous()):
                raise ValueError("backward needs contiguous CUDA tensors of one dtype")
            if t is not None and t.numel() < nel:
                raise ValueError("backward needs tensors of the packed sizes (U, gU, gA: sum m*n; sigma, gsigma: "
                                 "sum n; V, gV: sum n*n)")
        ptr = lambda t: t.data_ptr() if t is not None else None
        st = stream if stream is not None else torch.cuda.current_stream()
        fn = load().jsvd_plan_backward_f64 if dt == torch.float64 else load().jsvd_plan_backward_f32
        _lib.check(fn(self._handle, ptr(u), ptr(s), ptr(v), ptr(gu), ptr(gs), ptr(gv), ptr(ga), st.cuda_stream))

    def close(self):
        if self._handle:
            load().jsvd_plan_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_device_noalloc(ms, ns, a, u, s, v, sweeps=None, cfg: JsvdConfig | None = None, stream=None,
                         **overrides):
    """``jsvd_solve_device_noalloc_f64``: the solve of a BatchPlan of these
    shapes in one call that allocates nothing (static device work-lists; at
    most 32768 matrices that fit the LDS-resident kernel, FP64).  Tensors as
    in BatchPlan.execute; runs on the current torch stream."""
    import torch

    ms = np.ascontiguousarray(ms, dtype=np.int64)
    ns = np.ascontiguousarray(ns, dtype=np.int64)
    _, _, _, (a_elems, v_elems, s_elems) = pack_offsets(ms, ns)
    for t, nel in ((a, a_elems), (u, a_elems), (s, s_elems), (v, v_elems)):
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous() and t.numel() >= nel):
            raise ValueError("solve_device_noalloc needs contiguous float64 CUDA tensors of the packed sizes")
    if sweeps is not None and not (sweeps.is_cuda and sweeps.dtype == torch.int32 and sweeps.numel() >= len(ms)):
        raise ValueError("sweeps must be an int32 CUDA tensor with one slot per matrix")
    c = _as_cfg(cfg, overrides)
    st = stream if stream is not None else torch.cuda.current_stream()
    _lib.check(load().jsvd_solve_device_noalloc_f64(
        len(ms), ms.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), ns.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
        ctypes.byref(c), a.data_ptr(), u.data_ptr(), s.data_ptr(), v.data_ptr(),
        sweeps.data_ptr() if sweeps is not None else None, st.cuda_stream))


def tune_class(bucket: int, cls: int) -> int:
    """Set the shape class of a column-count bucket (n <= 8, 16, 32, 48, 64,
    96, 128: buckets 0..6; -1 = default) for plans created afterwards;
    returns the previous setting."""
    prev = load().jsvd_tune_class(bucket, cls)
    if prev == -2:
        raise JsvdInvalidArgument(_lib.JSVD_EINVAL, f"bucket {bucket} is out of range")
    return prev


def autotune_batch(ms, ns, a, u, s, v, reps: int = 3, cfg: JsvdConfig | None = None, stream=None, **overrides):
    """Time whole plan executions of this batch on device tensors of the
    packed layout (``jsvd_autotune_batch``) and keep, per column-count
    bucket, the fastest shape class; plans created afterwards use the table.
    Returns {"classes": [...], "samples": k, "ms_initial": t0, "ms_tuned": t1}."""
    import torch

    ms = np.asarray(ms, dtype=np.int64)
    ns = np.asarray(ns, dtype=np.int64)
    _, _, _, (a_elems, v_elems, s_elems) = pack_offsets(ms, ns)
    for t, nel in ((a, a_elems), (u, a_elems), (s, s_elems), (v, v_elems)):
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous() and t.numel() >= nel):
            raise ValueError("autotune_batch needs contiguous float64 CUDA tensors of the packed sizes")
    c = _as_cfg(cfg, overrides)
    st = stream if stream is not None else torch.cuda.current_stream()
    rep = _lib.JsvdBatchTuneReport()
    _lib.check(load().jsvd_autotune_batch(len(ms), ms.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                          ns.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), ctypes.byref(c),
                                          a.data_ptr(), u.data_ptr(), s.data_ptr(), v.data_ptr(), int(reps),
                                          st.cuda_stream, ctypes.byref(rep)))
    return {"classes": list(rep.classes), "samples": int(rep.samples), "ms_initial": float(rep.ms_initial),
            "ms_tuned": float(rep.ms_tuned)}
