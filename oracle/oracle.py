"""ctypes front end of the CPU oracle (oracle/jsvd_oracle.c).

TEST INFRASTRUCTURE ONLY.  The oracle restates the reference one-sided Jacobi
SVD (kleedaisuki/jacobi-svd-cuda-opt, src/domain/jacobi_svd.cu and
src/domain/jacobi_rotation_kernels.cuh) on the CPU in FP64.  tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg use it as the checker
and as the CPU baseline; the product never imports it.

Pinning (tests/test_oracle.py): every function here is checked against
LAPACK (numpy gesdd, scipy dgejsv), mpmath at 50 digits, closed forms
(2x2 analytic SVD, diagonal and orthogonal-column inputs, zero columns,
n = 1) and brute-force schedule invariants.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libjsvd_oracle.so")
_lib = None
_lock = threading.Lock()

_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


def build() -> str:
    """Compile the oracle with its Makefile (gcc; no GPU involved)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(_LIB_PATH):
                build()
            L = ctypes.CDLL(_LIB_PATH)
            L.jsvd_oracle_column_at.restype = ctypes.c_int32
            L.jsvd_oracle_column_at.argtypes = [ctypes.c_int32] * 4
            L.jsvd_oracle_schedule.restype = ctypes.c_int32
            L.jsvd_oracle_schedule.argtypes = [ctypes.c_int32, _i32p, _i32p]
            L.jsvd_oracle_rotation.restype = ctypes.c_int32
            L.jsvd_oracle_rotation.argtypes = [ctypes.c_double] * 4 + [_dp, _dp]
            L.jsvd_oracle_validate.restype = ctypes.c_int32
            L.jsvd_oracle_validate.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                               ctypes.c_double, ctypes.c_int32,
                                               ctypes.POINTER(ctypes.c_char_p)]
            L.jsvd_oracle_solve_f64.restype = ctypes.c_int32
            L.jsvd_oracle_solve_f64.argtypes = [_dp, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double,
                                                ctypes.c_int32, ctypes.c_int32, _dp, _dp, _dp, _i32p]
            L.jsvd_oracle_solve_batch_f64.restype = ctypes.c_int32
            L.jsvd_oracle_solve_batch_f64.argtypes = [ctypes.c_int64, _dp, _i64p, _i64p, _i64p,
                                                      ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                                                      _dp, _i64p, _dp, _i64p, _dp, _i64p, _i32p]
            _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(t)


def column_at(position: int, rnd: int, n: int) -> int:
    even = n + (n & 1)
    return lib().jsvd_oracle_column_at(position, rnd, n, even)


def schedule(n: int):
    """List of rounds, each a list of (p, q) pairs (reference host table order)."""
    even = n + (n & 1)
    if n < 2:
        return []
    pairs = np.zeros(((even - 1) * (even // 2) * 2,), dtype=np.int32)
    counts = np.zeros((even - 1,), dtype=np.int32)
    rounds = lib().jsvd_oracle_schedule(n, _ptr(pairs, _i32p), _ptr(counts, _i32p))
    out = []
    half = even // 2
    for r in range(rounds):
        base = r * half * 2
        out.append([(int(pairs[base + 2 * k]), int(pairs[base + 2 * k + 1])) for k in range(counts[r])])
    return out


def rotation(pp: float, qq: float, pq: float, eps: float):
    c = ctypes.c_double()
    s = ctypes.c_double()
    flag = lib().jsvd_oracle_rotation(pp, qq, pq, eps, ctypes.byref(c), ctypes.byref(s))
    return bool(flag), c.value, s.value


def validate(rows, cols, size, eps, max_sweeps):
    msg = ctypes.c_char_p()
    st = lib().jsvd_oracle_validate(rows, cols, size, eps, max_sweeps, ctypes.byref(msg))
    return st, (msg.value.decode() if msg.value else None)


def solve(a, eps: float = 1e-9, max_sweeps: int = 128, threads: int = 1):
    """FP64 one-sided Jacobi of a row-major m x n array.  Returns (U, sigma, V, sweeps)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    m, n = a.shape
    u = np.zeros((m, n))
    s = np.zeros((n,))
    v = np.zeros((n, n))
    sw = ctypes.c_int32()
    st = lib().jsvd_oracle_solve_f64(_ptr(a, _dp), m, n, eps, max_sweeps, threads, _ptr(u, _dp),
                                     _ptr(s, _dp), _ptr(v, _dp), ctypes.byref(sw))
    if st != 0:
        raise ValueError(validate(m, n, m * n, eps, max_sweeps)[1])
    return u, s, v, sw.value


def solve_batch(mats, eps: float = 1e-9, max_sweeps: int = 128, threads: int = 0):
    """Solve a list of row-major matrices, one per OpenMP thread."""
    if threads <= 0:
        threads = os.cpu_count() or 1
    count = len(mats)
    ms = np.array([x.shape[0] for x in mats], dtype=np.int64)
    ns = np.array([x.shape[1] for x in mats], dtype=np.int64)
    a_off = np.concatenate([[0], np.cumsum(ms * ns)[:-1]]).astype(np.int64)
    v_off = np.concatenate([[0], np.cumsum(ns * ns)[:-1]]).astype(np.int64)
    s_off = np.concatenate([[0], np.cumsum(ns)[:-1]]).astype(np.int64)
    a_all = np.concatenate([np.ascontiguousarray(x, dtype=np.float64).ravel() for x in mats])
    u_all = np.zeros_like(a_all)
    s_all = np.zeros(int(ns.sum()))
    v_all = np.zeros(int((ns * ns).sum()))
    sw = np.zeros(count, dtype=np.int32)
    st = lib().jsvd_oracle_solve_batch_f64(count, _ptr(a_all, _dp), _ptr(a_off, _i64p), _ptr(ms, _i64p),
                                           _ptr(ns, _i64p), eps, max_sweeps, threads, _ptr(u_all, _dp),
                                           _ptr(a_off, _i64p), _ptr(s_all, _dp), _ptr(s_off, _i64p),
                                           _ptr(v_all, _dp), _ptr(v_off, _i64p), _ptr(sw, _i32p))
    if st != 0:
        raise ValueError("invalid batch input")
    out = []
    for k in range(count):
        m, n = int(ms[k]), int(ns[k])
        out.append((u_all[a_off[k]:a_off[k] + m * n].reshape(m, n),
                    s_all[s_off[k]:s_off[k] + n].copy(),
                    v_all[v_off[k]:v_off[k] + n * n].reshape(n, n),
                    int(sw[k])))
    return out
