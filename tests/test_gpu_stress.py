"""GPU tests on the stress mix (jsvd.gen.stress_mix): the reference
generator's scenario kinds (dense uniform / normal / log-uniform,
diagonal-dominant, low-rank, ill-conditioned, sparse, zero columns) through
every solver path -- LDS work-list classes, the ring form of the wave
kernels (a lone 300 x 150 runs eight-wave workgroups with its column pairs
in registers), the grouped barrier form (JSVD_WAVE_RING=0: four-wave
workgroups with rows in registers) and the blocked path -- against the
oracle.

Bar: sigma sorted within 1e-12 of the oracle's largest singular value;
residual ||A - U S V^T|| / ||A|| <= 1e-12; ||V^T V - I||_F <= 1e-11; U
orthonormal over the singular values above 1e-6 sigma_max (the directions of
the noise-level singular values of the rank-deficient kinds are not
determined); zero columns keep sigma = 0 exactly and a zero U column, as in
the reference (src/infrastructure/jacobi_rotation_kernels.cuh:381-412).
"""
import sys

import numpy as np
import pytest

import jsvd
import oracle
from jsvd import gen, metrics

pytestmark = pytest.mark.gpu

EPS, SWEEPS = 1e-13, 30
LDS_SHAPES = [(12, 8), (40, 33), (64, 64), (100, 60), (24, 17)]


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _check(a, got, ref, kind):
    u, s, v, sw = got
    top = max(ref[1].max(), 1e-300)
    assert np.max(np.abs(np.sort(s) - np.sort(ref[1]))) <= 1e-12 * top, kind
    assert metrics.residual(a, u, s, v) <= 1e-12, kind
    assert metrics.orthogonality(v) <= 1e-11, kind
    assert metrics.orthogonality(u, s > 1e-6 * top) <= 1e-9, kind
    assert 1 <= sw <= SWEEPS, kind
    if kind == "zero-columns":
        zero = ~a.any(axis=0)
        assert (s[zero] == 0).all() and (u[:, zero] == 0).all(), kind


def test_stress_mix_lds_classes():
    mats, kinds = gen.stress_mix(8 * len(LDS_SHAPES) * 2, LDS_SHAPES, seed=77)
    _note(f"stress mix: {len(mats)} matrices through the work-list kernels")
    got = jsvd.svd_batched(mats, epsilon=EPS, max_sweeps=SWEEPS)
    refs = oracle.solve_batch(mats, EPS, SWEEPS)
    for a, g, r, kind in zip(mats, got, refs, kinds):
        _check(a, g, r, kind)


@pytest.mark.parametrize("bw,ring", [(0, "1"), (0, "0"), (16, "1")], ids=["ring", "grouped-barrier", "blocked16"])
def test_stress_mix_large_paths(bw, ring, monkeypatch):
    monkeypatch.setenv("JSVD_WAVE_RING", ring)
    mats, kinds = gen.stress_mix(8, [(300, 150)], seed=91)
    _note(f"stress mix 300x150, block_width={bw}, ring={ring}")
    refs = oracle.solve_batch(mats, EPS, SWEEPS)
    for a, r, kind in zip(mats, refs, kinds):
        g = jsvd.svd(a, epsilon=EPS, max_sweeps=SWEEPS, block_width=bw)
        _check(a, g, r, kind)
