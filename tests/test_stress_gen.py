"""CPU tests of the stress-mix generator (jsvd.gen.stress_matrix): each kind
has the structure the reference generator gives it
(psrc/generator/common.py:340-451), and streams are deterministic."""
import numpy as np
import pytest

from jsvd import gen


@pytest.mark.parametrize("kind", gen.STRESS_KINDS)
def test_kinds_are_deterministic_and_shaped(kind):
    a = gen.stress_matrix(kind, 40, 24, np.random.default_rng(3))
    b = gen.stress_matrix(kind, 40, 24, np.random.default_rng(3))
    assert a.shape == (40, 24) and np.isfinite(a).all()
    np.testing.assert_array_equal(a, b)


def test_kind_structure():
    rng = np.random.default_rng(1)
    z = gen.stress_matrix("zero-columns", 30, 20, rng)
    stride = max(2, 20 // 4)
    assert (z[:, ::stride] == 0).all() and (np.abs(z[:, 1]) > 0).any()
    d = gen.stress_matrix("diagonal-dominant", 30, 20, rng)
    off = d.copy()
    off[np.arange(20), np.arange(20)] = 0
    assert (np.abs(np.diag(d[:20, :20])) >= 1.0).all() and np.abs(off).max() <= 0.01
    s = gen.stress_matrix("sparse", 200, 100, rng, sparsity=0.5)
    assert 0.45 < np.mean(s == 0) < 0.55
    lr = gen.stress_matrix("low-rank", 60, 40, rng, rank=4)
    sv = np.linalg.svd(lr, compute_uv=False)
    assert sv[4] / sv[0] < 1e-5 and sv[3] / sv[0] > 1e-3
    ic = gen.stress_matrix("ill-conditioned", 60, 40, rng, rank=6, condition=1e8)
    sv = np.linalg.svd(ic, compute_uv=False)
    assert 1e6 < sv[0] / sv[5] < 1e10 and sv[6] / sv[0] < 1e-10
    lu = gen.stress_matrix("log-uniform", 50, 50, rng, lo=1e-3, hi=1e3)
    core = lu - 0.05 * np.sin(0.013 * np.arange(1, 51)[:, None] * np.arange(1, 51)[None, :])
    assert np.abs(core).min() >= 1e-3 * (1 - 1e-12) and np.abs(core).max() <= 1e3 * (1 + 1e-12)


def test_mix_cycles_kinds_and_shapes():
    mats, kinds = gen.stress_mix(24, [(12, 8), (40, 33)], seed=9)
    assert kinds[:8] == list(gen.STRESS_KINDS) and kinds[8:16] == list(gen.STRESS_KINDS)
    assert mats[0].shape == (12, 8) and mats[8].shape == (40, 33) and mats[16].shape == (12, 8)
