"""GPU parity of the FP64 path against the CPU oracle, through the C-ABI.

Tolerances (FP64 Jacobi in the reference's pair order; the dot products are
reduced in a different order, so agreement is to rounding, not bitwise):
  * sigma, compared position by position (the unsorted column order is part
    of the result): max |s - s_oracle| <= 1e-12 * max(s_oracle);
  * ||U^T U - I||_F, ||V^T V - I||_F and ||A - U S V^T||_F / ||A||_F
    <= 1e-12 * (n / 64 + 1);
  * sweep counts equal the oracle's within one sweep (rounding can move the
    last threshold decision by one sweep).
At eps = 1e-15 (the rounding floor of the cosine test) a rotation with
zeta = (qq - pp) / (2 pq) within rounding of 0 takes t = +1 or t = -1
depending on the last bit, which decides which of the two columns grows:
the multiset of sigma is the same but two columns may trade places.  Batches
at that floor therefore compare sigma sorted for every matrix and require
the positional match for at least 90% of them.
"""
import os
import sys

import numpy as np
import pytest

import jsvd
import oracle
from jsvd import gen, metrics

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "golden_small.npz")


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _assert_parity(a, got, ref, eps, tag="", allow_permutation=False):
    """Checks got against the oracle's ref; returns True when sigma matched position by position."""
    u, s, v, sw = got
    uo, so, vo, swo = ref
    n = a.shape[1]
    tol = 1e-12 * (n / 64 + 1)
    scale = max(so.max(), 1e-300)
    positional = bool(np.max(np.abs(s - so)) <= 1e-12 * scale)
    if not allow_permutation:
        assert positional, f"{tag} sigma (positional) differs"
    assert metrics.sigma_errors(s, so)[0] <= 1e-12, f"{tag} sigma (sorted) differs"
    mask = s > eps
    assert metrics.orthogonality(u, mask) <= tol, f"{tag} U orthogonality"
    assert metrics.orthogonality(v) <= tol, f"{tag} V orthogonality"
    if np.linalg.norm(a) > 0:
        assert metrics.residual(a, u, s, v) <= tol, f"{tag} residual"
    assert abs(sw - swo) <= 1, f"{tag} sweeps {sw} vs oracle {swo}"
    return positional


def test_golden_fixtures():
    d = np.load(GOLDEN)
    for k in range(int(d["count"])):
        a = d[f"a_{k}"]
        _note(f"golden case {k} shape {a.shape}")
        got = jsvd.svd(a, epsilon=1e-15, max_sweeps=30)
        ref = (d[f"u_{k}"], d[f"s_{k}"], d[f"v_{k}"], int(d[f"sweeps_{k}"]))
        _assert_parity(a, got, ref, 1e-15, f"golden{k}")
        assert np.max(np.abs(np.sort(got[1])[::-1] - d[f"sigma_mp_{k}"])) <= 4e-15 * d[f"sigma_mp_{k}"][0]


@pytest.mark.parametrize("shape", [(2, 2), (3, 3), (16, 8), (17, 9), (64, 33), (100, 1), (128, 64), (96, 64),
                                   (250, 63), (40, 40), (96, 96), (128, 128), (160, 120), (129, 127)])
def test_lds_path_against_oracle(shape):
    # (128, 128), (160, 120) and (129, 127) take the rotation-log classes (A in
    # LDS, V rebuilt from the logged tangents); (96, 96) the 512-thread class.
    _note(f"lds path shape {shape}")
    rng = np.random.default_rng(100 + shape[0] + 1000 * shape[1])
    a = rng.standard_normal(shape)
    got = jsvd.svd(a, epsilon=1e-13, max_sweeps=30)
    ref = oracle.solve(a, 1e-13, 30)
    _assert_parity(a, got, ref, 1e-13, str(shape))


@pytest.mark.parametrize("shape,threads", [((256, 256), 256), ((300, 150), 256), ((300, 151), 512),
                                           ((5000, 40), 256), ((1024, 129), 1024), ((1100, 520), 256),
                                           ((600, 513), 1024)])
def test_persistent_path_against_oracle(shape, threads):
    # These shapes exceed the LDS budget, so each runs alone in a wave
    # kernel: 256 x 256, 300 x 150 and 300 x 151 in the ring form with
    # eight-wave workgroups; 1024 x 129 in the ring form with four-wave ones
    # (16 A rows per lane); 5000 x 40 in four-wave grouped workgroups with rows
    # streamed (5120 held); 600 x 513 eight-wave grouped with rows in
    # registers and 1100 x 520 eight-wave grouped streamed (more than 512
    # columns rule out the ring, more than 256 column pairs four-wave groups).
    # For the two wide Gaussian inputs the final column order is chaotic
    # under rounding: a one-ulp perturbation of the 1100 x 520 input moves
    # about 480 of its 520 sigma positions in the oracle itself (sorted sigma
    # agree to 1.5e-14), so they compare sigma sorted.  They run at
    # eps = 1e-14: at 1e-13, 520 columns each within eps of orthogonal sum to
    # ||U^T U - I||_F above the 9.1e-12 bound.
    wide = shape[1] > 512
    eps = 1e-14 if wide else 1e-13
    _note(f"persistent path shape {shape} threads {threads}")
    rng = np.random.default_rng(7 + shape[0] + shape[1])
    a = rng.standard_normal(shape)
    got = jsvd.svd(a, epsilon=eps, max_sweeps=30, threads_per_block=threads)
    ref = oracle.solve(a, eps, 30, threads=8)
    _assert_parity(a, got, ref, eps, str(shape), allow_permutation=wide)


def test_grouped_wave_batch_against_oracle_and_single_solves():
    # Matrices too large for LDS with at most 512 rows go to the ring form of
    # the wave-per-pair kernels, one launch per row-register count (the 121
    # matrices of up to 192 rows in groups that take several matrices in
    # turn).  The pair arithmetic does not depend on the launch, so every
    # matrix must equal its single solve bitwise, and the oracle to tolerance.
    _note("grouped wave batch")
    rng = np.random.default_rng(515)
    shapes = [(256, 256), (300, 150), (200, 180), (512, 100), (257, 200), (192, 192), (300, 151)]
    shapes += [(160, 150)] * 120 + [(64, 40), (30, 20), (128, 64)]
    mats = [rng.standard_normal(sh) for sh in shapes]
    ms = [a.shape[0] for a in mats]
    ns = [a.shape[1] for a in mats]
    plan = jsvd.BatchPlan(ms, ns, epsilon=1e-13, max_sweeps=30)
    assert plan.info["large_matrices"] == len(shapes) - 3
    plan.close()
    got = jsvd.svd_batched(mats, epsilon=1e-13, max_sweeps=30)
    ref = oracle.solve_batch(mats, 1e-13, 30)
    for k, (a, g, r) in enumerate(zip(mats, got, ref)):
        _assert_parity(a, g, r, 1e-13, f"grouped{k} {a.shape}")
    for k in list(range(7)) + [7, 70, 126, 128]:
        single = jsvd.svd(mats[k], epsilon=1e-13, max_sweeps=30)
        for x, y in zip(single[:3], got[k][:3]):
            np.testing.assert_array_equal(x, y)
        assert single[3] == got[k][3]


def test_tall_batch_streamed_against_oracle():
    # Plain-path matrices taller than 512 rows (and at most 512 columns) have
    # their rows streamed in 256-row chunks.  22 uneven ones (widest 129
    # columns) share one grouped launch; 100 of 1500 x 24 exceed the
    # 64 + 12^2 / 100 matrices from which one workgroup per matrix is chosen.
    # Both forms use the same pair arithmetic, so each of the 100 equals its
    # single (grouped) solve bitwise.  A single solve of at most 1024 rows
    # holds its rows in registers instead of streaming them; it rounds like
    # the streamed batch, so those singles equal their batch entries too.
    _note("tall batch, streamed rows")
    rng = np.random.default_rng(611)
    shapes = [(1024, 64)] * 16 + [(700, 90), (513, 77), (2000, 40), (600, 129), (1500, 100), (777, 31)]
    mats = [rng.standard_normal(sh) for sh in shapes]
    got = jsvd.svd_batched(mats, epsilon=1e-13, max_sweeps=30)
    ref = oracle.solve_batch(mats, 1e-13, 30)
    for k, (a, g, r) in enumerate(zip(mats, got, ref)):
        _assert_parity(a, g, r, 1e-13, f"tall{k} {a.shape}")
    for k in (0, 16, 17, 19):
        single = jsvd.svd(mats[k], epsilon=1e-13, max_sweeps=30)
        for x, y in zip(single[:3], got[k][:3]):
            np.testing.assert_array_equal(x, y, err_msg=f"single vs batch {mats[k].shape}")
        assert single[3] == got[k][3]
    # 1500 x 24 needs (36000 + 576 + 24) * 8 = 292,800 bytes, beyond the LDS
    # budget, and 100 of them exceed the 64 + 12^2 / 100 matrices from which
    # a streamed launch runs one workgroup per matrix (wave_block_stream_f64).
    mats = [rng.standard_normal((1500, 24)) for _ in range(100)]
    plan = jsvd.BatchPlan([1500] * 100, [24] * 100, epsilon=1e-13, max_sweeps=30)
    assert plan.info["large_matrices"] == 100 and plan.info["lds_matrices"] == 0
    plan.close()
    got = jsvd.svd_batched(mats, epsilon=1e-13, max_sweeps=30)
    ref = oracle.solve_batch(mats, 1e-13, 30)
    for k, (a, g, r) in enumerate(zip(mats, got, ref)):
        _assert_parity(a, g, r, 1e-13, f"tall block{k}")
    for k in (0, 99):
        single = jsvd.svd(mats[k], epsilon=1e-13, max_sweeps=30)
        for x, y in zip(single[:3], got[k][:3]):
            np.testing.assert_array_equal(x, y)
        assert single[3] == got[k][3]


@pytest.fixture
def barrier_forms(monkeypatch):
    """Plans built while this fixture is active use the barrier forms of the
    wave solves (grouped flag barriers, one workgroup per matrix) instead of
    the ring form."""
    monkeypatch.setenv("JSVD_WAVE_RING", "0")
    yield
    monkeypatch.delenv("JSVD_WAVE_RING")


def test_ring_form_equals_barrier_forms(barrier_forms):
    # The ring form (columns held in registers, moved through LDS and tagged
    # mailbox words) does the grouped kernel's pair arithmetic in the same
    # order, so every result must equal the barrier forms' bitwise: a lone
    # 256 x 256, a lone odd-width 300 x 151 (a dummy column in the circle),
    # a mixed batch whose 120 matrices of 160 x 150 the barrier path solves
    # one workgroup per matrix, 40 of 512 x 512 (eight rows per lane, fewer
    # ring groups than matrices, so groups take several in turn), and tall
    # ones of 513 - 1024 rows (four-wave workgroups, 16 A rows and up to
    # three V rows per lane; the barrier path streams their rows).
    _note("ring form vs barrier forms")
    rng = np.random.default_rng(4242)
    batches = [[rng.standard_normal((256, 256))], [rng.standard_normal((300, 151))],
               [rng.standard_normal(sh) for sh in [(256, 256), (300, 151), (512, 100)] + [(160, 150)] * 120],
               [gen.gaussian_square(512, seed=gen.BASE_SEED + 100 + k) for k in range(40)],
               [rng.standard_normal(sh) for sh in [(1024, 129), (777, 31)] + [(700, 90)] * 16]]
    for mats in batches:
        os.environ["JSVD_WAVE_RING"] = "1"
        plan = jsvd.BatchPlan([a.shape[0] for a in mats], [a.shape[1] for a in mats], epsilon=1e-14, max_sweeps=30)
        os.environ["JSVD_WAVE_RING"] = "0"
        assert plan.info["ring_matrices"] == plan.info["large_matrices"] == len(mats)
        plan.close()
        barrier = jsvd.svd_batched(mats, epsilon=1e-14, max_sweeps=30, with_stop=True)
        os.environ["JSVD_WAVE_RING"] = "1"
        try:
            ring = jsvd.svd_batched(mats, epsilon=1e-14, max_sweeps=30, with_stop=True)
        finally:
            os.environ["JSVD_WAVE_RING"] = "0"
        for k, (x, y) in enumerate(zip(ring, barrier)):
            for a, b in zip(x[:3], y[:3]):
                np.testing.assert_array_equal(a, b, err_msg=f"ring vs barrier {mats[k].shape} #{k}")
            assert x[3] == y[3] and x[4] == y[4] == jsvd.STOP_CONVERGED
    a = batches[1][0]
    _assert_parity(a, jsvd.svd(a, epsilon=1e-14, max_sweeps=30)[:4], oracle.solve(a, 1e-14, 30), 1e-14, "300x151")


def test_ring_xcd_placement_is_bitwise_neutral(monkeypatch):
    # The ring form places each group's workgroups on one XCD and, after
    # checking the placement, stores the mailbox words at workgroup scope
    # (they meet the readers in that XCD's L2).  Only the grid and the store
    # scope change, so the results must equal the group-major grid's
    # bitwise: 16 of 256 x 256 (two groups per XCD), a lone 300 x 151 and 24
    # tall ones (four-wave groups).
    _note("ring form, XCD placement vs group-major grid")
    rng = np.random.default_rng(777)
    batches = [[rng.standard_normal((256, 256)) for _ in range(16)], [rng.standard_normal((300, 151))],
               [rng.standard_normal((700, 90)) for _ in range(24)]]
    for mats in batches:
        monkeypatch.setenv("JSVD_WAVE_RING_XCD", "1")
        spread = jsvd.svd_batched(mats, epsilon=1e-14, max_sweeps=30, with_stop=True)
        monkeypatch.setenv("JSVD_WAVE_RING_XCD", "0")
        major = jsvd.svd_batched(mats, epsilon=1e-14, max_sweeps=30, with_stop=True)
        for k, (x, y) in enumerate(zip(spread, major)):
            for a, b in zip(x[:3], y[:3]):
                np.testing.assert_array_equal(a, b, err_msg=f"xcd vs group-major {mats[k].shape} #{k}")
            assert x[3] == y[3] and x[4] == y[4] == jsvd.STOP_CONVERGED


def test_ring_slot_sets_are_bitwise_neutral(monkeypatch):
    # With room for twice its LDS, the ring form uses two slot sets by round
    # parity and one workgroup barrier per round instead of two.  Only the
    # synchronisation changes, so the results must equal the one-set form's
    # bitwise: a lone 256 x 256, 16 of them (one workgroup per CU either
    # way) and a lone tall 700 x 90 (four-wave groups).
    _note("ring form, two slot sets vs one")
    rng = np.random.default_rng(778)
    batches = [[rng.standard_normal((256, 256))], [rng.standard_normal((256, 256)) for _ in range(16)],
               [rng.standard_normal((700, 90))]]
    for mats in batches:
        monkeypatch.setenv("JSVD_WAVE_RING_DB", "1")
        two = jsvd.svd_batched(mats, epsilon=1e-14, max_sweeps=30, with_stop=True)
        monkeypatch.setenv("JSVD_WAVE_RING_DB", "0")
        one = jsvd.svd_batched(mats, epsilon=1e-14, max_sweeps=30, with_stop=True)
        for k, (x, y) in enumerate(zip(two, one)):
            for a, b in zip(x[:3], y[:3]):
                np.testing.assert_array_equal(a, b, err_msg=f"two vs one slot set {mats[k].shape} #{k}")
            assert x[3] == y[3] and x[4] == y[4] == jsvd.STOP_CONVERGED


def test_tall_single_streamed_held_rows():
    # A lone plain-path matrix of more than 1024 rows and at most 512 columns
    # runs four-wave groups with its rows streamed; each pair keeps its first
    # 20 chunks of 256 rows (5120 rows) in registers from the dots to the
    # update and streams the rest twice, the update pass in reverse.  6100 x 40
    # has a ragged streamed tail (980 rows beyond the held ones), 3000 x 60
    # is held whole.  Against the oracle, and bitwise against the same
    # matrices inside a batch of 100 that runs one workgroup per matrix
    # without held rows (the arithmetic must not depend on them).
    _note("tall singles, held + streamed rows")
    rng = np.random.default_rng(907)
    for shape in ((6100, 40), (3000, 60)):
        a = rng.standard_normal(shape)
        single = jsvd.svd(a, epsilon=1e-13, max_sweeps=30)
        _assert_parity(a, single, oracle.solve(a, 1e-13, 30), 1e-13, f"held {shape}")
        batch = [a] + [rng.standard_normal(shape) for _ in range(99)]
        got = jsvd.svd_batched(batch, epsilon=1e-13, max_sweeps=30)
        for x, y in zip(single[:3], got[0][:3]):
            np.testing.assert_array_equal(x, y, err_msg=f"single vs batch {shape}")
        assert single[3] == got[0][3]


def test_grouped_wave_launch_with_narrow_groups():
    # 40 matrices of 512 x 512 stay in one grouped launch (below the
    # 4 + 256^2 / 200 matrices of the one-workgroup-per-matrix form), and 40
    # groups of one wave per pair (16 workgroups of sixteen waves each)
    # exceed the at most 512 resident workgroups of sixteen waves, so the
    # groups get fewer workgroups and each wave takes several pairs per
    # round.  Each matrix must equal its single solve bitwise; sigma is
    # checked against LAPACK (the oracle needs seconds per matrix here).
    # eps = 1e-14 (above the rounding floor of 512-row dots, about 2.5e-15):
    # at 1e-13, 512 columns each within eps of orthogonal can sum to
    # ||U^T U - I||_F above the 9e-12 bound.
    _note("grouped wave launch, 40 x 512x512")
    mats = [gen.gaussian_square(512, seed=gen.BASE_SEED + k) for k in range(40)]
    got = jsvd.svd_batched(mats, epsilon=1e-14, max_sweeps=30)
    for k, (a, g) in enumerate(zip(mats, got)):
        u, s, v, sw = g
        ref = np.linalg.svd(a, compute_uv=False)
        assert np.max(np.abs(np.sort(s)[::-1] - ref)) <= 1e-12 * ref[0], f"sigma {k}"
        tol = 1e-12 * (512 / 64 + 1)
        assert metrics.orthogonality(u, s > 1e-14) <= tol, f"U orthogonality {k}"
        assert metrics.orthogonality(v) <= tol, f"V orthogonality {k}"
        assert metrics.residual(a, u, s, v) <= tol, f"residual {k}"
    for k in (0, 39):
        single = jsvd.svd(mats[k], epsilon=1e-14, max_sweeps=30)
        for x, y in zip(single[:3], got[k][:3]):
            np.testing.assert_array_equal(x, y)
        assert single[3] == got[k][3]


def test_widest_plain_path_on_grid_barrier_kernel():
    # More than 2048 columns: one workgroup per pair with a grid barrier
    # (persistent_sweeps_f64).  The oracle needs minutes here, so sigma is
    # checked against LAPACK and the invariants with the bounds of the
    # 8192 x 2048 blocked test (thousands of rotations touch every column).
    _note("grid-barrier kernel, 2100 x 2049")
    rng = np.random.default_rng(2049)
    a = rng.standard_normal((2100, 2049))
    u, s, v, sw = jsvd.svd(a, epsilon=1e-14, max_sweeps=30)
    ref = np.linalg.svd(a, compute_uv=False)
    assert np.max(np.abs(np.sort(s)[::-1] - ref)) <= 1e-11 * ref[0]
    assert metrics.orthogonality(u, s > 1e-10 * ref[0]) <= 1e-9
    assert metrics.orthogonality(v) <= 1e-9
    assert metrics.residual(a, u, s, v) <= 1e-11
    assert sw < 30


def test_every_shape_class_against_oracle():
    # The autotuner may assign any fitting class to a column-count bucket:
    # force each class in turn on a batch whose shapes it fits.
    lib = jsvd.load()
    rng = np.random.default_rng(31)
    shapes = [(12, 6), (16, 8), (30, 16), (40, 29), (64, 32), (90, 48), (128, 64), (128, 100), (128, 128),
              (200, 120)]
    mats = [rng.standard_normal(s) for s in shapes]
    refs = oracle.solve_batch(mats, 1e-13, 30)
    try:
        for cls in range(lib.jsvd_class_count()):
            for b in range(7):
                lib.jsvd_tune_class(b, cls)
            _note(f"forced class {cls}")
            got = jsvd.svd_batched(mats, epsilon=1e-13, max_sweeps=30)
            for a, g, r in zip(mats, got, refs):
                _assert_parity(a, g, r, 1e-13, f"class{cls}{a.shape}")
    finally:
        for b in range(7):
            lib.jsvd_tune_class(b, -1)


def test_gaussian_256_at_floor_tolerance():
    # The 256 x 256 N(0,1) workload with eps = 1e-15 and 30 sweeps.
    _note("gaussian 256 eps 1e-15")
    a = gen.gaussian_square(256)
    got = jsvd.svd(a, epsilon=1e-15, max_sweeps=30)
    ref = oracle.solve(a, 1e-15, 30, threads=8)
    _assert_parity(a, got, ref, 1e-15, "gaussian256")


def test_special_cases():
    _note("special cases")
    d = np.diag([2.0, -5.0, 1.0, 3.0])
    u, s, v, sw = jsvd.svd(d, epsilon=1e-9, max_sweeps=30)
    assert sw == 1
    np.testing.assert_array_equal(s, [2.0, 5.0, 1.0, 3.0])
    np.testing.assert_array_equal(v, np.eye(4))
    np.testing.assert_array_equal(u, np.diag([1.0, -1.0, 1.0, 1.0]))

    rng = np.random.default_rng(5)
    z = rng.standard_normal((7, 4))
    z[:, 1] = 0.0
    u, s, v, sw = jsvd.svd(z, epsilon=1e-9, max_sweeps=30)
    uo, so, vo, swo = oracle.solve(z, 1e-9, 30)
    assert np.sum(s == 0.0) == 1 and np.all(u[:, s == 0.0] == 0.0)
    np.testing.assert_allclose(s, so, rtol=0, atol=1e-13 * so.max())

    one = np.array([[3.0], [4.0]])
    u, s, v, sw = jsvd.svd(one)
    assert sw == 1 and abs(s[0] - 5.0) < 1e-15 and v[0, 0] == 1.0

    nan = np.eye(4, 3)
    nan[0, 1] = np.nan
    u, s, v, sw = jsvd.svd(nan, epsilon=1e-9, max_sweeps=5)
    assert s[1] == 0.0 and np.all(u[:, 1] == 0.0) and s[0] == 1.0 and s[2] == 1.0

    a = rng.standard_normal((30, 30))
    assert jsvd.svd(a, epsilon=1e-15, max_sweeps=1)[3] == 1

    tiny = np.vstack([np.diag([1.0, 5e-10]), np.zeros((1, 2))])
    u, s, v, sw = jsvd.svd(tiny, epsilon=1e-9)
    assert s[1] == 5e-10 and np.all(u[:, 1] == 0.0)

    # A NaN epsilon passes the reference's `epsilon <= 0.0` check
    # (src/domain/jacobi_svd.cu:211): nothing rotates, one sweep, sigma = the
    # column norms, U = 0, V = I -- on the LDS path and a large-matrix path.
    for shape in [(9, 5), (300, 150)]:
        a = rng.standard_normal(shape)
        u, s, v, sw = jsvd.svd(a, epsilon=float("nan"), max_sweeps=8)
        uo, so, vo, swo = oracle.solve(a, float("nan"), 8)
        assert sw == swo == 1, shape
        np.testing.assert_array_equal(v, np.eye(shape[1]))
        assert np.all(u == 0.0) and np.all(uo == 0.0)
        np.testing.assert_allclose(s, so, rtol=1e-14)


def test_repeat_runs_are_bitwise_identical():
    _note("determinism")
    rng = np.random.default_rng(9)
    for shape in [(90, 60), (300, 150)]:
        a = rng.standard_normal(shape)
        r1 = jsvd.svd(a, epsilon=1e-14, max_sweeps=30)
        r2 = jsvd.svd(a, epsilon=1e-14, max_sweeps=30)
        for x, y in zip(r1[:3], r2[:3]):
            np.testing.assert_array_equal(x, y)
        assert r1[3] == r2[3]


def test_batched_host_api_against_oracle():
    _note("batched host api, 256 uneven matrices")
    mats, sig = gen.uneven_batch(256)
    got = jsvd.svd_batched(mats, epsilon=1e-13, max_sweeps=30)
    # eps = 1e-13 is above the rounding floor: positions must agree exactly.
    ref = oracle.solve_batch(mats, 1e-13, 30)
    for k, (a, g, r) in enumerate(zip(mats, got, ref)):
        _assert_parity(a, g, r, 1e-13, f"batch{k}")
        assert metrics.sigma_errors(g[1], sig[k])[0] < 1e-13


def test_batch_plan_device_full_uneven_batch():
    # The 4096-matrix uneven batch on device buffers through jsvd_plan_*:
    # every matrix against its constructed spectrum and the invariants, and a
    # seeded subset against the oracle.
    import torch

    _note("device plan, 4096 uneven matrices")
    mats, sig = gen.uneven_batch(4096)
    flat, ms, ns = gen.pack(mats)
    plan = jsvd.BatchPlan(ms, ns, epsilon=1e-15, max_sweeps=30)
    info = plan.info
    assert info["lds_matrices"] == 4096 and info["large_matrices"] == 0
    dev = torch.device("cuda:0")
    a = torch.from_numpy(flat).to(dev)
    u = torch.empty(plan.a_elems, dtype=torch.float64, device=dev)
    s = torch.empty(plan.s_elems, dtype=torch.float64, device=dev)
    v = torch.empty(plan.v_elems, dtype=torch.float64, device=dev)
    sw = torch.zeros(len(ms), dtype=torch.int32, device=dev)
    plan.execute(a, u, s, v, sw)
    torch.cuda.synchronize()
    u, s, v, sw = (t.cpu().numpy() for t in (u, s, v, sw))
    worst = 0.0
    for k in range(len(mats)):
        m, n = int(ms[k]), int(ns[k])
        uk = u[plan.a_off[k]:plan.a_off[k] + m * n].reshape(m, n)
        sk = s[plan.s_off[k]:plan.s_off[k] + n]
        vk = v[plan.v_off[k]:plan.v_off[k] + n * n].reshape(n, n)
        rep = metrics.report(mats[k], uk, sk, vk, sig[k], 1e-15)
        worst = max(worst, rep["sigma_err_norm"])
        assert rep["sigma_err_norm"] < 1e-13, k
        assert rep["orth_v"] < 1e-12 and rep["residual"] < 1e-12, k
        assert 1 <= sw[k] <= 30
    _note(f"worst normwise sigma error vs constructed spectrum: {worst:.3e}")
    subset = list(range(0, 4096, 7))
    ref = oracle.solve_batch([mats[k] for k in subset], 1e-15, 30)
    positional = 0
    for k, r in zip(subset, ref):
        m, n = int(ms[k]), int(ns[k])
        g = (u[plan.a_off[k]:plan.a_off[k] + m * n].reshape(m, n), s[plan.s_off[k]:plan.s_off[k] + n],
             v[plan.v_off[k]:plan.v_off[k] + n * n].reshape(n, n), int(sw[k]))
        positional += _assert_parity(mats[k], g, r, 1e-15, f"plan{k}", allow_permutation=True)
    _note(f"positional sigma match at eps=1e-15: {positional}/{len(subset)}")
    assert positional >= 0.9 * len(subset)
    plan.close()


def test_autotune_batch_keeps_parity():
    # jsvd_autotune_batch times whole plans on the caller's batch and installs
    # the fastest class per bucket; a plan built afterwards must still match
    # the oracle, and the reported table must hold valid choices.
    import torch

    _note("autotune_batch, 300 uneven matrices")
    mats, _ = gen.uneven_batch(300, seed=91)
    flat, ms, ns = gen.pack(mats)
    _, _, _, (a_el, v_el, s_el) = jsvd.pack_offsets(ms, ns)
    dev = torch.device("cuda:0")
    a = torch.from_numpy(flat).to(dev)
    u = torch.empty(a_el, dtype=torch.float64, device=dev)
    s = torch.empty(s_el, dtype=torch.float64, device=dev)
    v = torch.empty(v_el, dtype=torch.float64, device=dev)
    try:
        rep = jsvd.autotune_batch(ms, ns, a, u, s, v, reps=1, epsilon=1e-13, max_sweeps=30)
        _note(f"tuned classes {rep['classes']} {rep['ms_initial']:.3f} -> {rep['ms_tuned']:.3f} ms")
        ncls = jsvd.load().jsvd_class_count()
        assert all(-1 <= c < ncls for c in rep["classes"])
        assert rep["samples"] >= 2 and 0 < rep["ms_tuned"] <= rep["ms_initial"]
        plan = jsvd.BatchPlan(ms, ns, epsilon=1e-13, max_sweeps=30)
        sw = torch.zeros(len(ms), dtype=torch.int32, device=dev)
        plan.execute(a, u, s, v, sw)
        torch.cuda.synchronize()
        uh, sh, vh, swh = (t.cpu().numpy() for t in (u, s, v, sw))
        subset = list(range(0, len(mats), 5))
        ref = oracle.solve_batch([mats[k] for k in subset], 1e-13, 30)
        for k, r in zip(subset, ref):
            m, n = int(ms[k]), int(ns[k])
            g = (uh[plan.a_off[k]:plan.a_off[k] + m * n].reshape(m, n), sh[plan.s_off[k]:plan.s_off[k] + n],
                 vh[plan.v_off[k]:plan.v_off[k] + n * n].reshape(n, n), int(swh[k]))
            _assert_parity(mats[k], g, r, 1e-13, f"tuned{k}")
        plan.close()
    finally:
        for b in range(7):
            jsvd.tune_class(b, -1)


def test_batch_plan_repeat_execution():
    # The work-list counters reset themselves at the end of each launch: a
    # second execution of the same plan must give identical results.
    import torch

    _note("device plan repeat")
    mats, _ = gen.uneven_batch(300, seed=77)
    flat, ms, ns = gen.pack(mats)
    plan = jsvd.BatchPlan(ms, ns, epsilon=1e-14, max_sweeps=30)
    dev = torch.device("cuda:0")
    a = torch.from_numpy(flat).to(dev)
    outs = []
    for _ in range(3):
        u = torch.empty(plan.a_elems, dtype=torch.float64, device=dev)
        s = torch.empty(plan.s_elems, dtype=torch.float64, device=dev)
        v = torch.empty(plan.v_elems, dtype=torch.float64, device=dev)
        plan.execute(a, u, s, v)
        torch.cuda.synchronize()
        outs.append((u.cpu(), s.cpu(), v.cpu()))
    for o in outs[1:]:
        for x, y in zip(outs[0], o):
            assert torch.equal(x, y)
    plan.close()


def test_device_out_of_memory_is_reported():
    # A plan whose workspace is far beyond HBM (a 2^24 x 2^14 matrix needs a
    # 2 TiB column-major copy): hipMalloc fails, the C-ABI returns
    # JSVD_ENOMEM with the message shape of the reference's CUDA check
    # (src/domain/cuda_check.cuh:26-28), and the device stays usable.
    _note("out-of-memory plan")
    with pytest.raises(jsvd.JsvdError) as e:
        jsvd.BatchPlan([1 << 24], [1 << 14])
    assert e.value.status == 3  # JSVD_ENOMEM
    msg = str(e.value)
    assert msg.startswith("HIP call failed: ") and " @ " in msg and ", code=" in msg and ", message=" in msg
    a = np.random.default_rng(3).standard_normal((40, 30))
    u, s, v, sw = jsvd.svd(a, epsilon=1e-13, max_sweeps=30)
    assert metrics.residual(a, u, s, v) <= 1e-13
