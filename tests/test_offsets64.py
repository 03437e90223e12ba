"""64-bit index math of the packed batch layout and of the blocked path, at
sizes past 2^31 elements (no device: both queries are host-only C-ABI calls).

The reference indexes a matrix with 32-bit ints
(src/domain/matrix_index.cuh:12-27), so row * columns + col overflows once a
matrix holds 2^31 elements.  Here every offset is int64_t; these tests check
the library's numbers against Python's unbounded integers, computed from the
layout's definition (matrix k starts where matrix k-1 ends) rather than by
restating the library's code.
"""
import numpy as np
import pytest

import jsvd
from jsvd import _lib

TWO31 = 1 << 31


def _reference_offsets(ms, ns):
    a, v, s = [0], [0], [0]
    for m, n in zip(ms, ns):
        a.append(a[-1] + m * n)
        v.append(v[-1] + n * n)
        s.append(s[-1] + n)
    return a[:-1], v[:-1], s[:-1], (a[-1], v[-1], s[-1])


def test_packed_offsets_past_2_31_elements():
    # 65537 x 32768 = 2^31 + 2^15 elements per matrix: the second matrix starts
    # past 2^31, the third past 2^32.
    ms, ns = [65537, 65537, 65537, 3], [32768, 32768, 32768, 2]
    a, v, s, tot = jsvd.pack_offsets(ms, ns)
    ra, rv, rs, rtot = _reference_offsets(ms, ns)
    assert a.tolist() == ra and v.tolist() == rv and s.tolist() == rs and tot == rtot
    assert a[1] == TWO31 + (1 << 15) and a[2] > 2 * TWO31
    assert a.dtype == np.int64


def test_packed_offsets_many_matrices_cross_2_31_and_2_32():
    rng = np.random.default_rng(3)
    ns = rng.integers(1000, 3000, size=600)
    ms = ns + rng.integers(0, 4000, size=600)
    a, v, s, tot = jsvd.pack_offsets(ms, ns)
    ra, rv, rs, rtot = _reference_offsets(ms.tolist(), ns.tolist())
    assert a.tolist() == ra and v.tolist() == rv and s.tolist() == rs and tot == rtot
    assert tot[0] > 2 * TWO31


def test_packed_layout_limits_and_messages():
    # One 2^40 x 2^20 matrix is 2^60 elements: the largest packed buffer.
    a, _, _, tot = jsvd.pack_offsets([1 << 40], [1 << 20])
    assert tot[0] == 1 << 60
    with pytest.raises(_lib.JsvdInvalidArgument, match="exceeds 2\\^60"):
        jsvd.pack_offsets([1 << 40, 1 << 40], [1 << 20, 1 << 20])
    with pytest.raises(_lib.JsvdInvalidArgument, match="exceeds 2\\^60"):
        jsvd.pack_offsets([1 << 40], [(1 << 20) + 1])
    with pytest.raises(_lib.JsvdInvalidArgument, match="requires rows >= columns"):
        jsvd.pack_offsets([3], [4])
    with pytest.raises(_lib.JsvdInvalidArgument, match="non-zero"):
        jsvd.pack_offsets([3, 0], [2, 2])
    with pytest.raises(_lib.JsvdInvalidArgument, match="too large"):
        jsvd.pack_offsets([(1 << 40) + 1], [1])


@pytest.mark.parametrize("m,n,b", [(65600, 32768, 32), (65601, 32770, 16), (TWO31 + 5, 64, 32), (8192, 2048, 16)])
def test_blocked_layout_past_2_31_elements(m, n, b):
    lay = jsvd.blocked_layout(m, n, b)
    lda = -(-m // 64) * 64                 # rows padded to a multiple of 64
    blocks = 2 * -(-n // (2 * b))          # an even number of column blocks
    nv = blocks * b
    assert (lay["lda"], lay["blocks"], lay["nv"], lay["pairs"]) == (lda, blocks, nv, blocks // 2)
    assert lay["a_elements"] == lda * nv and lay["v_elements"] == nv * nv
    assert lay["chunk_rows"] % 64 == 0 and lay["chunks"] * lay["chunk_rows"] >= lda
    assert (lay["chunks"] - 1) * lay["chunk_rows"] < lda
    assert lay["gram_elements"] == lay["pairs"] * lay["chunks"] * 4 * b * b
    if m * n > TWO31:
        assert lay["a_elements"] > TWO31


def test_blocked_layout_rejects_bad_width_and_narrow_matrix():
    with pytest.raises(_lib.JsvdInvalidArgument, match="block_width"):
        jsvd.blocked_layout(100, 80, 8)
    with pytest.raises(_lib.JsvdInvalidArgument, match="n >= 2"):
        jsvd.blocked_layout(100, 63, 32)
