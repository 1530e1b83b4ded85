"""The multigrid tail splits node and element indices with FP32 reciprocals instead of integer division
(small_divmod in mycel_amd/csrc/gmg.hip): q = trunc((n + 1/2) * fl(1/d)) in FP32 arithmetic. Exact for every
index n < 2^16 and divisor d <= 4096 (the tail's levels have <= 512 nodes, the level above <= 4096), checked
here exhaustively with IEEE FP32 products as the device computes them."""
import numpy as np


def test_reciprocal_division_is_exact_on_the_tail_ranges():
    n = np.arange(0, 1 << 16, dtype=np.int64)
    nf = n.astype(np.float32) + np.float32(0.5)
    for d in range(1, 4097):
        rd = np.float32(1.0) / np.float32(d)
        q = (nf * rd).astype(np.float32).astype(np.int64)
        assert np.array_equal(q, n // d), d


def test_the_check_fails_without_the_half_offset():
    """The offset is what makes it exact: truncating n * fl(1/d) is wrong for some (n, d) in the same range."""
    n = np.arange(0, 1 << 16, dtype=np.int64)
    wrong = 0
    for d in (3, 7, 11, 13, 25, 49, 97):
        q = (n.astype(np.float32) * (np.float32(1.0) / np.float32(d))).astype(np.int64)
        wrong += int(np.count_nonzero(q != n // d))
    assert wrong > 0
