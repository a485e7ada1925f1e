"""The row partition of the two-phase coarsest-level sweeps
(kh_resident_partition, host only): every row owned by exactly one thread,
in consecutive runs; rows longer than 16 entries alone in their thread, in
the leading wavefronts; 8 entries per thread in the parallel phase, 16 above
4,096 entries; refusals for levels that do not fit."""
import numpy as np
import pytest

import kutta_amd
from kutta_amd import RESIDENT_MAX_ROWS, RESIDENT_THREADS


def pattern(lengths):
    return np.concatenate([[0], np.cumsum(lengths)]).astype(np.int32)


def check(row_ptr, table, cap):
    n = len(row_ptr) - 1
    T = RESIDENT_THREADS
    first, last = table[0::2], table[1::2]
    nnz = row_ptr[-1]
    assert cap == (8 if nnz <= 8 * T else 16)
    owner = np.full(n, -1)
    seen_short = False
    for t in range(T):
        r0, r1 = first[t], last[t]
        assert 0 <= r0 <= r1 <= n
        lengths = np.diff(row_ptr[r0:r1 + 1])
        if r1 > r0 and lengths.max() > 16:
            assert r1 == r0 + 1 and not seen_short  # a long row: alone, before every short run
        elif r1 > r0:
            if not seen_short:
                assert t % 64 == 0  # the short runs start at a wavefront
            seen_short = True
        assert np.all(owner[r0:r1] == -1)
        owner[r0:r1] = t
    assert np.all(owner >= 0)


@pytest.mark.parametrize("n,top,long_rows,seed", [
    (25, 5, 0, 1), (1067, 5, 70, 2), (2500, 3, 12, 3), (4096, 1, 0, 4), (3000, 2, 0, 5), (400, 16, 0, 6),
])
def test_partition_invariants(n, top, long_rows, seed):
    rng = np.random.default_rng(seed)
    lengths = rng.integers(1, top + 1, n)
    if long_rows:
        lengths[rng.choice(n, long_rows, replace=False)] = rng.integers(17, 65, long_rows)
    row_ptr = pattern(lengths)
    out = kutta_amd.resident_partition(row_ptr)
    assert out is not None
    check(row_ptr, *out)


def test_partition_refusals():
    part = kutta_amd.resident_partition
    assert part(pattern([1] * (RESIDENT_MAX_ROWS + 1))) is None  # too many rows
    assert part(pattern([4] * 2049)) is None  # more entries than 16 per thread
    assert part(pattern([17] * 600)) is None  # more long rows than threads
    assert part(pattern([3, 2, 2]) + 1) is None  # row_ptr[0] != 0
    assert part(pattern([3, 0, 2])) is None  # an empty row
    assert part(pattern([4] * 2048)) is not None  # exactly 16 per thread
