"""The product's host aggregation (libkutta_amd.so) against the oracle's
restatement of linalg/amg.cpp:25-174: aggregates, member lists, canonical
Galerkin entry maps and coarse values must agree bit for bit."""
import numpy as np
import pytest

import kutta_amd
import oracle


def product_aggregate(a, passes, m_matrix):
    host = kutta_amd.CsrMatrixHost(a.row_ptr, a.col, a.diag, a.val)
    opts = kutta_amd.AmgOptions(aggregation_passes=passes,
                                structure=kutta_amd.OperatorStructure(int(m_matrix)))
    return kutta_amd.aggregate(host, opts)


CASES = [
    ("poisson 3D", lambda: oracle.poisson(17, 13, 11), 3, False),
    ("random coefficients", lambda: oracle.poisson(15, 14, 9, 1.0, random=True), 2, False),
    ("anisotropic", lambda: oracle.poisson(20, 20, 6, 1000.0), 3, False),
    ("upwind (nonsymmetric)", lambda: oracle.poisson(16, 15, 1, 1.0, wind=3.0), 2, False),
    ("neumann M-matrix", lambda: oracle.poisson(14, 12, 10, 1.0, random=True, neumann=True), 2, True),
    ("neumann M-matrix 3 passes", lambda: oracle.poisson(12, 12, 12, 1.0, neumann=True), 3, True),
]


@pytest.mark.parametrize("name,make,passes,m_matrix", CASES, ids=[c[0] for c in CASES])
def test_aggregation_bitwise(name, make, passes, m_matrix):
    a = make()
    want = oracle.aggregate(a, passes=passes, m_matrix=m_matrix)
    got = product_aggregate(a, passes, m_matrix)
    for key in ("aggregate", "member_offsets", "members", "entry_offsets", "entries",
                "row_ptr", "col", "diag"):
        assert np.array_equal(got[key], want[key]), key
    assert np.array_equal(got["val"].view(np.uint64), want["val"].view(np.uint64))


def test_isolated_rows_stay_unaggregated():
    """A row with no off-diagonal coupling keeps aggregate -1 (linalg/amg.cpp:61)."""
    a = oracle.poisson(6, 5, 1)
    row_ptr, col, diag, val = a.row_ptr.copy(), a.col.copy(), a.diag.copy(), a.val.copy()
    # Zero the couplings of row 7 in both directions.
    for i in range(a.n_rows):
        for j in range(row_ptr[i], row_ptr[i + 1]):
            if (i == 7) != (col[j] == 7):
                val[j] = 0.0
    b = oracle.Csr.from_arrays(row_ptr, col, diag, val)
    want = oracle.aggregate(b, passes=2)
    got = product_aggregate(b, 2, False)
    assert want["aggregate"][7] == -1
    assert np.array_equal(got["aggregate"], want["aggregate"])
