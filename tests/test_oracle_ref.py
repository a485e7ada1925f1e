"""Pins the oracle's kernel restatement against (a) the reference's own kernel
bodies compiled from its sources (oracle/_ref, built here only) and (b) the
reference's known-answer kernel tests (tests/linalg/test_kernels.cpp:11-170)."""
import ctypes as C

import numpy as np
import pytest

import oracle
from helpers import mixed_magnitude_vector, random_operator, random_vector

REF = oracle.lazy(oracle.ref_kernels)
needs_ref = pytest.mark.skipif(not oracle.ref_built("ref_kernels"), reason="oracle/_ref not built (reference tree absent)")


def p(a):
    return C.c_void_p(a.ctypes.data)


def both(name, n, *arrays_and_scalars, outputs):
    """Runs oracle_<name> and ref_<name> on copies of the same inputs and
    returns the two sets of output arrays."""
    results = []
    for lib_call in (lambda *a: oracle.kernel(name, *a), lambda *a: _ref_call(name, *a)):
        args = [x.copy() if isinstance(x, np.ndarray) else x for x in arrays_and_scalars]
        lib_call(n, *args) if n is not None else lib_call(*args)
        results.append([args[k] for k in outputs])
    return results


def _ref_call(name, *args):
    fn = getattr(REF, f"ref_{name}")
    cargs = []
    for a in args:
        if isinstance(a, np.ndarray):
            cargs.append(p(a))
        elif isinstance(a, (float, np.float32)):
            cargs.append(C.c_float(a))
        else:
            cargs.append(C.c_int(int(a)))
    fn(*cargs)


def assert_bitwise(got, want):
    for g, w in zip(got, want):
        assert g.dtype == w.dtype
        assert np.array_equal(g.view(np.uint32), w.view(np.uint32))


@needs_ref
@pytest.mark.parametrize("seed", [1, 2])
def test_row_kernels_bitwise_vs_reference(seed):
    a = random_operator(seed)
    rp, col, diag, vals = a.row_ptr, a.col, a.diag, a.val.astype(np.float32)
    n = a.n_rows
    x = mixed_magnitude_vector(n, seed)
    b = mixed_magnitude_vector(n, seed + 10)
    inv = np.abs(random_vector(n, seed + 20)) + np.float32(0.1)
    status = np.zeros(3, np.int32)
    y = np.zeros(n, np.float32)
    got, want = both("csr_spmv", n, rp, col, vals, x, y, outputs=[4])
    assert_bitwise(got, want)
    got, want = both("csr_residual", n, rp, col, vals, x, b, y, outputs=[5])
    assert_bitwise(got, want)
    got, want = both("amg_smooth", n, rp, col, vals, x, b, inv, y, status, outputs=[6])
    assert_bitwise(got, want)
    b2 = b.copy()
    b2[::17] = np.inf
    b2[5::23] = np.nan
    got, want = both("amg_smooth_zero", n, rp, col, vals, np.zeros(n, np.float32), b2, inv, y,
                     status, outputs=[6])
    assert_bitwise(got, want)
    for l1 in (0, 1):
        got, want = both("krylov_diagonal", n, rp, diag, vals, y, y.copy(), l1, outputs=[3, 4])
        assert_bitwise(got, want)


@needs_ref
def test_transfer_and_galerkin_bitwise_vs_reference():
    a = random_operator(3)
    agg = oracle.aggregate(a, passes=2, m_matrix=False)
    n, nc = a.n_rows, len(agg["member_offsets"]) - 1
    fine = mixed_magnitude_vector(n, 5)
    coarse = mixed_magnitude_vector(nc, 6)
    got, want = both("amg_restrict", nc, agg["member_offsets"], agg["members"], fine,
                     np.zeros(nc, np.float32), outputs=[3])
    assert_bitwise(got, want)
    got, want = both("amg_prolong", n, agg["aggregate"], coarse, fine, outputs=[2])
    assert_bitwise(got, want)
    vals = a.val.astype(np.float32)
    nnz_c = len(agg["entry_offsets"]) - 1
    got, want = both("amg_galerkin", nnz_c, agg["entry_offsets"], agg["entries"], vals,
                     np.zeros(nnz_c, np.float32), outputs=[3])
    assert_bitwise(got, want)


@needs_ref
def test_m_matrix_diagonal_bitwise_vs_reference():
    a = oracle.poisson(9, 8, 7, 1.0, random=True, neumann=True)
    agg = oracle.aggregate(a, passes=2, m_matrix=True)
    nc = len(agg["member_offsets"]) - 1
    coarse = agg["val"].astype(np.float32)
    got, want = both("amg_m_matrix_diagonal", nc, agg["member_offsets"], agg["members"],
                     a.row_ptr, a.diag, a.val.astype(np.float32), agg["row_ptr"], agg["diag"],
                     coarse, outputs=[7])
    assert_bitwise(got, want)
    # The host coarsening's M-matrix diagonal is the device kernel's, bit for bit.
    assert np.array_equal(got[0], coarse)


@needs_ref
@pytest.mark.parametrize("count", [1, 7, 17, 24])
def test_coarse_solve_bitwise_vs_reference(count):
    a = oracle.poisson(count, 1, 1, 1.0)
    vals = a.val.astype(np.float32)
    b = random_vector(count, count)
    inv = (1.0 / vals[a.diag]).astype(np.float32)
    status = np.zeros(3, np.int32)
    x1 = np.zeros(count, np.float32)
    x2 = np.zeros(count, np.float32)
    oracle.kernel("amg_coarse_solve", a.row_ptr, a.col, vals, x1, b, inv, status, count, 32)
    _ref_call("amg_coarse_solve", a.row_ptr, a.col, vals, x2, b, inv, status, count, 32)
    assert np.array_equal(x1.view(np.uint32), x2.view(np.uint32))


@needs_ref
def test_scalar_state_machine_bitwise_vs_reference():
    rng = np.random.default_rng(11)
    for trial in range(200):
        s = rng.standard_normal(24).astype(np.float32) * np.float32(10.0) ** rng.integers(-8, 8, 24)
        s[20:24] = np.abs(s[20:24])
        if trial % 7 == 0:
            s[0] = 0.0
        if trial % 11 == 0:
            s[2] = -abs(s[2])
        status = np.array([0, trial % 60, trial % 2], np.int32)
        for which in range(8):
            flag = trial % 2 if which in (2, 7) else 0
            s1, st1, s2, st2 = s.copy(), status.copy(), s.copy(), status.copy()
            oracle.lib().oracle_solver_scalar(which, p(s1), p(st1), flag)
            REF.ref_solver_scalar(which, p(s2), p(st2), flag)
            assert np.array_equal(s1.view(np.uint32), s2.view(np.uint32)), (trial, which)
            assert np.array_equal(st1, st2), (trial, which)


@needs_ref
def test_kadd_up_vs_reference():
    rng = np.random.default_rng(5)
    a = rng.standard_normal(1000).astype(np.float32)
    b = (rng.standard_normal(1000) * 1e-4).astype(np.float32)
    oracle.lib()
    REF.ref_kadd_up.argtypes = [C.c_float, C.c_float]
    REF.ref_kadd_up.restype = C.c_float
    for x, y in zip(a, b):
        assert oracle.lib().oracle_kadd_up(x, y) == REF.ref_kadd_up(x, y)


# ------------------------------------------------- known answers (no _ref needed)
def test_known_answers_krylov_diagonal():
    """tests/linalg/test_kernels.cpp:30-41"""
    rows = np.array([0, 2, 5, 7], np.int32)
    diag = np.array([0, 3, 6], np.int32)
    values = np.array([2, -1, -1, 4, -2, -2, 8], np.float32)
    out, invalid = np.zeros(3, np.float32), np.zeros(3, np.float32)
    oracle.kernel("krylov_diagonal", 3, rows, diag, values, out, invalid, 0)
    assert out.tolist() == [0.5, 0.25, 0.125]
    l1 = np.zeros(3, np.float32)
    oracle.kernel("krylov_diagonal", 3, rows, diag, values, l1, invalid, 1)
    assert l1.tolist() == out.tolist()
    assert oracle.max_abs(invalid) == 0
    oracle.kernel("krylov_diagonal", 3, rows, diag, np.zeros(7, np.float32), out, invalid, 0)
    assert invalid.tolist() == [1, 1, 1]


def test_known_answers_transfers():
    """tests/linalg/test_kernels.cpp:44-58"""
    offsets = np.array([0, 2, 5], np.int32)
    members = np.array([0, 2, 1, 3, 4], np.int32)
    agg = np.array([0, 1, 0, 1, 1], np.int32)
    fine = np.array([1, 2, 3, 4, 5], np.float32)
    coarse = np.zeros(2, np.float32)
    oracle.kernel("amg_restrict", 2, offsets, members, fine, coarse)
    assert coarse.tolist() == [4, 11]
    oracle.kernel("amg_prolong", 5, agg, coarse, fine)
    assert fine.tolist() == [5, 13, 7, 15, 16]
    oracle.kernel("amg_galerkin", 2, offsets, members, fine, coarse)
    assert coarse.tolist() == [12, 44]


def test_known_answer_compensated_residual():
    """tests/linalg/test_kernels.cpp:60-84: the compensated residual resolves
    product and cancellation error to FP64 accuracy."""
    rows = np.array([0, 3, 6, 9], np.int32)
    cols = np.array([0, 1, 2] * 3, np.int32)
    vals = np.array([6, -1, -1, -1, 6, -1, -1, -1, 6], np.float32)
    x = np.array([np.nextafter(np.float32(1), np.float32(2)), -1, 1], np.float32)
    b = np.array([6, -8, 6], np.float32)
    r = np.zeros(3, np.float32)
    oracle.kernel("csr_residual", 3, rows, cols, vals, x, b, r)
    ref = b.astype(np.float64) - (vals.reshape(3, 3).astype(np.float64) @ x.astype(np.float64))
    assert np.allclose(r, ref.astype(np.float32), rtol=1e-6, atol=0)


def test_coarse_solve_equals_ordinary_sweeps():
    """tests/linalg/test_kernels.cpp:86-120: the fused coarse Jacobi equals
    residual + smooth sweeps and amg_smooth ping-pong sweeps, bitwise."""
    for count in (1, 17, 24):
        a = oracle.poisson(count, 1)
        vals = a.val.astype(np.float32)
        b = (((oracle.hash32(np.arange(count) + 19) % 2001).astype(np.int64) - 1000) / 1024.0).astype(np.float32)
        inv = (np.float32(1.0) / vals[a.diag]).astype(np.float32)
        status = np.zeros(3, np.int32)
        for sweeps in (1, 5, 32):
            y = np.zeros(count, np.float32)
            r = np.zeros(count, np.float32)
            for _ in range(sweeps):
                oracle.kernel("csr_residual", count, a.row_ptr, a.col, vals, y, b, r)
                y = (y + np.float32(2.0 / 3.0) * inv * r).astype(np.float32)
            x = np.zeros(count, np.float32)
            oracle.kernel("amg_coarse_solve", a.row_ptr, a.col, vals, x, b, inv, status, count, sweeps)
            assert np.array_equal(x, y)
            xs, rs = np.zeros(count, np.float32), np.zeros(count, np.float32)
            for _ in range(sweeps):
                oracle.kernel("amg_smooth", count, a.row_ptr, a.col, vals, xs, b, inv, rs, status)
                xs, rs = rs, xs
            assert np.array_equal(xs, y)


def test_pair_reduction_exact_under_cancellation():
    """tests/runtime/test_device.cpp:36-75 / test_backends.cpp:162-181: the
    (hi, lo) pair carries what a float sum cancels."""
    n = 1 << 12
    x = np.ones(n, np.float32)
    x[1::2] = -1
    x[0] = np.float32(1 + 2 ** -23)
    pair = oracle.reduce_into("sum", x)
    assert pair[0] == np.float32(2 ** -23) and pair[1] == 0
    y = np.full(n, np.float32(2 ** -24), np.float32)
    y[0] = 1
    pair = oracle.reduce_into("sum", y)
    total = 1 + (n - 1) * 2.0 ** -24
    assert float(pair[0]) + float(pair[1]) == total


@needs_ref
@pytest.mark.parametrize("reference", [-1, -2, 5])
@pytest.mark.parametrize("tet", [True, False])
def test_pressure_assembly_bitwise_vs_reference(tet, reference):
    """fv_pressure_assemble (kernels/fv_flow.kern.h:164-183) on a perturbed,
    renumbered box with a random flux, graded coefficients and garbage initial
    values: the oracle's values and right-hand side equal the reference body's."""
    b = oracle.pressure_problem(7, 6, 5, tet=tet, amplitude=0.2)
    n = b["row_ptr"].size - 1
    pin = int(b["reference"][0]) if reference == -2 else reference
    rng = np.random.default_rng(11)
    coefficient = b["diffusion"] * np.exp(rng.uniform(-6, 6, b["diffusion"].size)).astype(np.float32)
    flux = rng.standard_normal(b["diffusion"].size).astype(np.float32)
    values = rng.standard_normal(b["col"].size).astype(np.float32)
    rhs = np.full(n, np.nan, np.float32)
    got, want = both("fv_pressure_assemble", n, b["cell_face_offsets"], b["cell_faces"],
                     b["cell_face_sign"], b["adjacent"], b["matrix_entry"], b["row_ptr"], b["diag"],
                     coefficient, flux, values, rhs, pin, outputs=[9, 10])
    assert_bitwise(got, want)
    assert np.all(np.isfinite(got[0])) and np.all(np.isfinite(got[1]))


def test_fma_product_error_window_equals_dekker():
    """The device computes the compensated residual's product error with one
    FMA where 2^-100 <= |a*b| <= 2^100 and |a|, |b| <= 1e30, and with the
    reference's Dekker split elsewhere (kernels/linalg.kern.h:21-30). Inside
    that window Dekker's split is exact, so both give the exact rounding error;
    40 million seeded pairs agree bit for bit."""
    L = oracle.lib()
    L.oracle_product_error_window.argtypes = [C.c_long, C.c_ulonglong, C.POINTER(C.c_long)]
    L.oracle_product_error_window.restype = C.c_long
    inside = C.c_long(0)
    differ = L.oracle_product_error_window(40_000_000, 20261202, C.byref(inside))
    assert inside.value > 25_000_000
    assert differ == 0
