"""Pins the CPU oracle to the reference's own golden results.

The x86 GCC CPU-oracle run of bench_linalg
(validation/results/colab_t4/logs/bench_linalg.log:1-12; reproduced bitwise by
CUDA with --fmad=false, bench_linalg_nofmad_cuda.log) prints iteration counts,
residuals to 9 significant digits and the hierarchy description. The oracle
must reproduce every printed digit: a dropped compensation term, a different
pairing tie-break, a different poll/restart sequence or a wrong pair rounding
changes them.
"""
import numpy as np
import pytest

import oracle


def run_bench(size, anisotropy, kind, passes=3):
    """bench/bench_linalg.cpp:42-83 for one solver kind."""
    a = oracle.poisson(size, size, size, anisotropy)
    b = oracle.bench_rhs(a.n_rows)
    if kind == "AmgPcg":
        s = oracle.Solver("AmgPcg", a, passes=passes)
    else:
        s = oracle.Solver("PcgJacobi", a)
    vals = a.val.astype(np.float32)
    s.refresh(vals * np.float32(1.25))  # scale_copy(1.25) then refresh
    s.refresh(vals)                     # restore, refresh again
    x, st = s.solve(b, rel_tol=1e-6, abs_tol=1e-12, max_iter=2000)
    return st, s.describe(), x


GOLDEN = [
    # size, anisotropy, kind, iterations, relative residual, initial, final, converged, describe tail
    (64, 1, "PcgJacobi", 278, "6.85768115e-07", "288.643143", "0.000197942264", True, "PCG, Jacobi"),
    (64, 1, "AmgPcg", 15, "7.91394974e-07", "288.643143", "0.000228430732", True,
     "6 levels, op complexity 1.140"),
    (64, 1000, "PcgJacobi", 354, "1.53025686e-06", "288.643143", "0.00044169815", False, "PCG, Jacobi"),
    (64, 1000, "AmgPcg", 38, "1.51536251e-06", "288.643143", "0.000437398999", False,
     "6 levels, op complexity 1.134"),
]


def g9(v):
    """C++ iostream setprecision(9) default-float formatting."""
    return f"{v:.9g}"


@pytest.mark.parametrize("size,aniso,kind,its,rel,initial,final,converged,desc", GOLDEN)
def test_bench_linalg_golden(size, aniso, kind, its, rel, initial, final, converged, desc):
    st, describe, _ = run_bench(size, aniso, kind)
    assert st["iterations"] == its
    assert g9(st["final_residual"] / st["initial_residual"]) == g9(float(rel))
    assert g9(st["initial_residual"]) == initial
    assert g9(st["final_residual"]) == final
    assert st["converged"] == converged
    assert describe.endswith(desc)


@pytest.mark.slow
def test_bench_linalg_golden_128():
    """128^3 AMG-PCG: 21 iterations, 8.66568009e-07, 7 levels, complexity 1.142
    (bench_linalg.log:7)."""
    st, describe, _ = run_bench(128, 1, "AmgPcg")
    assert st["iterations"] == 21
    assert g9(st["final_residual"] / st["initial_residual"]) == "8.66568009e-07"
    assert g9(st["initial_residual"]) == "816.827145"
    assert g9(st["final_residual"]) == "0.000707836272"
    assert describe.endswith("7 levels, op complexity 1.142")


def test_amg_solution_matches_independent_double_cg():
    """Independent FP64 CG (tests/linalg/matrices.hpp:89-105 style) on a small
    random-coefficient operator: the AMG-PCG solution agrees to its tolerance."""
    a = oracle.poisson(12, 10, 8, 1.0, random=True)
    n = a.n_rows
    rng = np.random.default_rng(7)
    xs = rng.uniform(-1, 1, n)
    dense = np.zeros((n, n))
    for i in range(n):
        for j in range(a.row_ptr[i], a.row_ptr[i + 1]):
            dense[i, a.col[j]] = a.val[j]
    b = (dense @ xs).astype(np.float32)
    s = oracle.Solver("AmgPcg", a)
    x, st = s.solve(b, rel_tol=1e-6, abs_tol=0, max_iter=200)
    ref = np.linalg.solve(dense, b.astype(np.float64))
    assert st["converged"]
    r = dense @ x.astype(np.float64) - b.astype(np.float64)
    assert np.linalg.norm(r) <= 1.01e-6 * st["initial_residual"]
    cond = np.linalg.cond(dense)
    assert np.max(np.abs(x - ref)) <= 2e-6 * cond * np.max(np.abs(ref))


def test_neumann_m_matrix_projection_solution():
    """Pure-Neumann M-matrix with a zero-mean RHS: the solution satisfies A x = b
    up to the tolerance and has zero mean (constant-nullspace projection)."""
    a = oracle.poisson(10, 9, 7, 1.0, neumann=True)
    n = a.n_rows
    rng = np.random.default_rng(3)
    b64 = rng.standard_normal(n)
    b64 -= b64.mean()
    b = b64.astype(np.float32)
    s = oracle.Solver("AmgPcg", a, passes=2, nullspace=True, m_matrix=True)
    x, st = s.solve(b, rel_tol=1e-7, abs_tol=0, max_iter=300)
    assert st["converged"]
    dense = np.zeros((n, n))
    for i in range(n):
        for j in range(a.row_ptr[i], a.row_ptr[i + 1]):
            dense[i, a.col[j]] = a.val[j]
    r = dense @ x.astype(np.float64) - b64
    assert np.linalg.norm(r) <= 1e-5 * np.linalg.norm(b64)
    assert abs(x.astype(np.float64).mean()) < 1e-5


@pytest.mark.parametrize("kind", ["BicgstabJacobi", "AmgBicgstab"])
def test_bicgstab_solution_matches_dense_solve(kind):
    """BiCGStab (linalg/solver.cpp:314-361) on a non-symmetric convection-
    diffusion operator (tests/linalg/matrices.hpp:17-62 with wind): the returned
    x satisfies the tolerance in an independent FP64 residual, the reported
    final residual is that residual, and x agrees with a dense FP64 solve to the
    conditioning bound. A wrong sign, index or dropped term in the recurrences
    stalls or converges elsewhere."""
    a = oracle.poisson(14, 12, 1, 1.0, wind=3.0)
    n = a.n_rows
    dense = np.zeros((n, n))
    for i in range(n):
        for j in range(a.row_ptr[i], a.row_ptr[i + 1]):
            dense[i, a.col[j]] = a.val[j]
    assert not np.allclose(dense, dense.T)
    rng = np.random.default_rng(5)
    b = rng.standard_normal(n).astype(np.float32)
    s = oracle.Solver(kind, a, pre=2, post=2, passes=2)
    x, st = s.solve(b, rel_tol=1e-6, abs_tol=0, max_iter=400)
    assert st["converged"]
    r = b.astype(np.float64) - dense @ x.astype(np.float64)
    assert np.linalg.norm(r) <= 1.05e-6 * st["initial_residual"]
    assert np.linalg.norm(r) == pytest.approx(st["final_residual"], rel=1e-3, abs=1e-7)
    ref = np.linalg.solve(dense, b.astype(np.float64))
    assert np.max(np.abs(x - ref)) <= 2e-6 * np.linalg.cond(dense) * np.max(np.abs(ref))


def test_bicgstab_first_iterates_match_textbook():
    """The first BiCGStab-Jacobi iterates equal van der Vorst's right-
    preconditioned BiCGStab (shadow residual r0, M = diag(A)) run in FP64
    (linalg/solver.cpp:314-361; kernels/scalars.kern.h:81-149), to FP32
    rounding. Each recurrence term enters by the second iterate."""
    a = oracle.poisson(9, 8, 1, 1.0, wind=3.0)
    n = a.n_rows
    dense = np.zeros((n, n))
    for i in range(n):
        for j in range(a.row_ptr[i], a.row_ptr[i + 1]):
            dense[i, a.col[j]] = a.val[j]
    minv = 1.0 / np.diag(dense)
    b = np.random.default_rng(8).standard_normal(n).astype(np.float32)
    x = np.zeros(n)
    r = b.astype(np.float64).copy()
    shadow = r.copy()
    rho_prev = alpha = omega = 1.0
    v = p = np.zeros(n)
    for k in range(1, 4):
        rho = shadow @ r
        beta = (rho / rho_prev) * (alpha / omega)
        p = r + beta * (p - omega * v)
        z = minv * p
        v = dense @ z
        alpha = rho / (shadow @ v)
        s = r - alpha * v
        zs = minv * s
        t = dense @ zs
        omega = (t @ s) / (t @ t)
        x = x + alpha * z + omega * zs
        r = s - omega * t
        rho_prev = rho
        xo, st = oracle.Solver("BicgstabJacobi", a).solve(b, rel_tol=1e-12, abs_tol=0, max_iter=k,
                                                          interval=k)
        assert st["iterations"] == k
        assert np.max(np.abs(xo - x)) <= 1e-4 * np.max(np.abs(x)), k
