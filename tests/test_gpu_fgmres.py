"""Flexible GMRES(m) on the device (kh_gmres.hip; no reference counterpart),
pinned by textbook GMRES in FP64, a dense FP64 solve (as the reference's
tests/linalg/matrices.hpp:108-134 reference_dense) and AMG-PCG on an SPD
operator."""
import numpy as np
import pytest

import kutta_amd
import oracle
from kutta_amd import AmgOptions, SolveControl, SolverKind

pytestmark = pytest.mark.gpu


def dense_of(a):
    n = a.n_rows
    m = np.zeros((n, n))
    for i in range(n):
        for j in range(a.row_ptr[i], a.row_ptr[i + 1]):
            m[i, a.col[j]] = a.val[j]
    return m


def setup(device, a, kind, restart=30, options=None):
    host = kutta_amd.CsrMatrixHost(a.row_ptr, a.col, a.diag, a.val)
    matrix = kutta_amd.upload(device, host)
    if kind == SolverKind.AmgFgmres:
        solver = kutta_amd.make_amg_solver(kind, device, host, options or AmgOptions(pre_sweeps=2, post_sweeps=2))
    else:
        solver = kutta_amd.make_solver(kind, device, host)
    solver.set_restart(restart)
    return matrix, solver


def solve(device, matrix, solver, b, control):
    x = device.alloc(len(b))
    x.fill(np.float32(0))
    st = solver.solve(matrix, x, device.upload(b), control)
    return st, x.download()


def textbook_gmres(a, b, minv, steps):
    """Saad's right-preconditioned GMRES in FP64 from x0 = 0 (M = diag^-1):
    the k-step iterate minimizing ||b - A x|| over x in M K_k(A M, b)."""
    beta = np.linalg.norm(b)
    v = [b / beta]
    h = np.zeros((steps + 1, steps))
    z = []
    for j in range(steps):
        z.append(minv * v[j])
        w = a @ z[j]
        for i in range(j + 1):
            h[i, j] = w @ v[i]
            w = w - h[i, j] * v[i]
        h[j + 1, j] = np.linalg.norm(w)
        v.append(w / h[j + 1, j])
    e1 = np.zeros(steps + 1)
    e1[0] = beta
    y, *_ = np.linalg.lstsq(h, e1, rcond=None)
    return np.array(z).T @ y


@pytest.mark.parametrize("steps", [1, 2, 3, 5, 8])
def test_fgmres_jacobi_iterates_match_textbook_gmres(device, steps):
    """With a fixed Jacobi preconditioner flexible GMRES is right-preconditioned
    GMRES: after k steps (max_iter = k) x is the textbook FP64 iterate to FP32
    rounding. A wrong rotation sign, Hessenberg index or back substitution
    breaks this at the second step."""
    a = oracle.poisson(9, 8, 1, 1.0, wind=3.0)
    A = dense_of(a)
    b = np.random.default_rng(8).standard_normal(a.n_rows).astype(np.float32)
    matrix, solver = setup(device, a, SolverKind.FgmresJacobi)
    st, x = solve(device, matrix, solver, b, SolveControl(1e-12, 0.0, steps, 8))
    assert st.iterations == steps
    ref = textbook_gmres(A, b.astype(np.float64), 1.0 / np.diag(A), steps)
    assert np.abs(x - ref).max() <= 2e-5 * np.abs(ref).max()


@pytest.mark.parametrize("kind", [SolverKind.FgmresJacobi, SolverKind.AmgFgmres])
@pytest.mark.parametrize("restart", [30, 7])
def test_fgmres_solves_nonsymmetric_to_dense_reference(device, kind, restart):
    """Convection-diffusion (tests/linalg/matrices.hpp with wind): FGMRES
    reaches the tolerance, its reported residual is the true FP64 residual,
    and x agrees with a partial-pivoted dense FP64 solve to the conditioning
    bound; a short restart still converges (more iterations)."""
    a = oracle.poisson(20, 18, 1, 1.0, wind=4.0)
    A = dense_of(a)
    b = np.random.default_rng(5).standard_normal(a.n_rows).astype(np.float32)
    matrix, solver = setup(device, a, kind, restart)
    st, x = solve(device, matrix, solver, b, SolveControl(1e-6, 0.0, 600, 8))
    assert st.converged
    r = b.astype(np.float64) - A @ x.astype(np.float64)
    assert np.linalg.norm(r) <= 1.05e-6 * st.initial_residual
    assert np.linalg.norm(r) == pytest.approx(st.final_residual, rel=1e-3, abs=1e-7)
    ref = np.linalg.solve(A, b.astype(np.float64))
    assert np.abs(x - ref).max() <= 2e-6 * np.linalg.cond(A) * np.abs(ref).max()
    if restart < 30:
        assert st.iterations > restart


def test_amg_fgmres_agrees_with_amg_pcg_on_spd(device):
    """On the SPD pinned pressure operator both reach 1e-6 and their solutions
    agree to the tolerance; flexible GMRES minimizes the residual, so it needs
    no more iterations than PCG."""
    problem = oracle.pressure_problem(12, 11, 10, tet=False, amplitude=0.0, reference=0)
    a = problem.csr()
    b = np.random.default_rng(2).standard_normal(a.n_rows).astype(np.float32)
    b[0] = 0
    opts = AmgOptions(aggregation_passes=2, structure=kutta_amd.OperatorStructure.SymmetricMMatrix)
    control = SolveControl(1e-6, 0.0, 200, 8)
    m1, gm = setup(device, a, SolverKind.AmgFgmres, options=opts)
    st_g, x_g = solve(device, m1, gm, b, control)
    host = kutta_amd.CsrMatrixHost(a.row_ptr, a.col, a.diag, a.val)
    m2 = kutta_amd.upload(device, host)
    pcg = kutta_amd.make_amg_solver(SolverKind.AmgPcg, device, host, opts)
    st_p, x_p = solve(device, m2, pcg, b, control)
    assert st_g.converged and st_p.converged
    assert st_g.iterations <= st_p.iterations
    A = dense_of(a)
    ref = np.linalg.solve(A, b.astype(np.float64))
    scale = np.abs(ref).max()
    cond = np.linalg.cond(A)
    for x in (x_g, x_p):
        assert np.abs(x - ref).max() <= 2e-6 * cond * scale


def test_restart_length_is_validated(device):
    a = oracle.poisson(6, 5, 1)
    matrix, solver = setup(device, a, SolverKind.FgmresJacobi)
    with pytest.raises(kutta_amd.KuttaError):
        solver.set_restart(0)
    with pytest.raises(kutta_amd.KuttaError):
        solver.set_restart(64)
