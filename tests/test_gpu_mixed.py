"""The mixed-precision AMG path (no reference counterpart): a coarse level
solved by an FP64 dense inverse (kh_dense.hip). Pinned by a dense FP64 solve
of the whole operator, by the plain FP32 path's solution, and by its effect on
convergence (an exact coarse solve cannot need more iterations)."""
import numpy as np
import pytest

import kutta_amd
import oracle
from kutta_amd import AmgOptions, OperatorStructure, SolveControl, SolverKind

pytestmark = pytest.mark.gpu


def dense_of(a):
    n = a.n_rows
    m = np.zeros((n, n))
    for i in range(n):
        for j in range(a.row_ptr[i], a.row_ptr[i + 1]):
            m[i, a.col[j]] = a.val[j]
    return m


def make(device, a, kind=SolverKind.AmgPcg, direct=0, passes=2):
    host = kutta_amd.CsrMatrixHost(a.row_ptr, a.col, a.diag, a.val)
    matrix = kutta_amd.upload(device, host)
    opts = AmgOptions(aggregation_passes=passes, structure=OperatorStructure.SymmetricMMatrix)
    solver = kutta_amd.make_amg_solver(kind, device, host, opts)
    level = solver.set_direct(direct) if direct else 0
    solver.refresh(matrix)
    return matrix, solver, level


def solve(device, matrix, solver, b, rel=1e-6):
    x = device.alloc(len(b))
    x.fill(np.float32(0))
    st = solver.solve(matrix, x, device.upload(b), SolveControl(rel, 0.0, 500, 8))
    return st, x.download()


@pytest.mark.parametrize("kind", [SolverKind.AmgPcg, SolverKind.AmgFgmres])
@pytest.mark.parametrize("direct", [400, 60])
def test_dense_coarse_solve_converges_to_dense_reference(device, kind, direct):
    """Pinned pressure operator of a 12x11x10 hex box: with the first coarse
    level of at most `direct` rows solved in FP64, PCG and FGMRES reach 1e-6
    in no more iterations than the plain FP32 V-cycle, and x agrees with a
    dense FP64 solve of the whole system to the conditioning bound."""
    a = oracle.pressure_problem(12, 11, 10, tet=False, amplitude=0.0, reference=0).csr()
    b = np.random.default_rng(4).standard_normal(a.n_rows).astype(np.float32)
    b[0] = 0
    m0, plain, _ = make(device, a, kind)
    st0, x0 = solve(device, m0, plain, b)
    m1, mixed, level = make(device, a, kind, direct)
    assert level >= 1 and mixed.levels()[level][0] <= direct
    assert "FP64 dense solve" in mixed.describe()
    st1, x1 = solve(device, m1, mixed, b)
    assert st0.converged and st1.converged
    assert st1.iterations <= st0.iterations
    A = dense_of(a)
    ref = np.linalg.solve(A, b.astype(np.float64))
    bound = 2e-6 * np.linalg.cond(A) * np.abs(ref).max()
    assert np.abs(x1 - ref).max() <= bound
    assert np.abs(x1 - x0).max() <= bound


def test_dense_coarse_solve_on_the_perturbed_tet_operator(device):
    """The benchmark's operator family (perturbed Freudenthal tets, origin
    pinned) at 16^3: the mixed V-cycle converges in fewer iterations than the
    plain one and both solutions satisfy the FP64 residual bound."""
    problem = oracle.pressure_problem(16, 16, 16)
    a = problem.csr()
    b = np.random.default_rng(1).standard_normal(a.n_rows).astype(np.float32)
    m0, plain, _ = make(device, a)
    st0, x0 = solve(device, m0, plain, b)
    m1, mixed, level = make(device, a, direct=2048)
    st1, x1 = solve(device, m1, mixed, b)
    assert st0.converged and st1.converged and st1.iterations <= st0.iterations
    for st in (st0, st1):
        assert st.final_residual <= 1e-6 * st.initial_residual


def test_dense_coarse_solve_is_validated(device):
    a = oracle.poisson(10, 9, 8)
    host = kutta_amd.CsrMatrixHost(a.row_ptr, a.col, a.diag, a.val)
    general = kutta_amd.make_amg_solver(SolverKind.AmgPcg, device, host, AmgOptions())
    with pytest.raises(kutta_amd.KuttaError):
        general.set_direct(100)  # needs the symmetric M-matrix structure
    p = oracle.pressure_problem(8, 8, 8, tet=False, amplitude=0.0, reference=0).csr()
    hp = kutta_amd.CsrMatrixHost(p.row_ptr, p.col, p.diag, p.val)
    s = kutta_amd.make_amg_solver(SolverKind.AmgPcg, device, hp,
                                  AmgOptions(structure=OperatorStructure.SymmetricMMatrix))
    with pytest.raises(kutta_amd.KuttaError):
        s.set_direct(9000)
    m = kutta_amd.upload(device, hp)
    s.set_direct(100)
    x = device.alloc(p.n_rows)
    x.fill(np.float32(0))
    with pytest.raises(kutta_amd.KuttaError):  # no refresh since set_direct
        s.solve(m, x, device.upload(np.ones(p.n_rows, np.float32)), SolveControl(1e-6, 0, 50, 8))
    assert s.set_direct(0) == 0
