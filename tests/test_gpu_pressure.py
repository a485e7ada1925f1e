"""The pressure-correction operator assembled on the MI355X by
fv_pressure_assemble (kernels/fv_flow.kern.h:164-183) against the oracle, and
the pinned-cell AMG-PCG pressure solve of the benchmark workload."""
import numpy as np
import pytest

import kutta_amd
import oracle
from kutta_amd import AmgOptions, OperatorStructure, SolveControl, SolverKind

pytestmark = pytest.mark.gpu


def bits(a):
    return a.view(np.uint32)


@pytest.mark.parametrize("shape,reference", [((13, 12, 11), -2), ((9, 8, 7), -1), ((10, 9, 8), 17)])
def test_device_assembly_bitwise_vs_oracle(device, shape, reference):
    print(f"[assembly] {shape} reference={reference}", flush=True)
    p = kutta_amd.PressureProblem(device, *shape, reference=reference)
    o = oracle.pressure_problem(*shape, reference=reference)
    assert p.reference == o["reference"][0]
    assert np.array_equal(bits(p["values"]), bits(o["values"]))
    assert np.array_equal(bits(p["rhs"]), bits(o["rhs"]))
    for name in ("row_ptr", "col", "diag", "adjacent", "matrix_entry", "cells_new_to_old"):
        assert np.array_equal(p[name], o[name]), name


def test_reassembly_is_idempotent(device):
    p = kutta_amd.PressureProblem(device, 8, 8, 8)
    first = p["values"].copy()
    p.assemble(-2)
    p.assemble(-2)
    assert np.array_equal(bits(p["values"]), bits(first))


def test_full_size_operator_bitwise(device):
    """The 998,250-cell benchmark operator, assembled on the GPU, equals the
    oracle's assembly bit for bit."""
    p = kutta_amd.PressureProblem(device, 55, 55, 55)
    o = oracle.pressure_problem(55, 55, 55)
    assert p.n_cells == 998_250 and p.nnz == 4_954_950
    assert p.reference == o["reference"][0]
    assert np.array_equal(bits(p["values"]), bits(o["values"]))


@pytest.mark.parametrize("sweeps", [2, 5])
def test_pinned_pressure_solve_matches_oracle(device, sweeps):
    p = kutta_amd.PressureProblem(device, 24, 23, 22)
    b = oracle.neumann_rhs(p.n_cells, reference=p.reference)
    options = AmgOptions(pre_sweeps=sweeps, post_sweeps=sweeps, aggregation_passes=2,
                         structure=OperatorStructure.SymmetricMMatrix)
    solver = kutta_amd.make_amg_solver(SolverKind.AmgPcg, device, p.host, options)
    solver.refresh(p)
    x = device.alloc(p.n_cells)
    x.fill(np.float32(0))
    control = SolveControl(1e-6, 0.0, 2000)
    st = solver.solve(p, x, device.upload(b), control)
    a = oracle.Csr.from_arrays(p["row_ptr"], p["col"], p["diag"], p["values"].astype(np.float64))
    so = oracle.Solver("AmgPcg", a, pre=sweeps, post=sweeps, passes=2, m_matrix=True)
    xo, sto = so.solve(b, rel_tol=1e-6, abs_tol=0.0, max_iter=2000)
    assert st.converged and sto["converged"]
    assert st.iterations == sto["iterations"]
    assert solver.describe() == so.describe()
    assert np.array_equal(bits(x.download()), bits(xo))


def test_full_size_pinned_solve_bitwise_vs_oracle(device):
    """The benchmark solve itself: the 998,250-row C2 operator assembled on the
    GPU, its seeded zero-mean Gaussian right-hand side (pinned cell zero),
    aggregation AMG with 2 passes, V(5,5), PCG to 1e-6 from x = 0: the same
    iteration count and x bit for bit as the oracle on the same bytes."""
    p = kutta_amd.PressureProblem(device, 55, 55, 55, reference=-2)
    rng = np.random.default_rng(20261202)
    b64 = rng.standard_normal(p.n_cells)
    b64 -= b64.mean()
    b = b64.astype(np.float32)
    b[p.reference] = 0
    options = AmgOptions(pre_sweeps=5, post_sweeps=5, aggregation_passes=2,
                         structure=OperatorStructure.SymmetricMMatrix)
    solver = kutta_amd.make_amg_solver(SolverKind.AmgPcg, device, p.host, options)
    solver.refresh(p)
    solver.autotune_tail(p, 2)  # whichever plan: the same iterates
    x = device.alloc(p.n_cells)
    x.fill(np.float32(0))
    st = solver.solve(p, x, device.upload(b), SolveControl(1e-6, 0.0, 2000))
    print(f"[c2 full] {p.n_cells} rows, {st.iterations} iterations, plan: {solver.plan()}", flush=True)
    a = oracle.Csr.from_arrays(p["row_ptr"], p["col"], p["diag"], p["values"].astype(np.float64))
    so = oracle.Solver("AmgPcg", a, pre=5, post=5, passes=2, m_matrix=True)
    xo, sto = so.solve(b, rel_tol=1e-6, abs_tol=0.0, max_iter=2000)
    assert st.converged and sto["converged"]
    assert st.iterations == sto["iterations"] == 26
    assert np.array_equal(bits(x.download()), bits(xo))
