"""The optional K-cycle (linalg/amg.cpp:283-336, kernels/scalars.kern.h:157-186)
with flexible PCG (linalg/solver.cpp:295-297) and with BiCGStab on the MI355X:
iteration counts and x bit for bit against the oracle, through every launch
form (fused or reference kernels, graph replay, coarse tail and coarse-sweep
plans picked by the autotuner), for the full K-cycle and for one stopped at
a coarse level."""
import numpy as np
import pytest

import kutta_amd
import oracle
from kutta_amd import AmgCycle, AmgOptions, NullSpace, OperatorStructure, SolveControl, SolverKind

pytestmark = pytest.mark.gpu


def bits(a):
    return a.view(np.uint32)


def solve_both(device, a, b, kind, options, control, fused=True, graphs=True, autotune=False):
    host = kutta_amd.CsrMatrixHost(a.row_ptr, a.col, a.diag, a.val)
    matrix = kutta_amd.upload(device, host)
    solver = kutta_amd.make_amg_solver(kind, device, host, options)
    solver.set_fused(fused)
    solver.set_graphs(graphs)
    solver.refresh(matrix)
    if autotune:
        solver.autotune_tail(matrix, 2)
    x = device.alloc(a.n_rows)
    x.fill(np.float32(0))
    st = solver.solve(matrix, x, device.upload(b), SolveControl(*control))
    ref = oracle.Solver(kind.name, a, pre=options.pre_sweeps, post=options.post_sweeps,
                        coarse_sweeps=options.coarse_sweeps, max_levels=options.max_levels,
                        passes=options.aggregation_passes, nullspace=options.nullspace == NullSpace.Constant,
                        m_matrix=options.structure == OperatorStructure.SymmetricMMatrix,
                        cycle=options.cycle.name, k_depth=options.k_depth, k_min_rows=options.k_min_rows)
    xo, sto = ref.solve(b, rel_tol=control[0], abs_tol=control[1], max_iter=control[2])
    assert solver.describe() == ref.describe()
    return st, x.download(), sto, xo, solver


CASES = {  # poisson(...), kind, sweeps, passes, k_depth, max_levels[, k_min_rows]
    "poisson3d": ((24, 22, 12, 1.0), SolverKind.AmgPcg, 2, 1, 0, 24),
    "poisson3d-rows": ((24, 22, 12, 1.0), SolverKind.AmgPcg, 1, 1, 0, 6, 1000),
    "poisson3d-depth1": ((24, 22, 12, 1.0), SolverKind.AmgPcg, 1, 1, 1, 24),
    "anisotropic-random": ((16, 15, 9, 1000.0, True), SolverKind.AmgPcg, 1, 1, 0, 24),
    "three-passes": ((30, 28, 20, 1.0), SolverKind.AmgPcg, 5, 3, 0, 24),
    "convection": ((20, 18, 1, 1.0, False, False, 30.0), SolverKind.AmgBicgstab, 1, 1, 0, 5),
    "convection-3d": ((16, 14, 10, 1.0, False, False, 3.0), SolverKind.AmgBicgstab, 2, 1, 2, 24),
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("form", ["fused+graphs", "fused", "reference-kernels", "autotuned"])
def test_k_cycle_bitwise_vs_oracle(device, name, form):
    problem, kind, sweeps, passes, k_depth, levels, *rows = CASES[name]
    print(f"[k-cycle] {name} {form}", flush=True)
    a = oracle.poisson(*problem)
    b = np.random.default_rng(3).standard_normal(a.n_rows).astype(np.float32)
    options = AmgOptions(pre_sweeps=sweeps, post_sweeps=sweeps, aggregation_passes=passes, max_levels=levels,
                         cycle=AmgCycle.K, k_depth=k_depth, k_min_rows=rows[0] if rows else 0)
    st, x, sto, xo, _ = solve_both(device, a, b, kind, options, (1e-6, 1e-12, 400), fused=form != "reference-kernels",
                                   graphs=form == "fused+graphs", autotune=form == "autotuned")
    assert st.converged and sto["converged"]
    assert st.iterations == sto["iterations"]
    assert st.final_residual == pytest.approx(sto["final_residual"], rel=1e-12)
    assert np.array_equal(bits(x), bits(xo))


def test_k_cycle_neumann_projection_bitwise(device):
    """Constant-nullspace operator: every coarse correction is projected."""
    a = oracle.poisson(12, 10, 4, 1.0, False, True)
    b = np.zeros(a.n_rows, np.float32)
    b[0], b[-1] = 1.0, -1.0
    options = AmgOptions(aggregation_passes=1, nullspace=NullSpace.Constant, cycle=AmgCycle.K)
    st, x, sto, xo, _ = solve_both(device, a, b, SolverKind.AmgPcg, options, (1e-6, 1e-12, 400))
    assert st.converged and st.iterations == sto["iterations"]
    assert np.array_equal(bits(x), bits(xo))


def test_k_cycle_pressure_problem_bitwise(device):
    """The benchmark's pinned pressure operator (24 x 23 x 22 perturbed tets,
    symmetric M-matrix coarsening, 2 passes) with the K-cycle stopped at
    level 2 and V(2,2): the solve the fast pressure variant uses."""
    p = kutta_amd.PressureProblem(device, 24, 23, 22)
    b = oracle.neumann_rhs(p.n_cells, reference=p.reference)
    a = oracle.Csr.from_arrays(p["row_ptr"], p["col"], p["diag"], p["values"].astype(np.float64))
    options = AmgOptions(pre_sweeps=2, post_sweeps=2, aggregation_passes=2,
                         structure=OperatorStructure.SymmetricMMatrix, cycle=AmgCycle.K, k_depth=2)
    st, x, sto, xo, solver = solve_both(device, a, b, SolverKind.AmgPcg, options, (1e-6, 0.0, 2000),
                                        autotune=True)
    print(f"[k-cycle] pressure {p.n_cells} rows: {st.iterations} iterations, plan {solver.plan()}", flush=True)
    assert st.converged and st.iterations == sto["iterations"]
    assert np.array_equal(bits(x), bits(xo))


def test_k_cycle_resolve_after_refresh_is_bitwise(device):
    """Refresh with new values, then solve twice: the K-cycle's per-level
    scalars and status carry nothing from one solve into the next."""
    a = oracle.poisson(24, 22, 12)
    host = kutta_amd.CsrMatrixHost(a.row_ptr, a.col, a.diag, a.val)
    matrix = kutta_amd.upload(device, host)
    options = AmgOptions(pre_sweeps=2, post_sweeps=2, aggregation_passes=1, cycle=AmgCycle.K)
    solver = kutta_amd.make_amg_solver(SolverKind.AmgPcg, device, host, options)
    solver.refresh(matrix)
    b = device.upload(np.random.default_rng(8).standard_normal(a.n_rows).astype(np.float32))
    outs = []
    for _ in range(2):
        x = device.alloc(a.n_rows)
        x.fill(np.float32(0))
        st = solver.solve(matrix, x, b, SolveControl(1e-6, 1e-12, 400))
        outs.append((st.iterations, x.download()))
    assert outs[0][0] == outs[1][0]
    assert np.array_equal(bits(outs[0][1]), bits(outs[1][1]))


def test_full_size_pressure_variant_bitwise(device):
    """The benchmark's K-cycle variant on the 998,250-row C2 operator:
    V(1,1), 4 coarsest sweeps, 5 levels, inner corrections on the coarse
    levels of at least 50,000 rows (levels 1 and 2): the same iteration count
    and x bit for bit as the oracle."""
    p = kutta_amd.PressureProblem(device, 55, 55, 55, reference=-2)
    rng = np.random.default_rng(20261202)
    b64 = rng.standard_normal(p.n_cells)
    b64 -= b64.mean()
    b = b64.astype(np.float32)
    b[p.reference] = 0
    a = oracle.Csr.from_arrays(p["row_ptr"], p["col"], p["diag"], p["values"].astype(np.float64))
    options = AmgOptions(pre_sweeps=1, post_sweeps=1, coarse_sweeps=4, max_levels=5, aggregation_passes=2,
                         structure=OperatorStructure.SymmetricMMatrix, cycle=AmgCycle.K, k_min_rows=50000)
    st, x, sto, xo, solver = solve_both(device, a, b, SolverKind.AmgPcg, options, (1e-6, 0.0, 2000),
                                        autotune=True)
    print(f"[k-cycle] C2 variant: {st.iterations} iterations, plan {solver.plan()}", flush=True)
    assert st.converged and st.iterations == sto["iterations"]
    assert np.array_equal(bits(x), bits(xo))
