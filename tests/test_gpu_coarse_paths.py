"""The pressure V-cycle's paths for hierarchies that stop coarsening early
(anisotropic meshes, where pairwise aggregation stalls): a coarsest level too
large for the one-workgroup coarse tail is smoothed by one launch of all its
sweeps (the operator read from global memory, or held in registers where the
level has at most 4,096 rows and 8,192 entries), aggregates of about two members are restricted 128 coarse rows per
workgroup (or as a residual, one thread or one wavefront per row, then the
member sums), the prolongation runs folded into the next sweep (one thread or
one wavefront per row) or in place, and levels with long rows are swept one
thread or one wavefront per row. All
must leave the iterates bitwise equal to the oracle's
reference sequence (linalg/amg.cpp:283-310)."""
import numpy as np
import pytest

import kutta_amd
import oracle
from kutta_amd import AmgOptions, SolveControl, SolverKind

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("size,aniso,levels", [(40, 1, 4), (32, 100, 3), (40, 1, 5)])
def test_large_coarsest_level_and_pair_aggregates_bitwise(device, size, aniso, levels):
    a = oracle.poisson(size, size, size, aniso)
    b = oracle.bench_rhs(a.n_rows)
    ref = oracle.Solver("AmgPcg", a, passes=1, max_levels=levels)
    x_ref, st_ref = ref.solve(b, rel_tol=1e-6, abs_tol=1e-12, max_iter=300)
    host = kutta_amd.CsrMatrixHost(a.row_ptr, a.col, a.diag, a.val)
    matrix = kutta_amd.upload(device, host)
    solver = kutta_amd.make_amg_solver(SolverKind.AmgPcg, device, host,
                                       AmgOptions(aggregation_passes=1, max_levels=levels))
    rows = [r for r, _ in solver.levels()]
    print(f"[coarse paths] {size}^3 aniso {aniso}: levels {rows}", flush=True)
    assert 2048 < rows[-1] <= kutta_amd.COARSE_SWEEPS_MAX_ROWS  # neither tail nor 24-row solve
    assert rows[0] <= 2 * rows[1]  # pair aggregates: the wide restriction
    resident = kutta_amd.resident_partition(ref.level_array(len(rows) - 1, "row_ptr"),
                                            ref.level_array(len(rows) - 1, "col")) is not None
    solver.set_tail(0)
    solver.refresh(matrix)
    xb = device.upload(b)
    plans = [(True, mode, sweeps, prolong, 0) for mode in (0, 1, 2, 3) for sweeps in (0, 1, 2) for prolong in (0, 1, 2)]
    plans += [(True, 0, 0, prolong, 1) for prolong in (0, 1, 2)] + [(False, 0, 0, 0, 0)]
    for fused, mode, sweeps, prolong, sweep in plans:
        solver.set_fused(fused)
        solver.set_plan(mode, sweeps, prolong, sweep)
        x = device.alloc(a.n_rows)
        x.fill(np.float32(0))
        st = solver.solve(matrix, x, xb, SolveControl(1e-6, 1e-12, 300))
        case = (fused, solver.plan())
        if fused and sweeps == 2:  # registers where the level's partition fits, else one workgroup
            assert ("register-resident sweeps" if resident else "one-workgroup sweeps") in case[1]
        assert st.iterations == st_ref["iterations"], case
        got = x.download()
        assert np.array_equal(got.view(np.uint32), x_ref.view(np.uint32)), case
    solver.set_fused(True)
    solver.autotune_tail(matrix, 2)
    print(f"[coarse paths] autotuned plan: {solver.plan()}", flush=True)


def test_register_resident_coarsest_sweeps_bitwise(device):
    """The tet pressure operator's coarsest level (522 rows, long rows where
    aggregation stalls) with the coarse tail off: its 32 sweeps a launch
    each, in one workgroup from global memory, and in one workgroup with the
    operator in registers all give the oracle's iterates bit for bit."""
    p = kutta_amd.PressureProblem(device, 40, 40, 40)
    a = oracle.Csr.from_arrays(p["row_ptr"], p["col"], p["diag"], p["values"].astype(np.float64))
    b = oracle.neumann_rhs(p.n_cells, reference=p.reference)
    ref = oracle.Solver("AmgPcg", a, passes=2, m_matrix=True)
    x_ref, st_ref = ref.solve(b, rel_tol=1e-6, abs_tol=0.0, max_iter=300)
    last = len(ref.levels()) - 1
    assert kutta_amd.resident_partition(ref.level_array(last, "row_ptr"), ref.level_array(last, "col")) is not None
    host = p.host
    solver = kutta_amd.make_amg_solver(SolverKind.AmgPcg, device, host,
                                       AmgOptions(aggregation_passes=2,
                                                  structure=kutta_amd.OperatorStructure.SymmetricMMatrix))
    solver.set_tail(0)
    solver.refresh(p)
    xb = device.upload(b)
    for sweeps, name in ((0, "sweep launches"), (1, "one-workgroup sweeps"), (2, "register-resident sweeps")):
        solver.set_plan(-1, sweeps)
        assert name in solver.plan()
        x = device.alloc(p.n_cells)
        x.fill(np.float32(0))
        st = solver.solve(p, x, xb, SolveControl(1e-6, 0.0, 300))
        print(f"[coarse paths] pressure 40^3, coarsest {name}: {st.iterations} iterations", flush=True)
        assert st.iterations == st_ref["iterations"]
        assert np.array_equal(x.download().view(np.uint32), x_ref.view(np.uint32)), name
