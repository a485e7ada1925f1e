"""The pressure V-cycle's paths for hierarchies that stop coarsening early
(anisotropic meshes, where pairwise aggregation stalls): a coarsest level too
large for the one-workgroup coarse tail is smoothed by one launch of all its
sweeps, aggregates of about two members are restricted 128 coarse rows per
workgroup, the prolongation runs folded into the next sweep or in place,
and levels with long rows are swept one thread or one wavefront per row. Both must leave the iterates bitwise equal to the oracle's
reference sequence (linalg/amg.cpp:283-310)."""
import numpy as np
import pytest

import kutta_amd
import oracle
from kutta_amd import AmgOptions, SolveControl, SolverKind

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("size,aniso,levels", [(40, 1, 4), (32, 100, 3)])
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
    solver.set_tail(0)
    solver.refresh(matrix)
    xb = device.upload(b)
    plans = [(True, mode, sweeps, prolong, 0) for mode in (0, 1, 2) for sweeps in (0, 1) for prolong in (0, 1)]
    plans += [(True, 0, 0, prolong, 1) for prolong in (0, 1)] + [(False, 0, 0, 0, 0)]
    for fused, mode, sweeps, prolong, sweep in plans:
        solver.set_fused(fused)
        solver.set_plan(mode, sweeps, prolong, sweep)
        x = device.alloc(a.n_rows)
        x.fill(np.float32(0))
        st = solver.solve(matrix, x, xb, SolveControl(1e-6, 1e-12, 300))
        case = (fused, solver.plan())
        assert st.iterations == st_ref["iterations"], case
        got = x.download()
        assert np.array_equal(got.view(np.uint32), x_ref.view(np.uint32)), case
    solver.set_fused(True)
    solver.autotune_tail(matrix, 2)
    print(f"[coarse paths] autotuned plan: {solver.plan()}", flush=True)
