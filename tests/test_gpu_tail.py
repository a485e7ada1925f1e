"""The one-workgroup coarse-grid tail (KH_FUSED_COARSE_TAIL) against the
per-level launch sequence it replaces: bitwise-identical iterates, iteration
counts and residuals for every admissible first level, on hierarchies ending
in the small direct coarse solve and in coarse sweeps, with and without the
constant-nullspace projection, for AMG-PCG and AMG-BiCGStab."""
import numpy as np
import pytest

import kutta_amd
import oracle
from kutta_amd import AmgOptions, NullSpace, OperatorStructure, SolveControl, SolverKind

pytestmark = pytest.mark.gpu

CASES = [
    ("poisson 3 passes", lambda: oracle.poisson(30, 28, 26), SolverKind.AmgPcg, AmgOptions()),
    ("random 2 passes V(2,2)", lambda: oracle.poisson(33, 31, 29, 1.0, random=True), SolverKind.AmgPcg,
     AmgOptions(pre_sweeps=2, post_sweeps=2, aggregation_passes=2)),
    ("pressure M-matrix, constant nullspace", lambda: oracle.pressure_problem(16, 15, 14, reference=-1).csr(),
     SolverKind.AmgPcg,
     AmgOptions(aggregation_passes=2, nullspace=NullSpace.Constant,
                structure=OperatorStructure.SymmetricMMatrix)),
    ("pressure pinned, coarse sweeps", lambda: oracle.pressure_problem(18, 17, 16).csr(), SolverKind.AmgPcg,
     AmgOptions(aggregation_passes=2, coarse_rows=400, structure=OperatorStructure.SymmetricMMatrix)),
    ("bicgstab V(2,3)", lambda: oracle.poisson(24, 22, 1, 1.0, wind=2.0), SolverKind.AmgBicgstab,
     AmgOptions(pre_sweeps=2, post_sweeps=3, aggregation_passes=2)),
]


def run(device, solver, matrix, b, tail):
    used = solver.set_tail(tail)
    x = device.alloc(b.size)
    x.fill(np.float32(0))
    st = solver.solve(matrix, x, device.upload(b), SolveControl(1e-6, 0.0, 300))
    return used, st, x.download()


@pytest.mark.parametrize("name,make,kind,opts", CASES, ids=[c[0] for c in CASES])
def test_tail_matches_level_launches(device, name, make, kind, opts):
    print(f"[tail] {name}", flush=True)
    a = make()
    host = kutta_amd.CsrMatrixHost(a.row_ptr, a.col, a.diag, a.val)
    matrix = kutta_amd.upload(device, host)
    solver = kutta_amd.make_amg_solver(kind, device, host, opts)
    rng = np.random.default_rng(29)
    b64 = rng.standard_normal(a.n_rows)
    if opts.nullspace == NullSpace.Constant:
        b64 -= b64.mean()
    b = b64.astype(np.float32)
    levels = solver.levels()
    assert len(levels) >= 3, levels
    used, st0, x0 = run(device, solver, matrix, b, 0)
    assert used == 0
    tried = 0
    for first in [-1] + list(range(1, len(levels))):
        try:
            used, st, x = run(device, solver, matrix, b, first)
        except kutta_amd.KuttaError:
            assert first > 0
            continue
        if used == 0:
            continue
        tried += 1
        print(f"[tail] {name}: first level {used} of {len(levels)}", flush=True)
        assert st.iterations == st0.iterations, used
        assert st.final_residual == st0.final_residual, used
        assert np.array_equal(x.view(np.uint32), x0.view(np.uint32)), used
    assert tried >= 2


def test_tail_levels_and_limits(device):
    a = oracle.poisson(30, 28, 26)
    host = kutta_amd.CsrMatrixHost(a.row_ptr, a.col, a.diag, a.val)
    solver = kutta_amd.make_amg_solver(SolverKind.AmgPcg, device, host, AmgOptions())
    levels = solver.levels()
    assert levels[0][0] > kutta_amd.TAIL_MAX_ROWS
    auto = solver.set_tail(-1)
    assert auto >= 1
    assert all(rows <= kutta_amd.TAIL_MAX_ROWS for rows, _ in levels[auto:])
    with pytest.raises(kutta_amd.KuttaError):
        solver.set_tail(len(levels))
    assert solver.set_tail(0) == 0
    # A failed request keeps the previous choice.
    assert solver.set_tail(auto) == auto
    with pytest.raises(kutta_amd.KuttaError):
        solver.set_tail(len(levels) + 3)
    assert solver.set_tail(-1) == auto
