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


def test_to_bf16_rounds_like_the_oracle(device):
    """KH_TO_BF16 (the coarse-level value copies) against the oracle's
    round-to-nearest-even (itself pinned by closed forms): ties to even,
    carries into the exponent, subnormals, infinities and signed zeros; NaN
    stays NaN."""
    from oracle import learned as ol
    rng = np.random.default_rng(3)
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)
    special = np.array([0.0, -0.0, 1.0, -1.0, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, 1.9999999, 3.0e38, -3.4e38,
                        1e-40, -1e-42, np.inf, -np.inf], np.float32)
    values = np.concatenate([special, (rng.standard_normal(4000) * 10.0 ** rng.integers(-30, 30, 4000))
                             .astype(np.float32), one + ulp * rng.integers(0, 128, 500).astype(np.float32) / 2])
    src = device.upload(values.astype(np.float32))
    dst = device.alloc((values.size + 1) // 2, np.uint32)  # room for n uint16 entries
    device.launch(kutta_amd.Kernel.to_bf16, values.size, src, dst)
    got16 = dst.download().view(np.uint16)[:values.size]
    got = (got16.astype(np.uint32) << 16).view(np.float32)
    expected = ol.bf16(values.astype(np.float64)).astype(np.float32)
    assert np.array_equal(got.view(np.uint32), expected.view(np.uint32))
    nan = device.upload(np.array([np.nan], np.float32))
    device.launch(kutta_amd.Kernel.to_bf16, 1, nan, dst)
    assert np.isnan((dst.download().view(np.uint16)[:1].astype(np.uint32) << 16).view(np.float32)[0])


@pytest.mark.parametrize("kind,direct", [(SolverKind.AmgPcg, 0), (SolverKind.AmgFgmres, 0),
                                         (SolverKind.AmgPcg, 2048), (SolverKind.AmgFgmres, 2048)])
def test_bf16_coarse_levels_converge_within_the_bound(device, kind, direct):
    """bfloat16 values on the coarse levels (FP32 arithmetic; the coarse tail
    and the dense level keep FP32): the solve still reaches 1e-6, in at most a
    few more iterations than the FP32 V-cycle, and x satisfies the FP64
    residual bound on the perturbed tet operator the benchmark uses."""
    problem = oracle.pressure_problem(18, 18, 18)
    a = problem.csr()
    b = np.random.default_rng(5).standard_normal(a.n_rows).astype(np.float32)
    m0, plain, _ = make(device, a, kind, direct)
    st0, x0 = solve(device, m0, plain, b)
    host = kutta_amd.CsrMatrixHost(a.row_ptr, a.col, a.diag, a.val)
    m1 = kutta_amd.upload(device, host)
    opts = AmgOptions(aggregation_passes=2, structure=OperatorStructure.SymmetricMMatrix)
    mixed = kutta_amd.make_amg_solver(kind, device, host, opts)
    if direct:
        mixed.set_direct(direct)
    mixed.set_coarse_bf16(True)
    x = device.alloc(a.n_rows)
    x.fill(np.float32(0))
    with pytest.raises(kutta_amd.KuttaError):  # the copies are made by refresh
        mixed.solve(m1, x, device.upload(b), SolveControl(1e-6, 0.0, 500, 8))
    mixed.refresh(m1)
    assert "bfloat16" in mixed.describe()
    st1, x1 = solve(device, m1, mixed, b)
    print(f"[bf16 coarse] {kind.name} direct {direct}: {st0.iterations} -> {st1.iterations} iterations", flush=True)
    assert st1.converged and st1.final_residual <= 1e-6 * st1.initial_residual
    assert st1.iterations <= st0.iterations + 3
    import scipy.sparse as sp
    A = sp.csr_matrix((a.val, a.col, a.row_ptr), shape=(a.n_rows, a.n_rows))
    res = np.linalg.norm(b - A @ x1.astype(np.float64)) / np.linalg.norm(b)
    assert res < 2e-6


@pytest.mark.parametrize("nx,ny,nz", [(1, 1, 1), (7, 9, 1), (8, 8, 1), (5, 13, 1), (10, 10, 2), (9, 9, 9)])
def test_blocked_gauss_jordan_inverse(device, nx, ny, nz):
    """KH_DENSE_INVERT (blocked Gauss-Jordan, 64-pivot blocks, FP64 MFMA
    products) against numpy's FP64 inverse of the same operator, on sizes
    below, at and across block boundaries (1, 63, 64, 65, 200 and 729 rows):
    entrywise within a conditioning bound, and A A^-1 = I."""
    a = oracle.poisson(nx, ny, nz, 1.0, random=True)
    n = a.n_rows
    print(f"[dense invert] {n} rows", flush=True)
    A = dense_of(a).astype(np.float32).astype(np.float64)  # the FP32 values the device expands
    rows = device.upload(np.asarray(a.row_ptr, np.int32), np.int32)
    cols = device.upload(np.asarray(a.col, np.int32), np.int32)
    vals = device.upload(np.asarray(a.val, np.float32))
    inverse = device.alloc(n * n * 2, np.float32)  # n*n doubles
    scratch = device.alloc(4096 * 2, np.float32)   # 4096 doubles
    device.launch(kutta_amd.Kernel.dense_invert, n, rows, cols, vals, inverse, scratch)
    got = inverse.download().view(np.float64).reshape(n, n)
    ref = np.linalg.inv(A)
    cond = np.linalg.cond(A)
    assert np.abs(got - ref).max() <= 1e-13 * cond * np.abs(ref).max()
    assert np.abs(A @ got - np.eye(n)).max() <= 1e-12 * cond


def test_lazy_dense_refresh(device):
    """Lazy refresh of the dense level's inverse: an unchanged operator keeps
    it (the solve bit for bit as before); values 0.1% away keep it too and
    the solve still converges, close to the eagerly refreshed solve; values
    5% away recompute it, and the solve then equals the eager solver's bit for
    bit (the same inverse of the same values)."""
    p = kutta_amd.PressureProblem(device, 24, 23, 22)
    b = device.upload(oracle.neumann_rhs(p.n_cells, reference=p.reference))
    opts = AmgOptions(aggregation_passes=2, structure=OperatorStructure.SymmetricMMatrix)
    control = SolveControl(1e-6, 0.0, 400)
    m = kutta_amd.upload(device, p.host)
    vals = m.val.download()
    lazy = kutta_amd.make_amg_solver(SolverKind.AmgPcg, device, p.host, opts)
    eager = kutta_amd.make_amg_solver(SolverKind.AmgPcg, device, p.host, opts)
    assert lazy.set_direct(8192) == eager.set_direct(8192) > 0
    with pytest.raises(kutta_amd.KuttaError):
        lazy.set_dense_tolerance(-1.0)
    lazy.set_dense_tolerance(1e-2)

    def solve(solver):
        x = device.alloc(p.n_cells)
        x.fill(np.float32(0))
        st = solver.solve(m, x, b, control)
        assert st.converged
        return st, x.download()

    lazy.refresh(m)
    assert lazy.dense_refreshes()[:2] == (1, 0)
    st0, x0 = solve(lazy)
    lazy.refresh(m)
    inv, kept, dist = lazy.dense_refreshes()
    assert (inv, kept) == (1, 1) and dist == 0.0
    st1, x1 = solve(lazy)
    assert st1.iterations == st0.iterations and np.array_equal(x1.view(np.uint32), x0.view(np.uint32))
    m.val.upload((vals * np.float32(1.001)).astype(np.float32))
    lazy.refresh(m)
    inv, kept, dist = lazy.dense_refreshes()
    assert (inv, kept) == (1, 2) and 5e-4 < dist < 2e-3
    st2, x2 = solve(lazy)
    eager.refresh(m)
    ste, xe = solve(eager)
    assert abs(st2.iterations - ste.iterations) <= 2
    assert np.abs(x2.astype(np.float64) - xe).max() <= 1e-4 * np.abs(xe).max()
    m.val.upload((vals * np.float32(1.05)).astype(np.float32))
    lazy.refresh(m)
    inv, kept, dist = lazy.dense_refreshes()
    assert (inv, kept) == (2, 2) and dist == 0.0
    st3, x3 = solve(lazy)
    eager.refresh(m)
    ste, xe = solve(eager)
    assert st3.iterations == ste.iterations and np.array_equal(x3.view(np.uint32), xe.view(np.uint32))
