"""The HIP-native Krylov/AMG solver end to end against the oracle and the
reference's golden bench_linalg results (bench_linalg.log:1-12)."""
import numpy as np
import pytest

import kutta_amd
import oracle
from kutta_amd import AmgOptions, NullSpace, OperatorStructure, SolveControl, SolverKind

pytestmark = pytest.mark.gpu


def g9(v):
    return f"{v:.9g}"


def product_setup(device, a, kind, options=None):
    host = kutta_amd.CsrMatrixHost(a.row_ptr, a.col, a.diag, a.val)
    matrix = kutta_amd.upload(device, host)
    if kind in (SolverKind.AmgPcg, SolverKind.AmgBicgstab):
        solver = kutta_amd.make_amg_solver(kind, device, host, options)
    else:
        solver = kutta_amd.make_solver(kind, device, host)
    return host, matrix, solver


def bench_flow(device, size, aniso, kind, fused=True):
    """bench/bench_linalg.cpp:42-83 on the HIP backend."""
    a = oracle.poisson(size, size, size, aniso)
    host, matrix, solver = product_setup(device, a, kind)
    solver.set_fused(fused)
    vals = a.val.astype(np.float32)
    matrix.val.upload(vals * np.float32(1.25))
    solver.refresh(matrix)
    matrix.val.upload(vals)
    solver.refresh(matrix)
    b = device.upload(oracle.bench_rhs(a.n_rows))
    x = device.alloc(a.n_rows)
    x.fill(np.float32(0))
    st = solver.solve(matrix, x, b, SolveControl(1e-6, 1e-12, 2000))
    return st, solver.describe(), x.download()


# validation/results/colab_t4/logs/bench_linalg.log:1-12 (every case the log holds).
GOLDEN = [
    (64, 1, SolverKind.PcgJacobi, 278, "6.85768115e-07", "288.643143", "0.000197942264", True, "PCG, Jacobi"),
    (64, 1, SolverKind.AmgPcg, 15, "7.91394974e-07", "288.643143", "0.000228430732", True,
     "6 levels, op complexity 1.140"),
    (128, 1, SolverKind.PcgJacobi, 640, "7.17042009e-07", "816.827145", "0.000585699377", True, "PCG, Jacobi"),
    (128, 1, SolverKind.AmgPcg, 21, "8.66568009e-07", "816.827145", "0.000707836272", True,
     "7 levels, op complexity 1.142"),
    (64, 1000, SolverKind.PcgJacobi, 354, "1.53025686e-06", "288.643143", "0.00044169815", False, "PCG, Jacobi"),
    (64, 1000, SolverKind.AmgPcg, 38, "1.51536251e-06", "288.643143", "0.000437398999", False,
     "6 levels, op complexity 1.134"),
]


@pytest.mark.parametrize("size,aniso,kind,its,rel,initial,final,conv,desc", GOLDEN)
def test_bench_linalg_golden_on_gpu(device, size, aniso, kind, its, rel, initial, final, conv, desc):
    print(f"[golden] {size}^3 aniso={aniso} {kind.name}", flush=True)
    st, describe, _ = bench_flow(device, size, aniso, kind)
    assert st.iterations == its
    assert g9(st.final_residual / st.initial_residual) == rel
    assert g9(st.initial_residual) == initial
    assert g9(st.final_residual) == final
    assert st.converged == conv
    assert describe.endswith(desc)


def test_fused_and_reference_sequences_agree(device):
    st1, _, x1 = bench_flow(device, 40, 1, SolverKind.AmgPcg, fused=True)
    st2, _, x2 = bench_flow(device, 40, 1, SolverKind.AmgPcg, fused=False)
    assert st1.iterations == st2.iterations
    assert st1.final_residual == st2.final_residual
    assert np.array_equal(x1.view(np.uint32), x2.view(np.uint32))


def oracle_solve(a, kind, b, control, **opts):
    s = oracle.Solver(kind, a, **opts)
    x, st = s.solve(b, rel_tol=control.rel_tol, abs_tol=control.abs_tol, max_iter=control.max_iter,
                    interval=control.residual_check_interval)
    return x, st, s


CASES = [
    ("poisson 3 passes", lambda: oracle.poisson(30, 28, 26), dict(), AmgOptions()),
    ("random 2 passes V(2,2)", lambda: oracle.poisson(33, 31, 29, 1.0, random=True),
     dict(pre=2, post=2, passes=2), AmgOptions(pre_sweeps=2, post_sweeps=2, aggregation_passes=2)),
    ("neumann M-matrix, constant nullspace",
     lambda: oracle.pressure_problem(16, 15, 14).csr(),
     dict(passes=2, nullspace=True, m_matrix=True),
     AmgOptions(aggregation_passes=2, nullspace=NullSpace.Constant,
                structure=OperatorStructure.SymmetricMMatrix)),
]


@pytest.mark.parametrize("name,make,oopts,popts", CASES, ids=[c[0] for c in CASES])
def test_amg_pcg_matches_oracle(device, name, make, oopts, popts):
    print(f"[parity] {name}", flush=True)
    a = make()
    rng = np.random.default_rng(17)
    b64 = rng.standard_normal(a.n_rows)
    if popts.nullspace == NullSpace.Constant:
        b64 -= b64.mean()
    b = b64.astype(np.float32)
    control = SolveControl(1e-6, 0.0, 300)
    xo, sto, so = oracle_solve(a, "AmgPcg", b, control, **oopts)
    host, matrix, solver = product_setup(device, a, SolverKind.AmgPcg, popts)
    assert solver.describe() == so.describe()
    assert solver.levels() == so.levels()
    db = device.upload(b)
    x = device.alloc(a.n_rows)
    x.fill(np.float32(0))
    st = solver.solve(matrix, x, db, control)
    assert st.iterations == sto["iterations"]
    assert st.converged == sto["converged"]
    xp = x.download()
    assert np.array_equal(xp.view(np.uint32), xo.view(np.uint32))
    assert st.final_residual == pytest.approx(sto["final_residual"], rel=1e-12)


def test_hierarchy_and_refresh_match_oracle(device):
    """Aggregates, transfers and the device-refreshed coarse operators
    (amg_galerkin + amg_m_matrix_diagonal) equal the oracle's bit for bit."""
    a = oracle.pressure_problem(13, 12, 11).csr()
    popts = AmgOptions(aggregation_passes=2, nullspace=NullSpace.Constant,
                       structure=OperatorStructure.SymmetricMMatrix)
    host, matrix, solver = product_setup(device, a, SolverKind.AmgPcg, popts)
    so = oracle.Solver("AmgPcg", a, passes=2, nullspace=True, m_matrix=True)
    vals = a.val.astype(np.float32)
    scaled = vals * np.float32(1.13)
    matrix.val.upload(scaled)
    solver.refresh(matrix)
    so.refresh(scaled)
    levels = solver.levels()
    assert len(levels) >= 3
    for lvl in range(len(levels) - 1):
        for which in ("aggregate", "member_offsets", "members", "entry_offsets", "entries"):
            assert np.array_equal(solver.transfer(lvl, which), so.transfer(lvl, which)), (lvl, which)
    for lvl in range(1, len(levels)):
        assert np.array_equal(solver.level_values(lvl).view(np.uint32),
                              so.level_values(lvl).view(np.uint32)), lvl


def test_bicgstab_matches_oracle(device):
    a = oracle.poisson(24, 22, 1, 1.0, wind=2.0)
    rng = np.random.default_rng(3)
    b = rng.standard_normal(a.n_rows).astype(np.float32)
    control = SolveControl(1e-5, 0.0, 200)
    for kind, ok in ((SolverKind.BicgstabJacobi, "BicgstabJacobi"), (SolverKind.AmgBicgstab, "AmgBicgstab")):
        xo, sto, _ = oracle_solve(a, ok, b, control, passes=2, pre=2, post=2)
        opts = AmgOptions(pre_sweeps=2, post_sweeps=2, aggregation_passes=2)
        host, matrix, solver = product_setup(device, a, kind, opts)
        x = device.alloc(a.n_rows)
        x.fill(np.float32(0))
        st = solver.solve(matrix, x, device.upload(b), control)
        assert st.iterations == sto["iterations"], kind
        assert np.array_equal(x.download().view(np.uint32), xo.view(np.uint32)), kind


def test_failure_paths_are_host_forced(device):
    """solver.cpp:97-176 failure handling without any GPU fault: a NaN RHS
    raises NonfiniteError; max_iter=0 returns the initial residual."""
    a = oracle.poisson(10, 10, 10)
    host, matrix, solver = product_setup(device, a, SolverKind.AmgPcg)
    x = device.alloc(a.n_rows)
    x.fill(np.float32(0))
    bad = np.ones(a.n_rows, np.float32)
    bad[3] = np.nan
    with pytest.raises(kutta_amd.NonfiniteError):
        solver.solve(matrix, x, device.upload(bad), SolveControl(1e-6, 0, 10))
    st = solver.solve(matrix, x, device.upload(np.ones(a.n_rows, np.float32)), SolveControl(1e-6, 0, 0))
    assert st.iterations == 0 and not st.converged
    assert st.final_residual == st.initial_residual


@pytest.mark.parametrize("nx,ny,nz,wind,rel,interval,max_iter", [
    (24, 22, 1, 2.0, 1e-5, 8, 200),
    (40, 30, 1, 5.0, 1e-7, 8, 400),    # long enough for the 50-update replacement
    (12, 10, 8, 1.0, 1e-6, 1, 200),
    (30, 30, 1, 3.0, 1e-2, 3, 100),    # the momentum tolerance of the SIMPLEC step
])
def test_bicgstab_jacobi_fused_equals_reference_sequence(device, nx, ny, nz, wind, rel, interval, max_iter):
    """The five-launch fused BiCGStab-Jacobi iteration gives the same
    iterations, final residual and x, bit for bit, as the reference kernel
    sequence (linalg/solver.cpp:314-361) and as the oracle."""
    print(f"[bicgstab] {nx}x{ny}x{nz} wind {wind} rel {rel} interval {interval}", flush=True)
    a = oracle.poisson(nx, ny, nz, 1.0, wind=wind)
    b = np.random.default_rng(nx * ny + nz).standard_normal(a.n_rows).astype(np.float32)
    control = SolveControl(rel, 0.0, max_iter, interval)
    xo, sto, _ = oracle_solve(a, "BicgstabJacobi", b, control)
    results = []
    for fused in (True, False):
        host, matrix, solver = product_setup(device, a, SolverKind.BicgstabJacobi, AmgOptions())
        solver.set_fused(fused)
        x = device.alloc(a.n_rows)
        x.fill(np.float32(0))
        st = solver.solve(matrix, x, device.upload(b), control)
        results.append((st, x.download()))
    for st, x in results:
        assert st.iterations == sto["iterations"]
        assert st.final_residual == pytest.approx(sto["final_residual"], rel=1e-12)
        assert np.array_equal(x.view(np.uint32), xo.view(np.uint32))
    assert results[0][0].final_residual == results[1][0].final_residual


@pytest.mark.parametrize("interval", [8, 3])
def test_batched_bicgstab_equals_individual_solves(device, interval):
    """Four systems in one batched BiCGStab-Jacobi solve (KH_BATCH_* launches,
    no reference counterpart): different operators (wind, anisotropy), right
    -hand sides and iteration counts, one that needs no iteration (b = 0), and
    a second solve of each that starts from the first's adaptive batch
    interval. Every x and every SolveStats equals its own solve's, bit for
    bit; the oracle pins the single solves."""
    n = 26 * 24
    ops = [oracle.poisson(26, 24, 1, 1.0, wind=2.0), oracle.poisson(26, 24, 1, 3.0, wind=5.0),
           oracle.poisson(26, 24, 1, 1.0, random=True, wind=1.0), oracle.poisson(26, 24, 1, 1.0, wind=0.5)]
    rng = np.random.default_rng(interval)
    rhs = [rng.standard_normal(n).astype(np.float32) for _ in range(3)] + [np.zeros(n, np.float32)]
    controls = SolveControl(1e-6, 0.0, 300, interval)
    print(f"[batched bicgstab] 4 lanes, interval {interval}", flush=True)

    def fresh():
        out = []
        for a in ops:
            _, matrix, solver = product_setup(device, a, SolverKind.BicgstabJacobi, AmgOptions())
            out.append((matrix, solver))
        return out

    single, batched = fresh(), fresh()
    xs_single = [device.alloc(n) for _ in ops]
    xs_batched = [device.alloc(n) for _ in ops]
    bs = [device.upload(b) for b in rhs]
    for round_ in range(2):
        for x in xs_single + xs_batched:
            x.fill(np.float32(0))
        st_single = [solver.solve(matrix, x, b, controls) for (matrix, solver), x, b in zip(single, xs_single, bs)]
        st_batched = kutta_amd.solve_batched([s for _, s in batched], [m for m, _ in batched], xs_batched, bs,
                                             controls)
        for k in range(len(ops)):
            a, b = st_single[k], st_batched[k]
            assert (a.iterations, a.initial_residual, a.final_residual, a.converged) == \
                   (b.iterations, b.initial_residual, b.final_residual, b.converged), (round_, k)
            assert np.array_equal(xs_single[k].download().view(np.uint32),
                                  xs_batched[k].download().view(np.uint32)), (round_, k)
        its = [s.iterations for s in st_single]
        assert its[3] == 0 and len(set(its[:3])) > 1  # the lanes really differ
    xo, sto, _ = oracle_solve(ops[0], "BicgstabJacobi", rhs[0], controls)
    assert st_single[0].iterations == sto["iterations"]
    assert np.array_equal(xs_batched[0].download().view(np.uint32), xo.view(np.uint32))


def test_pcg_graph_replay_equals_launches(device):
    """Fused PCG iterations replayed from a captured HIP graph (the default)
    against the same solver launching kernel by kernel: bitwise x, iterations
    and residuals over several solves, with the operator's values changed in
    place and refreshed between them (the graph is reused), a different
    coarse tail (the graph is recaptured), and a solve that stops at
    max_iter (its last iteration runs outside the graph). The oracle pins the
    first solve."""
    a = oracle.pressure_problem(20, 19, 18).csr()
    n = a.n_rows
    opts = AmgOptions(aggregation_passes=2, structure=OperatorStructure.SymmetricMMatrix)
    host = kutta_amd.CsrMatrixHost(a.row_ptr, a.col, a.diag, a.val)
    runs = []
    for graphs in (True, False):
        matrix = kutta_amd.upload(device, host)
        solver = kutta_amd.make_amg_solver(SolverKind.AmgPcg, device, host, opts)
        solver.set_graphs(graphs)
        solver.refresh(matrix)
        out = []
        rng = np.random.default_rng(9)
        for k in range(5):
            if k == 2:  # scale the values in place (same buffers) and refresh
                matrix.val.upload((a.val * 1.5).astype(np.float32))
                solver.refresh(matrix)
            if k == 3:
                solver.set_tail(0)
            b = rng.standard_normal(n).astype(np.float32)
            x = device.alloc(n)
            x.fill(np.float32(0))
            control = SolveControl(1e-6, 0.0, 7 if k == 4 else 300, 8)
            st = solver.solve(matrix, x, device.upload(b), control)
            out.append((st.iterations, st.final_residual, st.converged, x.download()))
        runs.append(out)
    for (i0, r0, c0, x0), (i1, r1, c1, x1) in zip(*runs):
        assert (i0, r0, c0) == (i1, r1, c1)
        assert np.array_equal(x0.view(np.uint32), x1.view(np.uint32))
    assert runs[0][4][0] == 7 and not runs[0][4][2]
