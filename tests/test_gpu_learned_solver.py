"""The learned coarse correction inside AMG-FGMRES (no reference
counterpart): pairs harvested by the solver itself, exact level corrections,
training on the device, and the comparison with the plain V-cycle on the same
right-hand sides. Every learned solve is judged by its true residual (FP32
product, FP64 norm, as the plain path) and by an FP64 host residual; the
fallback rule is pinned at its two extremes."""
import time

import numpy as np
import pytest
import scipy.sparse as sp

import kutta_amd
import oracle
from kutta_amd import AmgOptions, OperatorStructure, SolveControl, SolverKind

pytestmark = pytest.mark.gpu

REL = 1e-6


def rhs(n, seed):
    b = np.random.default_rng(seed).standard_normal(n).astype(np.float32)
    b[0] = 0
    return b


@pytest.fixture(scope="module")
def trained(device):
    a = oracle.pressure_problem(24, 24, 24, tet=True, amplitude=0.2, reference=0).csr()
    print(f"[learned solver] {a.n_rows} rows", flush=True)
    host = kutta_amd.CsrMatrixHost(a.row_ptr, a.col, a.diag, a.val)
    matrix = kutta_amd.upload(device, host)
    opts = AmgOptions(aggregation_passes=2, structure=OperatorStructure.SymmetricMMatrix)
    solver = kutta_amd.make_amg_solver(SolverKind.AmgFgmres, device, host, opts)
    solver.refresh(matrix)
    solver.set_learned(level=2, layers=3, seed=0, capacity=96)
    solver.set_learned_mode(1)
    for k in range(6):
        x = device.alloc(a.n_rows)
        x.fill(np.float32(0))
        st = solver.solve(matrix, x, device.upload(rhs(a.n_rows, 100 + k)), SolveControl(REL, 0.0, 300, 8))
        assert st.converged
    harvested = solver.learned_report()["harvested"]
    pairs = solver.learned_targets(1e-6)
    train_pairs = (4 * pairs) // 5  # the rest are held out
    losses, relative = solver.learned_train(400, 1e-3, train_pairs)
    held_out = solver.learned_evaluate(train_pairs, pairs - train_pairs)
    solver.set_learned_mode(0)
    A = sp.csr_matrix((a.val, a.col, a.row_ptr), shape=(a.n_rows, a.n_rows))
    return dict(a=a, A=A, matrix=matrix, solver=solver, harvested=harvested, pairs=pairs, losses=losses,
                relative=relative, train_pairs=train_pairs, held_out=held_out)


def run(device, t, b, mode, reduction=0.5, max_relative=np.inf):
    solver = t["solver"]
    solver.set_learned_mode(mode, reduction, max_relative)
    x = device.alloc(len(b))
    x.fill(np.float32(0))
    bb = device.upload(b)
    device.synchronize()
    t0 = time.perf_counter()
    st = solver.solve(t["matrix"], x, bb, SolveControl(REL, 0.0, 300, 8))
    device.synchronize()
    ms = (time.perf_counter() - t0) * 1e3
    solver.set_learned_mode(0)
    return st, x.download(), ms


def test_harvest_targets_and_training(trained):
    rep = trained["solver"].learned_report()
    assert rep["level"] == 2 and rep["rows"] > 0
    # Six solves of 10-30 FGMRES steps each: every built step is one pair.
    assert 30 <= trained["harvested"] <= 96
    assert trained["pairs"] == trained["harvested"]
    losses, relative = trained["losses"], trained["relative"]
    assert losses.shape == (400,) and np.all(np.isfinite(losses))
    assert relative.shape == (400,) and np.all(relative > 0)
    print(f"[learned solver] loss {losses[:20].mean():.4g} -> {losses[-20:].mean():.4g}, "
          f"unexplained {relative[:20].mean():.3f} -> {relative[-20:].mean():.3f}", flush=True)
    assert losses[-20:].mean() < losses[:20].mean()
    assert relative[-20:].mean() < relative[:20].mean()
    # The report's figure is the mean over the last epoch (one step per
    # training pair).
    k = trained["train_pairs"]
    assert rep["relative"] == pytest.approx(relative[-k:].mean(), rel=1e-12)


def test_held_out_pairs(trained):
    """Pairs the network never trained on: its unexplained fraction there is
    below one (it generalizes beyond its training residuals) and no better
    than a few times its training figure."""
    held, train = trained["held_out"], trained["solver"].learned_report()["relative"]
    print(f"[learned solver] unexplained: training {train:.3f}, held out {held:.3f} "
          f"({trained['pairs'] - trained['train_pairs']} pairs)", flush=True)
    assert 0 < held < 1
    assert held > 0.5 * train


def test_learned_solves_converge_and_compare_with_plain(device, trained):
    """A/B on three unseen right-hand sides: the plain V-cycle and the learned
    stage (never dropped: reduction 0, no gate) both reach 1e-6 by the solver's own
    true residual and by an FP64 host residual. On this operator (level 2:
    5k nodes) the network explains most of its targets, and the stage takes
    no more FGMRES steps than the plain V-cycle on any right-hand side and
    fewer in total; the times are printed (the stage's launches cost more
    than the steps it saves)."""
    A = trained["A"]
    before = trained["solver"].learned_report()
    plain_it, learned_it = [], []
    for seed in (7, 8, 9):
        b = rhs(A.shape[0], seed)
        st0, x0, ms0 = run(device, trained, b, 0)
        st1, x1, ms1 = run(device, trained, b, 2, reduction=0.0)
        for st, x in ((st0, x0), (st1, x1)):
            assert st.converged and st.final_residual <= REL * st.initial_residual
            res = np.linalg.norm(b - A @ x.astype(np.float64)) / np.linalg.norm(b)
            assert res < 2 * REL
        plain_it.append(st0.iterations)
        learned_it.append(st1.iterations)
        print(f"[learned A/B] seed {seed}: plain {st0.iterations} it {ms0:.2f} ms, "
              f"learned {st1.iterations} it {ms1:.2f} ms", flush=True)
    assert all(l <= p for l, p in zip(learned_it, plain_it)) and sum(learned_it) < sum(plain_it)
    after = trained["solver"].learned_report()
    assert after["solves"] == before["solves"] + 3
    assert after["fallbacks"] == before["fallbacks"]
    assert after["applications"] - before["applications"] >= sum(learned_it)


def test_fallback_rule_extremes(device, trained):
    """Requiring a 99% fall per step drops the stage at the first poll (step
    8) of every solve that lasts longer; requiring none never drops it. Either
    way the solve converges (FGMRES accepts a changing preconditioner)."""
    b = rhs(trained["A"].shape[0], 12)
    before = trained["solver"].learned_report()
    st, _, _ = run(device, trained, b, 2, reduction=0.99)
    assert st.converged and st.iterations > 8
    mid = trained["solver"].learned_report()
    assert mid["fallbacks"] == before["fallbacks"] + 1
    # The stage ran for the 8 steps before the poll and was then dropped.
    assert mid["applications"] - before["applications"] == 8
    st, _, _ = run(device, trained, b, 2, reduction=0.0)
    assert st.converged
    assert trained["solver"].learned_report()["fallbacks"] == mid["fallbacks"]


def test_relative_loss_gate(device, trained):
    """A network that left more of its targets unexplained than allowed is
    not applied: the solve is the plain V-cycle's, iteration for iteration."""
    b = rhs(trained["A"].shape[0], 13)
    st0, x0, _ = run(device, trained, b, 0)
    before = trained["solver"].learned_report()
    st1, x1, _ = run(device, trained, b, 2, reduction=0.0, max_relative=1e-9)
    after = trained["solver"].learned_report()
    assert after["gated"] == before["gated"] + 1
    assert after["applications"] == before["applications"]
    assert st1.iterations == st0.iterations and np.array_equal(x0, x1)


def test_learned_stage_needs_amg_fgmres(device):
    a = oracle.pressure_problem(6, 6, 6, tet=False, amplitude=0.0, reference=0).csr()
    host = kutta_amd.CsrMatrixHost(a.row_ptr, a.col, a.diag, a.val)
    pcg = kutta_amd.make_amg_solver(SolverKind.AmgPcg, device, host,
                                    AmgOptions(structure=OperatorStructure.SymmetricMMatrix))
    with pytest.raises(kutta_amd.KuttaError):
        pcg.set_learned(level=1)
    fg = kutta_amd.make_amg_solver(SolverKind.AmgFgmres, device, host,
                                   AmgOptions(structure=OperatorStructure.SymmetricMMatrix))
    with pytest.raises(kutta_amd.KuttaError):
        fg.set_learned_mode(2)  # not configured
    with pytest.raises(kutta_amd.KuttaError):
        fg.set_learned(level=99)
