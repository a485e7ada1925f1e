"""Every kernel id of the HIP backend against the serial oracle on the same
seeded inputs: bitwise for all per-element arithmetic (no FP64 reduction
order involved), and the fused kernels bitwise against the reference kernel
sequence they replace."""
import numpy as np
import pytest

import oracle
from helpers import mixed_magnitude_vector, random_operator, random_vector
from kutta_amd import Kernel

pytestmark = pytest.mark.gpu


def bits(a):
    return np.asarray(a).view(np.uint32)


def same(got, want):
    assert got.shape == want.shape
    assert np.array_equal(bits(got), bits(want)), np.flatnonzero(bits(got) != bits(want))[:10]


class Op:
    """A seeded operator on host and device."""

    def __init__(self, device, seed=1, shape=(23, 19, 17), neumann=False):
        self.a = oracle.poisson(*shape, 1.0, random=True, neumann=neumann)
        self.n = self.a.n_rows
        self.vals = self.a.val.astype(np.float32)
        self.rp = device.upload(self.a.row_ptr)
        self.col = device.upload(self.a.col)
        self.diag = device.upload(self.a.diag)
        self.v = device.upload(self.vals)


@pytest.fixture(scope="module")
def op(device):
    return Op(device)


@pytest.fixture(scope="module")
def status0(device):
    s = device.alloc(3, np.int32)
    s.fill(np.int32(0))
    return s


@pytest.fixture(scope="module")
def status1(device):
    s = device.alloc(3, np.int32)
    s.upload(np.array([1, 0, 0], np.int32))
    return s


def test_spmv_and_residual(device, op, status0, status1):
    n = op.n
    x = mixed_magnitude_vector(n, 1)
    b = mixed_magnitude_vector(n, 2)
    dx, db, dy = device.upload(x), device.upload(b), device.alloc(n)
    device.launch(Kernel.csr_spmv, n, op.rp, op.col, op.v, dx, dy)
    want = np.zeros(n, np.float32)
    oracle.kernel("csr_spmv", n, op.a.row_ptr, op.a.col, op.vals, x, want)
    same(dy.download(), want)
    device.launch(Kernel.csr_residual, n, op.rp, op.col, op.v, dx, db, dy)
    oracle.kernel("csr_residual", n, op.a.row_ptr, op.a.col, op.vals, x, b, want)
    same(dy.download(), want)
    # guarded forms: run when status[0] == 0, skip otherwise
    dy.fill(np.float32(7))
    device.launch(Kernel.csr_spmv_guarded, n, op.rp, op.col, op.v, dx, dy, status1)
    device.launch(Kernel.csr_residual_guarded, n, op.rp, op.col, op.v, dx, db, dy, status1)
    assert np.all(dy.download() == 7)
    device.launch(Kernel.csr_residual_guarded, n, op.rp, op.col, op.v, dx, db, dy, status0)
    same(dy.download(), want)


def test_krylov_diagonal(device, op):
    n = op.n
    inv, bad = device.alloc(n), device.alloc(n)
    for l1 in (0, 1):
        device.launch(Kernel.krylov_diagonal, n, op.rp, op.diag, op.v, inv, bad, l1)
        wi, wb = np.zeros(n, np.float32), np.zeros(n, np.float32)
        oracle.kernel("krylov_diagonal", n, op.a.row_ptr, op.a.diag, op.vals, wi, wb, l1)
        same(inv.download(), wi)
        same(bad.download(), wb)
    # invalid codes: NaN entry -> 2, zero diagonal -> 1, negative diagonal with l1 -> 3
    vals = op.vals.copy()
    vals[op.a.row_ptr[3] + 1] = np.nan
    vals[op.a.diag[5]] = 0
    vals[op.a.diag[9]] = -1
    dv = device.upload(vals)
    for l1 in (0, 1):
        device.launch(Kernel.krylov_diagonal, n, op.rp, op.diag, dv, inv, bad, l1)
        wi, wb = np.zeros(n, np.float32), np.zeros(n, np.float32)
        oracle.kernel("krylov_diagonal", n, op.a.row_ptr, op.a.diag, vals, wi, wb, l1)
        same(inv.download(), wi)
        same(bad.download(), wb)
        assert wb[3] == 2 and wb[5] == 1 and wb[9] == (3 if l1 else 0)


def test_elementwise_known_answers(device):
    """tests/linalg/test_kernels.cpp:11-28, exact values."""
    r = device.upload(np.array([2, -4, 6], np.float32))
    x = device.upload(np.array([1, 2, 3], np.float32))
    inv = device.upload(np.array([0.5, 0.25, 2], np.float32))
    out = device.alloc(3)
    device.launch(Kernel.krylov_multiply, 3, r, inv, out)
    assert out.download().tolist() == [1, -1, 12]
    device.launch(Kernel.krylov_smooth, 3, r, inv, x, np.float32(0.5))
    assert x.download().tolist() == [1.5, 1.5, 9]
    device.launch(Kernel.krylov_shift, 3, x, np.float32(1.0))
    assert x.download().tolist() == [0.5, 0.5, 8]
    device.launch(Kernel.krylov_direction, 3, r, x, np.float32(2))
    assert x.download().tolist() == [3, -3, 22]
    device.launch(Kernel.krylov_step, 3, x, out, r, inv, np.float32(2))
    assert x.download().tolist() == [7, -11, 34]
    assert out.download().tolist() == [0, -1.5, 8]
    device.launch(Kernel.krylov_bicg_direction, 3, r, inv, x, np.float32(0.5), np.float32(2))
    assert x.download().tolist() == [5, -9.75, 21]
    device.launch(Kernel.krylov_difference, 3, r, inv, out, np.float32(2))
    assert out.download().tolist() == [1, -4.5, 2]
    device.launch(Kernel.krylov_bicg_step, 3, x, r, inv, np.float32(2), np.float32(4))
    assert x.download().tolist() == [11, -16.75, 41]


def test_elementwise_random_bitwise(device):
    """BLAS-1 and Krylov elementwise kernels in float32 (blas.kern.h,
    linalg.kern.h:72-107), against numpy float32 evaluated in the same order."""
    n = 100003
    xs = [mixed_magnitude_vector(n, s) for s in range(4)]
    a, b = np.float32(0.37), np.float32(-1.9)
    d = [device.upload(v) for v in xs]
    f32 = np.float32

    def run(kernel, out_index, *args):
        device.launch(kernel, n, *args)
        return d[out_index].download()

    x0, x1, x2, x3 = xs
    same(run(Kernel.axpy, 0, d[0], d[1], a), (x0 + a * x1).astype(f32))
    x0 = (x0 + a * x1).astype(f32)
    same(run(Kernel.xpby, 0, d[0], d[1], b), (x1 + b * x0).astype(f32))
    x0 = (x1 + b * x0).astype(f32)
    same(run(Kernel.axpbyz, 2, d[2], d[0], d[1], a, b), (a * x0 + b * x1).astype(f32))
    x2 = (a * x0 + b * x1).astype(f32)
    same(run(Kernel.scale_copy, 3, d[3], d[2], a), (a * x2).astype(f32))
    x3 = (a * x2).astype(f32)
    same(run(Kernel.mul, 3, d[3], d[0], d[1]), (x0 * x1).astype(f32))


def test_smoothing_kernels(device, op, status0, status1):
    n = op.n
    x = mixed_magnitude_vector(n, 3)
    b = mixed_magnitude_vector(n, 4)
    b[::31] = np.inf
    b[7::37] = np.nan
    inv = (np.abs(random_vector(n, 5)) + 0.1).astype(np.float32)
    dx, db, di, dn = device.upload(x), device.upload(b), device.upload(inv), device.alloc(n)
    st = np.zeros(3, np.int32)
    for kernel, name, xin in ((Kernel.amg_smooth, "amg_smooth", x),
                              (Kernel.amg_smooth_zero, "amg_smooth_zero", np.zeros(n, np.float32))):
        dx.upload(xin)
        device.launch(kernel, n, op.rp, op.col, op.v, dx, db, di, dn, status0)
        want = np.zeros(n, np.float32)
        oracle.kernel(name, n, op.a.row_ptr, op.a.col, op.vals, xin, b, inv, want, st)
        same(dn.download(), want)
        dn.fill(np.float32(3))
        device.launch(kernel, n, op.rp, op.col, op.v, dx, db, di, dn, status1)
        assert np.all(dn.download() == 3)
    # the fused form reads the zero iterate as a constant
    device.launch(Kernel.fused_smooth_from_zero, n, op.rp, op.col, op.v, db, di, dn, status0)
    want = np.zeros(n, np.float32)
    oracle.kernel("amg_smooth_zero", n, op.a.row_ptr, op.a.col, op.vals, np.zeros(n, np.float32),
                  b, inv, want, st)
    same(dn.download(), want)


@pytest.mark.parametrize("count", [1, 5, 17, 24])
def test_coarse_solve(device, count, status0):
    a = oracle.poisson(count, 1, 1, 1.0, random=True)
    vals = a.val.astype(np.float32)
    b = random_vector(count, count)
    inv = (np.float32(1.0) / vals[a.diag]).astype(np.float32)
    for sweeps in (1, 5, 32):
        want = np.zeros(count, np.float32)
        oracle.kernel("amg_coarse_solve", a.row_ptr, a.col, vals, want, b, inv,
                      np.zeros(3, np.int32), count, sweeps)
        dx = device.alloc(count)
        device.launch(Kernel.amg_coarse_solve, 1, device.upload(a.row_ptr), device.upload(a.col),
                      device.upload(vals), dx, device.upload(b), device.upload(inv), status0,
                      count, sweeps)
        same(dx.download(), want)


def test_transfers_and_galerkin(device, op, status0, status1):
    for m_matrix in (False, True):
        a = oracle.poisson(21, 17, 13, 1.0, random=True, neumann=m_matrix)
        agg = oracle.aggregate(a, passes=2, m_matrix=m_matrix)
        n, nc = a.n_rows, len(agg["member_offsets"]) - 1
        fine = mixed_magnitude_vector(n, 8)
        coarse = mixed_magnitude_vector(nc, 9)
        d_off, d_mem, d_agg = (device.upload(agg[k]) for k in ("member_offsets", "members", "aggregate"))
        dfine, dc = device.upload(fine), device.alloc(nc)
        device.launch(Kernel.amg_restrict, nc, d_off, d_mem, dfine, dc)
        want = np.zeros(nc, np.float32)
        oracle.kernel("amg_restrict", nc, agg["member_offsets"], agg["members"], fine, want)
        same(dc.download(), want)
        dc.fill(np.float32(5))
        device.launch(Kernel.amg_restrict_guarded, nc, d_off, d_mem, dfine, dc, status1)
        assert np.all(dc.download() == 5)
        dc.upload(coarse)
        device.launch(Kernel.amg_prolong_guarded, n, d_agg, dc, dfine, status0)
        wf = fine.copy()
        oracle.kernel("amg_prolong", n, agg["aggregate"], coarse, wf)
        same(dfine.download(), wf)
        vals = a.val.astype(np.float32)
        nnz_c = len(agg["entry_offsets"]) - 1
        dcv = device.alloc(nnz_c)
        device.launch(Kernel.amg_galerkin, nnz_c, device.upload(agg["entry_offsets"]),
                      device.upload(agg["entries"]), device.upload(vals), dcv)
        wcv = np.zeros(nnz_c, np.float32)
        oracle.kernel("amg_galerkin", nnz_c, agg["entry_offsets"], agg["entries"], vals, wcv)
        same(dcv.download(), wcv)
        if m_matrix:
            device.launch(Kernel.amg_m_matrix_diagonal, nc, d_off, d_mem, device.upload(a.row_ptr),
                          device.upload(a.diag), device.upload(vals), device.upload(agg["row_ptr"]),
                          device.upload(agg["diag"]), dcv)
            oracle.kernel("amg_m_matrix_diagonal", nc, agg["member_offsets"], agg["members"],
                          a.row_ptr, a.diag, vals, agg["row_ptr"], agg["diag"], wcv)
            same(dcv.download(), wcv)
            # equals the host coarsening's stored values
            same(wcv, agg["val"].astype(np.float32))


SCALAR_KERNELS = [(Kernel.solver_reset, 0), (Kernel.solver_pcg_alpha, 1), (Kernel.solver_pcg_beta, 2),
                  (Kernel.solver_bicg_beta, 3), (Kernel.solver_bicg_alpha, 4), (Kernel.solver_bicg_s, 5),
                  (Kernel.solver_bicg_omega, 6), (Kernel.solver_check, 7)]


def test_scalar_state_machine(device):
    rng = np.random.default_rng(21)
    ds, dst = device.alloc(24), device.alloc(3, np.int32)
    for trial in range(120):
        s = (rng.standard_normal(24) * 10.0 ** rng.integers(-8, 8, 24)).astype(np.float32)
        s[20:24] = np.abs(s[20:24])
        if trial % 7 == 0:
            s[0] = 0
        status = np.array([trial % 3 == 0 and 1 or 0, trial % 60, trial % 2], np.int32)
        for kernel, which in SCALAR_KERNELS:
            flag = trial % 2 if which in (2, 7) else (trial if which == 0 else 0)
            ws, wst = s.copy(), status.copy()
            oracle.lib().oracle_solver_scalar(which, oracle.C.c_void_p(ws.ctypes.data),
                                              oracle.C.c_void_p(wst.ctypes.data), flag)
            ds.upload(s)
            dst.upload(status)
            args = (ds, dst, flag) if which in (0, 2, 7) else (ds, dst)
            device.launch(kernel, 1, *args)
            same(ds.download(), ws)
            assert np.array_equal(dst.download(), wst), (trial, which)


def test_solver_vector_kernels(device):
    n = 65537
    rng = np.random.default_rng(4)
    v = [mixed_magnitude_vector(n, 40 + k) for k in range(6)]
    s = (rng.standard_normal(24)).astype(np.float32)
    s[9] = s[8] * np.float32(1e-9)  # a nonzero low word
    s[11] = s[10] * np.float32(-3e-9)
    d = [device.upload(x) for x in v]
    ds = device.upload(s)
    st = device.alloc(3, np.int32)
    st.upload(np.array([0, 0, 0], np.int32))
    x, r = v[0].copy(), v[1].copy()
    oracle.kernel("solver_step", n, x, r, v[2], v[3], s, np.zeros(3, np.int32))
    device.launch(Kernel.solver_step, n, d[0], d[1], d[2], d[3], ds, st)
    same(d[0].download(), x)
    same(d[1].download(), r)
    p = v[4].copy()
    oracle.kernel("solver_direction", n, v[5], p, s, np.zeros(3, np.int32))
    device.launch(Kernel.solver_direction, n, d[5], d[4], ds, st)
    same(d[4].download(), p)
    pair = np.array([123.25, 1e-6], np.float32)
    x = v[5].copy()
    oracle.kernel("solver_project", n, x, pair, np.float32(n), np.float32(0))
    device.launch(Kernel.solver_project, n, d[5], device.upload(pair), np.float32(n), np.float32(0))
    same(d[5].download(), x)


def test_fused_kernels_equal_reference_sequences(device, op, status0):
    """Each fused kernel against the reference kernels it replaces, on device."""
    n = op.n
    rng = np.random.default_rng(9)
    p = mixed_magnitude_vector(n, 50)
    s = np.zeros(24, np.float32)
    s[0], s[1] = np.float32(3.5), np.float32(1e-8)
    s[20], s[22] = np.float32(1e-30), np.float32(1e30)

    def fresh():
        st = device.alloc(3, np.int32)
        st.upload(np.array([0, 3, 0], np.int32))
        return device.upload(s), st

    # spmv + dot + alpha
    sa, sta = fresh()
    sb, stb = fresh()
    qa, qb, dp = device.alloc(n), device.alloc(n), device.upload(p)
    device.launch(Kernel.fused_spmv_dot_alpha, n, op.rp, op.col, op.v, dp, qa, sa, sta)
    device.launch(Kernel.csr_spmv_guarded, n, op.rp, op.col, op.v, dp, qb, stb)
    device.dot_into(dp, qb, n, sb, 2)
    device.launch(Kernel.solver_pcg_alpha, 1, sb, stb)
    same(qa.download(), qb.download())
    ga, gb = sa.download(), sb.download()
    same(ga[[0, 1, 2, 8]], gb[[0, 1, 2, 8]])  # pair hi words and alpha
    assert np.array_equal(sta.download(), stb.download())
    # step + dot + check
    xa, xb = device.upload(p), device.upload(p)
    ra, rb = device.upload(mixed_magnitude_vector(n, 51)), device.upload(mixed_magnitude_vector(n, 51))
    device.launch(Kernel.fused_step_dot_check, n, xa, ra, dp, qa, sa, sta)
    device.launch(Kernel.solver_step, n, xb, rb, dp, qb, sb, stb)
    device.dot_into(rb, rb, n, sb, 18)
    device.launch(Kernel.solver_check, 1, sb, stb, 1)
    same(xa.download(), xb.download())
    same(ra.download(), rb.download())
    assert sa.download()[18] == sb.download()[18]
    assert np.array_equal(sta.download(), stb.download())
    # residual + restrict; prolong + smooth
    agg = oracle.aggregate(op.a, passes=2)
    nc = len(agg["member_offsets"]) - 1
    d_off, d_mem, d_agg = (device.upload(agg[k]) for k in ("member_offsets", "members", "aggregate"))
    x = mixed_magnitude_vector(n, 52)
    b = mixed_magnitude_vector(n, 53)
    inv = (np.abs(random_vector(n, 54)) + 0.2).astype(np.float32)
    dx, db, di = device.upload(x), device.upload(b), device.upload(inv)
    c1, c2, rr = device.alloc(nc), device.alloc(nc), device.alloc(n)
    device.launch(Kernel.fused_residual_restrict, nc, op.rp, op.col, op.v, dx, db, d_off, d_mem, c1, status0)
    device.launch(Kernel.csr_residual_guarded, n, op.rp, op.col, op.v, dx, db, rr, status0)
    device.launch(Kernel.amg_restrict_guarded, nc, d_off, d_mem, rr, c2, status0)
    same(c1.download(), c2.download())
    corr = device.upload(mixed_magnitude_vector(nc, 55))
    o1, o2 = device.alloc(n), device.alloc(n)
    device.launch(Kernel.fused_prolong_smooth, n, op.rp, op.col, op.v, dx, db, di, o1, d_agg, corr, status0)
    device.launch(Kernel.amg_prolong_guarded, n, d_agg, corr, dx, status0)
    device.launch(Kernel.amg_smooth, n, op.rp, op.col, op.v, dx, db, di, o2, status0)
    same(o1.download(), o2.download())
    # smooth + dot + beta ; smooth + sum ; project + dot + beta
    sa, sta = fresh()
    sb, stb = fresh()
    device.launch(Kernel.fused_smooth_dot_beta, n, op.rp, op.col, op.v, dx, db, di, o1, sta, sa)
    device.launch(Kernel.amg_smooth, n, op.rp, op.col, op.v, dx, db, di, o2, stb)
    device.dot_into(db, o2, n, sb, 4)
    device.launch(Kernel.solver_pcg_beta, 1, sb, stb, 0)
    same(o1.download(), o2.download())
    same(sa.download()[[0, 4, 10]], sb.download()[[0, 4, 10]])
    assert np.array_equal(sta.download(), stb.download())
    sum1, sum2 = device.alloc(2), device.alloc(2)
    device.launch(Kernel.fused_smooth_sum, n, op.rp, op.col, op.v, dx, db, di, o1, status0, sum1)
    device.sum_into(o2, n, sum2)
    same(o1.download(), o2.download())
    assert sum1.download()[0] == sum2.download()[0]
    sa, sta = fresh()
    sb, stb = fresh()
    cnt = np.float32(n)
    device.launch(Kernel.fused_project_dot_beta, n, o1, sum1, cnt, np.float32(0), db, sa, sta)
    device.launch(Kernel.solver_project, n, o2, sum1, cnt, np.float32(0))
    device.dot_into(db, o2, n, sb, 4)
    device.launch(Kernel.solver_pcg_beta, 1, sb, stb, 0)
    same(o1.download(), o2.download())
    same(sa.download()[[0, 4, 10]], sb.download()[[0, 4, 10]])


@pytest.mark.parametrize("lengths", ["short", "long"])
def test_long_row_sweep_forms(device, lengths, status0, status1):
    """The wide-chunk and wavefront-per-row sweeps equal the reference's
    amg_smooth bit for bit, and the wavefront-per-row residual and folded
    prolongation sweep equal csr_residual and amg_prolong + amg_smooth, on
    rows of 1 to 600 entries (crossing the wavefront form's 256-entry passes)
    with mixed-magnitude values; a frozen status leaves the output
    untouched."""
    rng = np.random.default_rng(21 if lengths == "short" else 22)
    n = 700
    top = 20 if lengths == "short" else 600
    row_len = rng.integers(1, top + 1, n)
    row_len[:3] = [1, 256, min(top, 257)]
    rows = np.concatenate([[0], np.cumsum(row_len)]).astype(np.int32)
    cols = np.concatenate([np.sort(rng.choice(n, size=k, replace=False)) for k in row_len]).astype(np.int32)
    vals = mixed_magnitude_vector(int(rows[-1]), 23)
    x = mixed_magnitude_vector(n, 24)
    b = mixed_magnitude_vector(n, 25)
    inv = (np.abs(random_vector(n, 26)) + 0.1).astype(np.float32)
    drows, dcols, dvals = device.upload(rows), device.upload(cols), device.upload(vals)
    dx, db, di, dn = device.upload(x), device.upload(b), device.upload(inv), device.alloc(n)
    want = np.zeros(n, np.float32)
    oracle.kernel("amg_smooth", n, rows, cols, vals, x, b, inv, want, np.zeros(3, np.int32))
    for kernel in (Kernel.amg_smooth, Kernel.amg_smooth_wide, Kernel.amg_smooth_rowwave):
        dn.fill(np.float32(0))
        device.launch(kernel, n, drows, dcols, dvals, dx, db, di, dn, status0)
        same(dn.download(), want)
        dn.fill(np.float32(3))
        device.launch(kernel, n, drows, dcols, dvals, dx, db, di, dn, status1)
        assert np.all(dn.download() == 3)
    # The residual one wavefront per row (csr_residual).
    want_r = np.zeros(n, np.float32)
    oracle.kernel("csr_residual", n, rows, cols, vals, x, b, want_r)
    dn.fill(np.float32(0))
    device.launch(Kernel.csr_residual_rowwave, n, drows, dcols, dvals, dx, db, dn, status0)
    same(dn.download(), want_r)
    dn.fill(np.float32(3))
    device.launch(Kernel.csr_residual_rowwave, n, drows, dcols, dvals, dx, db, dn, status1)
    assert np.all(dn.download() == 3)
    # The prolongation folded into a sweep one wavefront per row: amg_prolong
    # (x + coarse[aggregate], unaggregated rows unchanged), then amg_smooth.
    nc = 97
    agg = rng.integers(-1, nc, n).astype(np.int32)
    coarse = mixed_magnitude_vector(nc, 27)
    xp = x.copy()
    oracle.kernel("amg_prolong", n, agg, coarse, xp)
    want_p = np.zeros(n, np.float32)
    oracle.kernel("amg_smooth", n, rows, cols, vals, xp, b, inv, want_p, np.zeros(3, np.int32))
    dagg, dc = device.upload(agg), device.upload(coarse)
    for kernel in (Kernel.fused_prolong_smooth, Kernel.prolong_smooth_rowwave):
        dn.fill(np.float32(0))
        device.launch(kernel, n, drows, dcols, dvals, dx, db, di, dn, dagg, dc, status0)
        same(dn.download(), want_p)
    dn.fill(np.float32(3))
    device.launch(Kernel.prolong_smooth_rowwave, n, drows, dcols, dvals, dx, db, di, dn, dagg, dc, status1)
    assert np.all(dn.download() == 3)


RESIDENT_CASES = {  # rows, longest ordinary row, rows of 17-105 entries, non-finite b entries
    "tiny": (25, 5, 0, 0),
    "c2-coarsest": (1067, 5, 70, 0),
    "long-rows": (2000, 3, 12, 2),
    "most-rows": (4096, 1, 0, 0),
    "cap16": (3000, 3, 3, 1),
    "short-runs": (600, 12, 0, 1),
}


@pytest.mark.parametrize("case", sorted(RESIDENT_CASES))
def test_register_resident_coarse_sweeps(device, case, status0, status1):
    """The two-phase coarsest-level sweeps equal the reference's sequence
    (amg_smooth_zero, then 31 amg_smooth sweeps) bit for bit, as the
    one-workgroup global-memory form does, for both entry capacities, with
    long rows and non-finite right-hand sides; a frozen status leaves x
    untouched."""
    import kutta_amd
    n, top, long_rows, bad = RESIDENT_CASES[case]
    rng = np.random.default_rng(sum(map(ord, case)))
    lengths = rng.integers(1, top + 1, n)
    if long_rows:
        lengths[rng.choice(n, long_rows, replace=False)] = rng.integers(17, 106, long_rows)
    # Rows with a non-finite right-hand side couple only to themselves, so
    # the non-finite values stay in those rows.
    isolated = rng.choice(n, bad, replace=False)
    lengths[isolated] = 1
    others = np.setdiff1d(np.arange(n), isolated)
    rows = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int32)
    cols = np.concatenate([[i] if i in isolated else
                           np.sort(np.concatenate([[i], rng.choice(others[others != i], size=k - 1, replace=False)]))
                           for i, k in enumerate(lengths)]).astype(np.int32)
    vals = (random_vector(int(rows[-1]), 31) * 0.1).astype(np.float32)
    b = mixed_magnitude_vector(n, 32)
    b[isolated] = [np.inf, np.nan][:bad]
    inv = (np.abs(random_vector(n, 33)) * 0.5 + 0.1).astype(np.float32)
    table, cap = kutta_amd.resident_partition(rows, cols)
    print(f"[resident] {case}: {n} rows, {rows[-1]} entries, cap {cap}", flush=True)
    assert np.all((table[0::2] <= table[1::2]) & (table[1::2] <= n)) and rows[-1] <= cap * len(table) // 2
    sweeps = 32
    cur = np.zeros(n, np.float32)
    nxt = np.zeros(n, np.float32)
    oracle.kernel("amg_smooth_zero", n, rows, cols, vals, np.zeros(n, np.float32), b, inv, cur, np.zeros(3, np.int32))
    for _ in range(sweeps - 1):
        oracle.kernel("amg_smooth", n, rows, cols, vals, cur, b, inv, nxt, np.zeros(3, np.int32))
        cur, nxt = nxt, cur
    drows, dcols, dvals = device.upload(rows), device.upload(cols), device.upload(vals)
    db, di, dx, dt = device.upload(b), device.upload(inv), device.alloc(n), device.upload(table)
    got = []
    for kernel, extra in ((Kernel.coarse_sweeps_resident, (dt, cap)), (Kernel.fused_coarse_sweeps, ())):
        dx.fill(np.float32(7))
        device.launch(kernel, n, drows, dcols, dvals, db, di, dx, status0, sweeps, *extra)
        got.append(dx.download())
    # A NaN's sign and payload follow the processor and the operand order
    # the compiler picks for a commutative operation, so for NaNs only their
    # place must agree; every other value bit for bit.
    nan = np.isnan(cur)
    assert nan.sum() <= bad
    for g in got:
        assert np.array_equal(np.isnan(g), nan)
        same(g[~nan], cur[~nan])
    dx.fill(np.float32(3))
    device.launch(Kernel.coarse_sweeps_resident, n, drows, dcols, dvals, db, di, dx, status1, sweeps, dt, cap)
    assert np.all(dx.download() == 3)
