"""Device runtime semantics on the MI355X (tests/runtime/test_device.cpp,
tests/runtime/test_backends.cpp restated): buffers, ordering, fills, error
paths and FP64 reductions against the serial oracle."""
import numpy as np
import pytest

import kutta_amd
import oracle
from kutta_amd import Kernel, KuttaError

pytestmark = pytest.mark.gpu


def test_upload_download_copy_fill(device):
    x = np.arange(1000, dtype=np.float32)
    a = device.upload(x)
    assert np.array_equal(a.download(), x)
    b = device.alloc(1000)
    device.copy(b, a)
    assert np.array_equal(b.download(), x)
    b.fill(np.float32(-2.5))
    assert np.all(b.download() == np.float32(-2.5))
    i = device.alloc(7, np.int32)
    i.fill(np.int32(-3))
    assert i.download().tolist() == [-3] * 7
    z = device.alloc(0)
    assert z.download().size == 0


def test_ordering_of_launches_copies_and_fills(device):
    """test_backends.cpp:71-118: work runs in call order on one stream."""
    n = 1 << 16
    x = device.upload(np.ones(n, np.float32))
    y = device.alloc(n)
    y.fill(np.float32(0))
    for k in range(10):
        device.launch(Kernel.axpy, n, y, x, np.float32(1.0))
        if k == 4:
            device.copy(x, y)  # x = 5 from here on
    assert np.all(y.download() == np.float32(5 + 5 * 5))


def test_launch_errors_are_reported_not_run(device):
    """test_backends.cpp:120-140: zero-size ops, bad ids and bad arguments."""
    x = device.alloc(4)
    device.launch(Kernel.axpy, 0, x, x, np.float32(1.0))  # n == 0 is a no-op
    with pytest.raises(KuttaError, match="bad kernel id"):
        device.launch(9999, 4, x)
    with pytest.raises(KuttaError, match="expects"):
        device.launch(Kernel.axpy, 4, x, x)
    with pytest.raises(KuttaError, match="wrong kind"):
        device.launch(Kernel.axpy, 4, x, np.float32(1.0), x)
    with pytest.raises(KuttaError, match="INT_MAX"):
        device.launch(Kernel.axpy, 1 << 31, x, x, np.float32(1.0))
    with pytest.raises(KuttaError, match="count must be"):
        device.launch(Kernel.amg_coarse_solve, 1, x, x, x, x, x, x, x, 25, 1)


@pytest.mark.parametrize("n", [1, 31, 257, 100003, 3_000_001])
def test_host_reductions_vs_oracle(device, n):
    """test_backends.cpp:41-69 tail sizes; FP64 accumulation, fixed order."""
    rng = np.random.default_rng(n)
    xv = (rng.standard_normal(n) * 10.0 ** rng.integers(-3, 3, n)).astype(np.float32)
    yv = rng.standard_normal(n).astype(np.float32)
    x, y = device.upload(xv), device.upload(yv)
    s, d, m = device.reduce_many([("sum", x, None, n), ("dot", x, y, n), ("max_abs", x, None, n)])
    scale = np.sum(np.abs(xv.astype(np.float64)))
    assert abs(s - oracle.sum_(xv)) <= 1e-13 * scale
    assert abs(d - oracle.dot(xv, yv)) <= 1e-13 * np.sum(np.abs(xv.astype(np.float64) * yv))
    assert m == oracle.max_abs(xv)
    # bitwise determinism run to run (test_backends.cpp:199-220)
    assert device.reduce_many([("sum", x, None, n)])[0] == s
    assert device.dot(x, y) == d


def test_reductions_exact_cancellation_and_nan(device):
    """test_backends.cpp:162-181 and :222-233."""
    n = 1 << 20
    v = np.ones(n, np.float32)
    v[1::2] = -1
    v[0] = np.float32(1 + 2 ** -23)
    x = device.upload(v)
    assert device.sum(x) == 2.0 ** -23
    dst = device.alloc(4)
    device.sum_into(x, n, dst)
    assert dst.download()[:2].tolist() == [np.float32(2 ** -23), 0.0]
    w = np.full(n, np.float32(2 ** -24))
    w[0] = 1
    device.sum_into(device.upload(w), n, dst)
    pair = dst.download()
    assert float(pair[0]) + float(pair[1]) == 1 + (n - 1) * 2.0 ** -24
    v[12345] = np.nan
    assert np.isnan(device.max_abs(device.upload(v)))
    v[12345] = np.inf
    assert device.max_abs(device.upload(v)) == np.inf


@pytest.mark.parametrize("n", [1, 1000, 1 << 20])
def test_reduce_into_pairs_vs_oracle(device, n):
    rng = np.random.default_rng(n + 1)
    vs = [rng.standard_normal(n).astype(np.float32) for _ in range(4)]
    bufs = [device.upload(v) for v in vs]
    dst = device.alloc(8)
    device.dot2_into(*bufs, n, dst, offset=2)
    got = dst.download()[2:6]
    want = oracle.reduce_into("dot2", *vs)
    for k in (0, 2):
        # hi words agree; the pair value agrees to FP64 summation-order error
        g = float(got[k]) + float(got[k + 1])
        w = float(want[k]) + float(want[k + 1])
        assert abs(g - w) <= 1e-13 * np.sum(np.abs(vs[k // 2 * 2].astype(np.float64) * vs[k // 2 * 2 + 1]))


def test_graph_capture_replays_match_eager(device):
    """A captured launch sequence replays like eager execution, several times,
    with eager work in between (fills are kernels, never memset nodes)."""
    n = 4096
    x = device.upload(np.arange(n, dtype=np.float32))
    y = device.alloc(n)
    y.fill(np.float32(0))
    dst = device.alloc(2)
    lib = kutta_amd.hip_lib()
    import ctypes as C
    assert lib.kh_graph_begin(device.kh) == 0
    device.launch(Kernel.axpy, n, y, x, np.float32(0.5))
    device.sum_into(y, n, dst)
    g = C.c_void_p()
    assert lib.kh_graph_end(device.kh, C.byref(g)) == 0
    for rep in range(1, 5):
        assert lib.kh_graph_launch(device.kh, g) == 0
        device.synchronize()
        want = np.float32(0.5) * np.arange(n, dtype=np.float32) * rep
        assert np.array_equal(y.download(), want)
        assert float(dst.download()[0]) == float(np.sum(want.astype(np.float64)))
        device.launch(Kernel.axpy, n, y, x, np.float32(0.0))  # eager work between replays
    lib.kh_graph_destroy(g)


def test_profile_counts_launches(device):
    n = 1 << 16
    x = device.upload(np.ones(n, np.float32))
    device.profile_reset()
    device.profile(Kernel.axpy, True)
    for _ in range(5):
        device.launch(Kernel.axpy, n, x, x, np.float32(1.0))
    count, ms = device.profile_read(Kernel.axpy)
    device.profile(Kernel.axpy, False)
    assert count == 5 and ms > 0
