"""kh_reduce_into_static, the reduction entry point that allocates nothing:
bit for bit the output of kh_reduce_into (the device-handle entry point it
mirrors) on the same inputs, over two calls in a row, the second smaller, so
the shared partials array left by the first call cannot leak into the second."""
import ctypes as C

import numpy as np
import pytest

import kutta_amd

pytestmark = pytest.mark.gpu

MODES = {"sum": 0, "dot": 1, "dot2": 2}


def _hip_runtime():
    lib = C.CDLL("libamdhip64.so")
    lib.hipDeviceSynchronize.restype = C.c_int
    return lib


def _reduce(device, mode, bufs, n, dst, static):
    lib = kutta_amd.hip_lib()
    x, y, u, v = (b.ptr for b in bufs)
    if static:
        rc = lib.kh_reduce_into_static(None, MODES[mode], x, y, u, v, n, dst.ptr, 2)
        assert rc == 0, lib.kh_last_error().decode()
        assert _hip_runtime().hipDeviceSynchronize() == 0
    else:
        rc = lib.kh_reduce_into(device.kh, MODES[mode], x, y, u, v, n, dst.ptr, 2)
        assert rc == 0, lib.kh_last_error().decode()
    device.synchronize()
    return dst.download()


@pytest.mark.parametrize("mode", ["sum", "dot", "dot2"])
def test_static_reduce_matches_device_entry_point(device, mode):
    for n in (3_000_001, 40_000):  # the second call is smaller
        print(f"[no-alloc] mode={mode} n={n}", flush=True)
        rng = np.random.default_rng(n + MODES[mode])
        vs = [rng.standard_normal(n).astype(np.float32) for _ in range(4)]
        bufs = [device.upload(a) for a in vs]
        out = []
        for static in (False, True):
            dst = device.alloc(8)
            dst.fill(np.float32(7))
            out.append(_reduce(device, mode, bufs, n, dst, static))
        assert np.array_equal(out[0].view(np.uint32), out[1].view(np.uint32)), (mode, n, out)
        assert np.all(out[1][:2] == 7) and np.all(out[1][2 + (4 if mode == "dot2" else 2):] == 7)


def test_static_reduce_empty_and_rejected_sizes(device):
    lib = kutta_amd.hip_lib()
    x = device.upload(np.ones(16, np.float32))
    dst = device.alloc(8)
    dst.fill(np.float32(7))
    assert lib.kh_reduce_into_static(None, 2, x.ptr, x.ptr, x.ptr, x.ptr, 0, dst.ptr, 0) == 0
    assert _hip_runtime().hipDeviceSynchronize() == 0
    assert np.array_equal(dst.download(), np.array([0, 0, 0, 0, 7, 7, 7, 7], np.float32))
    # Host-side rejections: nothing is launched.
    assert lib.kh_reduce_into_static(None, 0, x.ptr, None, None, None, 1 << 31, dst.ptr, 0) != 0
    assert b"KH_STATIC_REDUCE_MAX_N" in lib.kh_last_error()
    assert lib.kh_reduce_into_static(None, 3, x.ptr, x.ptr, None, None, 16, dst.ptr, 0) != 0
    assert lib.kh_reduce_into_static(None, 1, x.ptr, x.ptr, None, None, 16, x.ptr, 0) != 0
    assert lib.kh_reduce_into_static(None, 2, x.ptr, x.ptr, None, None, 16, dst.ptr, 0) != 0
