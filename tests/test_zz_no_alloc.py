"""The allocation-free entry point (ssu_center_matrix_fp64_noalloc): Gower
centring of an FP64 matrix in device memory, with its row statistics in the
library's fixed device array instead of allocated buffers.  Its result must
equal the allocating entry point's (ssu_center_matrix_fp64) bit for bit, on
two calls in a row, the second on a smaller matrix (the array keeps the first
call's contents), and sizes past its maximum are rejected with an error."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


class Hip:
    """Device buffers for the test through the HIP runtime itself (the entry point under
    test allocates nothing; its caller does)."""

    def __init__(self):
        import os
        path = "/opt/rocm/lib/libamdhip64.so"
        self.L = C.CDLL(path if os.path.exists(path) else "libamdhip64.so")
        self.live = []

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), C.c_size_t(a.nbytes)) == 0
        self.live.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0   # host to device
        return p

    def download(self, p, shape):
        out = np.empty(shape, np.float64)
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), 2) == 0   # to host
        return out

    def free(self):
        self.L.hipDeviceSynchronize()
        for p in self.live:
            self.L.hipFree(p)
        self.live = []


@pytest.fixture(scope="module")
def hip(ssu):
    if not ssu.device_available():
        pytest.fail("no HIP device visible: the HIP path cannot run (no CPU fallback exists)")
    h = Hip()
    yield h
    h.free()


def distance_matrix(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.random((n, 5))
    d = np.sqrt(((x[:, None, :] - x[None, :, :]) ** 2).sum(-1))
    return np.ascontiguousarray(d)


def reference_centring(ssu, d):
    n = d.shape[0]
    out = np.zeros_like(d)
    st = ssu.lib().ssu_center_matrix_fp64(d.ctypes.data_as(C.c_void_p), C.c_uint(n), out.ctypes.data_as(C.c_void_p))
    assert st == ssu.OKAY, ssu.last_error()
    return out


def noalloc_centring(ssu, hip, d):
    n = d.shape[0]
    src = hip.upload(d)
    dst = hip.upload(np.full_like(d, np.nan))
    st = ssu.lib().ssu_center_matrix_fp64_noalloc(src, C.c_uint(n), dst, C.c_void_p(0))
    assert st == ssu.OKAY, ssu.last_error()
    return hip.download(dst, d.shape)


def test_noalloc_centring_matches_allocating_entry_point(ssu, hip):
    for n, seed in ((700, 1), (130, 2)):   # the second call on a smaller matrix
        say("noalloc centring", n)
        d = distance_matrix(n, seed)
        assert np.array_equal(noalloc_centring(ssu, hip, d), reference_centring(ssu, d))


def test_noalloc_centring_in_place_and_limits(ssu, hip):
    d = distance_matrix(257, 3)
    buf = hip.upload(d)
    say("noalloc centring in place")
    assert ssu.lib().ssu_center_matrix_fp64_noalloc(buf, C.c_uint(257), buf, C.c_void_p(0)) == ssu.OKAY
    assert np.array_equal(hip.download(buf, d.shape), reference_centring(ssu, d))
    # past the fixed array's size: an error before any device work
    assert ssu.lib().ssu_center_matrix_fp64_noalloc(buf, C.c_uint(65537), buf, C.c_void_p(0)) == ssu.OUTPUT_ERROR
    assert "exceeds" in ssu.last_error()
