"""accel_realfft_forward_noalloc: the forward real FFT on device buffers
with its work buffer in the library's static device scratch.  Its output
equals accel_realfft_forward's (the allocating entry point, same kernels)
bit for bit, on two calls in a row, the second with a smaller series (the
scratch keeps the first call's contents), and sizes beyond the scratch
are refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
A = pytest.importorskip("presto_amd.accel")


def _noalloc(x, device=0):
    n = len(x)
    out = np.zeros(n // 2, dtype=np.complex64)
    with A.DeviceBuffer(4 * n, device) as d_in, A.DeviceBuffer(8 * (n // 2), device) as d_out:
        d_in.upload(np.ascontiguousarray(x, dtype=np.float32))
        rc = A.lib.accel_realfft_forward_noalloc(C.c_void_p(d_in.ptr), n, C.c_void_p(d_out.ptr),
                                                 None, device)
        assert rc == 0, A.lib.accel_last_error()
        d_out.download(out)   # blocking copy on the default stream, after the FFT
    return out


def test_noalloc_equals_allocating_entry_point_twice():
    rng = np.random.default_rng(11)
    for n in (1 << 22, 1 << 16):
        print(f"case: noalloc realfft n={n}", flush=True)
        x = rng.standard_normal(n).astype(np.float32)
        want = A.realfft_gpu(x)
        got = _noalloc(x)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), n


def test_noalloc_refuses_sizes_beyond_the_scratch():
    n = 1 << 23
    rc = A.lib.accel_realfft_forward_noalloc(C.c_void_p(1), n, C.c_void_p(1), None, 0)
    assert rc == -1
    assert b"scratch" in A.lib.accel_last_error()
