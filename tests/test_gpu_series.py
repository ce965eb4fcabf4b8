"""Time-series input prepared on the device (accel_search_set_series):
GPU real FFT, GPU de-reddening, bin 0 := 1+1i.  The de-reddened spectrum
must equal, bit for bit, the host de-reddening of the same FFT output (the
host version is itself bit-equal to the oracle), and a search on it must
give the same candidates."""
import numpy as np
import pytest

from presto_amd import synth

pytestmark = pytest.mark.gpu
A = pytest.importorskip("presto_amd.accel")


@pytest.mark.parametrize("N,dt", [(1 << 16, 1e-3), (1 << 21, 64e-6)])
def test_device_preparation_is_bit_exact(N, dt):
    print(f"case: device series preparation N={N}", flush=True)
    x = synth.make_series(N, dt, 71, [dict(f0=51.7, z=6.0, duty=0.05, amp=0.3)])
    ref = A.realfft_gpu(x)
    A.lib.accel_prepare_dat_spectrum(ref.ctypes.data, len(ref), float(N), dt)
    p = A.make_params(N // 2, N, dt, zmax=20, numharm=8, sigma=2.0)
    with A.Search(p) as s:
        s.set_series(x)
        got = s.spectrum()
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
        s.run()
        c1 = s.candidates()
        s.set_spectrum(ref)
        s.run()
        assert np.array_equal(c1, s.candidates())
