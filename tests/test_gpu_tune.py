"""The correlation's measured template-group size (accel_search_tune, the
counterpart of the reference's FFTW_MEASURE planning at
src/accel_utils.c:1133-1135): every group count writes the same raw plane,
so a tuned handle gives the untuned handle's plane rows and candidates bit
for bit."""
import numpy as np
import pytest

from presto_amd import accel as A, synth

pytestmark = pytest.mark.gpu


def test_tuned_handle_equals_default():
    N, dt = 1 << 21, 64e-6
    print(f"case: tune, N={N}, zmax 200, numharm 8", flush=True)
    x = synth.make_series(N, dt, 91, [dict(f0=233.3, z=40.0, duty=0.05, amp=0.04)])
    fft = A.spectrum_from_series(x, dt)
    p = A.make_params(len(fft), N, dt, zmax=200, numharm=8, sigma=2.0)
    with A.Search(p) as s0:
        s0.set_spectrum(fft)
        s0.run()
        c0 = s0.candidates()
        g0 = s0.corr_shape()["groups"]
        rows0 = [s0.plane_row(0, zi, 1000, 4096) for zi in (0, 100, 200)]
    with A.Search(p) as s1:
        s1.set_spectrum(fft)
        kept, table = s1.tune()
        print(f"default {g0} groups; timed {table}; kept {kept}", flush=True)
        assert len(table) >= 2 and table[0][0] == g0
        assert all(gr % 8 == 0 and ms > 0 for gr, ms in table)
        assert kept == min(table, key=lambda t: t[1])[0]
        assert s1.corr_shape()["groups"] == kept
        s1.run()
        c1 = s1.candidates()
        rows1 = [s1.plane_row(0, zi, 1000, 4096) for zi in (0, 100, 200)]
    assert len(c0) > 0 and c0.tobytes() == c1.tobytes()
    for a, b in zip(rows0, rows1):
        assert np.array_equal(a, b)
