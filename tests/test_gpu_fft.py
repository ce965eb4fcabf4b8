"""The GPU real FFT at lengths that are not powers of two
(src/fastffts.c:163-243 takes any even length; FFTW transforms any length
for the reference, src/fftcalls.c:60-159): mixed-radix Stockham passes
(radix 8/4/2 with 3, 5, 7, 11, 13) and Bluestein's chirp-z transform for
lengths with a larger prime factor.

* forward vs numpy's FP64 rfft: worst bin within 1e-5 (direct) / 3e-5
  (Bluestein) of the rms bin amplitude, DC and Nyquist in bin 0;
* the inverse (realfft +1, scaled 2/n) restores the series;
* a search on a non-power-of-two .dat series prepared on the device: the
  device spectrum equals the host de-reddening of the same FFT output bit
  for bit, and the candidates equal the oracle's on that spectrum."""
import numpy as np
import pytest

from oracle import oracle as O
from presto_amd import synth

pytestmark = pytest.mark.gpu
A = pytest.importorskip("presto_amd.accel")

# (n, Bluestein): M = n / 2 with only 2, 3, 5, 7, 11, 13 is direct
LENGTHS = [(3 << 16, False), (2 * 5 ** 7, False), (2 * 8 * 7 * 11 * 13, False),
           (1_000_000, False), (6_000_000, False), (1 << 20, False),
           (2 * 65537, True), (2 * 3 * 17 * 1009, True), (2 * 5 * 838861, True)]


def _rms_err(got, x):
    n = len(x)
    ref = np.fft.rfft(x.astype(np.float64))
    rms = np.sqrt(np.mean(np.abs(ref[1:n // 2]) ** 2))
    err = np.max(np.abs(got[1:] - ref[1:n // 2])) / rms
    dc = abs(got[0].real - ref[0].real) / rms
    ny = abs(got[0].imag - ref[n // 2].real) / rms
    return err, dc, ny


@pytest.mark.parametrize("n,blue", LENGTHS)
def test_forward_matches_fp64_and_inverse_restores(n, blue):
    print(f"case: realfft n={n} bluestein={blue}", flush=True)
    x = synth.make_series(n, 1e-3, n % 9973, [dict(f0=37.3, duty=0.1, amp=0.5)])
    got = A.realfft_gpu(x)
    err, dc, ny = _rms_err(got, x)
    bar = 3e-5 if blue else 1e-5
    print(f"  worst bin / rms {err:.2e}, DC {dc:.2e}, Nyquist {ny:.2e}", flush=True)
    assert err < bar and dc < 1e-3 and ny < 1e-3
    back = A.realfft_inverse_gpu(got)
    rel = np.max(np.abs(back - x)) / np.sqrt(np.mean(x.astype(np.float64) ** 2))
    assert rel < (1e-4 if blue else 3e-5), rel


@pytest.mark.parametrize("N,dt", [(3 << 16, 1e-3), (2 * 3 * 17 * 1009, 5e-4)])
def test_search_on_a_non_power_of_two_series(N, dt):
    print(f"case: device series preparation and search N={N}", flush=True)
    x = synth.make_series(N, dt, 88, [dict(f0=41.7, z=5.0, duty=0.05, amp=0.35)])
    ref = A.realfft_gpu(x)
    A.lib.accel_prepare_dat_spectrum(ref.ctypes.data, len(ref), float(N), dt)
    p = A.make_params(N // 2, N, dt, zmax=20, numharm=4, sigma=2.0, force_standard=True)
    with A.Search(p) as s:
        s.set_series(x)
        got = s.spectrum()
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
        s.run()
        cands = s.candidates()
    o = O.make_obs(len(ref), N, dt, zmax=20, numharm=4, sigma=2.0, force_standard=True)
    _, want, _ = O.accelsearch(o, ref)
    assert len(cands) == len(want) and len(cands) >= 1
    for g, w in zip(cands, want):
        assert g["numharm"] == int(w[2]) and abs(g["r"] - w[3]) < 1e-3
        assert abs(g["z"] - w[4]) < 0.05 and abs(g["sigma"] - w[1]) < 1e-3
