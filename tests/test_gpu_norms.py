"""Block normalisation on the GPU: every window's norm equals the oracle's
(src/accel_utils.c:1087-1099: 1/sqrt(median(|x|^2)/ln 2) over the window's
bins, zeros outside the spectrum) exactly, for the wavefront-per-window
selection (windows of at most 1024 bins) and the workgroup radix select
(longer windows), on the standard and in-memory paths, with tied powers
and windows that start before bin 0."""
import numpy as np
import pytest

from oracle import oracle as O
from presto_amd import synth

pytestmark = pytest.mark.gpu
A = pytest.importorskip("presto_amd.accel")


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("N,zmax,numharm,quantize,inmem", [
    (1 << 21, 200, 16, False, False),   # the plain workload's windows (1024 bins)
    (1 << 20, 200, 8, True, False),     # powers on a coarse grid: many ties
    (1 << 20, 50, 8, False, True),      # in-memory path, pre-population windows
    (1 << 20, 600, 4, False, False),    # windows above 1024 bins: radix select
])
def test_every_window_norm_equals_oracle(N, zmax, numharm, quantize, inmem):
    dt = 64e-6
    _say(f"case: norms N={N} zmax={zmax} numharm={numharm} quantize={quantize} inmem={inmem}")
    x = synth.make_series(N, dt, 31, [dict(f0=123.456, z=20.0, duty=0.05, amp=0.05)])
    fft = A.spectrum_from_series(x, dt)
    if quantize:
        fft = (np.round(fft.real * 2) / 2 + 1j * np.round(fft.imag * 2) / 2).astype(np.complex64)
    kw = dict(zmax=zmax, numharm=numharm, sigma=2.0, inmem=inmem, force_standard=not inmem)
    p = A.make_params(len(fft), N, dt, **kw)
    o = O.make_obs(len(fft), N, dt, **kw)
    with A.Search(p, device=0) as s:
        s.set_spectrum(fft)
        s.run()
        norms = s.norms()
        lo, nd = s.windows()
    want = np.array([O.block_norm(o, fft, int(a), int(b)) for a, b in zip(lo, nd)])
    bad = int(np.count_nonzero(norms != want))
    _say(f"{len(norms)} windows of {nd.min()}..{nd.max()} bins (first at bin {lo.min()}): "
         f"{bad} differ")
    assert bool(s.geom.inmem) == inmem
    assert lo.min() < 0 or inmem
    assert (nd.max() > 1024) == (zmax == 600)
    assert bad == 0


@pytest.mark.parametrize("N,dt", [(1 << 21, 64e-6), (3 * 5 * 7 * 2 ** 13, 1e-4)])
def test_device_deredden_equals_oracle(N, dt):
    """The de-reddening of a time series on the device (block medians one
    wavefront per block, then the interpolated scaling) equals the
    oracle's restatement of src/misc_utils.c:912-1067 bit for bit."""
    _say(f"case: device deredden N={N} dt={dt}")
    x = synth.make_series_f32(N, dt, 43, [dict(f0=77.7, z=5.0, duty=0.05, amp=0.2)])
    p = A.make_params(N // 2, N, dt, zmax=20, numharm=4, sigma=2.0)
    with A.Search(p) as s:
        s.set_series(x)
        spec = s.spectrum()
    ref, _ = O.deredden(A.realfft_gpu(x), 6, 200, 0.5 / dt, N * dt)
    diff = np.flatnonzero(spec.view(np.uint64) != ref.view(np.uint64))
    _say(f"{len(spec)} bins: {len(diff)} differ, at {diff[:8].tolist()}")
    # bin 0 becomes (1, 1) after the de-reddening unless -photon
    # (src/accel_utils.c:1621-1624)
    assert spec[0] == np.complex64(1.0 + 1.0j) and ref[0] == np.complex64(1.0 + 0.0j)
    assert diff.tolist() == [0]
