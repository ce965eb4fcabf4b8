"""GPU parity: the HIP path (through the C ABI) against the oracle.

Bars (from the task statement for this path):
  * plane powers within 1e-4 of the FP64 oracle, measured as
    |gpu - fp64| / max(fp64, 1) (powers are normalised so that noise has
    mean 1; the error of each case is printed);
  * normalisation factors equal to the oracle's to FP64 rounding;
  * hit lists equal except for cells within 1e-5 of a threshold;
  * identical candidate lists: same numharm, r within 1e-3 bins, z within
    0.05, sigma within 1e-3.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from presto_amd import synth
from parity_util import compare_cands as _compare_cands, compare_hits as _compare_hits, \
    plane_error as _plane_error

pytestmark = pytest.mark.gpu

A = pytest.importorskip("presto_amd.accel")


def _say(msg):
    print(msg, flush=True)


def _case(N, dt, seed, pulsars, **kw):
    x = synth.make_series(N, dt, seed, pulsars)
    fft = A.spectrum_from_series(x, dt)
    p = A.make_params(len(fft), N, dt, **kw)
    o = O.make_obs(len(fft), N, dt, **kw)
    return fft, p, o


def _run_gpu(fft, p):
    s = A.Search(p, device=0)
    s.set_spectrum(fft)
    s.run()
    return s


# ------------------------------------------------------------------ standard path
@pytest.fixture(scope="module")
def std_case():
    _say("case: standard path, N=2^16 dt=64us zmax=200 numharm=16, 2 pulsars")
    fft, p, o = _case(1 << 16, 64e-6, 21,
                      [dict(f0=61.3, z=9.0, duty=0.05, amp=0.35),
                       dict(f0=377.7, z=-7.0, duty=0.08, amp=0.3)],
                      zmax=200, numharm=16, sigma=2.0, force_standard=True)
    s = _run_gpu(fft, p)
    oh, oc, nb = O.accelsearch(o, fft)
    yield fft, p, o, s, oh, oc
    s.close()


def test_std_geometry_matches(std_case):
    fft, p, o, s, oh, oc = std_case
    g = s.geom
    assert (g.nblocks, g.corr_uselen, g.numz, g.inmem) == (
        (int(o.highestbin - o.rlo) - 1) // (o.corr_uselen // 2), o.corr_uselen, o.numz, 0)


def test_std_plane_within_1e4_of_fp64(std_case):
    fft, p, o, s, oh, oc = std_case
    fr = [(1, 1), (2, 1), (4, 3), (8, 5), (16, 1), (16, 15)]
    err = _plane_error(s, o, fft, [0, 7, s.geom.nblocks - 1], fr)
    _say(f"standard path plane error vs FP64 oracle: {err:.3e}")
    assert err < 1e-4


def test_std_hits_match_oracle(std_case):
    fft, p, o, s, oh, oc = std_case
    gh = s.hits()
    assert len(gh) > 0
    worst = _compare_hits(gh, oh, list(o.powcut))
    _say(f"standard path: {len(gh)} hits, worst power rel. diff {worst:.2e}")
    assert worst < 1e-4


def test_std_candidates_identical(std_case):
    fft, p, o, s, oh, oc = std_case
    gc = s.candidates()
    _compare_cands(gc, oc)
    assert len(gc) >= 2


# ------------------------------------------------------------------ in-memory path
@pytest.fixture(scope="module")
def inmem_case():
    _say("case: in-memory path, c1 geometry (N=2^20 dt=64us zmax=50 numharm=8), amp 0.12")
    fft, p, o = _case(1 << 20, 64e-6, 1,
                      [dict(f0=123.456, z=20.0, duty=0.05, amp=0.12),
                       dict(f0=17.3, z=-3.0, duty=0.1, amp=0.08)],
                      zmax=50, numharm=8, sigma=2.0)
    assert o.inmem == 1
    s = _run_gpu(fft, p)
    oh, oc, nb = O.accelsearch(o, fft)
    yield fft, p, o, s, oh, oc
    s.close()


def test_inmem_plane_within_1e4_of_fp64(inmem_case):
    fft, p, o, s, oh, oc = inmem_case
    err = _plane_error(s, o, fft, [0, 100, s.geom.nblocks - 1], [(1, 1)])
    _say(f"in-memory path plane error vs FP64 oracle: {err:.3e}")
    assert err < 1e-4


def test_inmem_hits_and_candidates(inmem_case):
    fft, p, o, s, oh, oc = inmem_case
    worst = _compare_hits(s.hits(), oh, list(o.powcut))
    _say(f"in-memory path: {len(oh)} oracle hits, worst power rel. diff {worst:.2e}")
    assert worst < 1e-4
    _compare_cands(s.candidates(), oc)
    assert len(oc) >= 1


def test_config1_workload_identical():
    """The CPU-reference workload itself (2^20 samples, z=20 pulsar at
    amplitude 0.05, zmax 50, numharm 8)."""
    _say("case: config-1 workload, exact parameters")
    x = synth.config1_series()
    fft = A.spectrum_from_series(x, 64e-6)
    p = A.make_params(len(fft), len(x), 64e-6, zmax=50, numharm=8, sigma=2.0)
    o = O.make_obs(len(fft), len(x), 64e-6, zmax=50, numharm=8, sigma=2.0)
    with A.Search(p) as s:
        s.set_spectrum(fft)
        s.run()
        gh, gc = s.hits(), s.candidates()
    oh, oc, _ = O.accelsearch(o, fft)
    _compare_hits(gh, oh, list(o.powcut))
    _compare_cands(gc, oc)
    _say(f"config-1: {len(gh)} hits, {len(gc)} candidates, identical")


@pytest.mark.parametrize("zmax,numharm,inmem", [(0, 1, False), (0, 8, True), (2, 2, False),
                                               (30, 32, False), (100, 4, True)])
def test_edge_configurations(zmax, numharm, inmem):
    _say(f"case: zmax={zmax} numharm={numharm} inmem={inmem}")
    fft, p, o = _case(1 << 15, 5e-4, 33, [dict(f0=77.7, z=0.0, duty=0.05, amp=0.4)],
                      zmax=zmax, numharm=numharm, sigma=2.0, force_standard=not inmem)
    assert o.inmem == int(inmem)
    with _run_gpu(fft, p) as s:
        gh, gc = s.hits(), s.candidates()
    oh, oc, _ = O.accelsearch(o, fft)
    _compare_hits(gh, oh, list(o.powcut))
    _compare_cands(gc, oc)


def test_fft_input_with_lobin():
    """A .fft input whose first bin is lobin 1500 (-lobin): the spectrum
    holds bins [lobin, lobin + numbins) and every window reads it at
    bin - lobin (src/accel_utils.c:993-1022); candidates keep absolute r.
    rlo 6500 keeps the 1/4 subharmonic windows above lobin: below it the
    reference zero-pads, the window median is 0 and its normalisation
    infinite, so the powers there are inf/nan on both sides (and so is
    every in-memory search with lobin > 8, whose plane is filled from
    bin 8 up, src/accelsearch.c:166-181): no parity case."""
    _say("case: .fft input, lobin 1500, rlo 6500, zmax 30 numharm 4, standard path")
    N, dt, lobin = 1 << 16, 5e-4, 1500
    x = synth.make_series(N, dt, 35, [dict(f0=77.7, z=4.0, duty=0.05, amp=0.4)])
    spec = np.ascontiguousarray(A.spectrum_from_series(x, dt)[lobin:])
    kw = dict(zmax=30, numharm=4, sigma=2.0, lobin=lobin, rlo=6500.0, dat_input=False,
              force_standard=True)
    p = A.make_params(len(spec), N, dt, **kw)
    o = O.make_obs(len(spec), N, dt, **kw)
    assert o.inmem == 0 and o.lobin == lobin and o.highestbin == len(spec) - 1
    with _run_gpu(spec, p) as s:
        gh, gc = s.hits(), s.candidates()
    oh, oc, _ = O.accelsearch(o, spec)
    assert np.isfinite(oh["pow"]).all()
    _compare_hits(gh, oh, list(o.powcut))
    _compare_cands(gc, oc)
    # the injected pulsar (r = f0 T + z/2) as a 4-harmonic candidate
    assert any(c["numharm"] == 4 and abs(c["r"] - (77.7 * N * dt + 2.0)) < 1.0 for c in gc)
    _say(f"lobin: {len(gh)} hits, {len(gc)} candidates, identical")


@pytest.mark.parametrize("rlo,rhi,zmax,inmem", [
    (3000.0, 9000.0, 25, False), (1234.5, 5000.5, 7, True), (1234.5, 6000.0, 9, False),
    (1234.3, 6000.0, 9, False), (1234.8, 7000.0, 11, False), (1234.75, 6000.0, 5, True)])
def test_explicit_r_range_and_odd_zmax(rlo, rhi, zmax, inmem):
    """-rlo/-rhi bound the search (src/accel_utils.c:1686-1730); an odd zmax
    is rounded up to the next even value.  A fractional rlo starts each
    block's plane at floor(calc_required_r(1, startr)) while the standard
    path's sub-harmonic lookups are taken from startr + ii/2 itself
    (src/accel_utils.c:1035-1047): half-bin starts (x.5) shift them by one
    half-bin, other fractions by a fraction (the FP64 per-cell variant)."""
    _say(f"case: rlo={rlo} rhi={rhi} zmax={zmax} inmem={inmem}")
    fft, p, o = _case(1 << 16, 5e-4, 36, [dict(f0=77.7, z=3.0, duty=0.05, amp=0.4)],
                      zmax=zmax, numharm=8, sigma=2.0, rlo=rlo, rhi=rhi,
                      force_standard=not inmem)
    assert o.inmem == int(inmem) and o.zhi == zmax + 1
    with _run_gpu(fft, p) as s:
        assert (s.geom.rlo, s.geom.highestbin, s.geom.zhi) == (o.rlo, o.highestbin, o.zhi)
        gh, gc = s.hits(), s.candidates()
    oh, oc, _ = O.accelsearch(o, fft)
    _compare_hits(gh, oh, list(o.powcut))
    _compare_cands(gc, oc)
    assert len(gc) >= 1
    # candidate r is the summed plane's r over numharm (src/accel_utils.c:1431)
    assert all(rlo - 1 <= c["r"] * c["numharm"] <= rhi + 1 for c in gc)


def _poisson_counts(N, dt, seed, f0, fdot, mean, frac, width=0.04):
    """Photon counts per bin of a pulsed source: a Gaussian pulse of the
    given phase width carrying `frac` of the mean rate."""
    t = np.arange(N) * dt
    ph = np.mod(f0 * t + 0.5 * fdot * t * t, 1.0)
    d = np.minimum(ph, 1.0 - ph)
    prof = np.exp(-0.5 * (d / width) ** 2)
    rate = mean * ((1.0 - frac) + frac * prof / prof.mean())
    return np.random.default_rng(seed).poisson(rate).astype(np.float32)


@pytest.mark.parametrize("inmem", [False, True])
def test_photon_fft_input(inmem):
    """-photon on an .fft of photon counts: nph is bin 0's real part (the
    total count, src/accel_utils.c:1599-1605) and every block is normalised
    by 1/sqrt(nph) instead of its median (:1080-1086)."""
    _say(f"case: -photon, .fft of Poisson counts, zmax 20 numharm 8, inmem={inmem}")
    N, dt = 1 << 16, 5e-4
    x = _poisson_counts(N, dt, 38, 41.7, 0.02, 3.0, 0.08)
    spec = A.realfft_gpu(x)
    nph = float(spec[0].real)
    assert abs(nph - x.sum()) < 1e-3 * x.sum()
    kw = dict(zmax=20, numharm=8, sigma=2.0, dat_input=False, force_standard=not inmem,
              nph=nph)
    p = A.make_params(len(spec), N, dt, **kw)
    o = O.make_obs(len(spec), N, dt, **kw)
    assert o.inmem == int(inmem) and o.nph == nph
    with _run_gpu(spec, p) as s:
        assert s.geom.nph == nph
        assert np.all(s.norms() == 1.0 / np.sqrt(nph))
        gh, gc = s.hits(), s.candidates()
    oh, oc, _ = O.accelsearch(o, spec)
    worst = _compare_hits(gh, oh, list(o.powcut))
    _compare_cands(gc, oc)
    assert worst < 1e-4 and len(gc) >= 1
    r_exp = 41.7 * N * dt + 0.02 * (N * dt) ** 2 / 2
    assert any(abs(c["r"] - r_exp) < 3.0 for c in gc)
    _say(f"photon: {len(gh)} hits, {len(gc)} candidates, identical")


def test_photon_dat_series_keeps_bin0():
    """-photon on a .dat series: the de-reddened bin 0 (1, 0) is kept, not
    set to (1, 1) (src/accel_utils.c:1599-1624), so nph is 1."""
    _say("case: -photon, .dat series through the device preparation")
    N, dt = 1 << 16, 5e-4
    x = synth.make_series_f32(N, dt, 39, [dict(f0=41.7, z=3.0, duty=0.05, amp=0.3)])
    p = A.make_params(N // 2, N, dt, zmax=20, numharm=4, sigma=2.0, nph=1.0)
    with A.Search(p) as s:
        s.set_series(x)
        spec = s.spectrum()
        s.run()
        gh, gc = s.hits(), s.candidates()
    assert spec[0] == np.complex64(1.0 + 0.0j)
    ref, _ = O.deredden(A.realfft_gpu(x), 6, 200, 0.5 / dt, N * dt)
    assert np.array_equal(spec.view(np.uint32), ref.view(np.uint32))
    o = O.make_obs(len(spec), N, dt, zmax=20, numharm=4, sigma=2.0, nph=1.0)
    oh, oc, _ = O.accelsearch(o, spec)
    _compare_hits(gh, oh, list(o.powcut))
    _compare_cands(gc, oc)
    assert len(gc) >= 1


# ------------------------------------------------------------------ jerk (w) search
@pytest.fixture(scope="module")
def jerk_case():
    _say("case: jerk search, N=2^16 dt=64us zmax=40 wmax=60 numharm=4, pulsars w=+-40")
    fft, p, o = _case(1 << 16, 64e-6, 41,
                      [dict(f0=211.9, z=8.0, w=40.0, duty=0.06, amp=0.5),
                       dict(f0=97.3, z=0.0, w=-40.0, duty=0.04, amp=0.5)],
                      zmax=40, wmax=60, numharm=4, sigma=2.0, force_standard=True)
    assert o.numw == 7
    s = _run_gpu(fft, p)
    oh, oc, nb = O.accelsearch(o, fft)
    yield fft, p, o, s, oh, oc
    s.close()


def test_jerk_plane_within_1e4_of_fp64(jerk_case):
    fft, p, o, s, oh, oc = jerk_case
    assert s.geom.numw == o.numw
    err = _plane_error(s, o, fft, [0, s.geom.nblocks // 2, s.geom.nblocks - 1],
                       [(1, 1), (2, 1), (4, 1), (4, 3)])
    _say(f"jerk plane error vs FP64 oracle: {err:.3e}")
    assert err < 1e-4


def test_jerk_hits_and_candidates(jerk_case):
    fft, p, o, s, oh, oc = jerk_case
    gh = s.hits()
    assert len(gh) > 0 and gh["wi"].max() > 0
    worst = _compare_hits(gh, oh, list(o.powcut))
    _say(f"jerk: {len(gh)} hits, worst power rel. diff {worst:.2e}, {len(oc)} candidates")
    assert worst < 1e-4
    gc = s.candidates()
    _compare_cands(gc, oc)
    assert any(abs(c["w"]) > 0 for c in gc)


def test_repeat_runs_are_deterministic(std_case):
    fft, p, o, s, oh, oc = std_case
    c1 = s.candidates().copy()
    h1 = np.sort(s.hits(), order=["block", "stage", "wi", "zi", "ii"])
    s.run()
    h2 = np.sort(s.hits(), order=["block", "stage", "wi", "zi", "ii"])
    assert np.array_equal(h1, h2)
    assert np.array_equal(c1, s.candidates())


# ------------------------------------------------------------------ full size
CONFIG2_FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                               "config2_candidates.json")


def test_config2_full_size_properties():
    """2^23 samples, zmax 200, numharm 16 (standard path): the full
    candidate list equals the oracle's (tests/golden/config2_candidates.json,
    written by tests/golden/make_config2_candidates.py over all 5432
    blocks), the 3 injected pulsars are found, and plane spot checks and
    the first blocks' hits agree with the oracle."""
    _say("case: config-2 full size")
    x, ps = synth.config2_series()
    dt = 64e-6
    fft = A.spectrum_from_series(x, dt)
    p = A.make_params(len(fft), len(x), dt, zmax=200, numharm=16, sigma=2.0)
    o = O.make_obs(len(fft), len(x), dt, zmax=200, numharm=16, sigma=2.0)
    assert o.inmem == 0
    fix = json.load(open(CONFIG2_FIXTURE))
    same_input = hashlib.sha256(fft.tobytes()).hexdigest() == fix["spectrum_sha256"]
    with A.Search(p) as s:
        s.set_spectrum(fft)
        s.run()
        gc, gh = s.candidates(), s.hits()
        want = np.array([[c["power"], c["sigma"], c["numharm"], c["r"], c["z"], c["w"]]
                         for c in fix["candidates"]])
        _say(f"full candidate list: {len(gc)} here, {len(want)} from the oracle "
             f"(input spectrum {'identical' if same_input else 'DIFFERENT'})")
        _compare_cands(gc, want)
        T = len(x) * dt
        _say("injected: " + ", ".join(f"f0={q['f0']:.2f} z={q['z']:.1f} amp={q['amp']:.3f}"
                                       for q in ps))
        _say("top candidates: " + ", ".join(f"r={c['r']:.2f} z={c['z']:.1f} h={c['numharm']} "
                                             f"sig={c['sigma']:.1f}" for c in gc[:6]))
        found = 0
        for q in ps:
            r_exp = q["f0"] * T + q["z"] / 2
            near = [c for c in gc for m in (1, 0.5, 2, 1 / 3, 3, 0.25, 4)
                    if abs(c["r"] - m * r_exp) < 2.0]
            # predicted noise-free summed power of the best harmonic stage:
            # harmonic h of a Gaussian pulse of width sig carries
            # N (A sig sqrt(2 pi))^2 exp(-4 pi^2 sig^2 h^2); only harmonics
            # with |h z| <= zmax stay inside the searched plane
            sig = q["duty"] / (2 * np.sqrt(2 * np.log(2)))
            p1 = len(x) * (q["amp"] * sig * np.sqrt(2 * np.pi)) ** 2
            best = 0.0
            for st in range(o.numharmstages):
                hs = [h for h in range(1, (1 << st) + 1) if abs(h * q["z"]) <= 200]
                pw = sum(p1 * np.exp(-4 * np.pi ** 2 * sig ** 2 * h * h) for h in hs)
                best = max(best, (pw + (1 << st)) / o.powcut[st])
            _say(f"pulsar f0={q['f0']:.2f}: predicted best stage power/cut {best:.2f}, "
                 f"{'found' if near else 'not found'}")
            if best > 1.5:
                assert near, ("detectable pulsar not found", q)
            found += bool(near)
        assert found >= 1
        err = _plane_error(s, o, fft, [0, 2500, s.geom.nblocks - 1], [(1, 1), (2, 1), (16, 9)])
        _say(f"config-2 plane spot-check error vs FP64 oracle: {err:.3e}")
        assert err < 1e-4
        ohits, _, nb = O.accelsearch(o, fft, max_blocks=12)
        sub = gh[gh["block"] < 12]
        _compare_hits(sub, ohits, list(o.powcut))
