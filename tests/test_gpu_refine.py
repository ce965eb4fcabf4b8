"""GPU parity of the candidate refinement and the output files.

* The reference's own optimiser battery (tests/test_maximize_rz.py of the
  reference: injected (r, z), perturbed starts, its recovery tolerances) run
  through the HIP kernel.
* The HIP refinement against the oracle's restatement for the same starts:
  both run the same Nelder-Mead scheme on the same function, so they end
  at the same optimum up to the flatness of the peak.  The interpolations
  differ at the 1e-7 level (FP64 sums on the GPU, the reference's FP32
  running sum in the oracle), which moves the optimum of a noisy peak by
  up to a few 1e-4 bins (the reference's own harness allows ~1e-3 bins of
  FFTW jitter, tests/test_maximize_rz.py:101-104).  Bars: 2e-3 bins in r,
  8e-3 in z (z is searched as z/4), the summed normalised power of the
  derivatives within 1e-3 (first order in the endpoint difference: the
  simplex maximises a sum built with one kernel width for all harmonics,
  the derivatives use each harmonic's own, src/maximize_rz.c:214-226), and
  the derivatives at the GPU's endpoint equal to the oracle's at that same
  point within 1e-5.
* A search -> optimize -> ACCEL / .cand round trip, read back with the
  line rules of the reference's sifting.py (python/presto/sifting.py:40-41,
  1075-1177)."""
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
from presto_amd import synth

pytestmark = pytest.mark.gpu
A = pytest.importorskip("presto_amd.accel")

N = 2 ** 14
R_RZ, R_HARM = N / 4.0, N / 16.0
THEO_POW = N ** 2.0 / 4.0
RZ_OFFSETS = [(0.0, 0.0), (0.12, -0.8), (-0.2, 1.3), (0.3, 2.0), (-0.35, -2.5)]
HARM_OFFSETS = [(0.0, 0.0), (0.15, 1.0), (-0.25, -1.5)]
Z_VALUES = [0.0, 1.0, 2.5, 5.0, -5.0, 10.0, -10.0, 30.0, -30.0, 55.0]
HARM_Z_VALUES = [0.0, 5.0, -8.0, 20.0]
TOL = {0.0: dict(dr=0.01, dz=0.03, plo=0.99, phi=1.01),
       1.0: dict(dr=0.05, dz=0.20, plo=0.90, phi=1.12)}


def _say(msg):
    print(msg, flush=True)


def _signal(r, z, noise_amp, seed, numharm=1):
    us = np.arange(N, dtype=np.float64) / N
    sig = np.zeros(N)
    for h in range(1, numharm + 1):
        rh, zh = r * h, z * h
        sig += np.cos(2.0 * np.pi * (us * (us * (zh / 2.0) + (rh - 0.5 * zh))))
    if noise_amp:
        sig = sig + noise_amp * np.random.default_rng(seed).standard_normal(N)
    return O.realfft(sig.astype(np.float32)).view(np.complex64)


@pytest.mark.parametrize("na,seed", [(0.0, 0), (1.0, 12345)])
def test_gpu_refinement_passes_reference_battery(na, seed):
    _say(f"case: reference optimiser battery on the GPU, noise {na}")
    tol = TOL[na]
    for z in Z_VALUES:
        ft = _signal(R_RZ, z, na, seed)
        rs = np.array([R_RZ + d for d, _ in RZ_OFFSETS])
        zs = np.array([z + d for _, d in RZ_OFFSETS])
        ro, zo, pw, _ = A.max_rz_harmonics(ft, rs, zs, 1)
        assert np.all(np.abs(ro - R_RZ) <= tol["dr"]), (z, ro - R_RZ)
        assert np.all(np.abs(zo - z) <= tol["dz"]), (z, zo - z)
        pf = pw[:, 0] / THEO_POW
        assert np.all((pf >= tol["plo"]) & (pf <= tol["phi"])), (z, pf)
    for numharm in (2, 3, 4):
        for z in HARM_Z_VALUES:
            ft = _signal(R_HARM, z, na, seed, numharm)
            rs = np.array([R_HARM + d for d, _ in HARM_OFFSETS])
            zs = np.array([z + d for _, d in HARM_OFFSETS])
            ro, zo, pw, der = A.max_rz_harmonics(ft, rs, zs, numharm)
            total = np.array([sum(d.pow for d in dl) for dl in der])
            assert np.all(np.abs(ro - R_HARM) <= tol["dr"]), (numharm, z, ro - R_HARM)
            assert np.all(np.abs(zo - z) <= tol["dz"]), (numharm, z, zo - z)
            pf = total / (numharm * THEO_POW)
            assert np.all((pf >= tol["plo"]) & (pf <= tol["phi"])), (numharm, z, pf)


def _check_endpoint(spec, numharm, ro, zo, der, oro, ozo, ords, lobin=0.0):
    """The GPU optimum against the oracle's: the endpoint differences, the
    relative difference of the summed normalised derivative powers, and
    the worst difference of every harmonic's derivatives against the
    oracle's evaluated at the GPU's own endpoint (pow, locpow relative;
    phase in rad)."""
    dr, dz = abs(ro - oro), abs(zo - ozo)
    gobj = sum(d.pow / d.locpow for d in der)
    oobj = sum(d.pow / d.locpow for d in ords)
    worst_p = worst_ph = 0.0
    for k, d in enumerate(der):
        n = k + 1
        lp = O.localpower3d(spec, (ro - lobin) * n, zo * n, 0.0)
        od = O.derivs3d(spec, (ro - lobin) * n, zo * n, 0.0, lp)
        worst_p = max(worst_p, abs(d.pow - od.pow) / max(od.pow, 1e-30),
                      abs(d.locpow - od.locpow) / od.locpow)
        e = abs(d.phs - od.phs)
        worst_ph = max(worst_ph, min(e, abs(e - 2 * np.pi)))
    return dr, dz, abs(gobj - oobj) / oobj, worst_p, worst_ph


def test_gpu_refinement_matches_oracle():
    _say("case: GPU refinement vs oracle on noisy multi-harmonic starts")
    worst = np.zeros(5)
    for numharm, r, z, dr, dz in [(1, R_HARM, 7.0, 0.2, -1.0), (2, R_HARM, -12.0, -0.3, 2.0),
                                  (4, R_HARM, 20.0, 0.1, 1.5), (8, R_HARM / 2, 3.0, 0.05, -0.5)]:
        ft = _signal(r, z, 1.0, 99 + numharm, numharm)
        ro, zo, pw, der = A.max_rz_harmonics(ft, [r + dr], [z + dz], numharm)
        oro, ozo, ords, opw, _ = O.max_rz_arr_harmonics(ft, r + dr, z + dz, numharm)
        res = _check_endpoint(ft, numharm, ro[0], zo[0], der[0], oro, ozo, ords)
        worst = np.maximum(worst, res)
    _say("refinement vs oracle: |dr| %.2e |dz| %.2e objective %.2e derivs %.2e phase %.2e"
         % tuple(worst))
    assert worst[0] < 2e-3 and worst[1] < 8e-3 and worst[2] < 1e-3
    assert worst[3] < 1e-5 and worst[4] < 1e-4


FUND_RE = re.compile(r"^\d")        # python/presto/sifting.py:40
HARMS_RE = re.compile(r"^[ ]\d")    # python/presto/sifting.py:41


def _parse_accel(path):
    """The candidate and harmonic lines as sifting.py's
    candlist_from_candfile reads them."""
    cands, harms = [], {}
    cur, nh = None, 0
    for line in open(path):
        if FUND_RE.match(line):
            f = line.split()
            cands.append(dict(num=int(f[0]), sigma=float(f[1]), ipow=float(f[2]),
                              cpow=float(f[3]), numharm=int(f[4]),
                              r=float(f[7].split("(")[0]), z=float(f[9].split("(")[0])))
            continue
        if HARMS_RE.match(line):
            f = line.split()
            cur = int(f[0])
            harms[cur] = [(f[3], float(f[9].split("(")[0]))]
            continue
        if cur is not None and line.strip() and not line.startswith("-") and \
                len(harms[cur]) < [c["numharm"] for c in cands if c["num"] == cur][0]:
            f = line.split()
            harms[cur].append((f[2], float(f[8].split("(")[0])))
    return cands, harms


def test_search_optimize_and_outputs(tmp_path):
    _say("case: search -> optimize -> ACCEL/.cand, N=2^16 zmax 100 numharm 8")
    dt = 64e-6
    x = synth.make_series(1 << 16, dt, 57, [dict(f0=61.3, z=9.0, duty=0.05, amp=0.35),
                                            dict(f0=377.7, z=-7.0, duty=0.08, amp=0.3)])
    fft = A.spectrum_from_series(x, dt)
    p = A.make_params(len(fft), len(x), dt, zmax=100, numharm=8, sigma=2.0)
    with A.Search(p) as s:
        s.set_spectrum(fft)
        s.run()
        cands = s.candidates()
        assert len(cands) >= 2
        s.optimize()
        props = s.fund_props()
        acc = str(tmp_path / "test_ACCEL_100")
        cand = str(tmp_path / "test_ACCEL_100.cand")
        inf = tmp_path / "test.inf"
        inf.write_text(f" Number of bins in the time series   =  {len(x)}\n"
                       f" Width of each time series bin (sec) =  {dt}\n")
        s.write_outputs(acc, cand, str(inf))
        # refined harmonics of the two strongest candidates against the oracle
        for i in range(2):
            c = cands[i]
            hr, hz, hw, pw, der = s.harmonics(i)
            assert len(hr) == c["numharm"]
            oro, ozo, ords, opw, _ = O.max_rz_arr_harmonics(
                fft, c["r"] - p.lobin, c["z"], int(c["numharm"]))
            res = _check_endpoint(fft, int(c["numharm"]), hr[0], hz[0], der, oro + p.lobin,
                                  ozo, ords, lobin=p.lobin)
            _say(f"  cand {i}: |dr| {res[0]:.2e} |dz| {res[1]:.2e} objective {res[2]:.2e} "
                 f"derivs {res[3]:.2e} phase {res[4]:.2e}")
            assert res[0] < 2e-3 and res[1] < 8e-3 and res[2] < 1e-3
            assert res[3] < 1e-5 and res[4] < 1e-4
            assert np.allclose(hr, hr[0] * np.arange(1, len(hr) + 1))
    # the .cand file: one 88-byte fourierprops per candidate, as computed
    raw = open(cand, "rb").read()
    assert len(raw) == 88 * len(cands)
    dt_fp = np.dtype([("r", "f8"), ("rerr", "f4"), ("pad0", "V4"), ("z", "f8"), ("zerr", "f4"),
                      ("pad1", "V4"), ("w", "f8"), ("werr", "f4")] +
                     [(n, "f4") for n in ("pow", "powerr", "sig", "rawpow", "phs", "phserr",
                                          "cen", "cenerr", "pur", "purerr", "locpow")])
    assert dt_fp.itemsize == 88
    rec = np.frombuffer(raw, dtype=dt_fp)
    for i, pr in enumerate(props):
        assert rec["r"][i] == pr.r and rec["z"][i] == pr.z and rec["pow"][i] == np.float32(pr.pow)
        assert rec["rerr"][i] == np.float32(0.5 / cands[i]["numharm"])
    # the text file, read the way sifting.py reads it
    fc, fh = _parse_accel(acc)
    assert [c["num"] for c in fc] == list(range(1, len(cands) + 1))
    for c, g in zip(fc, cands):
        assert c["numharm"] == g["numharm"]
        assert abs(c["sigma"] - g["sigma"]) <= 0.006
        assert abs(c["r"] - g["r"]) <= 0.5 / g["numharm"] + 1e-6
        assert abs(c["z"] - g["z"]) <= 2.0 / g["numharm"] + 1e-6
        assert len(fh[c["num"]]) == g["numharm"]
    txt = open(acc).read()
    assert "Number of bins in the time series" in txt
    _say(f"ACCEL file: {len(fc)} candidates, harmonics tables complete")


def test_optimize_with_lobin():
    """Harmonic polishing of a .fft input whose first bin is lobin: the
    reference starts the optimiser at r - lobin in the array's own bins,
    evaluates harmonic k at k (r - lobin), and reports (r + lobin) k
    (src/accel_utils.c:563-581)."""
    _say("case: optimize, .fft input with lobin 1500, zmax 30 numharm 4")
    N2, dt, lobin = 1 << 16, 5e-4, 1500
    x = synth.make_series(N2, dt, 35, [dict(f0=77.7, z=4.0, duty=0.05, amp=0.4)])
    spec = np.ascontiguousarray(A.spectrum_from_series(x, dt)[lobin:])
    p = A.make_params(len(spec), N2, dt, zmax=30, numharm=4, sigma=2.0, lobin=lobin,
                      rlo=6500.0, dat_input=False, force_standard=True)
    with A.Search(p) as s:
        s.set_spectrum(spec)
        s.run()
        cands = s.candidates()
        assert len(cands) >= 2
        s.optimize()
        for i in range(2):
            c = cands[i]
            hr, hz, hw, pw, der = s.harmonics(i)
            oro, ozo, ords, opw, _ = O.max_rz_arr_harmonics(
                spec, c["r"] - lobin, c["z"], int(c["numharm"]))
            res = _check_endpoint(spec, int(c["numharm"]), hr[0], hz[0], der, oro + lobin,
                                  ozo, ords, lobin=lobin)
            _say(f"  cand {i} (r {c['r']:.2f}): |dr| {res[0]:.2e} |dz| {res[1]:.2e} "
                 f"objective {res[2]:.2e} derivs {res[3]:.2e} phase {res[4]:.2e}")
            assert res[0] < 2e-3 and res[1] < 8e-3 and res[2] < 1e-3
            assert res[3] < 1e-5 and res[4] < 1e-4
            assert np.allclose(hr, hr[0] * np.arange(1, len(hr) + 1))


def _jerk_signal(r, z, w, noise_amp, seed, numharm=1):
    us = np.arange(N, dtype=np.float64) / N
    sig = np.zeros(N)
    for h in range(1, numharm + 1):
        rh, zh, wh = r * h, z * h, w * h
        r0, z0 = rh - 0.5 * zh + wh / 12.0, zh - 0.5 * wh
        sig += np.cos(2.0 * np.pi * (us * (us * (us * wh / 6.0 + z0 / 2.0) + r0)))
    if noise_amp:
        sig = sig + noise_amp * np.random.default_rng(seed).standard_normal(N)
    return O.realfft(sig.astype(np.float32)).view(np.complex64)


def test_gpu_jerk_refinement_matches_oracle():
    """max_rzw_arr_harmonics is a grid search over an oversampled volume:
    the GPU and the oracle must pick the same grid point, and agree on the
    powers at it (derivative fields d2pow/d2phs are finite differences of
    FP32-level values and are compared loosely)."""
    _say("case: GPU jerk refinement vs oracle")
    cases = [(1, R_HARM + 0.37, 10.0, 60.0, 0.1, 0.6, 7.0),
             (2, R_HARM + 0.11, -6.0, -40.0, -0.05, -0.4, 5.0),
             (4, R_HARM + 0.62, 4.0, 20.0, 0.08, 0.3, -6.0),
             (2, R_HARM + 0.9, 0.0, 0.0, 0.1, 0.5, 0.0)]
    for nh, r, z, w, dr, dz, dw in cases:
        ft = _jerk_signal(r, z, w, 0.5, 7 + nh, nh)
        ro, zo, wo, pw, der = A.max_rzw_harmonics(ft, [r + dr], [z + dz], [w + dw], nh)
        oro, ozo, owo, ords, opw = O.max_rzw_arr_harmonics(ft, r + dr, z + dz, w + dw, nh)
        _say(f"  numharm {nh}: gpu ({ro[0]:.4f}, {zo[0]:.3f}, {wo[0]:.2f}) "
             f"oracle ({oro:.4f}, {ozo:.3f}, {owo:.2f})")
        assert abs(ro[0] - oro) < 1e-9 and abs(zo[0] - ozo) < 1e-9 and abs(wo[0] - owo) < 1e-9
        assert abs(ro[0] - r) <= 0.5 / (8 * nh) + 0.1 / nh
        for k in range(nh):
            assert abs(pw[0, k] - opw[k]) <= 2e-5 * opw[k]
            assert abs(der[0][k].locpow - ords[k].locpow) <= 2e-5 * ords[k].locpow
            dphs = abs(der[0][k].phs - ords[k].phs)
            assert min(dphs, abs(dphs - 2 * np.pi)) < 1e-3


def test_jerk_search_optimize_and_outputs(tmp_path):
    _say("case: jerk search -> optimize -> ACCEL, N=2^16 zmax 40 wmax 60 numharm 4")
    dt = 64e-6
    x = synth.make_series(1 << 16, dt, 41, [dict(f0=211.9, z=8.0, w=40.0, duty=0.06, amp=0.5),
                                            dict(f0=97.3, z=0.0, w=-40.0, duty=0.04, amp=0.5)])
    fft = A.spectrum_from_series(x, dt)
    p = A.make_params(len(fft), len(x), dt, zmax=40, wmax=60, numharm=4, sigma=2.0,
                      force_standard=True)
    with A.Search(p) as s:
        s.set_spectrum(fft)
        s.run()
        cands = s.candidates()
        assert len(cands) >= 1
        s.optimize()
        for i in range(min(2, len(cands))):
            c = cands[i]
            hr, hz, hw, pw, der = s.harmonics(i)
            oro, ozo, owo, ords, opw = O.max_rzw_arr_harmonics(
                fft, c["r"] - p.lobin, c["z"], c["w"], int(c["numharm"]))
            assert abs(hr[0] - (oro + p.lobin)) < 1e-9
            assert abs(hz[0] - ozo) < 1e-9 and abs(hw[0] - owo) < 1e-9
            assert np.allclose(pw, opw[:len(pw)], rtol=2e-5)
        acc = str(tmp_path / "jerk_ACCEL_40_JERK_60")
        s.write_outputs(acc, str(tmp_path / "jerk.cand"))
    lines = [l for l in open(acc) if FUND_RE.match(l)]
    assert len(lines) == len(cands)
    for l, c in zip(lines, cands):
        f = l.split()
        # jerk files carry FFT 'w' after FFT 'z' (src/accel_utils.c:825-831)
        assert abs(float(f[10].split("(")[0]) - c["w"]) <= 20.0 / c["numharm"] + 1e-6
    head = open(acc).readline()
    assert "FFT 'w'" in head


def test_unpolished_rz_optimize_matches_oracle():
    """-noharmpolish (src/accel_utils.c:582-622): harmonic k of a candidate
    is refined on its own from (k r - lobin, k z) by max_rz_arr, the raw
    power as the objective; the endpoints, powers and derivatives of each
    harmonic against the oracle's max_rz_arr from the same start."""
    _say("case: -noharmpolish rz optimize, N=2^16 zmax 100 numharm 8")
    dt = 64e-6
    x = synth.make_series(1 << 16, dt, 57, [dict(f0=61.3, z=9.0, duty=0.05, amp=0.35),
                                            dict(f0=377.7, z=-7.0, duty=0.08, amp=0.3)])
    fft = A.spectrum_from_series(x, dt)
    p = A.make_params(len(fft), len(x), dt, zmax=100, numharm=8, sigma=2.0, noharmpolish=True)
    with A.Search(p) as s:
        s.set_spectrum(fft)
        s.run()
        cands = s.candidates()
        assert len(cands) >= 2
        s.optimize()
        worst = [0.0] * 4
        for i in range(2):
            c = cands[i]
            hr, hz, hw, pw, der = s.harmonics(i)
            assert len(hr) == c["numharm"]
            for k in range(int(c["numharm"])):
                opw, oro, ozo, ord_ = O.max_rz_arr(fft, c["r"] * (k + 1) - p.lobin,
                                                   c["z"] * (k + 1))
                lp = O.localpower3d(fft, hr[k] - p.lobin, hz[k], 0.0)
                od = O.derivs3d(fft, hr[k] - p.lobin, hz[k], 0.0, lp)
                e = [abs(hr[k] - (oro + p.lobin)), abs(hz[k] - ozo),
                     abs(pw[k] - opw) / opw, abs(der[k].pow - od.pow) / max(od.pow, 1e-30)]
                worst = [max(a, b) for a, b in zip(worst, e)]
        _say(f"  worst |dr| {worst[0]:.2e} |dz| {worst[1]:.2e} power {worst[2]:.2e} "
             f"derivs {worst[3]:.2e}")
        assert worst[0] < 2e-3 and worst[1] < 8e-3 and worst[2] < 1e-3 and worst[3] < 1e-5
        # not the polished multiples
        hr, _, _, _, _ = s.harmonics(0)
        assert len(hr) > 1 and not np.allclose(hr, hr[0] * np.arange(1, len(hr) + 1),
                                              rtol=0, atol=1e-9)


def test_unpolished_rzw_matches_oracle():
    """-noharmpolish jerk refinement, max_rzw_arr (src/maximize_rzw.c:3-55):
    the same grid point as the oracle, the raw grid maximum as the power
    and the start's local power in the derivatives."""
    _say("case: GPU unpolished jerk refinement vs oracle")
    dt = 64e-6
    x = synth.make_series(1 << 16, dt, 41, [dict(f0=211.9, z=8.0, w=40.0, duty=0.06, amp=0.5)])
    fft = A.spectrum_from_series(x, dt)
    p = A.make_params(len(fft), len(x), dt, zmax=40, wmax=60, numharm=2, sigma=2.0,
                      force_standard=True, noharmpolish=True)
    with A.Search(p) as s:
        s.set_spectrum(fft)
        s.run()
        cands = s.candidates()
        assert len(cands) >= 1
        s.optimize()
        for i in range(min(2, len(cands))):
            c = cands[i]
            hr, hz, hw, pw, der = s.harmonics(i)
            for k in range(int(c["numharm"])):
                opw, oro, ozo, owo, ord_ = O.max_rzw_arr(fft, c["r"] * (k + 1) - p.lobin,
                                                         c["z"] * (k + 1), c["w"] * (k + 1))
                _say(f"  cand {i} harmonic {k + 1}: gpu ({hr[k]:.4f}, {hz[k]:.3f}, "
                     f"{hw[k]:.2f}) oracle ({oro + p.lobin:.4f}, {ozo:.3f}, {owo:.2f})")
                assert abs(hr[k] - (oro + p.lobin)) < 1e-9
                assert abs(hz[k] - ozo) < 1e-9 and abs(hw[k] - owo) < 1e-9
                assert abs(pw[k] - opw) <= 2e-5 * opw
                assert abs(der[k].locpow - ord_.locpow) <= 2e-5 * ord_.locpow
                assert abs(der[k].pow - ord_.pow) <= 2e-5 * ord_.pow


def test_refinement_retry_after_a_wait_runs_out():
    """A candidate's harmonics are refined by workgroups that wait for each
    other at every simplex step.  With the first attempt's wait bound set to
    one poll the groups give up, and each candidate is refined again on its
    own with the full bound: the refined properties equal a normal run's
    bit for bit."""
    _say("case: refinement retry, N=2^16 zmax 100 numharm 8, first wait 1 poll")
    dt = 64e-6
    x = synth.make_series(1 << 16, dt, 58, [dict(f0=61.3, z=9.0, duty=0.05, amp=0.35),
                                            dict(f0=377.7, z=-7.0, duty=0.08, amp=0.3)])
    fft = A.spectrum_from_series(x, dt)
    p = A.make_params(len(fft), len(x), dt, zmax=100, numharm=8, sigma=2.0)
    with A.Search(p) as s:
        s.set_spectrum(fft)
        s.run()
        c = s.candidates()
        assert len(c) >= 2 and max(c["numharm"]) > 1
        s.optimize()
        want = [bytes(v) for v in s.fund_props()]
        A.lib.accel_refine_set_first_wait(0)
        try:
            s.optimize()
            got = [bytes(v) for v in s.fund_props()]
        finally:
            A.lib.accel_refine_set_first_wait(24)
    assert got == want
