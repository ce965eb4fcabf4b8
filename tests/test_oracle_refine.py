"""Pins the oracle's candidate refinement (oracle/o_refine.c).

The reference's own battery for its optimisers (tests/test_maximize_rz.py
of the reference: signals at known (r, z), perturbed starting points and
the recovery tolerances written there) is restated here against the
oracle; the exact GSL simplex path is not reproducible without GSL, so the
tolerances are the pin.  Closed forms pin calc_props / calc_rzwerrs: a
constant-amplitude sinusoid has purity 1, centroid 1/2 and phase 0 at its
peak, and the period / frequency conversions are plain algebra."""
import numpy as np
import pytest

from oracle import oracle as O

N = 2 ** 14
R_RZ = N / 4.0
R_HARM = N / 16.0
THEO_POW = N ** 2.0 / 4.0
RZ_OFFSETS = [(0.0, 0.0), (0.12, -0.8), (-0.2, 1.3), (0.3, 2.0), (-0.35, -2.5)]
HARM_OFFSETS = [(0.0, 0.0), (0.15, 1.0), (-0.25, -1.5)]
Z_VALUES = [0.0, 1.0, 2.5, 5.0, -5.0, 10.0, -10.0, 30.0, -30.0, 55.0]
HARM_Z_VALUES = [0.0, 5.0, -8.0, 20.0]
NOISE_CASES = [(0.0, 0), (1.0, 12345)]
TOL = {0.0: dict(dr=0.01, dz=0.03, plo=0.99, phi=1.01),
       1.0: dict(dr=0.05, dz=0.20, plo=0.90, phi=1.12)}


def _rfft(sig):
    """PRESTO's forward real FFT (packed, unnormalised) of a float32 series."""
    return O.realfft(np.asarray(sig, dtype=np.float32)).view(np.complex64)


def make_signal(r, z, noise_amp=0.0, seed=0, w=0.0, numharm=1):
    us = np.arange(N, dtype=np.float64) / N
    sig = np.zeros(N)
    for h in range(1, numharm + 1):
        rh, zh, wh = r * h, z * h, w * h
        r0, z0 = rh - 0.5 * zh + wh / 12.0, zh - 0.5 * wh
        sig += np.cos(2.0 * np.pi * (us * (us * (us * wh / 6.0 + z0 / 2.0) + r0)))
    if noise_amp:
        sig = sig + noise_amp * np.random.default_rng(seed).standard_normal(N)
    return _rfft(sig)


def _check(tag, rmax, zmax, pw, theo, r, z, na):
    tol = TOL[na]
    pf = pw / theo
    assert abs(rmax - r) <= tol["dr"], (tag, rmax - r)
    assert abs(zmax - z) <= tol["dz"], (tag, zmax - z)
    assert tol["plo"] <= pf <= tol["phi"], (tag, pf)


@pytest.mark.parametrize("na,seed", NOISE_CASES)
@pytest.mark.parametrize("z", Z_VALUES)
def test_max_rz_arr_recovers_injected_peak(z, na, seed):
    # tests/test_maximize_rz.py:114-135 of the reference
    ft = make_signal(R_RZ, z, na, seed)
    for dr, dz in RZ_OFFSETS:
        pw, rmax, zmax, rd = O.max_rz_arr(ft, R_RZ + dr, z + dz)
        _check(("rz", z, dr, dz), rmax, zmax, pw, THEO_POW, R_RZ, z, na)
        assert rd.pow > 0 and rd.locpow > 0


@pytest.mark.parametrize("numharm", [2, 3, 4])
@pytest.mark.parametrize("na,seed", NOISE_CASES)
def test_max_rz_arr_harmonics_recovers_injected_peak(numharm, na, seed):
    for z in HARM_Z_VALUES:
        ft = make_signal(R_HARM, z, na, seed, numharm=numharm)
        for dr, dz in HARM_OFFSETS:
            rmax, zmax, rds, pw, nev = O.max_rz_arr_harmonics(ft, R_HARM + dr, z + dz, numharm)
            # the reference's Python wrapper sums rderivs.pow (norm = 1)
            total = sum(rd.pow for rd in rds)
            _check(("rzh", numharm, z, dr, dz), rmax, zmax, total, numharm * THEO_POW,
                   R_HARM, z, na)
            assert nev < 2 * 5000 + 10


@pytest.mark.parametrize("numharm", [1, 2, 4])
@pytest.mark.parametrize("z,w", [(0.0, 0.0), (10.0, 60.0), (-20.0, -100.0)])
def test_max_rzw_arr_harmonics_lands_on_nearest_grid_point(numharm, z, w):
    """The jerk optimiser is a grid search (src/maximize_rzw.c:84-191):
    r steps of 1/(8 numharm), z steps of 1/(2 numharm), w steps of
    5/numharm around the start.  Noise-free injections must come back
    within half a step (plus the start's own grid offset) with the summed
    power of numharm full harmonics."""
    r = R_HARM + 0.37
    ft = make_signal(r, z, 0.0, 0, w=w, numharm=numharm)
    rmax, zmax, wmax, rds, pw = O.max_rzw_arr_harmonics(ft, r + 0.1, z + 0.6, w + 7.0,
                                                        numharm)
    assert abs(rmax - r) <= 0.5 / (8 * numharm) + 1e-9
    assert abs(zmax - z) <= 0.25 / numharm + 1e-9 + 0.6 % (0.5 / numharm)
    assert abs(wmax - w) <= 2.5 / numharm + 1e-9 + 7.0 % (5.0 / numharm)
    assert 0.97 <= sum(pw) / (numharm * THEO_POW) <= 1.01


@pytest.mark.parametrize("z,w", [(0.0, 0.0), (10.0, 60.0), (-20.0, -100.0)])
def test_max_rzw_arr_single_harmonic_grid(z, w):
    """The unpolished jerk optimiser (src/maximize_rzw.c:3-55) searches the
    9 x 9 x 24 grid r = (int) rin - 1 + k/8, zin +- 2, win +- 20 in steps of
    1/2 and 5: a noise-free injection comes back within half a step of
    the grid point nearest to it, with its full power, and the returned
    power is the raw (not locally normalised) grid maximum."""
    r = R_HARM + 0.37
    ft = make_signal(r, z, 0.0, 0, w=w)
    zin, win = z + 0.6, w + 7.0
    pw, rmax, zmax, wmax, rd = O.max_rzw_arr(ft, r + 0.1, zin, win)
    # on the grid
    assert abs((rmax - (int(r + 0.1) - 1)) * 8 - round((rmax - (int(r + 0.1) - 1)) * 8)) < 1e-9
    assert abs((zmax - zin) * 2 - round((zmax - zin) * 2)) < 1e-9 and abs(zmax - zin) <= 2.0
    assert abs((wmax - win) / 5 - round((wmax - win) / 5)) < 1e-9 and abs(wmax - win) <= 20.0
    assert abs(rmax - r) <= 0.5 / 8 + 1e-9
    assert abs(zmax - z) <= 0.25 + 1e-9 + 0.1
    assert abs(wmax - w) <= 2.5 + 1e-9 + 2.0
    assert 0.97 <= pw / THEO_POW <= 1.01
    # the derivatives use the local power at the start, not at the peak
    lp0 = O.localpower3d(ft, r + 0.1, zin, win)
    assert abs(rd.locpow - np.float32(lp0)) <= 1e-6 * lp0
    # the same single-harmonic peak as the polished search's (one harmonic)
    r1, z1, w1, _, _ = O.max_rzw_arr_harmonics(ft, r + 0.1, zin, win, 1)
    assert abs(r1 - rmax) <= 1.0 / 8 + 1e-9 and abs(z1 - zmax) <= 0.5 + 1e-9


def test_calc_props_of_a_pure_sinusoid():
    """Middleditch, Deich & Kulkarni (1993): a constant-amplitude signal
    over the whole run has purity 1 and centroid 1/2; its phase at the peak
    is that of the injected cosine (0)."""
    for z, w in [(0.0, 0.0), (10.0, 0.0), (10.0, 60.0)]:
        ft = make_signal(R_RZ, z, w=w)
        rd = O.derivs3d(ft, R_RZ, z, w, 1.0)
        p = O.calc_props(rd, R_RZ, z, w)
        assert abs(p.cen - 0.5) < 1e-4
        assert abs(p.pur - 1.0) < 0.02
        assert min(p.phs, 2 * np.pi - p.phs) < 1e-3
        assert abs(rd.pow / THEO_POW - 1.0) < 1e-3
        # error formulas (src/characteristics.c:217-262) from pow and pur
        t = rd.pow / rd.locpow
        assert np.isclose(p.powerr, np.sqrt(2 * t), rtol=1e-6)
        assert np.isclose(p.rerr, 3.0 / (np.pi * p.pur * np.sqrt(6.0 * t)), rtol=1e-6)
        assert np.isclose(p.cenerr, 1.0 / np.sqrt(24.0 * t), rtol=1e-6)
        assert np.isclose(p.sig, np.sqrt(2 * t - np.log(np.pi * t)), rtol=1e-6)


def test_calc_rzwerrs_algebra():
    rd = O.ORDerivs(pow=400.0, phs=1.0, dpow=0.0, dphs=-np.pi, d2pow=-2000.0, d2phs=0.0,
                    locpow=2.0)
    p = O.calc_props(rd, 1234.5, 12.0, 30.0)
    T = 100.0
    e = O.calc_rzwerrs(p, T)
    assert np.isclose(e.f, 12.345) and np.isclose(e.p, 1.0 / 12.345)
    assert np.isclose(e.fd, 12.0 / T ** 2) and np.isclose(e.fdd, 30.0 / T ** 3)
    assert np.isclose(e.pd, -12.0 / 1234.5 ** 2)
    assert np.isclose(e.pdd, (2 * 144.0 - 1234.5 * 30.0) / (T * 1234.5 ** 3))
    assert np.isclose(e.ferr, p.rerr / T, rtol=1e-6)
    assert np.isclose(p.cen, 0.5)


def test_rz_interp_matches_plane_definition():
    """rz_interp at a half-bin grid point equals the direct correlation with
    the FP64 response (the plane the search builds), to FP32 rounding."""
    ft = make_signal(R_RZ, 10.0, 1.0, 7)
    for r, z in [(R_RZ, 10.0), (R_RZ + 0.5, 8.0), (R_RZ - 3.0, -6.0)]:
        khw = O.lib.o_z_resp_halfwidth(z, 0)
        got = O.rz_interp(ft, r, z, khw)
        nk = 2 * khw
        frac = r - np.floor(r)
        resp = O.gen_z_response(frac, 1, z, nk)
        k0 = int(np.floor(r)) - khw
        ref = np.sum(ft[k0:k0 + nk].astype(np.complex128) * np.conj(resp))
        assert abs(got - ref) < 1e-5 * abs(ref) + 1e-3
