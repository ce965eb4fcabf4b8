"""The .cand consumer's CPU side: the oracle's restatement of prepfold's
fold (oracle/o_fold.c) pinned by closed forms and by an independent
point-by-point restatement of src/fold.c's buffer and dump rules, and the
library's record reader and start-parameter conversion."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle as O

A = pytest.importorskip("presto_amd.accel")


def test_cand_fold_params_invert_the_average():
    """The .cand holds average r and z; prepfold starts the fold from r0,
    z0 with r = r0 + z0/2 + w/6 and z = z0 + w/2
    (src/prepfold_pipeline.c:759-764)."""
    T = 536.870912
    r0, z0, w = 81234.375, -17.25, 42.0
    r, z = r0 + z0 / 2 + w / 6, z0 + w / 2
    f, fd, fdd = O.cand_fold_params(r, z, w, T)
    assert f == pytest.approx(r0 / T, rel=1e-15)
    assert fd == pytest.approx(z0 / T ** 2, rel=1e-13)
    assert fdd == pytest.approx(w / T ** 3, rel=1e-15)
    props = A.FourierProps()
    props.r, props.z, props.w = r, z, w
    assert A.cand_fold_params(props, T) == (f, fd, fdd)


def test_read_rzw_cand_record(tmp_path):
    """get_rzw_cand reads record candnum - 1 of 88-byte fourierprops."""
    assert C.sizeof(A.FourierProps) == 88
    recs = (A.FourierProps * 3)()
    for k in range(3):
        recs[k].r, recs[k].z, recs[k].w, recs[k].sig = 100.5 + k, 2.0 * k, -20.0 * k, 7.0 + k
    path = tmp_path / "x_ACCEL_50.cand"
    path.write_bytes(bytes(recs))
    got = A.read_rzw_cand(path, 2)
    assert bytes(got) == bytes(recs[1])
    with pytest.raises(A.AccelError):
        A.read_rzw_cand(path, 4)
    with pytest.raises(A.AccelError):
        A.read_rzw_cand(path, 0)


def test_default_proflen():
    """One bin per sample of the period, at most 64
    (src/prepfold_pipeline.c:349-357)."""
    assert A.fold_proflen(1.0 / 0.005, 1e-4) == 50
    assert A.fold_proflen(123.4, 256e-6) == int(1 / 123.4 / 256e-6 + 0.5) == 32
    assert A.fold_proflen(123.4, 64e-6) == 64
    assert A.fold_proflen(1.0, 1e-3) == 64


def test_constant_series_fills_every_bin_equally():
    """A constant c drizzled at constant f: each completed turn puts
    c / (f dt proflen) in every bin, so every bin of the summed profile
    holds K c / (f dt proflen) for K completed turns."""
    dt, f, c, nb = 1e-3, 9.7, 2.5, 32
    x = np.full(16 * 1024, c, dtype=np.float32)
    prof, st, n = O.fold_dat(x, dt, f, 0.0, 0.0, npart=4, proflen=nb)
    K = math.floor(f * n * dt + 1e-9)
    tot = prof.sum(0)
    assert np.allclose(tot, K * c / (f * dt * nb), rtol=1e-9)
    assert n == 16 * 1024
    assert all(s["numdata"] == 4096 for s in st)
    assert st[0]["data_avg"] == pytest.approx(c)


def _fold_points(x, dt, f, fd, fdd, npart, nb):
    """Independent restatement, point by point in Python floats: records of
    1024 points, the phase added since the start tracked as a whole, each
    piece of a point (split at bin edges and at whole turns of the phase
    added) credited to the part whose end the turn's end does not exceed."""
    rpp = len(x) // 1024 // npart
    used = rpp * npart * 1024
    prof = np.zeros((npart, nb))
    ph = lambda t: t * (t * (t * fdd / 6 + fd / 2) + f)  # noqa: E731
    proftime = 1024 * dt
    # phase added at the end of each part
    pend = []
    for p in range(npart):
        tlo = (p * rpp) * proftime + (rpp - 1) * proftime
        pend.append(ph(tlo + 1024 * dt))
    for i in range(used):
        b = i // 1024
        p, jj, l = b // rpp, b % rpp, i % 1024
        tlo = (p * rpp) * proftime + jj * proftime
        a, e = ph(tlo + l * dt), ph(tlo + (l + 1) * dt)
        d = e - a
        if d <= 1e-12:
            continue
        vpp = float(x[i]) / d
        lo = a
        while lo < e - 1e-15:
            bin_lo = math.floor((lo - math.floor(lo)) * nb + 1e-12)
            edge = math.floor(lo) + (bin_lo + 1) / nb
            turn_end = math.floor(lo) + 1.0
            hi = min(e, edge, turn_end)
            tp = p
            while tp < npart and pend[tp] < turn_end:
                tp += 1
            if tp < npart:
                prof[tp, bin_lo % nb] += (hi - lo) * vpp
            lo = hi
    return prof


@pytest.mark.parametrize("f,fd,fdd,npart,nb", [(37.3, 0.0, 0.0, 4, 24),
                                               (5.21, 2e-3, 0.0, 3, 16),
                                               (0.61, 1e-4, 3e-5, 2, 20)])
def test_fold_matches_point_by_point_restatement(f, fd, fdd, npart, nb):
    rng = np.random.default_rng(11)
    dt = 1e-3
    x = rng.standard_normal(12 * 1024).astype(np.float32)
    want = _fold_points(x, dt, f, fd, fdd, npart, nb)
    got, st, n = O.fold_dat(x, dt, f, fd, fdd, npart=npart, proflen=nb)
    scale = np.abs(want).max()
    assert np.max(np.abs(got - want)) < 1e-9 * scale


def test_pulse_lands_at_its_phase():
    """A narrow pulse train at phase 0.25 of a known f, fd folds into the
    bin holding phase 0.25, and the profile stands far above noise."""
    from presto_amd import synth
    N, dt, f0, fd = 1 << 16, 1e-3, 3.7, 1e-4
    t = np.arange(N) * dt
    phase = f0 * t + 0.5 * fd * t * t
    d = (phase - 0.25) - np.floor(phase - 0.25 + 0.5)
    x = (np.exp(-0.5 * (d / 0.01) ** 2) + 0.05 * synth.make_series(N, dt, 5)).astype(np.float32)
    prof, st, n = O.fold_dat(x, dt, f0, fd, 0.0, npart=8, proflen=40)
    tot = prof.sum(0)
    assert np.argmax(tot) == int(0.25 * 40)
    assert st[3]["redchi"] > 100


# ---- prepfold's p / p-dot search (o_fold_search) ----

def _parts(profs, prof_var=1.0):
    """Per-part statistics as prepfold keeps them: prof_avg the part's mean
    bin, prof_var a given variance per part."""
    return [dict(numdata=1000.0, data_avg=0.0, data_var=1.0, numprof=float(profs.shape[1]),
                 prof_avg=float(p.mean()), prof_var=prof_var, redchi=0.0) for p in profs]


def _redchi(prof, avg, var):
    return float(np.sum((prof - avg) ** 2) / var / (len(prof) - 1))


def test_fold_search_centre_trial_is_the_plain_sum():
    """The centre trial moves no part: its value is the reduced chi-squared
    of the summed parts against the summed averages and variances."""
    rng = np.random.default_rng(3)
    npart, L, npfact = 6, 24, 1
    profs = rng.standard_normal((npart, L)) + 5.0
    st = _parts(profs, prof_var=1.7)
    T = 100.0
    pp, best = O.fold_search(profs, st, 9.5, 1e-4, 0.0, T, np.arange(npart) * T / npart,
                             npfact=npfact)
    nt = 2 * npfact * L + 1
    assert pp.shape == (nt, nt)
    c = (nt - 1) // 2
    want = _redchi(profs.sum(0), sum(s["prof_avg"] for s in st), 1.7 * npart)
    assert pp[c, c] == np.float32(want)


def _spike_parts(npart, L, pos, h=50.0, seed=0):
    rng = np.random.default_rng(seed)
    profs = 0.01 * rng.standard_normal((npart, L))
    for ii in range(npart):
        profs[ii, pos[ii] % L] += h
    return profs


def test_fold_search_realigns_a_period_drift():
    """Folding at f a pulsar of frequency f + df moves its pulse by
    -df t L bins at time t; the trial period 1/(f + df) realigns the parts,
    so it holds the plane's maximum, every spike in one bin.  (8 parts of
    32 bins: no other trial of the plane gives the same whole-bin shifts.)"""
    npart, L, pstep, f, T = 8, 32, 2, 5.0, 200.0
    s = 1                                          # +1 bin per part
    df = -s * npart / (L * T)
    pos = [3 + s * ii for ii in range(npart)]
    profs = _spike_parts(npart, L, pos)
    st = _parts(profs)
    pp, best = O.fold_search(profs, st, f, 0.0, 0.0, T, np.arange(npart) * T / npart,
                             npfact=2, pstep=pstep)
    nt = pp.shape[0]
    c = (nt - 1) // 2
    ip = c - s * npart // pstep                     # the trial at f + df
    assert pp[c, ip] == pp.max()
    assert (best["ipd"], best["ip"]) == np.unravel_index(np.argmax(pp), pp.shape)
    aligned = np.zeros(L)
    for ii in range(npart):
        aligned += np.roll(profs[ii], -(pos[ii] - pos[0]))
    want = _redchi(aligned, sum(x["prof_avg"] for x in st), npart * 1.0)
    assert pp[c, ip] == pytest.approx(want, rel=1e-6)
    assert best["ip"] == ip and best["ipd"] == c
    assert best["p"] == pytest.approx(1.0 / (f + df), rel=1e-14)
    assert best["pd"] == 0.0


def test_fold_search_realigns_a_pdot_drift():
    """An f-dot error dfd moves the pulse by -dfd t^2 / 2 L bins: spikes at
    b0 - ii^2 over parts at t = ii T / npart are realigned by the p-dot
    trial with dfd = 2 npart^2 / (L T^2), whose p-dot is -(fd + dfd) / f^2."""
    npart, L, pdstep, f, fd, T = 8, 32, 4, 5.0, 1e-6, 200.0
    pos = [10 - ii * ii for ii in range(npart)]
    profs = _spike_parts(npart, L, pos, seed=1)
    st = _parts(profs)
    pp, best = O.fold_search(profs, st, f, fd, 0.0, T, np.arange(npart) * T / npart,
                             npfact=2, pdstep=pdstep)
    nt = pp.shape[0]
    c = (nt - 1) // 2
    dfd = 2.0 * npart ** 2 / (L * T ** 2)
    ipd = c + npart ** 2 // pdstep
    assert pp[ipd, c] == pp.max()
    assert best["ip"] == c and best["ipd"] == ipd
    assert best["pd"] == pytest.approx(-(fd + dfd) / f ** 2, rel=1e-12)
    assert best["p"] == pytest.approx(1.0 / f, rel=1e-15)


def test_fold_search_common_rotation_changes_nothing():
    """Rotating every part by the same number of bins rotates every trial's
    summed profile, which leaves each reduced chi-squared unchanged."""
    rng = np.random.default_rng(8)
    npart, L = 5, 20
    profs = rng.standard_normal((npart, L)) ** 2
    st = _parts(profs)
    args = (7.0, 2e-5, 0.0, 300.0, np.arange(npart) * 60.0)
    pp, best = O.fold_search(profs, st, *args, npfact=1)
    pp2, best2 = O.fold_search(np.roll(profs, 7, axis=1), st, *args, npfact=1)
    assert np.allclose(pp, pp2, rtol=1e-5, atol=0)
    assert (best["ip"], best["ipd"]) == (best2["ip"], best2["ipd"])


def test_fold_search_nosearch_keeps_the_fold():
    """-nosearch: the fold's own parameters, the plain sum of the parts
    (host arithmetic only), and a 201-trial plane once proflen > 100
    (src/prepfold_pipeline.c:1648-1654)."""
    rng = np.random.default_rng(4)
    for L, nt in ((40, 161), (128, 201)):
        profs = rng.standard_normal((6, L)) + 3.0
        st = [A.FoldStats(1000.0, 3.0, 1.0, float(L), float(p.mean()), 2.0, 0.0) for p in profs]
        pp, prof, best = A.fold_search(profs, st, 11.0, 2e-6, 0.0, 90.0, np.arange(6) * 15.0,
                                       nosearch=True)
        assert pp is None and best.numtrials == nt
        assert best.ip == best.ipd == (nt - 1) // 2
        assert np.array_equal(prof, profs.sum(0))
        assert best.p == 1.0 / 11.0 and best.pd == pytest.approx(-2e-6 / 121.0, rel=1e-15)
        want = _redchi(profs.sum(0), sum(float(p.mean()) for p in profs), 12.0)
        assert best.stats.redchi == pytest.approx(want, rel=1e-13)
        assert best.stats.prof_var == 12.0 and best.stats.data_avg == 3.0


def test_fold_search_realigns_a_pdotdot_drift():
    """-searchpdd: an f-dotdot error moves the pulse by -dfdd t^3 / 6 L
    bins; spikes at b0 - ii^3 over parts at t = ii T / npart are realigned
    by the p-dotdot trial of delay ii^3 (tot pdstep = npart^3), whose cube
    entry is the maximum; the opposite trial is not."""
    npart, L, pdstep, f, T = 4, 32, 4, 5.0, 200.0
    pos = [20 - ii ** 3 for ii in range(npart)]
    profs = _spike_parts(npart, L, pos, seed=2)
    st = _parts(profs)
    pp, best = O.fold_search(profs, st, f, 0.0, 0.0, T, np.arange(npart) * T / npart,
                             npfact=1, pdstep=pdstep, searchpdd=True)
    nt = pp.shape[0]
    c = (nt - 1) // 2
    assert pp.shape == (nt, nt, nt)
    k = npart ** 3 // pdstep
    assert pp[c + k, c, c] == pp.max() > pp[c - k, c, c]
    first = np.unravel_index(np.argmax(pp), pp.shape)
    assert (best["ipdd"], best["ipd"], best["ip"]) == first
    # without the p-dotdot loop the search is the p / p-dot plane of the
    # centre p-dotdot trial
    plane, b2 = O.fold_search(profs, st, f, 0.0, 0.0, T, np.arange(npart) * T / npart,
                              npfact=1, pdstep=pdstep)
    assert np.array_equal(plane, pp[c]) and b2["ipdd"] == 0
