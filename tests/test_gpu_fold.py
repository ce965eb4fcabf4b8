"""The .cand consumer on the GPU (accel_fold, accel_fold_search, the
accelfold command): the GPU fold against the oracle's restatement of
prepfold's fold (profiles within 1e-9 of the largest bin: FP64 phases here,
long double in the oracle as in the reference; statistics within 1e-12);
the p / p-dot search's plane bit-equal to the oracle's, with the same best
trial; and the end-to-end path accelsearch -> .cand -> fold -> search on
the sifting fixtures' pulsar."""
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle as O
from presto_amd import synth

pytestmark = pytest.mark.gpu
A = pytest.importorskip("presto_amd.accel")

HERE = os.path.dirname(os.path.abspath(__file__))


def _say(msg):
    print(msg, flush=True)


def _total_redchi(prof, x_used):
    """redchi of the summed profile from the points folded (the formula of
    src/fold.c:627-640 over all parts)."""
    tot = prof.sum(0)
    n = len(x_used)
    var = float(np.var(x_used.astype(np.float64), ddof=1))
    pvar = var * n / len(tot)
    return float(np.sum((tot - tot.mean()) ** 2) / (pvar * (len(tot) - 1)))


@pytest.mark.parametrize("N,dt,f,fd,fdd,npart,nb", [
    (1 << 20, 64e-6, 123.456, 3.1e-5, 0.0, 64, 64),      # default shape
    (1 << 18, 1e-3, 7.31, -2e-4, 1e-6, 7, 37),           # ragged records, fdd
    (1 << 17, 1e-3, 0.61, 1e-5, 0.0, 16, 50),            # turns longer than a record
    (1 << 17, 1e-3, 0.2, 0.0, 0.0, 64, 20),              # turns longer than two parts
    (1 << 18, 64e-6, 33.3, 0.0, 0.0, 4, 3000),           # profile beyond the LDS path
    (3 * 1024 + 517, 1e-3, 11.0, 0.0, 0.0, 1, 16),       # one part, a partial record dropped
])
def test_gpu_fold_equals_oracle(N, dt, f, fd, fdd, npart, nb):
    _say(f"case: fold N={N} dt={dt} f={f} fd={fd} fdd={fdd} npart={npart} proflen={nb}")
    x = synth.make_series(N, dt, 17, [dict(f0=f, z=fd * (N * dt) ** 2, duty=0.1, amp=0.4)])
    prof, parts, total, n = A.fold(x, dt, f, fd, fdd, npart=npart, proflen=nb)
    want, wst, wn = O.fold_dat(x, dt, f, fd, fdd, npart=npart, proflen=nb)
    assert n == wn
    err = np.max(np.abs(prof - want)) / np.max(np.abs(want))
    _say(f"profile error vs oracle {err:.2e}, total redchi {total.redchi:.3f}")
    assert err < 1e-9
    for p, (g, w) in enumerate(zip(parts, wst)):
        assert g.numdata == w["numdata"]
        assert g.data_avg == pytest.approx(w["data_avg"], rel=1e-12, abs=1e-12)
        assert g.data_var == pytest.approx(w["data_var"], rel=1e-12)
        assert g.redchi == pytest.approx(w["redchi"], rel=1e-8), p
    assert total.redchi == pytest.approx(_total_redchi(want, x[:wn]), rel=1e-8)


def test_cand_file_to_fold_confirms_the_pulsar(tmp_path):
    """The sifting fixtures' DM-30 trial (a 123.4 Hz pulsar, z = 4): the
    accelsearch command writes its .cand, whose first record folds into a
    profile far above noise at the pulsar's frequency (a reduced
    chi-squared of 4 over 31 degrees of freedom has a chance probability
    below 1e-12) -- through the library and through the accelfold command --
    while a fold at an unrelated frequency stays near 1."""
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_sift_fixtures as F
    from presto_amd import surface as S
    dm, i = 30.0, 3
    root = tmp_path / ("sift_DM%.2f" % dm)
    x = F.trial(dm, i)
    x.tofile(str(root) + ".dat")
    d = S.readinf(os.path.join(HERE, "golden", "1937_DM71.02_zerodm"))
    d.name = root.name.encode()
    d.N, d.dt, d.dm = float(F.N), F.DT, dm
    d.numonoff, d.onoff[0], d.onoff[1] = 1, 0.0, float(F.N - 1)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        S.writeinf(d)
    finally:
        os.chdir(cwd)
    cli = os.path.join(os.path.dirname(A.__file__), "accelsearch")
    r = subprocess.run([cli, "-zmax", "50", "-numharm", "8", "-sigma", "3.0", root.name + ".dat"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    cand = str(root) + "_ACCEL_50.cand"
    c = A.read_rzw_cand(cand, 1)
    T = F.N * F.DT
    f, fd, fdd = A.cand_fold_params(c, T)
    _say(f"candidate 1: r={c.r:.3f} z={c.z:.3f} sigma={c.sig:.2f} -> f={f:.6f} Hz")
    assert abs(f - 123.4) < 2.0 / T or abs(f - 2 * 123.4) < 2.0 / T
    prof, parts, total, n = A.fold(x, F.DT, f, fd, fdd)
    want, wst, wn = O.fold_dat(x, F.DT, f, fd, fdd, npart=64, proflen=A.fold_proflen(f, F.DT))
    assert np.max(np.abs(prof - want)) < 1e-9 * np.max(np.abs(want))
    _say(f"folded reduced chi-squared {total.redchi:.2f}")
    assert total.redchi > 4.0
    _, _, off, _ = A.fold(x, F.DT, 97.13, 0.0, 0.0)
    _say(f"off-frequency reduced chi-squared {off.redchi:.2f}")
    assert off.redchi < 2.0
    # prepfold's search of the fold: the same best trial through the library
    # and the command, and a better profile than the candidate's own
    rpp = n // 1024 // 64
    pt = np.arange(64) * rpp * 1024 * F.DT
    _, best_prof, best = A.fold_search(prof, parts, f, fd, fdd, n * F.DT, pt)
    _say(f"searched {best.numtrials}^2 trials: best p {best.p:.12f} s, "
         f"reduced chi-squared {best.stats.redchi:.2f}")
    assert best.stats.redchi >= total.redchi * 0.99
    assert abs(1.0 / best.p - 123.4) < 2.0 / T or abs(1.0 / best.p - 2 * 123.4) < 2.0 / T
    fold_cli = os.path.join(os.path.dirname(A.__file__), "accelfold")
    r = subprocess.run([fold_cli, "-accelfile", cand, "-accelcand", "1", str(root) + ".dat"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = r.stdout.splitlines()
    line = [s for s in out if s.startswith("Fold reduced chi-squared")][0]
    assert float(line.split("=")[1]) == pytest.approx(total.redchi, rel=1e-4)
    line = [s for s in out if s.startswith("Reduced chi-squared")][0]
    assert float(line.split("=")[1]) == pytest.approx(best.stats.redchi, rel=1e-4)
    line = [s for s in out if s.startswith("Best p =")][0]
    assert float(line.split()[3]) == pytest.approx(best.p, rel=1e-13)
    r = subprocess.run([fold_cli, "-nosearch", "-accelfile", cand, "-accelcand", "1",
                        str(root) + ".dat"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    line = [s for s in r.stdout.splitlines() if s.startswith("Best p =")][0]
    assert float(line.split()[3]) == pytest.approx(1.0 / f, rel=1e-13)


def _parttimes(N, dt, npart):
    """prepfold's sub-integration start times: whole reads of 1024 points
    (src/prepfold_pipeline.c:1515)."""
    return np.arange(npart) * ((N // 1024) // npart) * 1024 * dt


@pytest.mark.parametrize("N,dt,f,off,npart,nb,npfact,pstep,pdstep", [
    (1 << 20, 64e-6, 123.456, 1.3, 64, 64, 2, 2, 4),     # prepfold's defaults
    (1 << 18, 1e-3, 7.31, -0.7, 7, 37, 3, 1, 2),         # ragged parts, other steps
    (1 << 18, 64e-6, 33.3, 0.4, 16, 200, 1, 2, 4),       # profile longer than the workgroup
    (1 << 16, 1e-3, 2.5, 0.0, 1, 16, 1, 2, 4),           # one part: every trial the same
])
def test_gpu_fold_search_equals_oracle(N, dt, f, off, npart, nb, npfact, pstep, pdstep):
    """A pulsar folded `off` Fourier bins away from its frequency: the
    GPU's plane of reduced chi-squared equals the oracle's bit for bit, and
    the best trial, its period, p-dot, profile and statistics agree."""
    _say(f"case: fold search N={N} f={f} off={off} npart={npart} proflen={nb} "
         f"npfact={npfact} pstep={pstep} pdstep={pdstep}")
    x = synth.make_series(N, dt, 23, [dict(f0=f, z=0.0, duty=0.08, amp=0.3)])
    ff = f + off / (N * dt)
    prof, parts, total, n = A.fold(x, dt, ff, 0.0, 0.0, npart=npart, proflen=nb)
    T = n * dt                     # the span folded, as prepfold takes it
    pt = _parttimes(N, dt, npart)
    pp, best_prof, best = A.fold_search(prof, parts, ff, 0.0, 0.0, T, pt, npfact=npfact,
                                        pstep=pstep, pdstep=pdstep)
    st = [{k: getattr(p, k) for k, _ in A.FoldStats._fields_} for p in parts]
    want, wb = O.fold_search(prof, st, ff, 0.0, 0.0, T, pt, npfact=npfact, pstep=pstep,
                             pdstep=pdstep)
    nt = 2 * npfact * nb + 1
    assert best.numtrials == nt and pp.shape == (nt, nt)
    bad = int(np.count_nonzero(pp.view(np.uint32) != want.view(np.uint32)))
    _say(f"plane {nt}x{nt}: {bad} trials differ; best (ip, ipd) = ({best.ip}, {best.ipd}) "
         f"oracle ({wb['ip']}, {wb['ipd']}); redchi {best.stats.redchi:.3f} "
         f"(plain fold {total.redchi:.3f})")
    assert bad == 0
    assert (best.ip, best.ipd) == (wb["ip"], wb["ipd"])
    assert best.stats.redchi == pytest.approx(wb["redchi"], rel=1e-13)
    assert best.p == wb["p"] and best.pd == wb["pd"] and best.pdd == wb["pdd"]
    assert best.stats.redchi >= total.redchi * (1 - 1e-12)
    assert np.sum(best_prof) == pytest.approx(prof.sum(), rel=1e-12)
    if npart == 1:
        # one part: every trial is the same sum, and the first one wins
        assert (best.ip, best.ipd) == (0, 0) and np.all(pp == pp[0, 0])
    if off != 0.0:
        # the search finds the pulse the off-frequency fold smeared; where
        # it keeps the fold's p-dot, its period is the pulsar's to within
        # one trial step (pstep / proflen turns over T)
        assert best.stats.redchi > 3.0 * total.redchi
        if best.pd == 0.0:
            assert abs(1.0 / best.p - f) <= 1.5 * pstep / (nb * T)


@pytest.mark.parametrize("npart,nb,fdd", [(6, 16, 2e-7), (5, 11, -1e-6)])
def test_gpu_fold_search_pdotdot_equals_oracle(npart, nb, fdd):
    """-searchpdd: the GPU's cube of reduced chi-squared over (p-dotdot,
    p-dot, period) equals the oracle's bit for bit, with the same best
    trial and p-dotdot."""
    N, dt, f = 1 << 18, 1e-3, 3.3
    _say(f"case: fold search -searchpdd N={N} f={f} fdd={fdd} npart={npart} proflen={nb}")
    x = synth.make_series(N, dt, 29, [dict(f0=f, z=0.0, duty=0.08, amp=0.3)])
    prof, parts, total, n = A.fold(x, dt, f, 0.0, fdd, npart=npart, proflen=nb)
    T = n * dt
    pt = _parttimes(N, dt, npart)
    pp, best_prof, best = A.fold_search(prof, parts, f, 0.0, fdd, T, pt, npfact=1,
                                        searchpdd=True)
    st = [{k: getattr(p, k) for k, _ in A.FoldStats._fields_} for p in parts]
    want, wb = O.fold_search(prof, st, f, 0.0, fdd, T, pt, npfact=1, searchpdd=True)
    nt = 2 * nb + 1
    assert pp.shape == want.shape == (nt, nt, nt)
    bad = int(np.count_nonzero(pp.view(np.uint32) != want.view(np.uint32)))
    _say(f"cube {nt}^3: {bad} trials differ; best (ip, ipd, ipdd) = "
         f"({best.ip}, {best.ipd}, {best.ipdd}) oracle ({wb['ip']}, {wb['ipd']}, {wb['ipdd']})")
    assert bad == 0
    assert (best.ip, best.ipd, best.ipdd) == (wb["ip"], wb["ipd"], wb["ipdd"])
    assert best.p == wb["p"] and best.pd == wb["pd"] and best.pdd == wb["pdd"]
    assert best.stats.redchi == pytest.approx(wb["redchi"], rel=1e-13)
