"""The .cand consumer on the GPU (accel_fold, the accelfold command): the
GPU fold against the oracle's restatement of prepfold's fold (profiles
within 1e-9 of the largest bin: FP64 phases here, long double in the
oracle as in the reference; statistics within 1e-12), and the end-to-end
path accelsearch -> .cand -> fold on the sifting fixtures' pulsar."""
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
    fold_cli = os.path.join(os.path.dirname(A.__file__), "accelfold")
    r = subprocess.run([fold_cli, "-accelfile", cand, "-accelcand", "1", str(root) + ".dat"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    line = [s for s in r.stdout.splitlines() if s.startswith("Reduced chi-squared")][0]
    assert float(line.split("=")[1]) == pytest.approx(total.redchi, rel=1e-4)
