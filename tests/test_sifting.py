"""The ACCEL sifter (the C sifter accel_sift_* behind presto_amd/sift.py and
the accel_sift command; the rules of the reference's
python/presto/sifting.py and examplescripts/ACCEL_sift.py) on the CPU:

* parsing of ACCEL files written by this package's accelsearch (fixtures
  under tests/golden/sift/, written by tests/golden/make_sift_fixtures.py),
  checked against an independent reading of the same text and against the
  closed form of candidate_sigma for one trial (norm.isf of the chi^2
  tail with 2 numharm degrees of freedom at twice the summed power);
* every rejection rule and each of the duplicate, DM-problem and harmonic
  passes on hand-built candidate lists whose outcome the rules fix;
* the examplescripts/ACCEL_sift.py flow (the accel_sift command) over the
  six-DM fixture set: the pulsar survives with its three DM hits, the
  interference line that peaks at DM 0 goes to the DM-problem list.

On the GPU, the fixture set is written again by the accelsearch command
and must equal the committed files byte for byte."""
import glob
import os
import re
import shutil
import subprocess

import numpy as np
import pytest
from scipy import stats

S = pytest.importorskip("presto_amd.sift")
HERE = os.path.dirname(os.path.abspath(__file__))
SIFT = os.path.join(HERE, "golden", "sift")
T_FIX = 262144 * 0.000256


@pytest.fixture(autouse=True)
def _defaults():
    """Each test starts from the module's defaults."""
    keep = {k: getattr(S, k) for k in ("r_err", "long_period", "short_period",
                                        "sigma_threshold", "c_pow_threshold",
                                        "harm_pow_cutoff", "known_birds_p", "known_birds_f")}
    yield
    for k, v in keep.items():
        setattr(S, k, v)


def _sigma1(power, numharm):
    """candidate_sigma(power, numharm, 1) in closed form."""
    return stats.norm.isf(stats.chi2.sf(2.0 * power, 2 * numharm))


def _cand(r, sigma=10.0, numharm=1, dm="10.00", pows=None, cpow=200.0, T=100.0, num=1,
          fname="x_DM10.00_ACCEL_0"):
    c = S.Candidate(num, sigma, numharm, 0.0, cpow, r, 0.0, dm, fname, T)
    c.harm_pows = np.asarray(pows if pows is not None else [20.0] * numharm, dtype=float)
    c.hits = [(c.DM, 5.0, sigma)]
    return c


def test_parse_fixture_matches_text_and_closed_form():
    fn = os.path.join(SIFT, "sift_DM30.00_ACCEL_50")
    cl = S.candlist_from_candfile(fn)
    text = open(fn).read().splitlines()
    funds = [l for l in text if re.match(r"^\d", l)]
    assert len(cl) == len(funds) == 1
    c = cl[0]
    f = funds[0].split()
    assert (c.candnum, c.numharm) == (int(f[0]), int(f[4]))
    assert c.r == float(f[7].split("(")[0]) and c.z == float(f[9].split("(")[0])
    assert c.cpow == float(f[3]) and c.DM == 30.0 and c.T == T_FIX
    # the harmonic table: 8 rows after the header, power/locpow in the
    # fourth (first row) or third column, phase in the tenth / ninth
    i0 = next(i for i, l in enumerate(text) if re.match(r"^[ ]\d", l))
    rows = text[i0:i0 + c.numharm]
    pows = [float(rows[0].split()[3].split("(")[0])] + \
        [float(r.split()[2].split("(")[0]) for r in rows[1:]]
    assert np.allclose(c.harm_pows, pows, rtol=0, atol=0)
    phs = [float(rows[0].split()[9].split("(")[0])] + \
        [float(r.split()[8].split("(")[0]) for r in rows[1:]]
    assert np.allclose(np.angle(c.harm_amps), np.angle(np.exp(1j * np.array(phs))), atol=1e-6)
    assert c.ipow_det == pytest.approx(sum(pows), abs=1e-9)
    assert c.sigma == pytest.approx(_sigma1(sum(pows), 8), rel=1e-6)
    assert c.snr == pytest.approx(sum(np.sqrt(max(p - 1.0, 0.0)) for p in pows), rel=1e-12)
    assert c.hits == [(30.0, c.snr, c.sigma)]
    # one harmonic: sigma from the single power
    c1 = S.candlist_from_candfile(os.path.join(SIFT, "sift_DM0.00_ACCEL_0"))[0]
    assert c1.numharm == 1
    # near 37 sigma candidate_sigma takes the reference's asymptotic
    # branches (src/characteristics.c:489-496, 464-470), within 1e-5 of the
    # exact tail there
    assert c1.sigma == pytest.approx(_sigma1(c1.harm_pows[0], 1), rel=2e-5)


def test_parse_power_exponent_notation():
    assert S.parse_power("12.5(3.1)") == 12.5
    assert S.parse_power("4.55(23)x10^5") == pytest.approx(455000.0)
    # an exponent cut off by the column width counts as 5
    assert S.parse_power("4.55(23)x10^") == pytest.approx(455000.0)


def test_rejection_rules():
    S.sigma_threshold, S.c_pow_threshold, S.harm_pow_cutoff = 6.0, 100.0, 8.0
    T = 100.0
    cl = S.Candlist([
        _cand(50.0, sigma=5.0, cpow=50.0, num=1),             # 1 harmonic, both low
        _cand(60.0, sigma=5.0, cpow=150.0, num=2),            # 1 harmonic, cpow saves it
        _cand(70.0, sigma=5.5, numharm=2, num=3),             # 2 harmonics, sigma low
        _cand(80.0, sigma=9.0, numharm=2, pows=[7.0, 7.9], num=4),   # no harmonic >= 8
        _cand(90.0, sigma=9.0, numharm=8, pows=[9, 9, 9, 9, 9, 30, 9, 9], num=5),  # rogue 8
        _cand(95.0, sigma=9.0, numharm=4, pows=[9, 9, 9, 40], num=6),  # rogue 4
        _cand(97.0, sigma=9.0, numharm=4, pows=[9, 40, 9, 9], num=7),  # strong 2nd: fine
        _cand(T / 20.0, num=8),                                 # P = 20 s
        _cand(T * 3000.0, num=9),                               # P = 0.33 ms
        _cand(123.0, num=10)], trackbad=True)
    cl.default_rejection()
    assert sorted(c.candnum for c in cl) == [2, 7, 10]
    bad = {k: sorted(c.candnum for c in v) for k, v in cl.badlists.items()}
    assert bad["threshold"] == [1, 3] and bad["harmpowcutoff"] == [4]
    assert bad["rogueharmpow"] == [5, 6]
    assert bad["longperiod"] == [8] and bad["shortperiod"] == [9]
    # the default call reads no birdie globals (its defaults are empty
    # lists); explicit lists reject by frequency and by period
    S.known_birds_f = [(1.23, 0.01)]
    cl2 = S.Candlist([_cand(123.0, num=1), _cand(250.0, num=2)], trackbad=True)
    cl2.reject_knownbirds()
    assert len(cl2) == 2
    cl2.reject_knownbirds(known_birds_f=[(1.23, 0.01)])
    assert [c.candnum for c in cl2] == [2]
    cl2.reject_knownbirds(known_birds_p=[(400.0, 1.0)])
    assert len(cl2) == 0 and len(cl2.badlists["knownbirds"]) == 2


def test_remove_duplicates_keeps_strongest_and_rescans():
    S.r_err = 1.1
    # 100 / 101 / 102: the first pair matches (|dr| < 1.1), the strongest
    # (101) stays where it was and is compared again, now with 102
    cl = S.Candlist([_cand(102.0, sigma=4.0, dm="30.00", num=3),
                     _cand(100.0, sigma=5.0, dm="10.00", num=1),
                     _cand(101.0, sigma=9.0, dm="20.00", num=2),
                     _cand(104.0, sigma=3.0, dm="20.00", num=4)], trackdupes=True)
    S.remove_duplicate_candidates(cl, verbosity=0)
    assert [c.candnum for c in cl] == [2, 4]
    assert sorted(h[0] for h in cl[0].hits) == [10.0, 20.0, 30.0]
    assert sorted(c.candnum for c in cl.duplicates) == [1, 3]
    assert "duplicate of x_DM10.00_ACCEL_0:2" in cl.duplicates[0].note


def test_remove_DM_problems():
    dms = ["%.2f" % d for d in (0, 10, 20, 30, 40)]
    few = _cand(100.0, num=1)
    low = _cand(200.0, num=2)
    low.hits = [(0.0, 9.0, 12.0), (10.0, 5.0, 6.0)]
    gap = _cand(300.0, num=3)
    gap.hits = [(10.0, 9.0, 12.0), (30.0, 5.0, 6.0)]
    good = _cand(400.0, num=4)
    good.hits = [(20.0, 9.0, 12.0), (30.0, 5.0, 6.0), (40.0, 3.0, 4.0)]
    same = _cand(500.0, num=5)
    same.hits = [(30.0, 9.0, 12.0), (30.0, 5.0, 6.0)]   # one DM twice: index step 0
    cl = S.Candlist([few, low, gap, good, same], trackbad=True)
    S.remove_DM_problems(cl, 2, dms, 2.0, verbosity=0)
    assert sorted(c.candnum for c in cl) == [4, 5]
    notes = {c.candnum: c.note for c in cl.badlists["dmproblem"]}
    assert "only 1 DM hits" in notes[1]
    assert "<= low DM cutoff" in notes[2] and "gaps" in notes[3]


def test_remove_harmonics():
    S.r_err = 1.1
    T = 100.0
    f0 = 10.0
    fund = _cand(f0 * T, sigma=20.0, num=1, T=T)
    second = _cand(2 * f0 * T + 0.5, sigma=8.0, num=2, T=T)     # 2nd harmonic (within r_err)
    sub = _cand(f0 * T / 3.0, sigma=7.0, num=3, T=T)            # a third: 1/3th
    frac = _cand(f0 * T * 2.0 / 3.0, sigma=6.0, num=4, T=T)     # 2/3 ratio
    other = _cand(f0 * T * 1.37, sigma=5.0, num=5, T=T)          # unrelated
    cl = S.Candlist([other, frac, second, fund, sub], trackbad=True)
    S.remove_harmonics(cl, verbosity=0)
    assert sorted(c.candnum for c in cl) == [1, 5]
    notes = {c.candnum: c.note for c in cl.badlists["harmonic"]}
    assert "(2nd)" in notes[2] and "(1/3th)" in notes[3] and "(2/3th)" in notes[4]


def test_accel_sift_flow_on_fixture(tmp_path, capsys):
    for fn in glob.glob(os.path.join(SIFT, "*")):
        shutil.copy(fn, tmp_path)
    cli = os.path.join(os.path.dirname(S.__file__), "accel_sift")
    r = subprocess.run([cli, "-o", "candlist.txt"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = open(tmp_path / "candlist.txt").read().splitlines()
    S.sigma_threshold = 4.0
    cands = S.read_candidates(sorted(glob.glob(str(tmp_path / "*ACCEL_*0"))), track=True)
    assert out[0] == ("#" + "file:candnum".center(66) + "DM".center(9) + "SNR".center(8) +
                      "sigma".center(8) + "numharm".center(9) + "ipow".center(9) +
                      "cpow".center(9) + "P(ms)".center(14) + "r".center(12) +
                      "z".center(8) + "numhits".center(9))
    first = out[1]
    assert first.startswith("sift_DM30.00_ACCEL_50:1") and first.endswith("(3)")
    assert [l.split()[0] for l in out[2:5]] == ["DM=", "DM=", "DM="]
    assert [float(l.split()[1]) for l in out[2:5]] == [20.0, 30.0, 40.0]
    assert re.match(r"^  DM= 30\.00 SNR=\d\d\.\d\d Sigma=\d\d\.\d\d   \*+$", out[3])
    # the interference line peaks at DM 0: a DM problem after the merge
    S.sigma_threshold = 4.0
    cands.remove_duplicate_candidates(verbosity=0)
    dmstrs = ["%.2f" % d for d in (0, 10, 20, 30, 40, 50)]
    cands.remove_DM_problems(2, dmstrs, 2.0, verbosity=0)
    bad = cands.badlists["dmproblem"]
    assert len(bad) == 1 and abs(bad[0].f - 47.296) < 0.01
    assert sorted(h[0] for h in bad[0].hits) == [0.0, 0.0, 10.0, 10.0]
    assert "<= low DM cutoff" in bad[0].note
    cands.print_cand_summary(str(tmp_path / "summary.txt"))
    summary = open(tmp_path / "summary.txt").read()
    assert "      # Candidates with DM problems: 1\n" in summary
    assert "   Candlist contains 2 'good' candidates\n" in summary


@pytest.mark.gpu
def test_fixture_set_is_reproduced(tmp_path):
    """The accelsearch command writes the committed fixture set again."""
    import importlib.util
    print("case: sifting fixture set written again on the GPU", flush=True)
    spec = importlib.util.spec_from_file_location(
        "make_sift_fixtures", os.path.join(HERE, "golden", "make_sift_fixtures.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    mod.main(str(tmp_path))
    for fn in sorted(glob.glob(os.path.join(SIFT, "*"))):
        name = os.path.basename(fn)
        assert open(tmp_path / name, "rb").read() == open(fn, "rb").read(), name
    made = {os.path.basename(p) for p in glob.glob(str(tmp_path / "*ACCEL_*0"))}
    assert made == {os.path.basename(p) for p in glob.glob(os.path.join(SIFT, "*ACCEL_*0"))}
