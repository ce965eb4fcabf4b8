"""Birdie zapping (-zaplist): the library against the numpy restatement in
oracle/zap.py, and the statistics the reference promises (src/zapping.c:
68-79): zapped bins carry chi^2(2) noise at the surrounding power level,
nothing else changes, and the same zaplist zaps alike."""
import numpy as np
import pytest

from oracle import zap as Z

A = pytest.importorskip("presto_amd.accel")


def test_mt19937_matches_numpy_legacy_stream():
    rs = np.random.RandomState(1)
    ref = rs.randint(0, 2 ** 32, size=2000, dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(A.mt19937_raw(1, 2000), ref)
    assert A.mt19937_raw(1, 1)[0] == 1791095845   # the generator's published first output


def _zapfile(tmp_path, T):
    p = tmp_path / "birds.zaplist"
    p.write_text("# freq width\n"
                 "50.0 0.2\n"
                 "B 120.0 0.0\n"      # barycentric, narrower than a bin
                 "# comment\n"
                 "33.3 1.0\n")
    return str(p)


def test_read_birdies_matches_definition(tmp_path):
    T, v = 100.0, 1e-4
    path = _zapfile(tmp_path, T)
    lo, hi = A.read_birdies(path, T, v)
    olo, ohi = Z.read_birdies(path, T, v)
    assert np.array_equal(lo, olo) and np.array_equal(hi, ohi)
    # hand values: sorted by lobin, Doppler factor on non-'B' lines
    assert np.allclose(lo, [(33.3 - 0.5) * T * (1 + v), (50.0 - 0.1) * T * (1 + v),
                            120.0 * T - 0.5])
    assert np.allclose(hi - lo, [1.0 * T * (1 + v), 0.2 * T * (1 + v), 1.0])


def test_zap_equals_restatement_and_keeps_statistics(tmp_path):
    rng = np.random.default_rng(5)
    n = 1 << 16
    fft = ((rng.standard_normal(n) + 1j * rng.standard_normal(n)) * np.sqrt(0.5)).astype(
        np.complex64)
    fft[5000:5040] *= 30.0          # a strong birdie
    N, T = 2.0 * n, 100.0
    lo, hi = A.read_birdies(_zapfile(tmp_path, T), T, 0.0)
    lo = np.append(lo, [4999.6, 200.0e3])   # a wide one, and one past Nyquist
    hi = np.append(hi, [5040.2, 200.5e3])
    order = np.argsort(lo)
    lo, hi = lo[order], hi[order]
    got = A.zap_birdies(fft, N, lo, hi)
    ref = Z.zap_birdies(fft, N, lo, hi)
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
    changed = np.nonzero(got != fft)[0]
    zapped = np.zeros(n, bool)
    for a, b in zip(lo, hi):
        if a < N / 2:
            zapped[int(np.floor(a)):int(np.ceil(min(b, N / 2 - 1)))] = True
    assert np.all(zapped[changed])
    pw = np.abs(got[4999:5041]) ** 2
    assert 0.5 < pw.mean() < 2.0      # the birdie (power ~900) is gone
    assert np.array_equal(A.zap_birdies(fft, N, lo, hi), got)
