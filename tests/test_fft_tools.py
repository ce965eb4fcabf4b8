""".inf reading and writing (src/ioinf.c) against the reference's own .inf
files, and the realfft / rednoise / zapbirds commands: argument handling on
the CPU; on the GPU, their outputs against numpy's FFT and the oracle's
de-reddening and zapping of the same data."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle as O, zap as Z
from presto_amd import synth

A = pytest.importorskip("presto_amd.accel")
S = pytest.importorskip("presto_amd.surface")
HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.dirname(A.__file__)
GOLD = os.path.join(HERE, "golden")


def _tool(name, args, cwd):
    return subprocess.run([os.path.join(BIN, name)] + args, cwd=cwd, capture_output=True,
                          text=True, timeout=300)


def _s(b):
    return b.decode() if isinstance(b, bytes) else b


def test_readinf_radio_fixture(tmp_path):
    shutil.copy(os.path.join(GOLD, "1937_DM71.02_zerodm.inf"), tmp_path)
    d = S.readinf(str(tmp_path / "1937_DM71.02_zerodm"))
    assert _s(d.name) == "1937_DM71.02_zerodm" and _s(d.telescope) == "GBT"
    assert _s(d.instrument) == "GUPPI" and _s(d.object) == "1937+21"
    assert (d.ra_h, d.ra_m, d.dec_d, d.dec_m) == (19, 39, 21, 35)
    assert abs(d.ra_s - 38.592) < 1e-9 and abs(d.dec_s - 0.24) < 1e-9
    assert d.mjd_i == 55267 and abs(d.mjd_f - 0.436250000000001) < 1e-15
    assert d.bary == 0 and d.N == 471040 and d.dt == 8.192e-05
    assert d.numonoff == 1 and d.onoff[0] == 0 and d.onoff[1] == 471039
    assert _s(d.band) == "Radio" and d.fov == 905 and d.dm == 71.02
    assert d.freq == 720.830078125 and d.freqband == 200 and d.num_chan == 128
    assert d.chan_wid == 1.5625 and _s(d.analyzer) == "sransom"
    assert _s(d.notes).startswith("Project ID AGBT10A_022_01")


def test_readinf_xray_fixture_negative_declination(tmp_path):
    shutil.copy(os.path.join(GOLD, "M28A_bary.inf"), tmp_path)
    d = S.readinf(str(tmp_path / "M28A_bary"))
    assert _s(d.telescope) == "Chandra" and _s(d.band) == "X-ray"
    assert (d.dec_d, d.dec_m) == (-24, 52) and abs(d.dec_s - 10.771) < 1e-9
    assert d.N == 6700000000 and d.dt == 0.0005 and d.bary == 1
    assert d.fov == 3.0 and d.energy == 1.0 and d.energyband == 5.0


@pytest.mark.parametrize("fixture", ["1937_DM71.02_zerodm", "M28A_bary"])
def test_writeinf_round_trip(tmp_path, fixture):
    shutil.copy(os.path.join(GOLD, fixture + ".inf"), tmp_path)
    d = S.readinf(str(tmp_path / fixture))
    d.name = str(tmp_path / "copy").encode()
    S.writeinf(d)
    e = S.readinf(str(tmp_path / "copy"))
    for f, _ in S.InfoData._fields_:
        if f in ("name", "onoff", "notes"):
            continue
        assert getattr(d, f) == getattr(e, f), f
    # as in the reference, each write adds the notes' closing blank lines
    assert _s(d.notes).rstrip() == _s(e.notes).rstrip()
    assert list(d.onoff[:2]) == list(e.onoff[:2])
    # the reference's writeinf line layout: the key/value '=' at column 40
    for line in open(tmp_path / "copy.inf"):
        if "=" in line:
            assert line[40] == "=", line


def test_tools_reject_bad_arguments(tmp_path):
    r = _tool("realfft", [], tmp_path)
    assert r.returncode == 1 and "Usage" in r.stdout
    r = _tool("realfft", ["-bogus", "x.dat"], tmp_path)
    assert r.returncode == 1 and "unknown option" in r.stderr
    r = _tool("rednoise", ["x.dat"], tmp_path)
    assert r.returncode == 1 and "must be a fourier transform" in r.stdout
    r = _tool("zapbirds", ["x.fft"], tmp_path)
    assert r.returncode == 0 and "automatically zapping" in r.stdout
    r = _tool("zapbirds", ["-in", "a", "-out", "b", "x.fft"], tmp_path)
    assert r.returncode == 1 and "interactive mode is not built" in r.stderr


def _dat(tmp_path, name, N, dt, seed):
    x = synth.make_series_f32(N, dt, seed, [dict(f0=50.3, duty=0.1, amp=0.3)])
    x = x + np.float32(0.5) * np.sin(np.arange(N) * 2e-4).astype(np.float32)  # red noise
    x = x.astype(np.float32)
    x.tofile(tmp_path / (name + ".dat"))
    S.write_inf(str(tmp_path / name), N, dt)
    return x


@pytest.mark.gpu
def test_realfft_forward_inverse_and_rednoise(tmp_path):
    print("case: realfft tool, 2^18 samples, forward / -inv / -rednoise", flush=True)
    N, dt = 1 << 18, 1e-3
    x = _dat(tmp_path, "ts", N, dt, 81)
    r = _tool("realfft", ["ts.dat"], tmp_path)
    assert r.returncode == 0, r.stderr
    spec = np.fromfile(tmp_path / "ts.fft", dtype=np.complex64)
    ref = np.fft.rfft(x.astype(np.float64))
    rms = np.sqrt(np.mean(np.abs(ref[1:N // 2]) ** 2))
    assert len(spec) == N // 2
    # FP32 radix-2 passes; the red noise puts 100x the rms amplitude in the
    # lowest bins, and the FP32 rounding error follows the largest terms
    assert np.max(np.abs(spec[1:] - ref[1:N // 2])) / rms < 5e-5
    assert abs(spec[0].real - ref[0].real) < 1e-3 * np.sqrt(N)
    assert abs(spec[0].imag - ref[N // 2].real) < 1e-3 * np.sqrt(N)
    # the inverse restores the series (realfft -inv writes ts.dat again)
    os.rename(tmp_path / "ts.dat", tmp_path / "orig.dat")
    r = _tool("realfft", ["-inv", "ts.fft"], tmp_path)
    assert r.returncode == 0, r.stderr
    back = np.fromfile(tmp_path / "ts.dat", dtype=np.float32)
    xr = np.sqrt(np.mean(x.astype(np.float64) ** 2))
    assert np.max(np.abs(back - x)) / xr < 1e-5
    assert np.max(np.abs(np.fft.irfft(ref, N) - back)) / xr < 1e-5
    # -rednoise -outdir: <outdir>/ts_red.fft and a .inf naming it
    os.mkdir(tmp_path / "out")
    r = _tool("realfft", ["-rednoise", "-outdir", "out", "orig.dat"], tmp_path)
    assert r.returncode != 0 and "-rednoise" in r.stdout    # orig.inf is missing
    shutil.copy(tmp_path / "ts.inf", tmp_path / "orig.inf")
    r = _tool("realfft", ["-rednoise", "-startwidth", "8", "-endwidth", "150", "-outdir", "out",
                          "orig.dat"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    red = np.fromfile(tmp_path / "out" / "orig_red.fft", dtype=np.complex64)
    oracle_red, _ = O.deredden(spec, 8, 150, 6.0, N * dt)
    assert np.array_equal(red.view(np.uint32), oracle_red.view(np.uint32))
    d = S.readinf(str(tmp_path / "out" / "orig_red"))
    assert _s(d.name) == str(tmp_path / "out" / "orig_red") or \
        _s(d.name) == "out/orig_red"
    assert d.N == N and d.dt == dt


@pytest.mark.gpu
def test_rednoise_tool_matches_oracle(tmp_path):
    print("case: rednoise tool, 2^17 samples, copy and -inplace", flush=True)
    N, dt = 1 << 17, 5e-4
    x = _dat(tmp_path, "rn", N, dt, 82)
    spec = A.realfft_gpu(x)
    spec.tofile(tmp_path / "rn.fft")
    r = _tool("rednoise", ["rn.fft"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(tmp_path / "rn_red.fft", dtype=np.complex64)
    ref, n = O.deredden(spec, 6, 100, 6.0, N * dt)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert f"Rednoise removed from {n} of {N // 2} points" in r.stdout
    assert _s(S.readinf(str(tmp_path / "rn_red")).name) == "rn_red"
    os.remove(tmp_path / "rn_red.fft")
    r = _tool("rednoise", ["-inplace", "rn.fft"], tmp_path)
    assert r.returncode == 0 and not os.path.exists(tmp_path / "rn.fft")
    assert np.array_equal(np.fromfile(tmp_path / "rn_red.fft", dtype=np.complex64)
                          .view(np.uint32), ref.view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("constamp", [False, True])
def test_zapbirds_tool_matches_oracle(tmp_path, constamp):
    print(f"case: zapbirds tool, constamp {constamp}", flush=True)
    N, dt = 1 << 17, 5e-4
    x = _dat(tmp_path, "zb", N, dt, 83)
    spec = A.realfft_gpu(x)
    spec.tofile(tmp_path / "zb.fft")
    T = N * dt
    with open(tmp_path / "birds.zap", "w") as f:
        f.write("# freq width\n60.0 0.5\n120.0 1.0\nB 180.25 0.02\n 300.0 2.0\n")
    args = ["-zap", "-zapfile", "birds.zap", "-baryv", "1e-4"] + \
        (["-constamp"] if constamp else []) + ["zb.fft"]
    r = _tool("zapbirds", args, tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(tmp_path / "zb.fft", dtype=np.complex64)
    lo, hi = Z.read_birdies(str(tmp_path / "birds.zap"), T, 1e-4)
    ref = Z.zapbirds_tool(spec, N, lo, hi, constamp=constamp)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert not np.array_equal(got, spec)
