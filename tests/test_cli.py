"""The accelsearch command: argument handling on the CPU, and (GPU) a full
run on a synthetic .dat/.inf pair whose ACCEL output must list the same
candidates as the oracle's search of the same spectrum."""
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from presto_amd import synth

A = pytest.importorskip("presto_amd.accel")
CLI = os.path.join(os.path.dirname(A.__file__), "accelsearch")


def _inf(path, N, dt, telescope="None (Artificial Data Set)"):
    # the layout of src/ioinf.c:268-349 for artificial data
    with open(path, "w") as f:
        f.write(f" Data file name without suffix          =  {os.path.basename(path)[:-4]}\n"
                f" Telescope used                         =  {telescope}\n"
                f" Number of bins in the time series      =  {N:<11d}\n"
                f" Width of each time series bin (sec)    =  {dt:.15g}\n"
                f" Any breaks in the data? (1 yes, 0 no)  =  0\n"
                f" Type of observation (EM band)          =  Radio\n"
                f" Any additional notes:\n    synthetic\n\n")


def _run(args, cwd):
    return subprocess.run([CLI] + args, cwd=cwd, capture_output=True, text=True, timeout=300)


def test_cli_rejects_bad_arguments(tmp_path):
    assert os.path.exists(CLI)
    r = _run(["-bogus", "x.dat"], tmp_path)
    assert r.returncode == 1 and "unknown option" in r.stderr
    # -locpow runs the reference's block loop over the accel.h surface,
    # whose create_accelobs reads the .inf first
    r = _run(["-locpow", "x.dat"], tmp_path)
    assert r.returncode != 0 and "Error opening file 'x.inf'" in r.stdout
    # the ranges of clig/accelsearch_cmd.cli:13-44, with the generated
    # parser's messages (src/accelsearch_cmd.c:246-335)
    for args, msg in [(["-zmax", "1300"], "`-zmax' greater than max=1200"),
                      (["-numharm", "64"], "`-numharm' greater than max=32"),
                      (["-numharm", "0"], "`-numharm' smaller than min=1"),
                      (["-wmax", "4020"], "`-wmax' greater than max=4000"),
                      (["-sigma", "0.5"], "`-sigma' smaller than min=1.000000"),
                      (["-sigma", "31"], "`-sigma' greater than max=30.000000"),
                      (["-baryv", "0.2"], "`-baryv' greater than max=0.100000"),
                      (["-lobin", "-3"], "`-lobin' smaller than min=0"),
                      (["-flo", "-1"], "`-flo' smaller than min=0.000000"),
                      (["-ncpus", "0"], "`-ncpus' smaller than min=1")]:
        r = _run(args + ["x.dat"], tmp_path)
        assert r.returncode == 1 and msg in r.stderr, (args, r.stderr)
    r = _run(["x.txt"], tmp_path)
    assert r.returncode == 1 and "must be an '.fft' or '.[s]dat' file" in r.stderr
    r = _run(["missing.dat"], tmp_path)
    assert r.returncode == 1 and "cannot read N and dt" in r.stderr


def test_accelfold_rejects_bad_arguments(tmp_path):
    """accelfold keeps prepfold's option ranges (clig/prepfold_cmd.cli:55-67)
    and refuses a file that is not an ACCEL candidate file."""
    fold = os.path.join(os.path.dirname(A.__file__), "accelfold")
    base = ["-accelfile", "x_ACCEL_50.cand", "-accelcand", "1", "x.dat"]
    for args in (["-npfact", "11"], ["-npfact", "0"], ["-pstep", "11"], ["-pdstep", "21"],
                 ["-pdstep", "0"], ["-npart", "5000"]):
        r = subprocess.run([fold] + args + base, cwd=tmp_path, capture_output=True, text=True,
                           timeout=60)
        assert r.returncode == 1 and "must be" in r.stderr, (args, r.stderr)
    r = subprocess.run([fold, "-bogus"] + base, cwd=tmp_path, capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "usage" in r.stderr
    r = subprocess.run([fold, "-accelfile", "x.cand", "-accelcand", "1", "x.dat"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "not an ACCEL candidate file" in r.stderr


@pytest.mark.gpu
def test_gpu_realfft_matches_fp64():
    for n in (1 << 10, 1 << 16, 1 << 20):
        x = synth.make_series(n, 1e-3, n, [dict(f0=37.3, duty=0.1, amp=0.5)])
        got = A.realfft_gpu(x)
        ref = np.fft.rfft(x.astype(np.float64))
        # FP32 radix-2 passes: worst bin within 1e-5 of the rms bin amplitude
        rms = np.sqrt(np.mean(np.abs(ref[1:n // 2]) ** 2))
        err = np.max(np.abs(got[1:] - ref[1:n // 2])) / rms
        print(f"realfft n={n}: worst error / rms amplitude {err:.2e}", flush=True)
        assert err < 1e-5, (n, err)
        assert abs(got[0].real - ref[0].real) < 1e-3 * np.sqrt(n)
        assert abs(got[0].imag - ref[n // 2].real) < 1e-3 * np.sqrt(n)


@pytest.mark.gpu
def test_cli_end_to_end_matches_oracle(tmp_path):
    print("case: accelsearch CLI on a 2^18-sample .dat", flush=True)
    N, dt = 1 << 18, 64e-6
    x = synth.make_series(N, dt, 77, [dict(f0=123.456, z=12.0, duty=0.05, amp=0.2),
                                      dict(f0=444.4, z=-30.0, duty=0.08, amp=0.15)])
    x.tofile(tmp_path / "syn.dat")
    _inf(str(tmp_path / "syn.inf"), N, dt)
    r = _run(["-zmax", "50", "-numharm", "8", "-sigma", "2.0", "syn.dat"], tmp_path)
    print(r.stdout[-600:], r.stderr[-600:], flush=True)
    assert r.returncode == 0
    acc = tmp_path / "syn_ACCEL_50"
    assert acc.exists() and (tmp_path / "syn_ACCEL_50.cand").exists()
    lines = [l.split() for l in open(acc) if re.match(r"^\d", l)]
    # the same spectrum through the oracle
    fft = A.realfft_gpu(x)
    lib_fft = fft.copy()
    A.lib.accel_prepare_dat_spectrum(lib_fft.ctypes.data, len(lib_fft), float(N), dt)
    o = O.make_obs(len(lib_fft), N, dt, zmax=50, numharm=8, sigma=2.0)
    _, oc, _ = O.accelsearch(o, lib_fft)
    assert len(lines) == len(oc) and len(oc) >= 2
    for f, c in zip(lines, oc):
        assert int(f[4]) == int(c[2])
        assert abs(float(f[1]) - c[1]) <= 0.006
        assert abs(float(f[7].split("(")[0]) - c[3]) <= 0.5 / c[2] + 1e-6
    first = open(acc).read()
    assert "Number of bins in the time series" in first
    # -flo/-fhi always carry their defaults, which win over -rlo/-rhi
    # (src/accelsearch_cmd.c:49-61, src/accel_utils.c:1686-1730)
    r = _run(["-zmax", "50", "-numharm", "8", "-sigma", "2.0", "-rlo", "5000", "-rhi", "6000",
              "syn.dat"], tmp_path)
    assert r.returncode == 0
    assert open(acc).read() == first
