"""Writes the ACCEL files of a small synthetic DM-trial set with this
package's accelsearch command (GPU), for the sifting tests.

Six DM trials, DM 0, 10, ..., 50, of 2^18 samples at dt 256 us: seeded
noise, a pulsar (123.4 Hz, z = 4, 5 % duty) whose amplitude peaks at DM 30
and falls off as a Gaussian of width 10 in DM, and a narrow-band
interference line (47.3 Hz) that is strongest at DM 0.  Each trial gets a
.dat, a .inf carrying its DM, and `accelsearch -zmax 0` and `-zmax 50`
runs (numharm 8, sigma 3).

    python tests/golden/make_sift_fixtures.py OUTDIR

The committed fixtures under tests/golden/sift/ are the .inf and ACCEL
text files this wrote."""
import os
import subprocess
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", ".."))
from presto_amd import accel as A  # noqa: E402
from presto_amd import surface as S  # noqa: E402
from presto_amd import synth  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
N, DT = 1 << 18, 256e-6
DMS = [0.0, 10.0, 20.0, 30.0, 40.0, 50.0]


def trial(dm, i):
    amp = 0.18 * float(np.exp(-0.5 * ((dm - 30.0) / 10.0) ** 2))
    rfi = 0.25 * float(np.exp(-0.5 * (dm / 8.0) ** 2))
    pulsars = [dict(f0=123.4, z=4.0, duty=0.05, amp=amp),
               dict(f0=47.3, z=0.0, duty=0.5, amp=rfi)]
    return synth.make_series_f32(N, DT, 7000 + i, pulsars)


def main(out):
    os.makedirs(out, exist_ok=True)
    cli = os.path.join(os.path.dirname(A.__file__), "accelsearch")
    for i, dm in enumerate(DMS):
        root = os.path.join(out, "sift_DM%.2f" % dm)
        trial(dm, i).tofile(root + ".dat")
        # a radio .inf (it carries the DM), from the GBT fixture's fields
        d = S.readinf(os.path.join(HERE, "1937_DM71.02_zerodm"))
        d.name = os.path.basename(root).encode()
        d.object = d.observer = d.analyzer = b"synthetic"
        d.N, d.dt, d.dm = float(N), DT, dm
        d.numonoff, d.onoff[0], d.onoff[1] = 1, 0.0, float(N - 1)
        d.notes = b"    synthetic DM trial\n"
        cwd = os.getcwd()
        os.chdir(out)
        try:
            S.writeinf(d)
        finally:
            os.chdir(cwd)
        for z in (0, 50):
            r = subprocess.run([cli, "-zmax", str(z), "-numharm", "8", "-sigma", "3.0",
                                os.path.basename(root) + ".dat"], cwd=out,
                               capture_output=True, text=True, timeout=300)
            if r.returncode:
                sys.exit("accelsearch failed at DM %.2f zmax %d: %s" % (dm, z, r.stderr))
        os.remove(root + ".dat")
        print("DM %.2f done" % dm, flush=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "sift_fixtures")
