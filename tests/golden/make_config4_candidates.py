"""Writes the oracle's full candidate list for one spectrum of the jerk
workload (synth.config4_trials, trial 0: 2^22 samples, dt 64 us, one
binary-like signal, zmax 200, wmax 300, numharm 8, sigma 2.0, standard
path) as tests/golden/config4_candidates.json.

The spectrum is the one the GPU test searches: numpy's FP64 real FFT
rounded to FP32 and the library's host de-reddening
(accel.spectrum_from_series).  The oracle (the CPU restatement of
src/accelsearch.c:161-241) runs all 2920 blocks with OpenMP threads over
each block's templates; the candidate merge runs in the reference's block
order, so the list does not depend on the thread count.

    python tests/golden/make_config4_candidates.py

Takes about half an hour on 8 cores.  The fixture also records a checksum
of the spectrum, so that a test can tell a different input from a
different search."""
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from presto_amd import accel as A, synth  # noqa: E402

OUT = os.path.join(HERE, "config4_candidates.json")


def main(out=OUT):
    i, N, dt, seed, ps, expect = next(synth.config4_trials(1))
    x = synth.make_series(N, dt, seed, ps)
    fft = A.spectrum_from_series(x, dt)
    o = O.make_obs(len(fft), N, dt, zmax=200, wmax=300, numharm=8, sigma=2.0)
    t0 = time.time()
    _, cands, nb = O.accelsearch(o, fft)
    el = time.time() - t0
    rows = [dict(power=float(c[0]), sigma=float(c[1]), numharm=int(c[2]), r=float(c[3]),
                 z=float(c[4]), w=float(c[5])) for c in cands]
    doc = dict(workload="2^22 samples, dt 64 us, zmax 200, wmax 300, numharm 8, sigma 2.0",
               spectrum_sha256=hashlib.sha256(fft.tobytes()).hexdigest(),
               blocks=int(nb), oracle_seconds=round(el, 1), pulsars=ps,
               expected_r_z_w=list(expect), candidates=rows)
    with open(out, "w") as f:
        json.dump(doc, f, indent=0)
        f.write("\n")
    print(f"{len(rows)} candidates over {nb} blocks in {el:.0f} s -> {out}")


if __name__ == "__main__":
    main(*sys.argv[1:])
