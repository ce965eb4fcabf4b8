"""Writes the oracle's full candidate list for the plain 2^23-sample
workload (dt 64 us, 3 injected pulsars, zmax 200, numharm 16, sigma 2.0,
standard path) as tests/golden/config2_candidates.json.

The spectrum is the one the GPU test searches: synth.config2_series(),
numpy's FP64 real FFT rounded to FP32 and the library's host de-reddening
(accel.spectrum_from_series).  The oracle (the CPU restatement of
src/accelsearch.c:161-241) runs every block with OpenMP threads; the
candidate merge is replayed in the reference's block order, so the list
does not depend on the thread count.

    python tests/golden/make_config2_candidates.py

Takes a few minutes on 8 cores.  The fixture also records a checksum of
the spectrum, so that a test can tell a different input from a different
search."""
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

OUT = os.path.join(HERE, "config2_candidates.json")


def main(out=OUT):
    x, ps = synth.config2_series()
    dt = 64e-6
    fft = A.spectrum_from_series(x, dt)
    o = O.make_obs(len(fft), len(x), dt, zmax=200, numharm=16, sigma=2.0)
    t0 = time.time()
    _, cands, nb = O.accelsearch(o, fft)
    el = time.time() - t0
    rows = [dict(power=float(c[0]), sigma=float(c[1]), numharm=int(c[2]), r=float(c[3]),
                 z=float(c[4]), w=float(c[5])) for c in cands]
    doc = dict(workload="2^23 samples, dt 64 us, zmax 200, numharm 16, sigma 2.0",
               spectrum_sha256=hashlib.sha256(fft.tobytes()).hexdigest(),
               blocks=int(nb), oracle_seconds=round(el, 1), pulsars=ps, candidates=rows)
    with open(out, "w") as f:
        json.dump(doc, f, indent=0)
        f.write("\n")
    print(f"{len(rows)} candidates over {nb} blocks in {el:.0f} s -> {out}")


if __name__ == "__main__":
    main(*sys.argv[1:])
