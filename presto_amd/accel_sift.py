"""The reference's examplescripts/ACCEL_sift.py over this package's ACCEL
files: read every ACCEL file, drop duplicates across DMs, candidates with
DM problems and harmonically related ones, and print (or write) the
sorted candidate list.

    python -m presto_amd.accel_sift [-o candlist.txt] [--accel GLOB] [--inf GLOB]

Defaults as in the example script: ACCEL files '*ACCEL_*0', .inf files
'*DM*.inf', at least 2 DMs, low-DM cutoff 2.0, sigma threshold 4.0,
coherent-power threshold 100, r_err 1.1, periods 0.5 ms .. 15 s,
harmonic power cutoff 8."""
import argparse
import glob
import re
import sys
from operator import attrgetter

from . import sifting


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--accel", default="*ACCEL_*0", help="glob of the ACCEL files")
    ap.add_argument("--inf", default="*DM*.inf", help="glob of the .inf files")
    ap.add_argument("--min-num-dms", type=int, default=2)
    ap.add_argument("--low-dm-cutoff", type=float, default=2.0)
    ap.add_argument("--sigma", type=float, default=4.0)
    ap.add_argument("-o", "--output", default=None, help="candidate list file (stdout)")
    a = ap.parse_args(argv)
    sifting.sigma_threshold = a.sigma
    sifting.c_pow_threshold = 100.0
    sifting.known_birds_p = []
    sifting.known_birds_f = []
    sifting.r_err = 1.1
    sifting.short_period = 0.0005
    sifting.long_period = 15.0
    sifting.harm_pow_cutoff = 8.0

    inffiles = glob.glob(a.inf)
    candfiles = glob.glob(a.accel)
    if not inffiles:
        print("no .inf files match '%s'" % a.inf, file=sys.stderr)
        return 1
    # the DM of each trial from its file name (ACCEL_sift.py:52-56)
    if len(re.findall("_[0-9][0-9][0-9]M_", inffiles[0])):
        dmstrs = [x.split("DM")[-1].split("_")[0] for x in candfiles]
    else:
        dmstrs = [x.split("DM")[-1].split(".inf")[0] for x in inffiles]
    dms = sorted(map(float, dmstrs))
    dmstrs = ["%.2f" % x for x in dms]

    cands = sifting.read_candidates(candfiles)
    if len(cands):
        cands = sifting.remove_duplicate_candidates(cands)
    if len(cands):
        cands = sifting.remove_DM_problems(cands, a.min_num_dms, dmstrs, a.low_dm_cutoff)
    if len(cands):
        cands = sifting.remove_harmonics(cands)
    if len(cands):
        cands.sort(key=attrgetter('sigma'), reverse=True)
        sifting.write_candlist(cands, a.output)
    return 0


if __name__ == "__main__":
    sys.exit(main())
