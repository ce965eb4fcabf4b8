"""Sifting of accelsearch candidates across DM trials: the API of the
reference's python/presto/sifting.py (module globals, Candidate, Candlist,
read_candidates, remove_duplicate_candidates, remove_DM_problems,
remove_harmonics, write_candlist), written over the ACCEL files this
package produces.

The rules are the reference's, step for step (file:line of
python/presto/sifting.py at each function): the same globals and
defaults, the same parsing of the ACCEL text (fund_re / harms_re, :40-42,
:1075-1176), the same rejection tests and their order, the same
duplicate, DM-problem and harmonic passes with their list surgery, and
the same candlist text.  candidate_sigma is the library's restatement of
src/characteristics.c:526-546.  The plots of the reference
(plot_summary / plot_rejects / plot_goodcands) are not part of this
module.
"""
import copy
import glob
import os
import re
import sys
from operator import attrgetter

import numpy as np

from . import accel as _accel

# Defaults of python/presto/sifting.py:14-37; importers set them.
# How close two candidates are (in Fourier bins) to count as the same one
r_err = 1.1
# Longest period candidates to consider (s)
long_period = 15.0
# Shortest period candidates to consider (s)
short_period = 0.0005
# Ignore candidates with a sigma (from incoherent power summation) below this
sigma_threshold = 6.0
# Ignore candidates with a coherent power below this
c_pow_threshold = 100.0
# Ignore candidates none of whose harmonics reaches this power
harm_pow_cutoff = 8.0
# (ms, err) and (Hz, err) birdies
known_birds_p = []
known_birds_f = []

fund_re = re.compile(r"^\d")
harms_re = re.compile(r"^[ ]\d")
DM_re = re.compile(r"DM(\d+\.\d{2})")


def candidate_sigma(power, numsum, numtrials):
    """src/characteristics.c:526-546 (presto.candidate_sigma)."""
    return _accel.lib.accel_candidate_sigma(float(power), int(numsum), float(numtrials))


def remove_duplicate_candidates(candlist, *args, **kwargs):
    """sifting.py:46-62: calls Candlist.remove_duplicate_candidates."""
    candlist.remove_duplicate_candidates(*args, **kwargs)
    return candlist


def remove_DM_problems(candlist, *args, **kwargs):
    """sifting.py:65-81: calls Candlist.remove_DM_problems."""
    candlist.remove_DM_problems(*args, **kwargs)
    return candlist


def remove_harmonics(candlist, *args, **kwargs):
    """sifting.py:84-100: calls Candlist.remove_harmonics."""
    candlist.remove_harmonics(*args, **kwargs)
    return candlist


def write_candlist(candlist, *args, **kwargs):
    """sifting.py:103-104."""
    candlist.to_file(*args, **kwargs)


def sigma_to_size(sigmas):
    """sifting.py:107-118: plot marker sizes."""
    return np.clip(20 ** (sigmas / 6), 5, 400)


def print_sift_globals():
    """sifting.py:121-129."""
    print("r_err =", r_err)
    print("short_period =", short_period)
    print("long_period =", long_period)
    print("sigma_threshold =", sigma_threshold)
    print("c_pow_threshold =", c_pow_threshold)
    print("harm_pow_cutoff =", harm_pow_cutoff)
    print("known_birds_p =", known_birds_p)
    print("known_birds_f =", known_birds_f)


def parse_power(pow):
    """sifting.py:132-140: a power field, with the '^' exponent notation of
    large values."""
    power = float(pow.split("(")[0])
    if "^" in pow:
        try:
            expon = float(pow.split("^")[1])
        except ValueError:
            expon = 5
        power *= 10.0 ** expon
    return power


class Candidate(object):
    """sifting.py:143-181: one ACCEL candidate of one DM trial."""

    def __init__(self, candnum, sigma, numharm, ipow, cpow, bin, z, DMstr, filename, T):
        self.path, self.filename = os.path.split(filename)
        self.candnum = candnum
        self.sigma = sigma
        self.numharm = numharm
        self.ipow_det = ipow
        self.cpow = cpow
        self.r = bin
        self.f = bin / T
        self.z = z
        self.T = T
        self.p = 1.0 / self.f
        self.DMstr = DMstr
        self.DM = float(DMstr)
        self.harm_pows = None
        self.harm_amps = None
        self.snr = 0.0
        self.hits = []
        self.note = ""

    def add_as_hit(self, other):
        self.hits.extend(other.hits)

    def __str__(self):
        cand = self.filename + ':' + repr(self.candnum)
        return "%-65s   %7.2f  %6.2f  %6.2f  %s   %7.1f  %7.1f  %12.6f  %10.2f  %8.2f " % \
            (cand, self.DM, self.snr, self.sigma, "%2d".center(7) % self.numharm,
             self.ipow_det, self.cpow, self.p * 1000, self.r, self.z)

    def harms_to_snr(self):
        # the harmonics' amplitudes above the mean power level of 1
        harmamps = np.asarray(self.harm_pows) - 1.0
        harmamps[harmamps < 0.0] = 0.0
        self.snr = np.sum(np.sqrt(harmamps))


_HEADER = ("#" + "file:candnum".center(66) + "DM".center(9) + "SNR".center(8) +
           "sigma".center(8) + "numharm".center(9) + "ipow".center(9) + "cpow".center(9) +
           "P(ms)".center(14) + "r".center(12) + "z".center(8) + "numhits".center(9) + "\n")


class Candlist(object):
    """sifting.py:184-1072: the good candidates, the bad lists and the
    duplicates."""

    def __init__(self, cands=None, trackbad=False, trackdupes=False):
        self.cands = [] if cands is None else cands
        self.trackbad = trackbad
        self.trackdupes = trackdupes
        self.badlists = {'knownbirds': [], 'longperiod': [], 'shortperiod': [],
                         'threshold': [], 'harmpowcutoff': [], 'rogueharmpow': [],
                         'harmonic': [], 'dmproblem': []}
        self.duplicates = []

    def __iter__(self):
        return iter(self.cands)

    def __getitem__(self, key):
        return self.cands[key]

    def __delitem__(self, key):
        del self.cands[key]

    def sort(self, *args, **kwargs):
        self.cands.sort(*args, **kwargs)

    def mark_as_bad(self, icand, badlistname):
        cand = self.cands.pop(icand)
        if self.trackbad:
            self.badlists.setdefault(badlistname, []).append(cand)

    def mark_as_duplicate(self, icand):
        cand = self.cands.pop(icand)
        if self.trackdupes:
            self.duplicates.append(cand)

    def get_all_cands(self):
        return self.get_all_goodcands() + self.get_all_badcands()

    def get_all_goodcands(self):
        return self.cands + self.duplicates

    def get_all_badcands(self):
        cands = []
        for key in list(self.badlists.keys()):
            cands += self.badlists[key]
        return cands

    def reject_longperiod(self, long_period=None):
        """sifting.py:512-530."""
        if long_period is None:
            long_period = globals()['long_period']
        for ii in reversed(list(range(len(self.cands)))):
            cand = self.cands[ii]
            if cand.p > long_period:
                cand.note = "Period is too long (%g ms > %g ms)" % \
                    (cand.p * 1000, long_period * 1000)
                self.mark_as_bad(ii, 'longperiod')

    def reject_shortperiod(self, short_period=None):
        """sifting.py:532-550."""
        if short_period is None:
            short_period = globals()['short_period']
        for ii in reversed(list(range(len(self.cands)))):
            cand = self.cands[ii]
            if cand.p < short_period:
                cand.note = "Period is too short (%g ms < %g ms)" % \
                    (cand.p * 1000, short_period * 1000)
                self.mark_as_bad(ii, 'shortperiod')

    def reject_knownbirds(self, known_birds_f=[], known_birds_p=[]):
        """sifting.py:552-594.  As there, the globals are read only when None
        is passed: the defaults are empty lists."""
        if known_birds_f is None:
            known_birds_f = globals()['known_birds_f']
        if known_birds_p is None:
            known_birds_p = globals()['known_birds_p']
        for ii in reversed(list(range(len(self.cands)))):
            cand = self.cands[ii]
            known_bird = 0
            for bird, err in known_birds_f:
                if np.fabs(cand.f - bird) < err:
                    known_bird = 1
                    cand.note = "Freq (%.2f Hz) is within %g Hz of a known birdie " \
                                "centred at %.2f Hz" % (cand.f, err, bird)
                    break
            if known_bird:
                self.mark_as_bad(ii, 'knownbirds')
                continue
            for bird, err in known_birds_p:
                if np.fabs(cand.p * 1000.0 - bird) < err:
                    known_bird = 1
                    cand.note = "Period (%.2f ms) is within %g ms of a known birdie " \
                                "centred at %.2f ms" % (cand.f * 1000, err, bird)
                    break
            if known_bird:
                self.mark_as_bad(ii, 'knownbirds')
                continue

    def reject_threshold(self, sigma_threshold=None, c_pow_threshold=None):
        """sifting.py:596-634: one harmonic needs sigma or coherent power,
        several harmonics need sigma."""
        if sigma_threshold is None:
            sigma_threshold = globals()['sigma_threshold']
        if c_pow_threshold is None:
            c_pow_threshold = globals()['c_pow_threshold']
        for ii in reversed(list(range(len(self.cands)))):
            cand = self.cands[ii]
            if cand.numharm == 1:
                if cand.sigma < sigma_threshold and cand.cpow < c_pow_threshold:
                    cand.note = "Only 1 harmonic and both sigma (%g < %g) and coherent " \
                                "power (%g < %g) are too low." % \
                                (cand.sigma, sigma_threshold, cand.cpow, c_pow_threshold)
                    self.mark_as_bad(ii, 'threshold')
            else:
                if cand.sigma < sigma_threshold:
                    cand.note = "%d harmonics and sigma (%g < %g) is too low." % \
                                (cand.numharm, cand.sigma, sigma_threshold)
                    self.mark_as_bad(ii, 'threshold')

    def reject_harmpowcutoff(self, harm_pow_cutoff=None):
        """sifting.py:636-655."""
        if harm_pow_cutoff is None:
            harm_pow_cutoff = globals()['harm_pow_cutoff']
        for ii in reversed(list(range(len(self.cands)))):
            cand = self.cands[ii]
            maxharm = np.argmax(cand.harm_pows)
            maxpow = cand.harm_pows[maxharm]
            if maxpow < harm_pow_cutoff:
                cand.note = "All harmonics have power < %g" % harm_pow_cutoff
                self.mark_as_bad(ii, 'harmpowcutoff')

    def reject_rogueharmpow(self):
        """sifting.py:657-690: a single dominant high-numbered harmonic."""
        for ii in reversed(list(range(len(self.cands)))):
            cand = self.cands[ii]
            maxharm = np.argmax(cand.harm_pows)
            maxpow = cand.harm_pows[maxharm]
            sortedpows = np.sort(cand.harm_pows)
            if cand.numharm >= 8 and maxharm > 4 and maxpow > 2 * sortedpows[-2]:
                cand.note = "High-numbered harmonic (%d) has too much power" % maxharm
                self.mark_as_bad(ii, 'rogueharmpow')
            elif cand.numharm >= 4 and maxharm > 2 and maxpow > 3 * sortedpows[-2]:
                cand.note = "High-numbered harmonic (%d) has too much power" % maxharm
                self.mark_as_bad(ii, 'rogueharmpow')

    def default_rejection(self):
        """sifting.py:692-706."""
        self.reject_longperiod()
        self.reject_shortperiod()
        self.reject_knownbirds()
        self.reject_threshold()
        self.reject_harmpowcutoff()
        self.reject_rogueharmpow()

    def remove_duplicate_candidates(self, verbosity=1):
        """sifting.py:708-767: runs of candidates within r_err bins (in r
        order) collapse onto their highest-sigma member, which collects the
        others' DM hits; the scan stays at a run's start until no match."""
        if verbosity >= 1:
            print("  Sorting the %d candidates by frequency..." % self.get_numcands())
        self.cands.sort(key=attrgetter('r'))
        if verbosity >= 1:
            print("  Searching for dupes...")
        ii = 0
        while ii < self.get_numcands():
            jj = ii + 1
            if jj < self.get_numcands() and \
                    np.fabs(self.cands[ii].r - self.cands[jj].r) < r_err:
                jj += 1
                while jj < self.get_numcands() and \
                        np.fabs(self.cands[ii].r - self.cands[jj].r) < r_err:
                    jj += 1
                matches = self.cands[ii:jj]
                matches.sort(key=attrgetter('sigma'), reverse=True)
                bestindex = self.cands.index(matches[0])
                bestcand = self.cands[bestindex]
                for matchind in reversed(list(range(ii, jj))):
                    if matchind == bestindex:
                        continue
                    match = self.cands[matchind]
                    bestcand.add_as_hit(match)
                    match.note = "This candidate is a duplicate of %s:%d" % \
                        (bestcand.filename, bestcand.candnum)
                    self.mark_as_duplicate(matchind)
                    if verbosity >= 2:
                        print("Removing %s:%d (index: %d)" %
                              (match.filename, match.candnum, matchind))
                        print("    %s" % match.note)
            else:
                ii += 1
        if verbosity >= 1:
            print("Found %d candidates.\n" % self.get_numcands())
        self.cands.sort(key=attrgetter('sigma'), reverse=True)

    def remove_harmonics(self, verbosity=1):
        """sifting.py:769-857: in sigma order, every candidate against every
        weaker one: integer ratios 1..16 either way (within r_err/T scaled
        by the ratio), then ten small fractions.  As there, a list that ends
        up with one candidate compares it with itself (ratio 1) and fails on
        the then empty list."""
        numremoved = 0
        self.cands.sort(key=attrgetter('sigma'), reverse=True)
        f_err = r_err / self.cands[0].T
        if verbosity >= 1:
            print("\nSearching for duplicate harmonics...")
        ii = 0
        while 1:
            fundcand = self.cands[ii]
            jj = len(self.cands) - 1
            zapj = 0
            while 1:
                harmcand = self.cands[jj]
                if zapj:
                    print("Hey!")
                for factor in np.arange(1.0, 17.0):
                    if np.fabs(fundcand.f - harmcand.f * factor) < f_err * factor:
                        zapj = 1
                        harmstr = "1/%dth" % factor
                    elif np.fabs(fundcand.f - harmcand.f / factor) < f_err / factor:
                        zapj = 1
                        if factor == 2.0:
                            harmstr = "%dnd" % factor
                        else:
                            harmstr = "%dth" % factor
                    if zapj:
                        if verbosity >= 2:
                            print("Removing %s:%d (%.2f Hz) because it is a harmonic (%s) "
                                  "of %s:%d (%.2f Hz)" %
                                  (harmcand.filename, harmcand.candnum, harmcand.f, harmstr,
                                   fundcand.filename, fundcand.candnum, fundcand.f))
                        break
                for numer, denom in zip([3.0, 5.0, 2.0, 4.0, 5.0, 3.0, 5.0, 2.0, 3.0, 4.0],
                                        [2.0, 2.0, 3.0, 3.0, 3.0, 4.0, 4.0, 5.0, 5.0, 5.0]):
                    factor = numer / denom
                    if np.fabs(fundcand.f - harmcand.f * factor) < f_err * factor:
                        if verbosity >= 2:
                            print("Removing %s:%d (%.2f Hz) because it is a harmonic "
                                  "(%d/%dth) of %s:%d (%.2f Hz)" %
                                  (harmcand.filename, harmcand.candnum, harmcand.f, denom,
                                   numer, fundcand.filename, fundcand.candnum, fundcand.f))
                        harmstr = "%d/%dth" % (denom, numer)
                        zapj = 1
                        break
                if zapj:
                    harmcand.note = "This candidate (P=%.4f s, DM=%.2f) is a harmonic (%s) " \
                                    "of %s:%d (P=%.4f s, DM=%.2f)." % \
                                    (harmcand.p, harmcand.DM, harmstr, fundcand.filename,
                                     fundcand.candnum, fundcand.p, fundcand.DM)
                    numremoved += 1
                    self.mark_as_bad(jj, 'harmonic')
                    if verbosity >= 2:
                        print("Removing %s:%d (index: %d)" %
                              (harmcand.filename, harmcand.candnum, jj))
                        print("    %s" % harmcand.note)
                    zapj = 0
                jj -= 1
                if jj == ii:
                    break
            ii += 1
            if ii >= len(self.cands) - 1:
                break
        if verbosity >= 1:
            print("Removed a total of %d harmonics.\n" % numremoved)

    def remove_DM_problems(self, numdms, dmlist, low_DM_cutoff, verbosity=1):
        """sifting.py:859-942: too few DM hits, the best hit at or below
        low_DM_cutoff, or no two hits at neighbouring DMs of dmlist."""
        dms = np.unique([float(dm) for dm in dmlist])
        dmstrs = ['%.2f' % dm for dm in dms]
        dmdict = dict(list(zip(dmstrs, list(range(len(dms))))))
        numremoved = num_toofew = num_toolow = num_gaps = 0
        self.cands.sort(key=attrgetter('sigma'), reverse=True)
        for ii in reversed(list(range(len(self.cands)))):
            currcand = self.cands[ii]
            if len(currcand.hits) < numdms:
                numremoved += 1
                num_toofew += 1
                currcand.note = "Candidate has only %d DM hits. This is less than minimum " \
                                "for 'good' cands (%d hits)" % (len(currcand.hits), numdms)
                self.mark_as_bad(ii, 'dmproblem')
                if verbosity >= 2:
                    print("Removing %s:%d (index: %d)" %
                          (currcand.filename, currcand.candnum, ii))
                    print("    %s" % currcand.note)
                continue
            # a hit is (DM, SNR, sigma)
            imax = np.argmax(np.array([hit[2] for hit in currcand.hits]))
            hitdm, hitsnr, hitsigma = currcand.hits[imax]
            if float(hitdm) <= low_DM_cutoff:
                numremoved += 1
                num_toolow += 1
                currcand.note = "Hit with max sigma (%g) has dm (%.2f) <= low DM cutoff " \
                                "(%.2f) " % (hitsigma, hitdm, low_DM_cutoff)
                self.mark_as_bad(ii, 'dmproblem')
                if verbosity >= 2:
                    print("Removing %s:%d (index: %d)" %
                          (currcand.filename, currcand.candnum, ii))
                    print("    %s" % currcand.note)
                continue
            if len(currcand.hits) > 1:
                currcand.hits.sort(key=lambda cand: float(cand[0]))
                dm_indices = np.asarray([dmdict["%.2f" % currcand.hits[jj][0]]
                                         for jj in range(len(currcand.hits))])
                min_dmind_diff = min(dm_indices[1:] - dm_indices[:-1])
                if min_dmind_diff > 1:
                    numremoved += 1
                    num_gaps += 1
                    currcand.note = "DM list of hits has gaps (i.e. consecutive DMs don't " \
                                    "have hits)."
                    self.mark_as_bad(ii, 'dmproblem')
                    if verbosity >= 2:
                        print("Removing %s:%d (index: %d)" %
                              (currcand.filename, currcand.candnum, ii))
                        print("    %s" % currcand.note)
                    continue
        if verbosity >= 1:
            print("Removed %d candidates with DM problems.\n" % numremoved)
        if verbosity >= 2:
            print("  # with too few hits:", num_toofew)
            print("  # with peak SNR too low:", num_toolow)
            print("  # with gaps in DM hits:", num_gaps)

    def print_cand_summary(self, summaryfilenm=None):
        """sifting.py:944-981."""
        if summaryfilenm is None:
            summaryfile = sys.stdout
        elif summaryfilenm in [sys.stdout, sys.stderr]:
            summaryfile = summaryfilenm
        else:
            summaryfile = open(summaryfilenm, "w")
        b = self.badlists
        summaryfile.write("   Candlist contains %d 'good' candidates\n" % len(self.cands))
        summaryfile.write("      # Known RFI rejects:           %d\n" % len(b['knownbirds']))
        summaryfile.write("      # Short period rejects:        %d\n" % len(b['shortperiod']))
        summaryfile.write("      # Long period rejects:         %d\n" % len(b['longperiod']))
        summaryfile.write("      # Missed threshold:            %d\n" % len(b['threshold']))
        summaryfile.write("      # No good harmonics:           %d\n" % len(b['harmpowcutoff']))
        summaryfile.write("      # One bad harmonic:            %d\n" % len(b['rogueharmpow']))
        summaryfile.write("      # Duplicate candidates:        %d\n" % len(self.duplicates))
        summaryfile.write("      # Harmonic candidates:         %d\n" %
                          len(self.badlists['harmonic']))
        summaryfile.write("      # Candidates with DM problems: %d\n" %
                          len(self.badlists['dmproblem']))
        if summaryfilenm not in [None, sys.stdout, sys.stderr]:
            summaryfile.close()

    def write_cand_report(self, reportfilenm=None):
        """sifting.py:983-1007: every bad candidate with its note."""
        reportfile = sys.stdout if reportfilenm is None else open(reportfilenm, "w")
        reportfile.write(_HEADER)
        for badcand in self.get_all_badcands():
            reportfile.write("%s (%d)\n" % (str(badcand), len(badcand.hits)))
            reportfile.write("    Note: %s\n\n" % badcand.note)
        if reportfilenm is not None:
            reportfile.close()

    def __add__(self, other):
        copy_of_self = copy.deepcopy(self)
        copy_of_self.extend(other)
        return copy_of_self

    def get_numcands(self):
        return len(self)

    def __len__(self):
        return len(self.cands)

    def extend(self, other):
        """sifting.py:1029-1043."""
        self.cands.extend(other.cands)
        self.duplicates.extend(other.duplicates)
        for key in other.badlists:
            self.badlists.setdefault(key, []).extend(other.badlists[key])

    def to_file(self, candfilenm=None):
        """sifting.py:1045-1072: the candlist text, the DM hits of each
        multi-hit candidate below it."""
        candfile = sys.stdout if candfilenm is None else open(candfilenm, "w")
        candfile.write(_HEADER)
        for goodcand in self.cands:
            candfile.write("%s (%d)\n" % (str(goodcand), len(goodcand.hits)))
            if len(goodcand.hits) > 1:
                goodcand.hits.sort(key=lambda cand: float(cand[0]))
                for hit in goodcand.hits:
                    numstars = int(hit[2] / 3.0)
                    candfile.write("  DM=%6.2f SNR=%5.2f Sigma=%5.2f   " % hit +
                                   numstars * '*' + '\n')
        if candfilenm is not None:
            candfile.close()


def candlist_from_candfile(filename, trackbad=False, trackdupes=False):
    """sifting.py:1075-1176: the candidates of one ACCEL file, their
    harmonic powers and phases, S/N, and sigma re-computed from the summed
    harmonic powers for one trial."""
    with open(filename, 'r') as candfile:
        lines = candfile.readlines()
    for line in lines:
        if line.startswith(" Number of bins in the time series"):
            numsamp = int(line.split()[-1])
        if line.startswith(" Width of each time series bin (sec)"):
            dt = float(line.split()[-1])
    tobs = numsamp * dt

    cands = []
    candnums = []
    current_goodcandnum = 0
    last_candnum = 0
    last_goodcandnum = 0
    for line in lines:
        if fund_re.match(line):
            split_line = line.split()
            candnum = int(split_line[0])
            if len(cands) and len(split_line[0]) == 4 and cands[-1].candnum >= 9999:
                candnum = last_candnum + 1
            sigma = float(split_line[1])
            i_pow_det = float(split_line[2])
            c_pow = float(split_line[3])
            numharm = int(split_line[4])
            bin = float(split_line[7].split("(")[0])
            z = float(split_line[9].split("(")[0])
            DMstr = DM_re.search(filename).groups()[0]
            cands.append(Candidate(candnum, sigma, numharm, i_pow_det, c_pow, bin, z, DMstr,
                                   filename, tobs))
            candnums.append(candnum)
            last_candnum = candnum
            continue
        elif harms_re.match(line):
            split_line = line.split()
            candnum = int(split_line[0])
            # more than 10K candidates: the ACCEL file's numbers wrap
            if len(split_line[0]) == 4 and last_goodcandnum >= 9999:
                candnum = last_goodcandnum + 1
            if candnum in candnums:
                cand = cands[candnums.index(candnum)]
                cand.harm_pows = np.zeros(cand.numharm, dtype=np.float64)
                cand.harm_amps = np.zeros(cand.numharm, dtype=np.complex64)
                power = parse_power(split_line[3])
                phase = float(split_line[9].split("(")[0])
                cand.harm_pows[0] = power
                cand.harm_amps[0] = np.sqrt(power) * np.exp(phase * 1.0j)
                if cand.numharm > 1:
                    current_goodcandnum = candnum
                    current_harmnum = 1
                else:
                    last_goodcandnum = candnum
                    current_goodcandnum = 0
                    cand.harms_to_snr()
                    opt_ipow = cand.harm_pows[0]
                    cand.sigma = candidate_sigma(opt_ipow, 1, 1)
                    cand.hits = [(cand.DM, cand.snr, cand.sigma)]
                    cand.ipow_det = opt_ipow
            continue
        if current_goodcandnum:
            cand = cands[candnums.index(current_goodcandnum)]
            power = parse_power(line.split()[2])
            phase = float(line.split()[8].split("(")[0])
            cand.harm_pows[current_harmnum] = power
            cand.harm_amps[current_harmnum] = np.sqrt(power) * np.exp(phase * 1.0j)
            current_harmnum += 1
            if current_harmnum == cand.numharm:
                cand.harms_to_snr()
                opt_ipow = sum(cand.harm_pows)
                cand.sigma = candidate_sigma(opt_ipow, cand.numharm, 1)
                cand.hits = [(cand.DM, cand.snr, cand.sigma)]
                cand.ipow_det = opt_ipow
                last_goodcandnum = candnum
                current_goodcandnum = 0
    return Candlist(cands, trackbad=trackbad, trackdupes=trackdupes)


def read_candidates(filenms, prelim_reject=True, track=False):
    """sifting.py:1179-1205: every file's candidates, each list through
    default_rejection first."""
    candlist = Candlist(trackbad=track, trackdupes=track)
    numfiles = len(filenms)
    if filenms:
        print("\nReading candidates from %d files...." % len(filenms))
        for ii, filenm in enumerate(filenms):
            curr_candlist = candlist_from_candfile(filenm, trackbad=track, trackdupes=track)
            if prelim_reject:
                curr_candlist.default_rejection()
            candlist.extend(curr_candlist)
            sys.stdout.write(" Read %d of %d files (%d cands)\r" %
                             (ii + 1, numfiles, len(candlist)))
            sys.stdout.flush()
        print("\nDone")
    else:
        print("Error:  There are no candidate files to read!")
    return candlist


def sift_directory(dir, outbasenm):
    """sifting.py:1208-1270 without its plots: the ACCEL_0 and ACCEL_50
    files of a directory, their DMs from the .inf files, writes
    <outbasenm>.accelcands and .accelcands.report."""
    global sigma_threshold
    sigma_threshold = 5.0
    print_sift_globals()
    inffns = glob.glob(os.path.join(dir, '*.inf'))
    from . import surface as _surface
    dmstrs = ['%.2f' % _surface.readinf(fn[:-4]).dm for fn in inffns]
    results = []
    for suffix in ('*ACCEL_0', '*ACCEL_50'):
        fns = glob.glob(os.path.join(dir, suffix))
        cands = read_candidates(fns, track=True)
        print("Read %d candidates from %d files" % (len(cands), len(fns)))
        print("%d candidates passed default rejection" % len(cands))
        if len(cands):
            cands.remove_duplicate_candidates()
        if len(cands):
            cands.remove_DM_problems(2, dmstrs, 2.0)
        cands.print_cand_summary()
        results.append(cands)
    all_accel_cands = results[0] + results[1]
    if len(all_accel_cands):
        all_accel_cands.remove_harmonics()
        all_accel_cands.cands.sort(key=attrgetter('sigma'), reverse=True)
        print("Found %d good candidates" % len(all_accel_cands))
        all_accel_cands.to_file(outbasenm + ".accelcands")
    all_accel_cands.write_cand_report(outbasenm + ".accelcands.report")
    all_accel_cands.print_cand_summary()
    return all_accel_cands
