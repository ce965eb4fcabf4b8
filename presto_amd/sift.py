"""Sifting of accelsearch candidates across DM trials: a ctypes view of the
C sifter in libpresto_accel.so (accel_sift_*, include/presto_accel.h).

The names follow the reference's python/presto/sifting.py so that scripts
written against it read the same: module-level thresholds (r_err,
long_period, short_period, sigma_threshold, c_pow_threshold,
harm_pow_cutoff, known_birds_f, known_birds_p) read when a rule runs;
Candidate; Candlist with its reject_* methods, duplicate, DM-problem and
harmonic removal and text outputs; read_candidates and
candlist_from_candfile.  Every rule runs in the C library; a Candlist owns
a C sift set and its lists are read back from it.  Plots are not built."""
import ctypes as C

import numpy as np

from presto_amd.accel import lib, AccelError

r_err = 1.1
long_period = 15.0
short_period = 0.0005
sigma_threshold = 6.0
c_pow_threshold = 100.0
harm_pow_cutoff = 8.0
known_birds_p = []
known_birds_f = []

_d, _i, _p = C.c_double, C.c_int, C.c_void_p

GOOD, DUPLICATE = 0, 1
BADLISTS = ["knownbirds", "longperiod", "shortperiod", "threshold", "harmpowcutoff",
            "rogueharmpow", "harmonic", "dmproblem"]
R_LONG, R_SHORT, R_BIRDS, R_THRESH, R_HPOW, R_ROGUE = 1, 2, 4, 8, 16, 32
R_DEFAULT = 0x3F
MAXHITS = 4096


class _Params(C.Structure):
    _fields_ = [("r_err", _d), ("long_period", _d), ("short_period", _d),
                ("sigma_threshold", _d), ("c_pow_threshold", _d), ("harm_pow_cutoff", _d),
                ("birds_f", _p), ("nbirds_f", _i), ("birds_p", _p), ("nbirds_p", _i)]


class _Cand(C.Structure):
    _fields_ = [("path", C.c_char * 512), ("dmstr", C.c_char * 32), ("note", C.c_char * 512),
                ("candnum", _i), ("numharm", _i), ("nhits", _i), ("pad", _i),
                ("sigma", _d), ("ipow", _d), ("cpow", _d), ("r", _d), ("z", _d), ("T", _d),
                ("snr", _d)]


def _sig(name, res, *args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = list(args)


_sig("accel_sift_create", _p, _i, _i)
_sig("accel_sift_destroy", None, _p)
_sig("accel_sift_add", _i, _p, C.POINTER(_Cand), _p, _p, _p, _i)
_sig("accel_sift_read", _i, _p, C.c_char_p)
_sig("accel_sift_parse_power", _d, C.c_char_p)
_sig("accel_sift_reject", _i, _p, C.POINTER(_Params), C.c_uint)
_sig("accel_sift_remove_duplicates", _i, _p, C.POINTER(_Params))
_sig("accel_sift_remove_dm_problems", _i, _p, _i, _p, _i, _d)
_sig("accel_sift_remove_harmonics", _i, _p, C.POINTER(_Params))
_sig("accel_sift_sort_sigma", _i, _p)
_sig("accel_sift_extend", _i, _p, _p)
_sig("accel_sift_count", _i, _p, _i)
_sig("accel_sift_get", _i, _p, _i, _i, C.POINTER(_Cand), _p, _p, _p, _i)
_sig("accel_sift_write", _i, _p, C.c_char_p)
_sig("accel_sift_write_report", _i, _p, C.c_char_p)
_sig("accel_sift_write_summary", _i, _p, C.c_char_p)


def _check(rc, what):
    if rc < 0:
        raise AccelError(f"{what}: {lib.accel_last_error().decode()}")
    return rc


def _bytes(s):
    return s.encode() if isinstance(s, str) else s


def _params(birds_f=(), birds_p=(), **given):
    """The module thresholds, or the values given, as the C parameter block
    (birdie arrays kept alive on the returned object)."""
    g = dict(r_err=r_err, long_period=long_period, short_period=short_period,
             sigma_threshold=sigma_threshold, c_pow_threshold=c_pow_threshold,
             harm_pow_cutoff=harm_pow_cutoff)
    g.update({k: v for k, v in given.items() if v is not None})
    bf = np.ascontiguousarray(np.asarray(birds_f, dtype=np.float64).reshape(-1))
    bp = np.ascontiguousarray(np.asarray(birds_p, dtype=np.float64).reshape(-1))
    p = _Params(g["r_err"], g["long_period"], g["short_period"], g["sigma_threshold"],
                g["c_pow_threshold"], g["harm_pow_cutoff"],
                bf.ctypes.data if bf.size else None, bf.size // 2,
                bp.ctypes.data if bp.size else None, bp.size // 2)
    p._keep = (bf, bp)
    return p


def parse_power(txt):
    """"value(err)" or "value(err)x10^e" as a float (an exponent cut off by
    the column width counts as 5)."""
    return lib.accel_sift_parse_power(_bytes(txt))


class Candidate:
    """A candidate of an ACCEL file (the reference's Candidate fields)."""

    def __init__(self, candnum, sigma, numharm, ipow, cpow, bin, z, DMstr, filename, T):
        self.path, _, self.filename = filename.rpartition("/")
        self.fullname = filename
        self.candnum, self.sigma, self.numharm = candnum, sigma, numharm
        self.ipow_det, self.cpow, self.r, self.z, self.T = ipow, cpow, bin, z, T
        self.DMstr = DMstr
        self.harm_pows = None
        self.harm_amps = None
        self.snr = 0.0
        self.hits = []
        self.note = ""

    @property
    def f(self):
        return self.r / self.T

    @property
    def p(self):
        return 1.0 / self.f

    @property
    def DM(self):
        return float(self.DMstr)

    def harms_to_snr(self):
        a = np.maximum(np.asarray(self.harm_pows, dtype=np.float64) - 1.0, 0.0)
        self.snr = float(np.sum(np.sqrt(a)))

    def __str__(self):
        return "%-65s   %7.2f  %6.2f  %6.2f    %2d     %7.1f  %7.1f  %12.6f  %10.2f  %8.2f " % (
            f"{self.filename}:{self.candnum}", self.DM, self.snr, self.sigma, self.numharm,
            self.ipow_det, self.cpow, self.p * 1000, self.r, self.z)


def _from_c(h, lst, i):
    c = _Cand()
    pows = np.zeros(64)
    phs = np.zeros(64)
    hits = np.zeros((MAXHITS, 3))
    nh = _check(lib.accel_sift_get(h, lst, i, C.byref(c), pows.ctypes.data, phs.ctypes.data,
                                   hits.ctypes.data, MAXHITS), "accel_sift_get")
    out = Candidate(c.candnum, c.sigma, c.numharm, c.ipow, c.cpow, c.r, c.z,
                    c.dmstr.decode(), c.path.decode(), c.T)
    out.snr = c.snr
    out.note = c.note.decode()
    out.harm_pows = pows[:c.numharm].copy()
    out.harm_amps = (np.sqrt(out.harm_pows) * np.exp(1j * phs[:c.numharm])).astype(np.complex64)
    out.hits = [tuple(x) for x in hits[:min(nh, MAXHITS)]]
    return out


class Candlist:
    """A set of candidates in the C sifter: the good list, the duplicates
    and the reject lists (badlists)."""

    def __init__(self, cands=None, trackbad=False, trackdupes=False):
        self.trackbad, self.trackdupes = trackbad, trackdupes
        self.h = lib.accel_sift_create(int(trackbad), int(trackdupes))
        if not self.h:
            raise AccelError("cannot create a sift set")
        for c in cands or []:
            self.append(c)

    def append(self, c):
        cc = _Cand()
        cc.path = _bytes(c.fullname)[:511]
        cc.dmstr = _bytes(c.DMstr)
        cc.candnum, cc.numharm = int(c.candnum), int(c.numharm)
        cc.sigma, cc.ipow, cc.cpow = float(c.sigma), float(c.ipow_det), float(c.cpow)
        cc.r, cc.z, cc.T, cc.snr = float(c.r), float(c.z), float(c.T), float(c.snr)
        pows = np.ascontiguousarray(c.harm_pows if c.harm_pows is not None
                                    else np.zeros(c.numharm), dtype=np.float64)
        phs = (np.angle(np.asarray(c.harm_amps)).astype(np.float64) if c.harm_amps is not None
               else np.zeros(c.numharm))
        phs = np.ascontiguousarray(phs)
        hits = np.ascontiguousarray(np.asarray(c.hits, dtype=np.float64).reshape(-1, 3))
        _check(lib.accel_sift_add(self.h, C.byref(cc), pows.ctypes.data, phs.ctypes.data,
                                  hits.ctypes.data if len(hits) else None, len(hits)),
               "accel_sift_add")

    def _list(self, lst):
        n = _check(lib.accel_sift_count(self.h, lst), "accel_sift_count")
        return [_from_c(self.h, lst, i) for i in range(n)]

    @property
    def cands(self):
        return self._list(GOOD)

    @property
    def duplicates(self):
        return self._list(DUPLICATE)

    @property
    def badlists(self):
        return {name: self._list(2 + k) for k, name in enumerate(BADLISTS)}

    def __len__(self):
        return lib.accel_sift_count(self.h, GOOD)

    def get_numcands(self):
        return len(self)

    def __iter__(self):
        return iter(self.cands)

    def __getitem__(self, i):
        return self.cands[i]

    def get_all_goodcands(self):
        return self.cands + self.duplicates

    def get_all_badcands(self):
        out = []
        for v in self.badlists.values():
            out += v
        return out

    def get_all_cands(self):
        return self.get_all_goodcands() + self.get_all_badcands()

    def _reject(self, rules, **given):
        return _check(lib.accel_sift_reject(self.h, C.byref(_params(**given)), rules),
                      "accel_sift_reject")

    def reject_longperiod(self, long_period=None):
        self._reject(R_LONG, long_period=long_period)

    def reject_shortperiod(self, short_period=None):
        self._reject(R_SHORT, short_period=short_period)

    def reject_knownbirds(self, known_birds_f=[], known_birds_p=[]):
        """As the reference, the birdie lists are the arguments (empty by
        default), not the module's."""
        self._reject(R_BIRDS, birds_f=known_birds_f, birds_p=known_birds_p)

    def reject_threshold(self, sigma_threshold=None, c_pow_threshold=None):
        self._reject(R_THRESH, sigma_threshold=sigma_threshold, c_pow_threshold=c_pow_threshold)

    def reject_harmpowcutoff(self, harm_pow_cutoff=None):
        self._reject(R_HPOW, harm_pow_cutoff=harm_pow_cutoff)

    def reject_rogueharmpow(self):
        self._reject(R_ROGUE)

    def default_rejection(self):
        self._reject(R_DEFAULT)

    def remove_duplicate_candidates(self, verbosity=1):
        n = _check(lib.accel_sift_remove_duplicates(self.h, C.byref(_params())),
                   "accel_sift_remove_duplicates")
        if verbosity >= 1:
            print("Found %d candidates (%d duplicates merged).\n" % (len(self), n))

    def remove_DM_problems(self, numdms, dmlist, low_DM_cutoff, verbosity=1):
        dms = np.ascontiguousarray([float(d) for d in dmlist], dtype=np.float64)
        n = _check(lib.accel_sift_remove_dm_problems(self.h, int(numdms), dms.ctypes.data,
                                                     len(dms), float(low_DM_cutoff)),
                   "accel_sift_remove_dm_problems")
        if verbosity >= 1:
            print("Removed %d candidates with DM problems.\n" % n)

    def remove_harmonics(self, verbosity=1):
        n = _check(lib.accel_sift_remove_harmonics(self.h, C.byref(_params())),
                   "accel_sift_remove_harmonics")
        if verbosity >= 1:
            print("Removed a total of %d harmonics.\n" % n)

    def sort_by_sigma(self):
        lib.accel_sift_sort_sigma(self.h)

    def extend(self, other):
        """Moves every candidate of other into this set."""
        _check(lib.accel_sift_extend(self.h, other.h), "accel_sift_extend")

    def to_file(self, candfilenm=None):
        _check(lib.accel_sift_write(self.h, _bytes(candfilenm) if candfilenm else None),
               "accel_sift_write")

    def write_cand_report(self, reportfilenm=None):
        _check(lib.accel_sift_write_report(self.h, _bytes(reportfilenm) if reportfilenm else None),
               "accel_sift_write_report")

    def print_cand_summary(self, summaryfilenm=None):
        _check(lib.accel_sift_write_summary(self.h,
                                            _bytes(summaryfilenm) if summaryfilenm else None),
               "accel_sift_write_summary")

    def close(self):
        if getattr(self, "h", None):
            lib.accel_sift_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def candlist_from_candfile(filename, trackbad=False, trackdupes=False):
    cl = Candlist(trackbad=trackbad, trackdupes=trackdupes)
    _check(lib.accel_sift_read(cl.h, _bytes(filename)), "accel_sift_read")
    return cl


def read_candidates(filenms, prelim_reject=True, track=False):
    """Candidates of every ACCEL file, each file's preliminary rejection
    (default_rejection) applied when prelim_reject."""
    out = Candlist(trackbad=track, trackdupes=track)
    for fn in filenms:
        cl = candlist_from_candfile(fn, trackbad=track, trackdupes=track)
        if prelim_reject:
            cl.default_rejection()
        out.extend(cl)
    return out


def remove_duplicate_candidates(candlist, *args, **kwargs):
    candlist.remove_duplicate_candidates(*args, **kwargs)
    return candlist


def remove_DM_problems(candlist, *args, **kwargs):
    candlist.remove_DM_problems(*args, **kwargs)
    return candlist


def remove_harmonics(candlist, *args, **kwargs):
    candlist.remove_harmonics(*args, **kwargs)
    return candlist


def write_candlist(candlist, *args, **kwargs):
    candlist.to_file(*args, **kwargs)
