"""ctypes view of liboracle.so, the CPU restatement of PRESTO's accelsearch
path.  Test infrastructure only: imported by tests/, by
__graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
product package."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    return C.CDLL(LIB_PATH)


lib = _load()

_d, _i, _f, _ll, _p = C.c_double, C.c_int, C.c_float, C.c_longlong, C.c_void_p


class OObs(C.Structure):
    _fields_ = [
        ("N", _d), ("dt", _d), ("numbins", _ll),
        ("numharm", _i), ("zmax", _i), ("wmax", _i), ("wmaxP", _i),
        ("sigma", _f),
        ("flo", _d), ("fhi", _d), ("rlo_arg", _d), ("rhi_arg", _d),
        ("floP", _i), ("fhiP", _i), ("rloP", _i), ("rhiP", _i),
        ("lobin_arg", _ll), ("inmemP", _i), ("force_standard", _i), ("dat_input", _i),
        ("lobin", _ll), ("highestbin", _ll),
        ("maxkernlen", _i), ("corr_uselen", _i), ("fftlen", _i), ("numharmstages", _i),
        ("numz", _i), ("numw", _i), ("numbetween", _i), ("inmem", _i),
        ("T", _d), ("rlo", _d), ("rhi", _d), ("dr", _d), ("zlo", _d), ("zhi", _d),
        ("dz", _d), ("wlo", _d), ("whi", _d), ("dw", _d),
        ("powcut", _f * 6), ("numindep", _ll * 6), ("nph", _d), ("norm_type", _i),
    ]


class OHit(C.Structure):
    _fields_ = [("block", _i), ("stage", _i), ("wi", _i), ("zi", _i), ("ri", _i),
                ("pow", _f), ("sigma", _f)]


class OResult(C.Structure):
    _fields_ = [("hits", C.POINTER(OHit)), ("nhits", _ll), ("caphits", _ll),
                ("cands", _p), ("nblocks", _i)]


class OFfdot(C.Structure):
    _fields_ = [("rlo", _ll), ("zlo", _i), ("wlo", _i), ("numrs", _i), ("numzs", _i),
                ("numws", _i), ("powers", C.POINTER(_f)), ("dpowers", C.POINTER(_d)),
                ("rinds", C.POINTER(C.c_ushort)), ("zinds", C.POINTER(C.c_ushort))]


def _sig(name, res, *args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = list(args)
    return fn


_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_cf = np.ctypeslib.ndpointer(dtype=np.complex64, flags="C_CONTIGUOUS")
_cd = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")

_sig("o_fresnl", _i, _d, C.POINTER(_d), C.POINTER(_d))
_sig("o_r_resp_halfwidth", _i, _i)
_sig("o_z_resp_halfwidth", _i, _d, _i)
_sig("o_w_resp_halfwidth", _i, _d, _d, _i)
_sig("o_gen_r_response", None, _d, _i, _i, _cd)
_sig("o_gen_z_response", None, _d, _i, _d, _i, _cd)
_sig("o_gen_w_response", None, _d, _i, _d, _d, _i, _p, _p)
_sig("o_next2_to_n", _ll, _ll)
_sig("o_next_good_fftlen", _i, _i)
_sig("o_fftlen_from_kernwidth", _i, _i)
for _n in ("o_gauss_Q", "o_gauss_Qinv", "o_equivalent_gaussian_sigma"):
    _sig(_n, _d, _d)
for _n in ("o_chisq_Q_even", "o_chisq_Qinv_even", "o_chi2_logp"):
    _sig(_n, _d, _d, _d)
_sig("o_candidate_sigma", _d, _d, _i, _d)
_sig("o_power_for_sigma", _d, _d, _i, _d)
_sig("o_median", _f, _fp, _i)
_sig("o_invsqrtf", _f, _f)
_sig("o_cfft_f", None, _cf, C.c_long, _i)
_sig("o_cfft_d", None, _cd, C.c_long, _i)
_sig("o_realfft_f", None, _fp, C.c_long, _i)
_sig("o_deredden", C.c_long, _cf, C.c_long, _i, _i, _d, _d)
_sig("o_create_accelobs", _i, C.POINTER(OObs))
_sig("o_accelsearch", _i, C.POINTER(OObs), _cf, _i, _i, C.POINTER(OResult))
_sig("o_time_cfft_ns", C.c_double, C.c_long, _i)
_sig("o_accelsearch_range", _i, C.POINTER(OObs), _cf, _i, _i, _i, C.POINTER(OResult))
_sig("o_free_result", None, C.POINTER(OResult))
_sig("o_set_threads", None, _i)
_sig("o_result_ncands", _i, C.POINTER(OResult))
_sig("o_result_cands", None, C.POINTER(OResult), _dp)
_sig("o_prepare_dat", _ll, _fp, _ll, _d)
_sig("o_create_subharminfos", _p, C.POINTER(OObs), _i)
_sig("o_free_subharminfos", None, C.POINTER(OObs), _p)
_sig("o_subharm_fderivs_vol", C.POINTER(OFfdot), _i, _i, _d, _d, _p, C.POINTER(OObs), _cf, _i)
_sig("o_free_ffdot", None, C.POINTER(OFfdot))
_sig("o_block_norm", _d, C.POINTER(OObs), _cf, _ll, _i)


# ---------------------------------------------------------------- helpers
def fresnl(x):
    s, c = _d(), _d()
    lib.o_fresnl(float(x), C.byref(s), C.byref(c))
    return s.value, c.value


def gen_r_response(roffset, numbetween, numkern):
    out = np.zeros(numkern, np.complex128)
    lib.o_gen_r_response(roffset, numbetween, numkern, out)
    return out


def gen_z_response(roffset, numbetween, z, numkern):
    out = np.zeros(numkern, np.complex128)
    lib.o_gen_z_response(roffset, numbetween, z, numkern, out)
    return out


def gen_w_response(roffset, numbetween, z, w, numkern, fp64=False):
    if fp64:
        out = np.zeros(numkern, np.complex128)
        lib.o_gen_w_response(roffset, numbetween, z, w, numkern, None,
                             out.ctypes.data_as(_p))
    else:
        out = np.zeros(numkern, np.complex64)
        lib.o_gen_w_response(roffset, numbetween, z, w, numkern,
                             out.ctypes.data_as(_p), None)
    return out


def cfft(x, isign=-1):
    x = np.ascontiguousarray(x)
    y = x.copy()
    if y.dtype == np.complex64:
        lib.o_cfft_f(y, len(y), isign)
    else:
        y = y.astype(np.complex128)
        lib.o_cfft_d(y, len(y), isign)
    return y


def realfft(x, isign=-1):
    y = np.ascontiguousarray(x, dtype=np.float32).copy()
    lib.o_realfft_f(y, len(y), isign)
    return y


def median(x):
    y = np.ascontiguousarray(x, dtype=np.float32).copy()
    return float(lib.o_median(y, len(y)))


def deredden(fft, startwidth, endwidth, endfreq, T):
    y = np.ascontiguousarray(fft, dtype=np.complex64).copy()
    n = lib.o_deredden(y, len(y), startwidth, endwidth, endfreq, T)
    return y, n


def prepare_dat(series, dt):
    y = np.ascontiguousarray(series, dtype=np.float32).copy()
    nb = lib.o_prepare_dat(y, len(y), dt)
    return y.view(np.complex64)[:nb].copy()


def make_obs(numbins, N, dt, zmax=200, numharm=8, sigma=2.0, wmax=None, flo=1.0,
             fhi=10000.0, inmem=False, force_standard=False, lobin=0, dat_input=True,
             rlo=None, rhi=None, nph=0.0, locpow=False):
    """Fill an o_obs the way accelsearch's command line does (flo/fhi default
    to 'given', clig/accelsearch_cmd.cli) and run o_create_accelobs."""
    o = OObs()
    o.N, o.dt, o.numbins = float(N), float(dt), int(numbins)
    o.numharm, o.zmax, o.sigma = int(numharm), int(zmax), float(sigma)
    if wmax is not None:
        o.wmax, o.wmaxP = int(wmax), 1
    if rlo is not None:
        o.floP, o.rloP, o.rlo_arg = 0, 1, float(rlo)
    else:
        o.floP, o.flo = 1, float(flo)
    if rhi is not None:
        o.fhiP, o.rhiP, o.rhi_arg = 0, 1, float(rhi)
    else:
        o.fhiP, o.fhi = 1, float(fhi)
    o.inmemP = int(inmem)
    o.force_standard = int(force_standard)
    o.lobin_arg = int(lobin)
    o.dat_input = int(dat_input)
    o.nph = float(nph)
    o.norm_type = 1 if locpow else 0
    rc = lib.o_create_accelobs(C.byref(o))
    if rc:
        raise ValueError(f"o_create_accelobs failed ({rc})")
    return o


def accelsearch(obs, fft, fp64=False, max_blocks=-1, first_block=0):
    """Run the reference block loop over blocks [first_block, first_block +
    max_blocks) (max_blocks < 0: to the end); returns (hits structured
    array, cands, blocks run)."""
    fft = np.ascontiguousarray(fft, dtype=np.complex64)
    res = OResult()
    lib.o_accelsearch_range(C.byref(obs), fft, int(fp64), int(first_block), int(max_blocks),
                            C.byref(res))
    n = res.nhits
    hits = np.zeros(n, dtype=[("block", "i4"), ("stage", "i4"), ("wi", "i4"),
                              ("zi", "i4"), ("ri", "i4"), ("pow", "f4"), ("sigma", "f4")])
    if n:
        raw = np.ctypeslib.as_array(C.cast(res.hits, C.POINTER(C.c_byte)),
                                    shape=(n * C.sizeof(OHit),))
        hits = np.frombuffer(raw.tobytes(), dtype=hits.dtype).copy()
    nc = lib.o_result_ncands(C.byref(res))
    cands = np.zeros((nc, 7))
    if nc:
        lib.o_result_cands(C.byref(res), cands)
    nblocks = res.nblocks
    lib.o_free_result(C.byref(res))
    return hits, cands[:, :6], nblocks


class SubharmBanks:
    """The reference's template banks (create_subharminfos) for one obs."""

    def __init__(self, obs, fp64=False):
        self.obs = obs
        self.ptr = lib.o_create_subharminfos(C.byref(obs), int(fp64))
        self.fp64 = fp64
        self.shi_size = 7 * 4 + 4 + 3 * 8  # int fields + pad + 3 pointers

    def shi(self, stage, harm):
        """Pointer to shis[stage][harm-1]."""
        arr = C.cast(self.ptr, C.POINTER(C.c_void_p))
        base = arr[stage]
        return C.c_void_p(base + (harm - 1) * self.shi_size)

    def close(self):
        if self.ptr:
            lib.o_free_subharminfos(C.byref(self.obs), self.ptr)
            self.ptr = None


def subharm_fderivs_vol(numharm, harmnum, fullrlo, fullrhi, banks, obs, fft, fp64=False):
    """One block's plane [numws][numzs][numrs] plus its rlo/zlo/wlo."""
    fft = np.ascontiguousarray(fft, dtype=np.complex64)
    stage = int(np.log2(numharm))
    p = lib.o_subharm_fderivs_vol(numharm, harmnum, fullrlo, fullrhi,
                                  banks.shi(stage, harmnum), C.byref(obs), fft, int(fp64))
    f = p.contents
    n = f.numws * f.numzs * f.numrs
    src = f.dpowers if fp64 else f.powers
    arr = np.ctypeslib.as_array(src, shape=(n,)).copy().reshape(f.numws, f.numzs, f.numrs)
    meta = dict(rlo=f.rlo, zlo=f.zlo, wlo=f.wlo, numrs=f.numrs, numzs=f.numzs, numws=f.numws)
    lib.o_free_ffdot(p)
    return arr, meta


def block_norm(obs, fft, lobin, numdata):
    fft = np.ascontiguousarray(fft, dtype=np.complex64)
    return lib.o_block_norm(C.byref(obs), fft, int(lobin), int(numdata))


class OCand(C.Structure):
    pass


OCand._fields_ = [("power", _f), ("sigma", _f), ("numharm", _i), ("r", _d), ("z", _d),
                  ("w", _d), ("next", C.POINTER(OCand))]
_sig("o_insert_new_accelcand", C.POINTER(OCand), C.POINTER(OCand), _f, _f, _i, _d, _d, _d,
     C.POINTER(_i))
_sig("o_sort_accelcands", C.POINTER(OCand), C.POINTER(OCand))
_sig("o_eliminate_harmonics", C.POINTER(OCand), C.POINTER(OCand), C.POINTER(_i))
_sig("o_free_cands", None, C.POINTER(OCand))


class CandList:
    """The reference's candidate list (GSList of accelcand) in the oracle."""

    def __init__(self):
        self.head = C.POINTER(OCand)()

    def insert(self, power, sigma, numharm, r, z=0.0, w=0.0):
        added = _i(0)
        self.head = lib.o_insert_new_accelcand(self.head, power, sigma, numharm, r, z, w,
                                               C.byref(added))
        return added.value

    def _items(self):
        out, p = [], self.head
        while p:
            out.append(p.contents)
            p = p.contents.next
        return out

    def rs(self):
        return [c.r for c in self._items()]

    def sigmas(self):
        return [c.sigma for c in self._items()]

    def sort(self):
        self.head = lib.o_sort_accelcands(self.head)

    def eliminate(self):
        n = _i(len(self._items()))
        self.head = lib.o_eliminate_harmonics(self.head, C.byref(n))
        return n.value

    def free(self):
        lib.o_free_cands(self.head)
        self.head = C.POINTER(OCand)()


# ---------------------------------------------------------------- refinement
class ORDerivs(C.Structure):
    _fields_ = [(n, _d) for n in ("pow", "phs", "dpow", "dphs", "d2pow", "d2phs", "locpow")]


class OFourierProps(C.Structure):
    _fields_ = [("r", _d), ("rerr", _f), ("z", _d), ("zerr", _f), ("w", _d), ("werr", _f)] + \
               [(n, _f) for n in ("pow", "powerr", "sig", "rawpow", "phs", "phserr", "cen",
                                  "cenerr", "pur", "purerr", "locpow")]


class ORzwErrs(C.Structure):
    _fields_ = [(n, _d) for n in ("p", "pd", "pdd", "f", "fd", "fdd")] + \
               [(n, _f) for n in ("perr", "pderr", "pdderr", "ferr", "fderr", "fdderr")]


_sig("o_rz_interp", None, _cf, C.c_long, _d, _d, _i, C.POINTER(_f), C.POINTER(_f))
_sig("o_rzw_interp", None, _cf, C.c_long, _d, _d, _d, _i, C.POINTER(_f), C.POINTER(_f))
_sig("o_localpower3d", _d, _cf, C.c_long, _d, _d, _d)
_sig("o_derivs3d", None, _cf, C.c_long, _d, _d, _d, _d, C.POINTER(ORDerivs))
_sig("o_calc_props", _i, C.POINTER(ORDerivs), _d, _d, _d, C.POINTER(OFourierProps))
_sig("o_calc_rzwerrs", _i, C.POINTER(OFourierProps), _d, C.POINTER(ORzwErrs))
_sig("o_max_rz_arr", _d, _cf, C.c_long, _d, _d, C.POINTER(_d), C.POINTER(_d),
     C.POINTER(ORDerivs))
_sig("o_max_rz_arr_harmonics", C.c_long, _cf, C.c_long, _i, _d, _d, C.POINTER(_d),
     C.POINTER(_d), C.POINTER(ORDerivs), _dp)
_sig("o_max_rzw_arr_harmonics", None, _cf, C.c_long, _i, _d, _d, _d, C.POINTER(_d),
     C.POINTER(_d), C.POINTER(_d), C.POINTER(ORDerivs), _dp)
_sig("o_max_rzw_arr", _d, _cf, C.c_long, _d, _d, _d, C.POINTER(_d), C.POINTER(_d),
     C.POINTER(_d), C.POINTER(ORDerivs))


def _c64(data):
    return np.ascontiguousarray(data, dtype=np.complex64)


def rz_interp(data, r, z, khw):
    d = _c64(data)
    a, b = _f(), _f()
    lib.o_rz_interp(d, len(d), r, z, khw, C.byref(a), C.byref(b))
    return complex(a.value, b.value)


def rzw_interp(data, r, z, w, khw):
    d = _c64(data)
    a, b = _f(), _f()
    lib.o_rzw_interp(d, len(d), r, z, w, khw, C.byref(a), C.byref(b))
    return complex(a.value, b.value)


def localpower3d(data, r, z, w=0.0):
    d = _c64(data)
    return lib.o_localpower3d(d, len(d), r, z, w)


def derivs3d(data, r, z, w, locpow):
    d = _c64(data)
    out = ORDerivs()
    lib.o_derivs3d(d, len(d), r, z, w, locpow, C.byref(out))
    return out


def calc_props(rd, r, z, w):
    out = OFourierProps()
    if lib.o_calc_props(C.byref(rd), r, z, w, C.byref(out)):
        raise ValueError("calc_props: power or local power out of range")
    return out


def calc_rzwerrs(props, T):
    out = ORzwErrs()
    if lib.o_calc_rzwerrs(C.byref(props), T, C.byref(out)):
        raise ValueError("calc_rzwerrs: T out of range")
    return out


def max_rz_arr(data, r, z):
    """[maxpow, rmax, zmax, rderivs] as src/maximize_rz.c:103-140."""
    d = _c64(data)
    ro, zo, rd = _d(), _d(), ORDerivs()
    p = lib.o_max_rz_arr(d, len(d), r, z, C.byref(ro), C.byref(zo), C.byref(rd))
    return p, ro.value, zo.value, rd


def max_rz_arr_harmonics(data, r, z, numharm):
    """[rmax, zmax, [rderivs], powers, nevals] as src/maximize_rz.c:171-228."""
    d = _c64(data)
    ro, zo = _d(), _d()
    rds = (ORDerivs * numharm)()
    pw = np.zeros(numharm)
    nev = lib.o_max_rz_arr_harmonics(d, len(d), numharm, r, z, C.byref(ro), C.byref(zo),
                                     rds, pw)
    return ro.value, zo.value, list(rds), pw, nev


def max_rzw_arr(data, r, z, w):
    """[maxpow, rmax, zmax, wmax, rderivs] as src/maximize_rzw.c:3-55."""
    d = _c64(data)
    ro, zo, wo, rd = _d(), _d(), _d(), ORDerivs()
    p = lib.o_max_rzw_arr(d, len(d), r, z, w, C.byref(ro), C.byref(zo), C.byref(wo),
                          C.byref(rd))
    return p, ro.value, zo.value, wo.value, rd


def max_rzw_arr_harmonics(data, r, z, w, numharm):
    """[rmax, zmax, wmax, [rderivs], powers] as src/maximize_rzw.c:84-191."""
    d = _c64(data)
    ro, zo, wo = _d(), _d(), _d()
    rds = (ORDerivs * numharm)()
    pw = np.zeros(numharm)
    lib.o_max_rzw_arr_harmonics(d, len(d), numharm, r, z, w, C.byref(ro), C.byref(zo),
                                C.byref(wo), rds, pw)
    return ro.value, zo.value, wo.value, list(rds), pw


# ---------------------------------------------------------------- de-dispersion
_sig("o_delay_from_dm", _d, _d, _d)
_sig("o_dedisp_plan", _d, _i, _d, _d, _d, _i, _i, _i, _i, _d, _d, C.POINTER(_i),
     C.POINTER(_i), C.POINTER(_i))
_sig("o_prepsubband", _ll, C.POINTER(_f), _ll, _i, _d, _d, _d, _i, _i, _i, _i, _d, _d, _ll,
     C.POINTER(_f), C.POINTER(_d))


def dedisp_plan(numchan, dt, lofreq, chanwidth, spectra_per_subint, nsub, downsamp, numdms,
                lodm, dmstep):
    """(avgdm, idispdt[numchan], offsets[numdms][nsub], blocksperread) as
    src/prepsubband.c:255-306, 364-387 compute them."""
    idd = (_i * numchan)()
    off = (_i * (numdms * nsub))()
    bpr = _i()
    avg = lib.o_dedisp_plan(numchan, dt, lofreq, chanwidth, spectra_per_subint, nsub, downsamp,
                            numdms, lodm, dmstep, idd, off, C.byref(bpr))
    return avg, np.array(idd[:]), np.array(off[:]).reshape(numdms, nsub), bpr.value


def prepsubband(raw, dt, lofreq, chanwidth, spectra_per_subint, nsub, downsamp, numdms, lodm,
                dmstep, numout):
    """prepsubband -nobary -noclip over raw[nspec][numchan] (float32,
    ascending frequency): (out[numdms][numout], points written, pad value)."""
    raw = np.ascontiguousarray(raw, dtype=np.float32)
    nspec, numchan = raw.shape
    out = np.zeros((numdms, numout), dtype=np.float32)
    pad = _d()
    n = lib.o_prepsubband(raw.ctypes.data_as(C.POINTER(_f)), nspec, numchan, dt, lofreq,
                          chanwidth, spectra_per_subint, nsub, downsamp, numdms, lodm, dmstep,
                          numout, out.ctypes.data_as(C.POINTER(_f)), C.byref(pad))
    if n < 0:
        raise ValueError("prepsubband: delays exceed the working blocks")
    return out, n, pad.value


_sig("o_corr_loc_pow", None, C.POINTER(_f), _i, C.POINTER(_f))


def corr_loc_pow(powers):
    p = np.ascontiguousarray(powers, dtype=np.float32)
    out = np.zeros_like(p)
    lib.o_corr_loc_pow(p.ctypes.data_as(C.POINTER(_f)), len(p), out.ctypes.data_as(C.POINTER(_f)))
    return out


class OFoldStats(C.Structure):
    _fields_ = [(n, _d) for n in ("numdata", "data_avg", "data_var", "numprof", "prof_avg",
                                  "prof_var", "redchi")]


_sig("o_cand_fold_params", None, _d, _d, _d, _d, C.POINTER(_d), C.POINTER(_d), C.POINTER(_d))
_sig("o_fold_dat", _ll, C.POINTER(_f), _ll, _d, _d, _d, _d, _i, _i, C.POINTER(_d),
     C.POINTER(OFoldStats))


_sig("o_fold_search", _i, C.POINTER(_d), C.POINTER(OFoldStats), _i, _i, _d, _d, _d, _d,
     C.POINTER(_d), _i, _i, _i, _i, C.POINTER(_f), C.POINTER(_d))


def fold_search(profs, stats, f, fd, fdd, T, parttimes, npfact=2, pstep=2, pdstep=4,
                searchpdd=False):
    """prepfold's p / p-dot search over the sub-integrations: (ppdot
    [numtrials][numtrials] of reduced chi-squared -- [numtrials]^3, indexed
    [ipdd][ipd][ip], with searchpdd -- and a dict of the best trial)."""
    profs = np.ascontiguousarray(profs, dtype=np.float64)
    npart, proflen = profs.shape
    st = (OFoldStats * npart)()
    for s_, d in zip(st, stats):
        for k, _ in OFoldStats._fields_:
            setattr(s_, k, d[k])
    nt = 2 * npfact * proflen + 1
    pp = np.zeros((nt, nt, nt) if searchpdd else (nt, nt), dtype=np.float32)
    best = np.zeros(7)
    pt = np.ascontiguousarray(parttimes, dtype=np.float64)
    lib.o_fold_search(profs.ctypes.data_as(C.POINTER(_d)), st, npart, proflen, f, fd, fdd, T,
                      pt.ctypes.data_as(C.POINTER(_d)), npfact, pstep, pdstep, int(searchpdd),
                      pp.ctypes.data_as(C.POINTER(_f)), best.ctypes.data_as(C.POINTER(_d)))
    return pp, dict(ip=int(best[0]), ipd=int(best[1]), redchi=best[2], p=best[3], pd=best[4],
                    pdd=best[5], ipdd=int(best[6]))


def cand_fold_params(r, z, w, T):
    """prepfold's start f, fdot, fdotdot of a candidate's average r, z, w."""
    f, fd, fdd = _d(), _d(), _d()
    lib.o_cand_fold_params(r, z, w, T, C.byref(f), C.byref(fd), C.byref(fdd))
    return f.value, fd.value, fdd.value


def fold_dat(series, dt, f, fd, fdd, npart=64, proflen=64):
    """prepfold's fold of a .dat: (rawfolds [npart][proflen], per-part
    stats as dicts, points folded)."""
    x = np.ascontiguousarray(series, dtype=np.float32)
    prof = np.zeros((npart, proflen), dtype=np.float64)
    st = (OFoldStats * npart)()
    n = lib.o_fold_dat(x.ctypes.data_as(C.POINTER(_f)), len(x), dt, f, fd, fdd, npart, proflen,
                       prof.ctypes.data_as(C.POINTER(_d)), st)
    stats = [{k: getattr(s, k) for k, _ in OFoldStats._fields_} for s in st]
    return prof, stats, int(n)
