"""ctypes view of the accel.h surface of libpresto_accel.so
(include/accel_surface.h): the reference's structs with their member order,
the accel.h functions, and `drive()`, the reference driver's block loop
(src/accelsearch.c:82-290) written against them.

Like presto_amd.accel, it loads only the built HIP library (no CPU
fallback)."""
import ctypes as C
import os

from presto_amd.accel import lib, FourierProps, RDerivs, AccelError

_d, _i, _f, _ll, _p, _cp = C.c_double, C.c_int, C.c_float, C.c_longlong, C.c_void_p, C.c_char_p

MAXNUMONOFF = 40


class GSList(C.Structure):
    pass


GSList._fields_ = [("data", _p), ("next", C.POINTER(GSList))]


class InfoData(C.Structure):
    _fields_ = [(n, _d) for n in ("ra_s", "dec_s", "N", "dt", "fov", "mjd_f", "dm", "freq",
                                  "freqband", "chan_wid", "wavelen", "waveband", "energy",
                                  "energyband")] + \
               [("onoff", _d * (2 * MAXNUMONOFF))] + \
               [(n, _i) for n in ("num_chan", "mjd_i", "ra_h", "ra_m", "dec_d", "dec_m", "bary",
                                  "numonoff")] + \
               [("notes", C.c_char * 500), ("name", C.c_char * 200), ("object", C.c_char * 100),
                ("instrument", C.c_char * 100), ("observer", C.c_char * 100),
                ("analyzer", C.c_char * 100), ("telescope", C.c_char * 40),
                ("band", C.c_char * 40), ("filt", C.c_char * 7)]


def _opt(name, typ):
    return [(name + "P", _i), (name, typ), (name + "C", _i)]


class Cmdline(C.Structure):
    _fields_ = (_opt("ncpus", _i) + _opt("lobin", _i) + _opt("numharm", _i) + _opt("zmax", _i)
                + _opt("wmax", _i) + _opt("sigma", _f) + _opt("rlo", _d) + _opt("rhi", _d)
                + _opt("flo", _d) + _opt("fhi", _d)
                + [("inmemP", _i), ("photonP", _i), ("medianP", _i), ("locpowP", _i)]
                + _opt("zaplist", _cp) + _opt("baryv", _d)
                + [("otheroptP", _i), ("noharmpolishP", _i), ("noharmremoveP", _i),
                   ("argc", _i), ("argv", C.POINTER(_cp)), ("full_cmd_line", _cp)])


class AccelObs(C.Structure):
    _fields_ = ([(n, _ll) for n in ("N", "numbins", "lobin", "highestbin")]
                + [(n, _i) for n in ("maxkernlen", "corr_uselen", "fftlen", "numharmstages",
                                     "numz", "numw", "numbetween", "numzap", "dat_input",
                                     "mmap_file", "inmem", "norm_type")]
                + [(n, _d) for n in ("dt", "T", "rlo", "rhi", "dr", "zlo", "zhi", "dz", "wlo",
                                     "whi", "dw", "baryv")]
                + [("nph", _f), ("sigma", _f), ("powcut", C.POINTER(_f)),
                   ("ffdotplane", C.POINTER(_f)), ("lobins", C.POINTER(_d)),
                   ("hibins", C.POINTER(_d)), ("numindep", C.POINTER(_ll)),
                   ("fftfile", _p), ("workfile", _p), ("fft", _p), ("rootfilenm", _cp),
                   ("candnm", _cp), ("accelnm", _cp), ("workfilenm", _cp),
                   ("use_harmonic_polishing", _i)])


class AccelCand(C.Structure):
    _fields_ = [("power", _f), ("sigma", _f), ("numharm", _i), ("r", _d), ("z", _d),
                ("w", _d), ("pows", C.POINTER(_d)), ("hirs", C.POINTER(_d)),
                ("hizs", C.POINTER(_d)), ("hiws", C.POINTER(_d)),
                ("derivs", C.POINTER(RDerivs))]


class Kernel(C.Structure):
    _fields_ = [(n, _i) for n in ("z", "w", "fftlen", "numgoodbins", "numbetween",
                                  "kern_half_width")] + [("data", _p)]


class SubharmInfo(C.Structure):
    _fields_ = [(n, _i) for n in ("numharm", "harmnum", "zmax", "wmax", "numkern_zdim",
                                  "numkern_wdim", "numkern")] + \
               [("kern", C.POINTER(C.POINTER(Kernel))), ("rinds", C.POINTER(C.c_ushort)),
                ("zinds", C.POINTER(C.c_ushort))]


class FFDotPows(C.Structure):
    _fields_ = [("rlo", _ll), ("zlo", _i), ("wlo", _i), ("numrs", _i), ("numzs", _i),
                ("numws", _i), ("powers", C.POINTER(C.POINTER(C.POINTER(_f)))),
                ("rinds", C.POINTER(C.c_ushort)), ("zinds", C.POINTER(C.c_ushort))]


LAYOUT_TYPES = {"accelobs": AccelObs, "accelcand": AccelCand, "kernel": Kernel,
                "subharminfo": SubharmInfo, "ffdotpows": FFDotPows, "infodata": InfoData,
                "Cmdline": Cmdline, "fourierprops": FourierProps, "rderivs": RDerivs}

_OP, _SP, _FP, _LP = (C.POINTER(AccelObs), C.POINTER(C.POINTER(SubharmInfo)),
                      C.POINTER(FFDotPows), C.POINTER(GSList))


def _sig(name, res, *args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = list(args)


_sig("create_accelobs", None, _OP, C.POINTER(InfoData), C.POINTER(Cmdline), _i)
_sig("free_accelobs", None, _OP)
_sig("create_subharminfos", _SP, _OP)
_sig("free_subharminfos", None, _OP, _SP)
_sig("subharm_fderivs_vol", _FP, _i, _i, _d, _d, C.POINTER(SubharmInfo), _OP)
_sig("copy_ffdotpows", _FP, _FP)
_sig("free_ffdotpows", None, _FP)
_sig("add_ffdotpows", None, _FP, _FP, _i, _i)
_sig("add_ffdotpows_ptrs", None, _FP, _FP, _i, _i)
_sig("fund_to_ffdotplane", None, _FP, _OP)
_sig("fund_to_ffdotplane_trans", None, _FP, _OP)
_sig("inmem_add_ffdotpows", None, _FP, _OP, _i, _i)
_sig("inmem_add_ffdotpows_trans", None, _FP, _OP, _i, _i)
_sig("search_ffdotpows", _LP, _FP, _i, _OP, _LP)
_sig("sort_accelcands", _LP, _LP)
_sig("eliminate_harmonics", _LP, _LP, C.POINTER(_i))
_sig("optimize_accelcand", None, C.POINTER(AccelCand), _OP)
_sig("output_fundamentals", None, C.POINTER(FourierProps), _LP, _OP, C.POINTER(InfoData))
_sig("output_harmonics", None, _LP, _OP, C.POINTER(InfoData))
_sig("free_accelcand", None, _p, _p)
_sig("print_accelcand", None, _p, _p)
_sig("get_fourier_amplitudes", _p, _ll, _i, _OP)
_sig("accel_surface_set_exit", _i, _i)
_sig("accel_surface_failed", _i)
_sig("accel_surface_clear", None)
_sig("accel_surface_set_device", _i, _i)
_sig("accel_ffdotpows_sync", _i, _FP)
_sig("accel_gslist_free", None, _LP)
_sig("accel_surface_layout", _i, _cp, C.POINTER(_ll), _i)
_sig("accel_readinf", _i, C.POINTER(InfoData), _cp)
_sig("accel_writeinf", _i, C.POINTER(InfoData))
_sig("accel_calc_props", _i, C.POINTER(RDerivs), _d, _d, _d, C.POINTER(FourierProps))
_sig("accel_last_error", _cp)

EXPORTED = ["create_accelobs", "free_accelobs", "create_subharminfos", "free_subharminfos",
            "subharm_fderivs_vol", "copy_ffdotpows", "free_ffdotpows", "add_ffdotpows",
            "add_ffdotpows_ptrs", "fund_to_ffdotplane", "fund_to_ffdotplane_trans",
            "inmem_add_ffdotpows", "inmem_add_ffdotpows_trans", "search_ffdotpows",
            "sort_accelcands", "eliminate_harmonics", "optimize_accelcand",
            "output_fundamentals", "output_harmonics", "free_accelcand", "print_accelcand",
            "get_fourier_amplitudes", "accel_surface_set_exit", "accel_surface_failed",
            "accel_surface_clear", "accel_surface_set_device", "accel_ffdotpows_sync",
            "accel_gslist_free", "accel_surface_layout", "accel_readinf", "accel_writeinf"]

# failures raise in Python instead of calling exit()
lib.accel_surface_set_exit(0)


def _check():
    if lib.accel_surface_failed():
        lib.accel_surface_clear()
        raise AccelError(lib.accel_last_error().decode())


def layout(name):
    out = (_ll * 64)()
    n = lib.accel_surface_layout(name.encode(), out, 64)
    if n < 0:
        raise KeyError(name)
    return list(out[:n])


def readinf(root):
    d = InfoData()
    if lib.accel_readinf(C.byref(d), root.encode()):
        raise AccelError(lib.accel_last_error().decode())
    return d


def writeinf(d):
    if lib.accel_writeinf(C.byref(d)):
        raise AccelError(lib.accel_last_error().decode())


def write_inf(root, N, dt, telescope="None (Artificial Data Set)"):
    """A .inf in the layout of src/ioinf.c writeinf (artificial data)."""
    with open(root + ".inf", "w") as f:
        f.write(f" Data file name without suffix          =  {os.path.basename(root)}\n")
        f.write(f" Telescope used                         =  {telescope}\n")
        f.write(f" Number of bins in the time series      =  {N:<11.0f}\n")
        f.write(f" Width of each time series bin (sec)    =  {dt:.15g}\n")
        f.write(" Any breaks in the data? (1 yes, 0 no)  =  0\n")
        f.write(" Data analyzed by                       =  synthetic\n")
        f.write(" Any additional notes:\n    synthetic data\n\n")


def make_cmdline(infile, zmax=200, numharm=8, sigma=2.0, wmax=None, flo=1.0, fhi=10000.0,
                 inmem=False, noharmremove=False, lobin=0, rlo=None, rhi=None, photon=False,
                 noharmpolish=False, locpow=False):
    """Cmdline with the clig defaults of clig/accelsearch_cmd.cli:13-48.
    rlo / rhi clear floP / fhiP, which the command line itself never does
    (an embedding caller can)."""
    c = Cmdline()
    c.ncpusP, c.ncpus = 1, 1
    c.lobinP, c.lobin = 1, lobin
    c.numharmP, c.numharm = 1, numharm
    c.zmaxP, c.zmax = 1, zmax
    c.wmaxP, c.wmax = (1, wmax) if wmax is not None else (0, 0)
    c.sigmaP, c.sigma = 1, sigma
    c.floP, c.flo = 1, flo
    c.fhiP, c.fhi = 1, fhi
    c.rloP = c.rhiP = 0
    if rlo is not None:
        c.floP, c.rloP, c.rlo = 0, 1, float(rlo)
    if rhi is not None:
        c.fhiP, c.rhiP, c.rhi = 0, 1, float(rhi)
    c.inmemP = 1 if inmem else 0
    c.photonP = 1 if photon else 0
    c.locpowP = 1 if locpow else 0
    c.noharmpolishP = 1 if noharmpolish else 0
    c.noharmremoveP = 1 if noharmremove else 0
    c._argv = (_cp * 1)(infile.encode())
    c.argc, c.argv = 1, C.cast(c._argv, C.POINTER(_cp))
    return c


def _list(head):
    out = []
    p = head
    while p:
        out.append(C.cast(p.contents.data, C.POINTER(AccelCand)).contents)
        p = p.contents.next
    return out


def drive(infile, cmd, optimize=True, outputs=True, stop_after_blocks=None):
    """The reference driver (src/accelsearch.c:82-290) against the surface:
    returns (obs fields, candidate dicts).  Writes the ACCEL and .cand
    files next to infile when outputs is set."""
    obs, idata = AccelObs(), InfoData()
    lib.create_accelobs(C.byref(obs), C.byref(idata), C.byref(cmd), 1)
    _check()
    rstep = obs.corr_uselen * 0.5
    shis = lib.create_subharminfos(C.byref(obs))
    _check()
    cands = _LP()
    try:
        if obs.inmem:
            startr = 8.0
            while startr < obs.rlo:
                nextr = startr + rstep
                f = lib.subharm_fderivs_vol(1, 1, startr, nextr - 0.5, shis[0], C.byref(obs))
                _check()
                lib.fund_to_ffdotplane(f, C.byref(obs))
                lib.free_ffdotpows(f)
                _check()
                startr = nextr
        startr, nb = obs.rlo, 0
        while startr + rstep < obs.highestbin:
            if stop_after_blocks is not None and nb >= stop_after_blocks:
                break
            nextr = startr + rstep
            lastr = nextr - 0.5
            fund = lib.subharm_fderivs_vol(1, 1, startr, lastr, shis[0], C.byref(obs))
            _check()
            cands = lib.search_ffdotpows(fund, 1, C.byref(obs), cands)
            if obs.numharmstages > 1:
                if obs.inmem:
                    lib.fund_to_ffdotplane(fund, C.byref(obs))
                for stage in range(1, obs.numharmstages):
                    h = 1 << stage
                    for harm in range(1, h, 2):
                        if obs.inmem:
                            lib.inmem_add_ffdotpows(fund, C.byref(obs), h, harm)
                        else:
                            sub = lib.subharm_fderivs_vol(h, harm, startr, lastr,
                                                          C.byref(shis[stage][harm - 1]),
                                                          C.byref(obs))
                            _check()
                            lib.add_ffdotpows(fund, sub, h, harm)
                            lib.free_ffdotpows(sub)
                    cands = lib.search_ffdotpows(fund, h, C.byref(obs), cands)
            lib.free_ffdotpows(fund)
            _check()
            startr = nextr
            nb += 1
        lib.free_subharminfos(C.byref(obs), shis)
        shis = None
        numcands = C.c_int(0)
        p = cands
        while p:
            numcands.value += 1
            p = p.contents.next
        if numcands.value:
            cands = lib.sort_accelcands(cands)
            if cmd.numharm > 1 and not cmd.noharmremoveP:
                lib.eliminate_harmonics(cands, C.byref(numcands))
            if optimize:
                for c in _list(cands):
                    lib.optimize_accelcand(C.byref(c), C.byref(obs))
                    _check()
            if optimize and outputs:
                lst = _list(cands)
                props = (FourierProps * len(lst))()
                for i, c in enumerate(lst):
                    lib.accel_calc_props(c.derivs, c.r, c.z, c.w, C.byref(props[i]))
                    props[i].rerr = 0.5 / c.numharm
                    props[i].zerr = 2.0 / c.numharm
                    props[i].werr = 20.0 / c.numharm
                lib.output_fundamentals(props, cands, C.byref(obs), C.byref(idata))
                lib.output_harmonics(cands, C.byref(obs), C.byref(idata))
                _check()
                with open(obs.candnm.decode(), "wb") as f:
                    f.write(bytes(props))
        out = [dict(power=c.power, sigma=c.sigma, numharm=c.numharm, r=c.r, z=c.z, w=c.w,
                    pows=[c.pows[k] for k in range(c.numharm)] if c.pows else None)
               for c in _list(cands)]
        info = dict(inmem=obs.inmem, numharmstages=obs.numharmstages, corr_uselen=obs.corr_uselen,
                    rlo=obs.rlo, highestbin=obs.highestbin, numz=obs.numz, numw=obs.numw,
                    accelnm=obs.accelnm.decode(), candnm=obs.candnm.decode(),
                    workfilenm=obs.workfilenm.decode())
        return info, out
    finally:
        if shis:
            lib.free_subharminfos(C.byref(obs), shis)
        p = cands
        while p:
            lib.free_accelcand(p.contents.data, None)
            p = p.contents.next
        lib.accel_gslist_free(cands)
        lib.free_accelobs(C.byref(obs))
