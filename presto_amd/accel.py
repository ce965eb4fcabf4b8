"""ctypes binding of libpresto_accel.so (include/presto_accel.h).

The binding only loads the built HIP library; there is no CPU fallback.  A
missing library raises at import, and every entry point reports the C
library's error text on failure."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpresto_accel.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built: run __graft_entry__.build() "
                      "or python presto_amd/build.py")
lib = C.CDLL(LIB_PATH)

_d, _i, _f, _ll, _p = C.c_double, C.c_int, C.c_float, C.c_longlong, C.c_void_p


class Params(C.Structure):
    _fields_ = [("N", _d), ("dt", _d), ("numbins", _ll), ("lobin", _ll),
                ("numharm", _i), ("zmax", _i), ("wmax", _i), ("wmaxP", _i),
                ("sigma", _f), ("flo", _d), ("fhi", _d), ("rlo", _d), ("rhi", _d),
                ("floP", _i), ("fhiP", _i), ("rloP", _i), ("rhiP", _i),
                ("inmem", _i), ("force_standard", _i), ("dat_input", _i), ("noharmremove", _i),
                ("nph", _d), ("noharmpolish", _i)]


class Geom(C.Structure):
    _fields_ = [("N", _ll), ("numbins", _ll), ("lobin", _ll), ("highestbin", _ll),
                ("maxkernlen", _i), ("corr_uselen", _i), ("fftlen", _i),
                ("numharmstages", _i), ("numz", _i), ("numw", _i), ("numbetween", _i),
                ("inmem", _i),
                ("dt", _d), ("T", _d), ("rlo", _d), ("rhi", _d), ("dr", _d), ("zlo", _d),
                ("zhi", _d), ("dz", _d), ("wlo", _d), ("whi", _d), ("dw", _d),
                ("sigma", _f), ("powcut", _f * 6), ("numindep", _ll * 6), ("nblocks", _i),
                ("nph", _d)]


class Cand(C.Structure):
    _fields_ = [("power", _f), ("sigma", _f), ("numharm", _i), ("pad", _i),
                ("r", _d), ("z", _d), ("w", _d)]


class Hit(C.Structure):
    _fields_ = [("block", _i), ("stage", _i), ("wi", _i), ("zi", _i), ("ii", _i),
                ("pow", _f)]


HIT_DTYPE = np.dtype([("block", "i4"), ("stage", "i4"), ("wi", "i4"), ("zi", "i4"),
                      ("ii", "i4"), ("pow", "f4")])
CAND_DTYPE = np.dtype([("power", "f4"), ("sigma", "f4"), ("numharm", "i4"), ("pad", "i4"),
                       ("r", "f8"), ("z", "f8"), ("w", "f8")])


def _sig(name, res, *args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = list(args)


_PP, _GP = C.POINTER(Params), C.POINTER(Geom)
_sig("accel_params_default", None, _PP)
_sig("accel_geom_create", _i, _PP, _GP)
_sig("accel_search_create", _p, _PP, _i)
_sig("accel_search_destroy", None, _p)
_sig("accel_search_geom", _GP, _p)
_sig("accel_search_set_spectrum", _i, _p, _p)
_sig("accel_search_set_device_spectrum", _i, _p, _p)
_sig("accel_search_set_series", _i, _p, _p, _ll)
_sig("accel_search_spectrum", _i, _p, _p, _ll)
_sig("accel_dered_blocks", _i, C.c_long, _i, _i, _d, _d, _p, _p, _i)
_sig("accel_search_run", _i, _p)
_sig("accel_search_set_txtcand", _i, _p, C.c_char_p)
_sig("accel_search_launch", _i, _p)
_sig("accel_search_collect", _i, _p)
_sig("accel_search_ncands", _i, _p)
_sig("accel_search_cands", _i, _p, _p, _i)
_sig("accel_search_nhits", _ll, _p)
_sig("accel_search_host_timings", _i, _p, _p)
_sig("accel_search_hits", _i, _p, _p, _ll)
_sig("accel_search_plane_row", _i, _p, _i, _i, _ll, _i, _p)
_sig("accel_search_nwindows", _i, _p)
_sig("accel_search_norms", _i, _p, _p, _i)
_sig("accel_search_timings", _i, _p, _p)
_sig("accel_search_tune", _i, _p, _i, _p, _p, _p)
_sig("accel_refine_set_first_wait", None, _i)
_sig("accel_search_corr_shape", _i, _p, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i),
     C.POINTER(_ll))
_sig("accel_search_plane_cells", _ll, _p)
_sig("accel_search_windows", _i, _p, _p, _p, _i)
_sig("accel_deredden", C.c_long, _p, C.c_long, _i, _i, _d, _d)
_sig("accel_prepare_dat_spectrum", C.c_long, _p, C.c_long, _d, _d)
_sig("accel_candidate_sigma", _d, _d, _i, _d)
_sig("accel_power_for_sigma", _d, _d, _i, _d)
_sig("accel_last_error", C.c_char_p)


class RDerivs(C.Structure):
    _fields_ = [(n, _d) for n in ("pow", "phs", "dpow", "dphs", "d2pow", "d2phs", "locpow")]


class FourierProps(C.Structure):
    _fields_ = [("r", _d), ("rerr", _f), ("z", _d), ("zerr", _f), ("w", _d), ("werr", _f)] + \
               [(n, _f) for n in ("pow", "powerr", "sig", "rawpow", "phs", "phserr", "cen",
                                  "cenerr", "pur", "purerr", "locpow")]


class RzwErrs(C.Structure):
    _fields_ = [(n, _d) for n in ("p", "pd", "pdd", "f", "fd", "fdd")] + \
               [(n, _f) for n in ("perr", "pderr", "pdderr", "ferr", "fderr", "fdderr")]


MAXHARM = 32
_sig("accel_search_optimize", _i, _p)
_sig("accel_search_cand_harmonics", _i, _p, _i, _p, _p, _p, _p, _p, _i)
_sig("accel_search_fund_props", _i, _p, _p, _i)
_sig("accel_search_write_outputs", _i, _p, C.c_char_p, C.c_char_p, C.c_char_p)
_sig("accel_max_rz_harmonics", _i, _p, _ll, _i, _p, _p, _p, _p, _p, _p, _p, _i)
_sig("accel_max_rzw_harmonics", _i, _p, _ll, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i)
_sig("accel_calc_props", _i, C.POINTER(RDerivs), _d, _d, _d, C.POINTER(FourierProps))
_sig("accel_calc_rzwerrs", _i, C.POINTER(FourierProps), _d, C.POINTER(RzwErrs))
_sig("accel_nice_output", _i, C.c_char_p, _d, _d, _i, _i)
_sig("accel_read_birdies", _i, C.c_char_p, _d, _d, _p, _p, _i)
_sig("accel_zap_birdies", _i, _p, _ll, _d, _p, _p, _i)
_sig("accel_mt19937_raw", _i, C.c_uint, _p, _i)
_sig("accel_realfft_forward", _i, _p, _ll, _p, _i)
_sig("accel_realfft_inverse", _i, _p, _ll, _p, _i)
_sig("accel_deredden_gpu", C.c_long, _p, _ll, _i, _i, _d, _d, _i)
_sig("accel_zapbirds_fft", _i, _p, _ll, _d, _p, _p, _i, _i)

EXPORTED = ["accel_params_default", "accel_geom_create", "accel_search_create",
            "accel_search_set_series", "accel_search_spectrum", "accel_dered_blocks",
            "accel_search_destroy", "accel_search_geom", "accel_search_set_spectrum",
            "accel_search_set_device_spectrum", "accel_search_run", "accel_search_launch",
            "accel_search_set_txtcand",
            "accel_search_collect", "accel_search_ncands", "accel_search_cands",
            "accel_search_nhits", "accel_search_hits", "accel_search_plane_row",
            "accel_search_nwindows", "accel_search_norms", "accel_search_timings",
            "accel_search_tune", "accel_refine_set_first_wait",
            "accel_search_host_timings",
            "accel_search_corr_shape", "accel_search_plane_cells",
            "accel_search_windows", "accel_deredden", "accel_prepare_dat_spectrum",
            "accel_candidate_sigma", "accel_power_for_sigma", "accel_search_optimize",
            "accel_search_cand_harmonics", "accel_search_fund_props",
            "accel_search_write_outputs", "accel_max_rz_harmonics",
            "accel_max_rzw_harmonics", "accel_calc_props",
            "accel_calc_rzwerrs", "accel_nice_output", "accel_read_birdies",
            "accel_zap_birdies", "accel_mt19937_raw", "accel_realfft_forward",
            "accel_realfft_inverse", "accel_deredden_gpu", "accel_zapbirds_fft",
            "accel_batch_create", "accel_batch_destroy", "accel_batch_handles",
            "accel_batch_run", "accel_search_footprint", "accel_template_cache_clear",
            "accel_server_create",
            "accel_server_submit", "accel_server_wait", "accel_server_cands",
            "accel_server_props", "accel_server_memory", "accel_server_destroy",
            "accel_server_timing",
            "accel_last_error", "accel_search_set_device_series", "accel_dedisp_plan",
            "accel_dedisperse_dev", "accel_dedisperse", "accel_device_alloc", "accel_device_free",
            "accel_device_upload", "accel_device_download", "accel_realfft_forward_noalloc",
            "accel_sift_create", "accel_sift_destroy", "accel_sift_add", "accel_sift_read",
            "accel_sift_parse_power", "accel_sift_reject", "accel_sift_remove_duplicates",
            "accel_sift_remove_dm_problems", "accel_sift_remove_harmonics",
            "accel_sift_sort_sigma", "accel_sift_extend", "accel_sift_count", "accel_sift_get",
            "accel_sift_write", "accel_sift_write_report", "accel_sift_write_summary",
            "accel_read_rzw_cand", "accel_cand_fold_params", "accel_fold_proflen", "accel_fold",
            "accel_fold_search"]


class AccelError(RuntimeError):
    pass


def _err():
    return AccelError(lib.accel_last_error().decode())


def make_params(numbins, N, dt, zmax=200, numharm=8, sigma=2.0, wmax=None, flo=1.0,
                fhi=10000.0, inmem=False, force_standard=False, lobin=0, dat_input=True,
                rlo=None, rhi=None, nph=0.0, noharmpolish=False):
    """accelsearch command-line semantics (flo/fhi default to given)."""
    p = Params()
    lib.accel_params_default(C.byref(p))
    p.N, p.dt, p.numbins, p.lobin = float(N), float(dt), int(numbins), int(lobin)
    p.numharm, p.zmax, p.sigma = int(numharm), int(zmax), float(sigma)
    if wmax is not None:
        p.wmax, p.wmaxP = int(wmax), 1
    if rlo is not None:
        p.floP, p.rloP, p.rlo = 0, 1, float(rlo)
    else:
        p.floP, p.flo = 1, float(flo)
    if rhi is not None:
        p.fhiP, p.rhiP, p.rhi = 0, 1, float(rhi)
    else:
        p.fhiP, p.fhi = 1, float(fhi)
    p.inmem, p.force_standard, p.dat_input = int(inmem), int(force_standard), int(dat_input)
    p.nph = float(nph)
    p.noharmpolish = int(noharmpolish)
    return p


def geometry(params):
    g = Geom()
    if lib.accel_geom_create(C.byref(params), C.byref(g)):
        raise _err()
    return g


class Search:
    """One accelsearch on the GPU: create (templates resident in HBM), set a
    spectrum, run, read candidates."""

    def __init__(self, params, device=0):
        self.params = params
        self.h = lib.accel_search_create(C.byref(params), int(device))
        if not self.h:
            raise _err()
        # a copy: the handle's own struct is freed with the handle
        self.geom = Geom.from_buffer_copy(lib.accel_search_geom(self.h).contents)

    def close(self):
        if self.h:
            lib.accel_search_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        self.close()

    def set_spectrum(self, fft):
        fft = np.ascontiguousarray(fft, dtype=np.complex64)
        if len(fft) != self.params.numbins:
            raise ValueError("spectrum length differs from params.numbins")
        if lib.accel_search_set_spectrum(self.h, fft.ctypes.data):
            raise _err()

    def set_series(self, x):
        """Time series in; realfft + deredden on the GPU."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        if lib.accel_search_set_series(self.h, x.ctypes.data, len(x)):
            raise _err()

    def spectrum(self):
        out = np.zeros(self.geom.numbins, dtype=np.complex64)
        if lib.accel_search_spectrum(self.h, out.ctypes.data, len(out)):
            raise _err()
        return out

    def set_device_series(self, dptr, n):
        """A time series already in device memory (e.g. one DM trial)."""
        if lib.accel_search_set_device_series(self.h, C.c_void_p(int(dptr)), int(n)):
            raise _err()

    def set_device_spectrum(self, dptr):
        if lib.accel_search_set_device_spectrum(self.h, C.c_void_p(int(dptr))):
            raise _err()

    def run(self):
        if lib.accel_search_run(self.h):
            raise _err()

    def set_txtcand(self, path):
        """Open (or with None close) the .txtcand work file of an .fft input."""
        if lib.accel_search_set_txtcand(self.h, None if path is None else str(path).encode()):
            raise _err()

    def launch(self):
        if lib.accel_search_launch(self.h):
            raise _err()

    def collect(self):
        if lib.accel_search_collect(self.h):
            raise _err()

    def candidates(self):
        n = lib.accel_search_ncands(self.h)
        out = np.zeros(n, dtype=CAND_DTYPE)
        if n:
            lib.accel_search_cands(self.h, out.ctypes.data, n)
        return out

    def hits(self):
        n = lib.accel_search_nhits(self.h)
        out = np.zeros(n, dtype=HIT_DTYPE)
        if n:
            lib.accel_search_hits(self.h, out.ctypes.data, n)
        return out

    def plane_row(self, wi, zi, hb0, n):
        out = np.zeros(n, dtype=np.float32)
        if lib.accel_search_plane_row(self.h, wi, zi, int(hb0), int(n), out.ctypes.data):
            raise _err()
        return out

    def norms(self):
        n = lib.accel_search_nwindows(self.h)
        out = np.zeros(max(n, 1), dtype=np.float64)
        lib.accel_search_norms(self.h, out.ctypes.data, n)
        return out[:n]

    def windows(self):
        n = lib.accel_search_nwindows(self.h)
        lo = np.zeros(max(n, 1), dtype=np.int64)
        nd = np.zeros(max(n, 1), dtype=np.int32)
        if lib.accel_search_windows(self.h, lo.ctypes.data, nd.ctypes.data, n) < 0:
            raise _err()
        return lo[:n], nd[:n]

    def timings(self):
        ms = np.zeros(5, dtype=np.float32)
        lib.accel_search_timings(self.h, ms.ctypes.data)
        return dict(norms=float(ms[0]), corr=float(ms[1]), halo=float(ms[2]),
                    harmsum=float(ms[3]), total=float(ms[4]))

    def host_timings(self):
        ms = np.zeros(4, dtype=np.float64)
        lib.accel_search_host_timings(self.h, ms.ctypes.data)
        return dict(launch=float(ms[0]), wait=float(ms[1]), copy=float(ms[2]),
                    replay=float(ms[3]))

    def tune(self, ncand=8):
        """Measured template-group size of the correlation (accel_search_tune):
        returns (groups kept, [(groups, ms), ...])."""
        g = np.zeros(ncand, dtype=np.int32)
        ms = np.zeros(ncand, dtype=np.float32)
        best = _i()
        n = lib.accel_search_tune(self.h, ncand, g.ctypes.data, ms.ctypes.data, C.byref(best))
        if n < 0:
            raise _err()
        return best.value, [(int(g[i]), float(ms[i])) for i in range(n)]

    def corr_shape(self):
        L, u, g, ns = _i(), _i(), _i(), _ll()
        lib.accel_search_corr_shape(self.h, C.byref(L), C.byref(u), C.byref(g), C.byref(ns))
        return dict(L=L.value, uselen=u.value, groups=g.value, nseg=ns.value)

    def plane_cells(self):
        return lib.accel_search_plane_cells(self.h)

    def optimize(self):
        """optimize_accelcand for every candidate (FP64, on the GPU)."""
        if lib.accel_search_optimize(self.h):
            raise _err()

    def harmonics(self, i):
        """(hirs, hizs, hiws, pows, [RDerivs]) of refined candidate i."""
        hr, hz, hw, pw = (np.zeros(MAXHARM) for _ in range(4))
        der = (RDerivs * MAXHARM)()
        n = lib.accel_search_cand_harmonics(self.h, i, hr.ctypes.data, hz.ctypes.data,
                                            hw.ctypes.data, pw.ctypes.data, der, MAXHARM)
        if n < 0:
            raise _err()
        return hr[:n], hz[:n], hw[:n], pw[:n], list(der)[:n]

    def fund_props(self):
        n = lib.accel_search_ncands(self.h)
        out = (FourierProps * max(n, 1))()
        if n and lib.accel_search_fund_props(self.h, out, n) != n:
            raise _err()
        return list(out)[:n]

    def write_outputs(self, accelnm, candnm=None, inf_path=None):
        enc = (lambda x: x.encode() if x is not None else None)
        if lib.accel_search_write_outputs(self.h, enc(accelnm), enc(candnm), enc(inf_path)):
            raise _err()


class BatchItem(C.Structure):
    _fields_ = [("params", Params), ("fft", _p), ("optimize", _i), ("accelnm", C.c_char_p),
                ("candnm", C.c_char_p), ("inf_path", C.c_char_p), ("cands", _p),
                ("maxcands", _i), ("ncands", _i), ("status", _i), ("gpu_ms", _f),
                ("d_series", _p), ("nseries", _ll)]


_sig("accel_batch_create", _p, _i)
_sig("accel_batch_destroy", None, _p)
_sig("accel_batch_handles", _i, _p)
_sig("accel_batch_run", _i, _p, _p, _i)


class Batch:
    """A pool of search handles on one device; run() pipelines many
    spectra through search, refinement and output."""

    def __init__(self, device=0):
        self.h = lib.accel_batch_create(int(device))

    def close(self):
        if self.h:
            lib.accel_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        self.close()

    def handles(self):
        return lib.accel_batch_handles(self.h)

    def run(self, jobs, optimize=True, maxcands=4096):
        """jobs: list of (params, spectrum, accelnm or None, candnm or None).
        Returns a list of candidate arrays (CAND_DTYPE), one per job, and
        the per-job device times (ms)."""
        n = len(jobs)
        items = (BatchItem * n)()
        keep = []
        outs = []
        for it, (p, fft, acc, cand) in zip(items, jobs):
            fft = np.ascontiguousarray(fft, dtype=np.complex64)
            if len(fft) != p.numbins:
                raise ValueError("spectrum length differs from params.numbins")
            out = np.zeros(maxcands, dtype=CAND_DTYPE)
            keep.append(fft)
            outs.append(out)
            it.params = p
            it.fft = fft.ctypes.data
            it.optimize = int(optimize)
            it.accelnm = acc.encode() if acc else None
            it.candnm = cand.encode() if cand else None
            it.inf_path = None
            it.cands = out.ctypes.data
            it.maxcands = maxcands
        rc = lib.accel_batch_run(self.h, items, n)
        if rc:
            bad = [i for i in range(n) if items[i].status]
            raise AccelError(f"batch items {bad[:8]} failed: {lib.accel_last_error().decode()}")
        return ([o[:min(it.ncands, maxcands)].copy() for o, it in zip(outs, items)],
                [it.gpu_ms for it in items])

    def run_device(self, params, series_ptrs, nseries, optimize=False, maxcands=4096):
        """Pipelined searches of time series already in device memory (one
        device pointer of nseries floats per job, e.g. the DM trials of
        dedisperse_dev), all with the same params.  Returns the candidate
        arrays and per-job device times, as run()."""
        n = len(series_ptrs)
        items = (BatchItem * n)()
        outs = []
        for it, ptr in zip(items, series_ptrs):
            out = np.zeros(maxcands, dtype=CAND_DTYPE)
            outs.append(out)
            it.params = params
            it.fft = None
            it.d_series = int(ptr)
            it.nseries = int(nseries)
            it.optimize = int(optimize)
            it.cands = out.ctypes.data
            it.maxcands = maxcands
        rc = lib.accel_batch_run(self.h, items, n)
        if rc:
            bad = [i for i in range(n) if items[i].status]
            raise AccelError(f"batch items {bad[:8]} failed: {lib.accel_last_error().decode()}")
        return ([o[:min(it.ncands, maxcands)].copy() for o, it in zip(outs, items)],
                [it.gpu_ms for it in items])


_sig("accel_search_footprint", _ll, _PP)
_sig("accel_template_cache_clear", None)
_sig("accel_server_create", _p, _i, _ll, _i)
_sig("accel_server_submit", _i, _p, _PP, _p, _ll, _i)
_sig("accel_server_wait", _i, _p, _i, C.POINTER(_d), C.POINTER(_d))
_sig("accel_server_cands", _i, _p, _i, _p, _i)
_sig("accel_server_props", _i, _p, _i, _p, _i)
_sig("accel_server_memory", _ll, _p)
_sig("accel_server_timing", _i, _p, _i, _p)
_sig("accel_server_destroy", None, _p)


def footprint(params):
    """Device bytes a search with these parameters needs."""
    return lib.accel_search_footprint(C.byref(params))


class Server:
    """The request loop: submit time series as they arrive, wait for each;
    requests run concurrently when their memory fits the budget."""

    def __init__(self, device=0, budget_bytes=200 << 30, workers=4):
        self.h = lib.accel_server_create(int(device), int(budget_bytes), int(workers))
        if not self.h:
            raise AccelError("cannot start the server")

    def submit(self, params, series, optimize=True):
        x = np.ascontiguousarray(series, dtype=np.float32)
        t = lib.accel_server_submit(self.h, C.byref(params), x.ctypes.data, len(x), int(optimize))
        if t < 0:
            raise AccelError("submit failed (bad parameters)")
        return t

    def wait(self, ticket):
        """(status, latency_ms, service_ms)"""
        lat, svc = _d(), _d()
        st = lib.accel_server_wait(self.h, ticket, C.byref(lat), C.byref(svc))
        return st, lat.value, svc.value

    def candidates(self, ticket):
        n = lib.accel_server_cands(self.h, ticket, None, 0)
        out = np.zeros(max(n, 0), dtype=CAND_DTYPE)
        if n > 0:
            lib.accel_server_cands(self.h, ticket, out.ctypes.data, n)
        return out

    def props(self, ticket):
        n = lib.accel_server_props(self.h, ticket, None, 0)
        if n <= 0:
            return []
        out = (FourierProps * n)()
        lib.accel_server_props(self.h, ticket, out, n)
        return list(out)

    def memory(self):
        return lib.accel_server_memory(self.h)

    def timing(self, ticket):
        """dict of queued, evict, create, prep, search, refine, work, service
        (ms) and footprint bytes"""
        v = np.zeros(9)
        if lib.accel_server_timing(self.h, ticket, v.ctypes.data):
            raise AccelError("request not finished")
        return dict(zip(("queued", "evict", "create", "prep", "search", "refine", "work",
                         "service", "bytes"), v))

    def close(self):
        if self.h:
            lib.accel_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        self.close()


def max_rz_harmonics(spec, r, z, numharm, device=0):
    """max_rz_arr_harmonics for arrays of starts on the GPU.  Returns
    rout, zout, powers [n, MAXHARM], derivs (list of lists of RDerivs)."""
    spec = np.ascontiguousarray(spec, dtype=np.complex64)
    r = np.ascontiguousarray(np.atleast_1d(r), dtype=np.float64)
    z = np.ascontiguousarray(np.atleast_1d(z), dtype=np.float64)
    nh = np.ascontiguousarray(np.broadcast_to(numharm, r.shape), dtype=np.int32)
    n = len(r)
    ro, zo = np.zeros(n), np.zeros(n)
    pw = np.zeros((n, MAXHARM))
    der = (RDerivs * (n * MAXHARM))()
    if lib.accel_max_rz_harmonics(spec.ctypes.data, len(spec), n, nh.ctypes.data,
                                  r.ctypes.data, z.ctypes.data, ro.ctypes.data, zo.ctypes.data,
                                  pw.ctypes.data, der, int(device)):
        raise _err()
    dl = [list(der)[i * MAXHARM:i * MAXHARM + nh[i]] for i in range(n)]
    return ro, zo, pw, dl


def max_rzw_harmonics(spec, r, z, w, numharm, device=0):
    """max_rzw_arr_harmonics for arrays of starts on the GPU.  Returns
    rout, zout, wout, powers [n, MAXHARM], derivs (lists of RDerivs)."""
    spec = np.ascontiguousarray(spec, dtype=np.complex64)
    r = np.ascontiguousarray(np.atleast_1d(r), dtype=np.float64)
    z = np.ascontiguousarray(np.atleast_1d(z), dtype=np.float64)
    w = np.ascontiguousarray(np.atleast_1d(w), dtype=np.float64)
    nh = np.ascontiguousarray(np.broadcast_to(numharm, r.shape), dtype=np.int32)
    n = len(r)
    ro, zo, wo = np.zeros(n), np.zeros(n), np.zeros(n)
    pw = np.zeros((n, MAXHARM))
    der = (RDerivs * (n * MAXHARM))()
    if lib.accel_max_rzw_harmonics(spec.ctypes.data, len(spec), n, nh.ctypes.data,
                                   r.ctypes.data, z.ctypes.data, w.ctypes.data, ro.ctypes.data,
                                   zo.ctypes.data, wo.ctypes.data, pw.ctypes.data, der,
                                   int(device)):
        raise _err()
    dl = [list(der)[i * MAXHARM:i * MAXHARM + nh[i]] for i in range(n)]
    return ro, zo, wo, pw, dl


def calc_props(rd, r, z, w):
    out = FourierProps()
    if lib.accel_calc_props(C.byref(rd), r, z, w, C.byref(out)):
        raise AccelError("calc_props: power or local power out of range")
    return out


def calc_rzwerrs(props, T):
    out = RzwErrs()
    if lib.accel_calc_rzwerrs(C.byref(props), T, C.byref(out)):
        raise AccelError("calc_rzwerrs: T out of range")
    return out


def read_birdies(zaplist, T, baryv=0.0):
    """get_birdies: sorted (lobins, hibins) of a zaplist file."""
    n = lib.accel_read_birdies(zaplist.encode(), T, baryv, None, None, 0)
    if n < 0:
        raise AccelError(f"cannot read zaplist {zaplist}")
    lo, hi = np.zeros(max(n, 1)), np.zeros(max(n, 1))
    lib.accel_read_birdies(zaplist.encode(), T, baryv, lo.ctypes.data, hi.ctypes.data, n)
    return lo[:n], hi[:n]


def zap_birdies(fft, N, lobins, hibins):
    """The -zaplist step on a spectrum (in place on a complex64 copy,
    returned)."""
    out = np.ascontiguousarray(fft, dtype=np.complex64).copy()
    lo = np.ascontiguousarray(lobins, dtype=np.float64)
    hi = np.ascontiguousarray(hibins, dtype=np.float64)
    if lib.accel_zap_birdies(out.ctypes.data, len(out), float(N), lo.ctypes.data,
                             hi.ctypes.data, len(lo)):
        raise AccelError("zapping failed (birdie outside the spectrum)")
    return out


class DedispParams(C.Structure):
    _fields_ = [("numchan", _i), ("dt", _d), ("lofreq", _d), ("chanwidth", _d),
                ("spectra_per_subint", _i), ("nsub", _i), ("downsamp", _i), ("numdms", _i),
                ("lodm", _d), ("dmstep", _d), ("numout", _ll)]


_sig("accel_dedisp_plan", _i, _p, _ll, _p, _p, _p, _p)
_sig("accel_search_set_device_series", _i, _p, _p, _ll)
_sig("accel_dedisperse_dev", _i, _p, _p, _ll, _p, _p, _p, _i)
_sig("accel_device_alloc", _p, C.c_size_t, _i)
_sig("accel_device_free", _i, _p)
_sig("accel_device_upload", _i, _p, _p, C.c_size_t)
_sig("accel_device_download", _i, _p, _p, C.c_size_t)
_sig("accel_dedisperse", _ll, _p, _p, _ll, _p, _p, _i)
_sig("accel_realfft_forward_noalloc", _i, _p, _ll, _p, _p, _i)


def dedisp_params(numchan, dt, lofreq, chanwidth, spectra_per_subint, nsub, downsamp, numdms,
                  lodm, dmstep, numout):
    return DedispParams(numchan, dt, lofreq, chanwidth, spectra_per_subint, nsub, downsamp,
                        numdms, lodm, dmstep, numout)


def dedisp_plan(p, nspec):
    """(idispdt[numchan], offsets[numdms][nsub], blocksperread, written)."""
    idd = np.zeros(p.numchan, dtype=np.int32)
    off = np.zeros((p.numdms, p.nsub), dtype=np.int32)
    bpr, nw = C.c_int(), C.c_longlong()
    if lib.accel_dedisp_plan(C.byref(p), int(nspec), idd.ctypes.data, off.ctypes.data,
                             C.byref(bpr), C.byref(nw)):
        raise _err()
    return idd, off, bpr.value, nw.value


def dedisperse(p, raw, device=0):
    """DM trials of raw[nspec][numchan] (float32): (out[numdms][numout],
    written, pad value)."""
    raw = np.ascontiguousarray(raw, dtype=np.float32)
    out = np.zeros((p.numdms, p.numout), dtype=np.float32)
    pad = C.c_double()
    n = lib.accel_dedisperse(C.byref(p), raw.ctypes.data, raw.shape[0], out.ctypes.data,
                             C.byref(pad), int(device))
    if n < 0:
        raise _err()
    return out, n, pad.value


class DeviceBuffer:
    """Device memory from the library (accel_device_alloc), freed on close."""

    def __init__(self, nbytes, device=0):
        self.ptr = lib.accel_device_alloc(int(nbytes), int(device))
        if not self.ptr:
            raise _err()
        self.nbytes = int(nbytes)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        if arr.nbytes > self.nbytes or lib.accel_device_upload(self.ptr, arr.ctypes.data,
                                                               arr.nbytes):
            raise _err()

    def download(self, arr):
        if arr.nbytes > self.nbytes or lib.accel_device_download(arr.ctypes.data, self.ptr,
                                                                 arr.nbytes):
            raise _err()
        return arr

    def close(self):
        if self.ptr:
            lib.accel_device_free(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def dedisperse_dev(p, d_raw, nspec, d_out, device=0):
    """DM trials from a device filterbank into a device buffer: (pad value,
    [subband kernel ms, DM-sum kernel ms])."""
    pad = C.c_double()
    ms = (C.c_float * 2)()
    if lib.accel_dedisperse_dev(C.byref(p), C.c_void_p(int(d_raw)), int(nspec),
                                C.c_void_p(int(d_out)), C.byref(pad), ms, int(device)):
        raise _err()
    return pad.value, [ms[0], ms[1]]


def realfft_gpu(x, device=0):
    """PRESTO-packed forward real FFT of a float32 series on the GPU."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.zeros(len(x) // 2, dtype=np.complex64)
    if lib.accel_realfft_forward(x.ctypes.data, len(x), out.ctypes.data, int(device)):
        raise _err()
    return out


def realfft_inverse_gpu(spec, device=0):
    """realfft(data, n, +1) on the GPU: the series whose PRESTO-packed
    spectrum is spec (n/2 complex bins)."""
    spec = np.ascontiguousarray(spec, dtype=np.complex64)
    out = np.zeros(2 * len(spec), dtype=np.float32)
    if lib.accel_realfft_inverse(spec.ctypes.data, len(out), out.ctypes.data, int(device)):
        raise _err()
    return out


def deredden_gpu(fft, startwidth, endwidth, endfreq, T, device=0):
    """deredden of a spectrum computed on the GPU (a de-reddened copy)."""
    out = np.ascontiguousarray(fft, dtype=np.complex64).copy()
    n = lib.accel_deredden_gpu(out.ctypes.data, len(out), int(startwidth), int(endwidth),
                               float(endfreq), float(T), int(device))
    if n < 0:
        raise _err()
    return out, n


def spectrum_from_series_gpu(x, dt, device=0):
    """spectrum_from_series with the real FFT on the GPU (FP32, as the
    accelsearch command does it)."""
    out = realfft_gpu(x, device)
    if lib.accel_prepare_dat_spectrum(out.ctypes.data, len(out), float(len(x)), float(dt)) < 0:
        raise AccelError("deredden failed (spectrum too short)")
    return out


def mt19937_raw(seed, n):
    out = np.zeros(n, dtype=np.uint32)
    lib.accel_mt19937_raw(seed, out.ctypes.data, n)
    return out


def nice_output(val, err, numerr, length=0):
    buf = C.create_string_buffer(96)
    if lib.accel_nice_output(buf, val, err, length, numerr) < 0:
        raise AccelError("nice_output: undefined case")
    return buf.value.decode()


def spectrum_from_series(x, dt):
    """Spectrum of a float32 time series the way accelsearch prepares a .dat:
    forward real FFT in PRESTO's packing (bin 0 holds DC and Nyquist), then
    accel_prepare_dat_spectrum.  The real FFT here is numpy's (FP64, rounded
    to FP32)."""
    x = np.asarray(x, dtype=np.float32)
    X = np.fft.rfft(x.astype(np.float64))
    n2 = len(x) // 2
    out = np.empty(n2, dtype=np.complex64)
    out[1:] = X[1:n2]
    out[0] = complex(X[0].real, X[n2].real)
    if lib.accel_prepare_dat_spectrum(out.ctypes.data, n2, float(len(x)), float(dt)) < 0:
        raise AccelError("deredden failed (spectrum too short)")
    return out


class FoldStats(C.Structure):
    """foldstats (include/presto.h:261-269)"""
    _fields_ = [(n, _d) for n in ("numdata", "data_avg", "data_var", "numprof", "prof_avg",
                                  "prof_var", "redchi")]


_sig("accel_read_rzw_cand", _i, C.c_char_p, _i, C.POINTER(FourierProps))
_sig("accel_cand_fold_params", None, C.POINTER(FourierProps), _d, C.POINTER(_d), C.POINTER(_d),
     C.POINTER(_d))
_sig("accel_fold_proflen", _i, _d, _d)
_sig("accel_fold", _ll, _p, _ll, _d, _d, _d, _d, _i, _i, _p, C.POINTER(FoldStats),
     C.POINTER(FoldStats), _i)


class FoldBest(C.Structure):
    """accel_fold_best: the best trial of prepfold's p / p-dot search."""
    _fields_ = [("numtrials", _i), ("ip", _i), ("ipd", _i), ("ipdd", _i), ("p", _d), ("pd", _d),
                ("pdd", _d), ("stats", FoldStats)]


_sig("accel_fold_search", _i, _p, C.POINTER(FoldStats), _i, _i, _d, _d, _d, _d, _p, _i, _i, _i,
     _i, _i, _p, _p, C.POINTER(FoldBest), _i)


def read_rzw_cand(candfile, candnum):
    """Record candnum (from 1) of a .cand file (get_rzw_cand)."""
    out = FourierProps()
    if lib.accel_read_rzw_cand(str(candfile).encode(), int(candnum), C.byref(out)):
        raise _err()
    return out


def cand_fold_params(props, T):
    """prepfold's start f, fdot, fdotdot for a candidate's average r, z, w."""
    f, fd, fdd = _d(), _d(), _d()
    lib.accel_cand_fold_params(C.byref(props), T, C.byref(f), C.byref(fd), C.byref(fdd))
    return f.value, fd.value, fdd.value


def fold_proflen(f, dt):
    return lib.accel_fold_proflen(f, dt)


def fold(series, dt, f, fd=0.0, fdd=0.0, npart=64, proflen=None, device=0):
    """prepfold's fold of a time series on the GPU: (rawfolds [npart][proflen],
    per-part stats, total stats, points folded); stats are FoldStats."""
    x = np.ascontiguousarray(series, dtype=np.float32)
    if proflen is None:
        proflen = fold_proflen(f, dt)
    prof = np.zeros((npart, proflen), dtype=np.float64)
    parts = (FoldStats * npart)()
    total = FoldStats()
    n = lib.accel_fold(x.ctypes.data, len(x), dt, f, fd, fdd, npart, proflen, prof.ctypes.data,
                       parts, C.byref(total), device)
    if n < 0:
        raise _err()
    return prof, list(parts), total, int(n)


def fold_search(rawfolds, parts, f, fd, fdd, T, parttimes, npfact=2, pstep=2, pdstep=4,
                nosearch=False, searchpdd=False, device=0):
    """prepfold's period / p-dot (and with searchpdd p-dotdot) search over
    the sub-integrations of a fold on the GPU: (ppdot [numtrials][numtrials]
    float32 reduced chi-squared indexed [ipd][ip] -- [numtrials]^3,
    [ipdd][ipd][ip], with searchpdd -- or None with nosearch; best profile;
    FoldBest)."""
    prof = np.ascontiguousarray(rawfolds, dtype=np.float64)
    npart, proflen = prof.shape
    pt = np.ascontiguousarray(parttimes, dtype=np.float64)
    if len(pt) != npart or len(parts) != npart:
        raise ValueError("one part time and one FoldStats per sub-integration")
    st = (FoldStats * npart)(*parts)
    nt = 2 * npfact * proflen + 1
    pp = None if nosearch else np.zeros((nt, nt, nt) if searchpdd else (nt, nt),
                                        dtype=np.float32)
    bestprof = np.zeros(proflen, dtype=np.float64)
    best = FoldBest()
    if lib.accel_fold_search(prof.ctypes.data, st, npart, proflen, f, fd, fdd, T, pt.ctypes.data,
                             npfact, pstep, pdstep, int(bool(nosearch)), int(bool(searchpdd)),
                             None if pp is None else pp.ctypes.data, bestprof.ctypes.data,
                             C.byref(best), device):
        raise _err()
    return pp, bestprof, best
