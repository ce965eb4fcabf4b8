/*
 * presto_accel.h -- C ABI of the MI355X accelsearch search path
 * (libpresto_accel.so).
 *
 * The reference path is PRESTO's accelsearch: src/accelsearch.c:161-246
 * drives the functions of include/accel.h:134-164 over overlap-save blocks
 * of the spectrum.  This library keeps that behaviour -- block geometry,
 * normalisation, harmonic index rules, thresholds, the order-dependent
 * candidate merge, sorting and harmonic removal -- and computes the planes
 * and harmonic sums on the GPU.  Each entry point below names the
 * reference interface it replaces.
 *
 * Conventions: plain C types, host pointers unless a name says "device",
 * return 0 on success and a negative code on failure (accel_last_error()
 * describes it), where the reference printed and called exit().
 */
#ifndef PRESTO_ACCEL_H
#define PRESTO_ACCEL_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* include/ransomfft.h:13-18 */
typedef struct {
    float r, i;
} accel_fcomplex;

/* The search inputs: the accelsearch command line
 * (clig/accelsearch_cmd.cli:13-48) and the .inf values that
 * create_accelobs reads (src/accel_utils.c:1497, 1616, 1676). */
typedef struct {
    double N;            /* samples in the time series (.inf N) */
    double dt;           /* sample time (s) (.inf dt) */
    long long numbins;   /* complex bins in the spectrum */
    long long lobin;     /* -lobin: bin number of the first spectrum element */
    int numharm;         /* -numharm: 1, 2, 4, 8, 16 or 32 */
    int zmax;            /* -zmax */
    int wmax, wmaxP;     /* -wmax (jerk search when wmaxP) */
    float sigma;         /* -sigma */
    double flo, fhi;     /* -flo / -fhi (Hz) */
    double rlo, rhi;     /* -rlo / -rhi (bins) */
    int floP, fhiP, rloP, rhiP;
    int inmem;           /* -inmem */
    int force_standard;  /* never pick the in-memory path (testing) */
    int dat_input;       /* the spectrum came from a .dat time series */
    int noharmremove;    /* -noharmremove: keep harmonically related candidates */
    double nph;          /* -photon: the photon count (obs->nph, bin 0's real part,
                          * src/accel_utils.c:1599-1605); every block is then
                          * normalised by 1/sqrt(nph) (:1080-1086) instead of its
                          * median.  0: block medians.  A .dat spectrum keeps bin 0
                          * as de-reddened, (1, 0), instead of (1, 1) (:1606-1624) */
    int noharmpolish;    /* -noharmpolish: refine every harmonic on its own, without
                          * polishing (src/accel_utils.c:582-622, max_rz_arr /
                          * max_rzw_arr) */
} accel_params;

/* Fills the command-line defaults: numharm 8, zmax 200, sigma 2.0,
 * flo 1 Hz and fhi 10 kHz counted as given (src/accelsearch_cmd.c:30-74). */
void accel_params_default(accel_params *p);

/* The numeric part of accelobs (include/accel.h:37-81) as computed by
 * create_accelobs (src/accel_utils.c:1596-1804). */
typedef struct {
    long long N, numbins, lobin, highestbin;
    int maxkernlen, corr_uselen, fftlen, numharmstages, numz, numw, numbetween;
    int inmem;
    double dt, T, rlo, rhi, dr, zlo, zhi, dz, wlo, whi, dw;
    float sigma;
    float powcut[6];
    long long numindep[6];
    int nblocks;         /* blocks of the search loop (src/accelsearch.c:189) */
    double nph;          /* photon normalisation in use (0 when lobin > 0,
                          * src/accel_utils.c:1626-1628) */
} accel_geom;

/* Replaces the numeric part of create_accelobs (src/accel_utils.c:1455). */
int accel_geom_create(const accel_params *p, accel_geom *g);

/* One entry of the final candidate list (accelcand, include/accel.h:83-95,
 * without the refinement fields). */
typedef struct {
    float power, sigma;
    int numharm;
    int pad;
    double r, z, w;
} accel_cand;

/* One threshold crossing, in the reference's search order
 * (src/accel_utils.c:1419-1450). */
typedef struct {
    int block, stage, wi, zi, ii;
    float pow;
} accel_hit;

typedef struct accel_search accel_search;

/* Creates the device state of one search: geometry, FP64-generated
 * templates resident in HBM, normalisation windows, index maps, the raw
 * plane and the hit buffer.  Replaces create_subharminfos
 * (src/accel_utils.c:240) and the allocations of create_accelobs.
 * device: HIP device ordinal.  Returns NULL on failure. */
accel_search *accel_search_create(const accel_params *p, int device);
void accel_search_destroy(accel_search *s);   /* free_subharminfos / free_accelobs */
const accel_geom *accel_search_geom(const accel_search *s);

/* Copies the spectrum (numbins complex values, first one at bin lobin) to
 * HBM.  Replaces the mmap / .dat read of create_accelobs. */
int accel_search_set_spectrum(accel_search *s, const accel_fcomplex *fft);
/* Takes a time series of n samples (n even, n/2 = numbins, lobin 0) and prepares its spectrum on the GPU as accelsearch prepares a
 * .dat: realfft, deredden and bin 0 := 1+1i (src/accel_utils.c:1507-1562,
 * 1621-1624). */
int accel_search_set_series(accel_search *s, const float *series, long long n);
/* The same for a series already in device memory on the handle's device
 * (a DM trial of accel_dedisperse_dev): no host round trip. */
int accel_search_set_device_series(accel_search *s, const float *d_series, long long n);
/* Copies the first n bins of the search's current spectrum to the host. */
int accel_search_spectrum(const accel_search *s, accel_fcomplex *out, long long n);
/* The median blocks of deredden (src/misc_utils.c:940-1046): starts and
 * lengths tiling [1, numbins); fills at most max entries, returns the
 * block count or -1. */
int accel_dered_blocks(long numbins, int startwidth, int endwidth, double endfreq, double T,
                       long long *start, int *len, int max);
/* Uses a spectrum already in device memory (not copied, not owned). */
int accel_search_set_device_spectrum(accel_search *s, const void *d_fft);

/* Runs the block loop of src/accelsearch.c:161-228: normalisation of every
 * block (subharm_fderivs_vol's median, src/accel_utils.c:1087-1099), the
 * fused correlation (subharm_fderivs_vol), the harmonic sums
 * (add_ffdotpows / inmem_add_ffdotpows) and the thresholds
 * (search_ffdotpows), then replays the hits through the reference's
 * candidate merge, sort_accelcands and eliminate_harmonics
 * (src/accelsearch.c:232-241). */
int accel_search_run(accel_search *s);
/* The .txtcand work file of an .fft input (src/accel_utils.c:1666-1684):
 * opens path for writing ("w"); every later run then writes one line per
 * candidate that the merge adds or replaces, in the reference's format
 * and order (search_ffdotpows, src/accel_utils.c:1443-1447).  NULL closes
 * the file.  Returns 0, or -1 if the file cannot be opened or closed. */
int accel_search_set_txtcand(accel_search *s, const char *path);
/* Only the device part (norms, correlation, harmonic sums); no hit replay.
 * Leaves the hit count on the device; accel_search_collect() finishes. */
int accel_search_launch(accel_search *s);
int accel_search_collect(accel_search *s);

int accel_search_ncands(const accel_search *s);
int accel_search_cands(const accel_search *s, accel_cand *out, int n);
long long accel_search_nhits(const accel_search *s);
int accel_search_hits(const accel_search *s, accel_hit *out, long long n);

/* Inspection (tests): raw plane values for row (wi, zi), half-bins
 * [hb0, hb0+n); normalisation factors (FP64) of every window. */
int accel_search_plane_row(const accel_search *s, int wi, int zi, long long hb0, int n,
                           float *out);
int accel_search_nwindows(const accel_search *s);
int accel_search_norms(const accel_search *s, double *out, int n);
/* Milliseconds of the last run (HIP events): block norms (on a second
 * stream, beside the correlation), correlation, the gap between the
 * correlation and the harmonic sum (the plane's tile halos are written by
 * the correlation, so about 0), harmonic sum (including any wait for the
 * norms), total. */
int accel_search_timings(const accel_search *s, float *ms5);
/* Host wall-clock milliseconds of the last run: enqueueing the kernels,
 * waiting for them, copying the hits back, replaying the hits into the
 * candidate list (merge, sort, harmonic removal). */
int accel_search_host_timings(const accel_search *s, double *ms4);
/* Launch shape of the correlation kernel: segment length, used half-bins
 * per segment, template groups, segments. */
int accel_search_corr_shape(const accel_search *s, int *L, int *uselen, int *groups,
                            long long *nseg);
/* Measured choice of the correlation's template-group size, the
 * counterpart of the reference's FFTW_MEASURE planning of its correlation
 * FFTs (src/accel_utils.c:1133-1135, src/fftcalls.c:130-138): times the
 * correlation kernel over the current spectrum for each group count that
 * is a multiple of 8 giving 4 to 64 templates per group (at most ncand of
 * them, nearest the default first), keeps the fastest for this handle and
 * returns how many were timed; groups[i] and ms[i] (best of 3 rounds)
 * describe each, *best the one kept.  Every group count writes the same
 * raw plane values, so results do not change. */
int accel_search_tune(accel_search *s, int ncand, int *groups, float *ms, int *best);
/* Raw plane half-bins written per run (templates x half-bins). */
long long accel_search_plane_cells(const accel_search *s);

/* Normalisation windows of the last plan: bins [lobin, lobin+numdata). */
int accel_search_windows(const accel_search *s, long long *lobin, int *numdata, int n);

/* realfft(data, n, -1) (src/fastffts.c:163-243) of a host series on the
 * GPU: n even (any such length); out receives n/2 complex bins, out[0] =
 * (DC, Nyquist).  Replaces the FFT of create_accelobs's .dat input
 * (src/accel_utils.c:1547). */
int accel_realfft_forward(const float *series, long long n, accel_fcomplex *out, int device);

/* accel_realfft_forward on device buffers that allocates nothing: no
 * counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  d_series (n floats) and d_out (n/2 complex) are
 * device memory on `device`; the work runs on `stream` (a hipStream_t,
 * NULL for the default stream) and is complete when that stream's work
 * is; the call itself does not wait.  The FFT's work buffer is a 24 MiB
 * array each device holds from its first call until the process ends: an
 * n whose work does not fit it is refused (powers of two fit up to 2^22).
 * Calls that overlap on one
 * device share that buffer and race: serialise them.  The results equal
 * accel_realfft_forward's bit for bit. */
int accel_realfft_forward_noalloc(const float *d_series, long long n, accel_fcomplex *d_out,
                                  void *stream, int device);

/* realfft(data, n, +1) (src/fastffts.c:163-243): the inverse of the above,
 * scaled by 2/n so that it returns the original series; on the GPU, n
 * even.  Replaces the in-core inverse of realfft -inv
 * (src/realfft.c:331). */
int accel_realfft_inverse(const accel_fcomplex *spec, long long n, float *out, int device);

/* deredden (src/misc_utils.c:1054-1067) of a host spectrum, computed on
 * the GPU (exact block medians, the reference's FP32 interpolation and
 * invsqrtf); bin 0 becomes (1, 0).  Returns the bins written or -1.  The
 * de-reddening of realfft -rednoise and rednoise (src/realfft.c:342,
 * src/rednoise.c:86). */
long accel_deredden_gpu(accel_fcomplex *fft, long long numbins, int startwidth, int endwidth,
                        double endfreq, double T, int device);

/* Running-median de-reddening of a spectrum in place, on the host (the
 * restatement the GPU version is checked against); replaces deredden
 * (src/misc_utils.c:1054).  Returns the bins written or -1. */
long accel_deredden(accel_fcomplex *fft, long numbins, int startwidth, int endwidth,
                    double endfreq, double T);
/* The .dat preparation after the real FFT: accel_deredden(fft, numbins, 6,
 * 200, 0.5/dt, N*dt) then bin 0 := 1+1i (src/accel_utils.c:1561,1621-1624). */
long accel_prepare_dat_spectrum(accel_fcomplex *fft, long numbins, double N, double dt);

/* get_birdies (src/birdzap.c:6-90): reads a zaplist (frequency and width
 * in Hz per line, '#' comments, 'B' prefix = already barycentric) into
 * sorted bin ranges for an observation of T seconds at radial velocity
 * baryv (v/c).  Fills at most maxbirds entries; returns the number of
 * birdies in the file or -1. */
int accel_read_birdies(const char *zaplist, double T, double baryv, double *lobins,
                       double *hibins, int maxbirds);
/* The -zaplist step of src/accelsearch.c:87-110 on an in-memory spectrum
 * of an N-sample series: zapbirds (src/zapping.c:58-160) over the sorted
 * birdies, replacing each with chi^2(2) noise at the local power level
 * (MT19937 seeded with 1, as the reference's ZAPSEED). */
int accel_zap_birdies(accel_fcomplex *fft, long long numbins, double N, const double *lobins,
                      const double *hibins, int nbirds);
/* The automatic mode of the zapbirds command (src/zapbirds.c:273-303) on a
 * whole .fft held in memory: birdies are zapped while lobin < N/2 - 1,
 * hibin clipped to N/2 - 1, over numbins = min(N/2, bins held);
 * constamp: the constant-amplitude replacement of -constamp
 * (src/zapping.c:124-128) instead of noise. */
int accel_zapbirds_fft(accel_fcomplex *fft, long long numbins, double N, const double *lobins,
                       const double *hibins, int nbirds, int constamp);
/* n raw 32-bit outputs of MT19937 seeded with init_genrand(seed), the
 * generator behind the zapping noise (for tests). */
int accel_mt19937_raw(unsigned int seed, unsigned int *out, int n);

/* Significance (src/characteristics.c:526-559), for callers and tests. */
double accel_candidate_sigma(double power, int numsum, double numtrials);
double accel_power_for_sigma(double sigma, int numsum, double numtrials);

/* ---- refinement and output ---- */

/* include/presto.h:160-168 */
typedef struct {
    double pow, phs, dpow, dphs, d2pow, d2phs, locpow;
} accel_rderivs;

/* include/presto.h:170-188: the record of the .cand file (88 bytes) */
typedef struct {
    double r;
    float rerr;
    double z;
    float zerr;
    double w;
    float werr;
    float pow, powerr, sig, rawpow, phs, phserr, cen, cenerr, pur, purerr, locpow;
} accel_fourierprops;

/* include/presto.h:233-246 */
typedef struct {
    double p, pd, pdd, f, fd, fdd;
    float perr, pderr, pdderr, ferr, fderr, fdderr;
} accel_rzwerrs;

/* Highest harmonic count a refined candidate may carry. */
#define ACCEL_MAXHARM 32

/* optimize_accelcand (src/accel_utils.c:552-626) for every candidate of
 * the last run, as src/accelsearch.c:250-257 calls it, with harmonic
 * polishing (the default), in FP64 on the GPU: max_rz_arr_harmonics
 * (src/maximize_rz.c:171) or, for jerk searches, max_rzw_arr_harmonics
 * (src/maximize_rzw.c:84). */
int accel_search_optimize(accel_search *s);
/* No counterpart in the reference.  The refinement runs a candidate's
 * harmonics as workgroups that wait for each other; a wait is bounded and
 * a candidate whose wait runs out is refined again on its own.  This sets
 * the first attempt's bound to 2^log2_polls polls (0..24, default 24) for
 * the process, so that tests can make the first attempt fail and check the
 * retry's results; nothing else should lower it. */
void accel_refine_set_first_wait(int log2_polls);
/* Refined harmonics of candidate i (after accel_search_optimize): the
 * accelcand fields hirs, hizs, hiws, pows and derivs
 * (include/accel.h:83-95).  Returns the harmonic count. */
int accel_search_cand_harmonics(const accel_search *s, int i, double *hirs, double *hizs,
                                double *hiws, double *pows, accel_rderivs *derivs, int nmax);
/* The fundamentals' fourierprops written to the .cand file
 * (src/accelsearch.c:262-275).  Returns the count written. */
int accel_search_fund_props(const accel_search *s, accel_fourierprops *props, int n);
/* The ACCEL text file (output_fundamentals + output_harmonics,
 * src/accel_utils.c:684-979) and, if candnm is not NULL, the .cand binary
 * (src/accelsearch.c:281-284).  inf_path: the .inf to append to the text
 * file as the reference does (src/accel_utils.c:975-977), or NULL. */
int accel_search_write_outputs(const accel_search *s, const char *accelnm, const char *candnm,
                               const char *inf_path);

/* max_rz_arr_harmonics (src/maximize_rz.c:171-228) for n (r, z) starts on
 * a host spectrum of numdata bins (r relative to its first bin), on the
 * GPU.  pows and derivs hold n x ACCEL_MAXHARM entries. */
int accel_max_rz_harmonics(const accel_fcomplex *spec, long long numdata, int n,
                           const int *numharm, const double *rin, const double *zin,
                           double *rout, double *zout, double *pows, accel_rderivs *derivs,
                           int device);

/* max_rzw_arr_harmonics (src/maximize_rzw.c:84-191) for n (r, z, w)
 * starts on a host spectrum, on the GPU; outputs as above plus wout. */
int accel_max_rzw_harmonics(const accel_fcomplex *spec, long long numdata, int n,
                            const int *numharm, const double *rin, const double *zin,
                            const double *win, double *rout, double *zout, double *wout,
                            double *pows, accel_rderivs *derivs, int device);

/* src/characteristics.c:195-262 and :330-396 */
int accel_calc_props(const accel_rderivs *d, double r, double z, double w,
                     accel_fourierprops *p);
int accel_calc_rzwerrs(const accel_fourierprops *p, double T, accel_rzwerrs *e);

/* The value(error) text of src/output.c:6-113 (numerr 1) and :114-216
 * (numerr 2); returns the string length or -1. */
int accel_nice_output(char *out, double val, double err, int len, int numerr);

/* Template banks are generated once per (z, w) grid and shared by all
 * handles; this frees the banks no handle uses. */
void accel_template_cache_clear(void);

/* Device memory (bytes) a search handle for p needs, including series
 * input and refinement buffers (for admission by memory footprint). */
long long accel_search_footprint(const accel_params *p);

/* ---- batches of DM-trial spectra ---- */

/* One spectrum of a batch: inputs, what to do with it, and its results. */
typedef struct {
    accel_params params;          /* search parameters of this spectrum */
    const accel_fcomplex *fft;    /* params.numbins host bins */
    int optimize;                 /* refine the candidates (and write files) */
    const char *accelnm, *candnm, *inf_path;   /* outputs, or accelnm NULL */
    accel_cand *cands;            /* caller's array for the candidates, or NULL */
    int maxcands;
    int ncands;                   /* out: candidates found */
    int status;                   /* out: 0 or -1 */
    float gpu_ms;                 /* out: device time of its search */
    /* a time series already in device memory instead of fft (fft NULL):
     * nseries floats, prepared on the device as accel_search_set_device_series
     * does; it must stay valid until accel_batch_run returns */
    const float *d_series;
    long long nseries;
} accel_batch_item;

typedef struct accel_batch accel_batch;

/* A pool of search handles on one device, reused by every spectrum of the
 * same parameters.  No counterpart in the reference, which runs one
 * accelsearch process per spectrum (src/accelsearch.c:82-313). */
accel_batch *accel_batch_create(int device);
void accel_batch_destroy(accel_batch *b);
/* Search handles the pool holds (each with its resident templates). */
int accel_batch_handles(const accel_batch *b);
/* Runs every item as accelsearch would (search, candidates, and when
 * optimize is set the refinement and output files), pipelining the items:
 * the upload and search of item i+1 overlap the hit replay, refinement and
 * output of item i.  Results equal separate runs.  Returns 0, or -1 when
 * any item failed (its status says which). */
int accel_batch_run(accel_batch *b, accel_batch_item *items, int n);

/* ---- a request loop (serving) ---- */

typedef struct accel_server accel_server;

/* Starts `workers` threads serving requests on one device within
 * hbm_budget bytes of device memory.  No counterpart in the reference
 * (one accelsearch process per input, src/accelsearch.c:82-313). */
accel_server *accel_server_create(int device, long long hbm_budget, int workers);
/* Queues a time series of n samples (copied) for an accelsearch with
 * parameters p (numbins = n/2); optimize: refine the candidates.  Requests
 * are admitted in arrival order when their memory fits beside the running
 * ones, and run concurrently on their own streams.  Returns a ticket. */
int accel_server_submit(accel_server *sv, const accel_params *p, const float *series,
                        long long n, int optimize);
/* Waits for a request; latency = submit to done, service = start to done
 * (milliseconds).  Returns its status (0 or -1). */
int accel_server_wait(accel_server *sv, int ticket, double *latency_ms, double *service_ms);
/* Results of a finished request: candidates (returns the count) and, when
 * refined, the fundamentals' fourierprops. */
int accel_server_cands(accel_server *sv, int ticket, accel_cand *out, int n);
int accel_server_props(accel_server *sv, int ticket, accel_fourierprops *out, int n);
/* Where a finished request's time went (ms): queued, evicting cached
 * handles, creating its handle, preparation (upload, real FFT,
 * de-reddening), search, refinement, the last three together, the whole
 * service; then its device-memory footprint (bytes). */
int accel_server_timing(accel_server *sv, int ticket, double *ms9);
/* Device memory held by live handles (running and cached). */
long long accel_server_memory(accel_server *sv);
/* Finishes every queued request, then stops the workers. */
void accel_server_destroy(accel_server *sv);

/* ---- DM trials from a filterbank: subband de-dispersion ---- */

/* prepsubband's parameters for its topocentric path
 * (clig/prepsubband_cmd.cli: -nsub, -lodm, -dmstep, -numdms, -downsamp,
 * -numout, with -nobary and -noclip) and the filterbank's (.inf freq,
 * chan_wid, dt; spectra_per_subint of the raw reader). */
typedef struct {
    int numchan;            /* channels, in ascending frequency */
    double dt;              /* sample time (s) */
    double lofreq;          /* centre of the lowest channel (MHz) */
    double chanwidth;       /* channel width (MHz), > 0 */
    int spectra_per_subint; /* spectra per raw block */
    int nsub;               /* -nsub: divides numchan */
    int downsamp;           /* -downsamp: divides spectra_per_subint */
    int numdms;             /* -numdms */
    double lodm, dmstep;    /* -lodm, -dmstep */
    long long numout;       /* -numout: samples per DM trial */
} accel_dedisp_params;

/* The plan of src/prepsubband.c:255-306, 364-387: channel delays inside
 * each subband (idispdt[numchan], samples of dt), subband offsets of each
 * DM (offsets[numdms][nsub], samples of downsamp dt), blocksperread, and
 * the samples the reference writes from nspec spectra before it pads
 * (*nwritten).  Any output may be NULL.  -1 when the channels do not split
 * into subbands, or when a delay exceeds the reference's working blocks
 * (src/backend_common.c:785-789, src/prepsubband.c:937-943). */
int accel_dedisp_plan(const accel_dedisp_params *p, long long nspec, int *idispdt, int *offsets,
                      int *blocksperread, long long *nwritten);
/* The DM trials prepsubband -nobary -noclip writes from a filterbank
 * (no mask, no ignored channels): raw [nspec][numchan] FP32 and out
 * [numdms][numout] both in device memory on `device`; the samples after
 * the written ones hold the running mean of the first trial's
 * (src/prepsubband.c:399-431, 667-718), returned in *padval.  ms (or
 * NULL): device time of the subband kernel and of the DM-sum kernel. */
int accel_dedisperse_dev(const accel_dedisp_params *p, const float *d_raw, long long nspec,
                         float *d_out, double *padval, float *ms, int device);
/* The same from and to host memory; returns the written count or -1. */
long long accel_dedisperse(const accel_dedisp_params *p, const float *raw, long long nspec,
                           float *out, double *padval, int device);
/* Device buffers for the device-resident path (filterbank in, DM trials
 * out, accel_search_set_device_series): no counterpart in the reference,
 * whose stages exchange files. */
void *accel_device_alloc(size_t bytes, int device);
int accel_device_free(void *d);
int accel_device_upload(void *d_dst, const void *src, size_t bytes);
int accel_device_download(void *dst, const void *d_src, size_t bytes);

/* ---- sifting candidates across DM trials ---- */

/* The sifting of PRESTO's python/presto/sifting.py:14-1206 (driven as in
 * examplescripts/ACCEL_sift.py) over ACCEL files, on the host.  A sift set
 * holds candidates in lists: the good list, the duplicates and eight
 * reject lists (numbered in the order of the reference's report). */
typedef struct accel_sift accel_sift;
enum {
    ACCEL_SIFT_GOOD = 0,
    ACCEL_SIFT_DUPLICATE,
    ACCEL_SIFT_BAD_KNOWNBIRDS,
    ACCEL_SIFT_BAD_LONGPERIOD,
    ACCEL_SIFT_BAD_SHORTPERIOD,
    ACCEL_SIFT_BAD_THRESHOLD,
    ACCEL_SIFT_BAD_HARMPOWCUTOFF,
    ACCEL_SIFT_BAD_ROGUEHARMPOW,
    ACCEL_SIFT_BAD_HARMONIC,
    ACCEL_SIFT_BAD_DMPROBLEM,
    ACCEL_SIFT_NLISTS
};

/* The rules of accel_sift_reject (Candlist.reject_*,
 * python/presto/sifting.py:504-647); a candidate failing several goes to
 * the first in this order. */
#define ACCEL_SIFT_R_LONGPERIOD 1u
#define ACCEL_SIFT_R_SHORTPERIOD 2u
#define ACCEL_SIFT_R_KNOWNBIRDS 4u
#define ACCEL_SIFT_R_THRESHOLD 8u
#define ACCEL_SIFT_R_HARMPOWCUTOFF 16u
#define ACCEL_SIFT_R_ROGUEHARMPOW 32u
#define ACCEL_SIFT_R_DEFAULT 0x3fu   /* Candlist.default_rejection */

/* The sifting module's thresholds (python/presto/sifting.py:14-36). */
typedef struct {
    double r_err;             /* bins: one signal when |r1 - r2| < r_err (1.1) */
    double long_period;       /* s (15.0) */
    double short_period;      /* s (0.0005) */
    double sigma_threshold;   /* (6.0) */
    double c_pow_threshold;   /* coherent power (100.0) */
    double harm_pow_cutoff;   /* (8.0) */
    const double *birds_f;    /* known birdies: (Hz, half width Hz) pairs */
    int nbirds_f;
    const double *birds_p;    /* (ms, half width ms) pairs */
    int nbirds_p;
} accel_sift_params;

/* One candidate (Candidate, python/presto/sifting.py:135-171). */
typedef struct {
    char path[512];           /* ACCEL file it came from */
    char dmstr[32];           /* DM text of the file name ("30.00") */
    char note[512];           /* why it was set aside */
    int candnum, numharm, nhits, pad;
    double sigma, ipow, cpow, r, z, T, snr;
} accel_sift_cand;

/* A sift set; removed candidates are kept in their lists only when
 * track_bad / track_dupes are set. */
accel_sift *accel_sift_create(int track_bad, int track_dupes);
void accel_sift_destroy(accel_sift *s);
/* Adds a candidate to the good list: harm_pows / harm_phases numharm
 * values (or NULL), hits nhits (DM, S/N, sigma) triples. */
int accel_sift_add(accel_sift *s, const accel_sift_cand *c, const double *harm_pows,
                   const double *harm_phases, const double *hits, int nhits);
/* candlist_from_candfile (python/presto/sifting.py:1075-1176): the
 * candidates of one ACCEL file, re-scored from their refined harmonics
 * (sigma for one trial, S/N, the file's DM as their one hit).  Returns the
 * number read or -1. */
int accel_sift_read(accel_sift *s, const char *accel_path);
/* parse_power (python/presto/sifting.py:124-132) */
double accel_sift_parse_power(const char *txt);
/* Candlist.reject_* / default_rejection: moves the good candidates that
 * fail one of `rules`; returns how many. */
int accel_sift_reject(accel_sift *s, const accel_sift_params *p, unsigned rules);
/* Candlist.remove_duplicate_candidates (python/presto/sifting.py:660-721):
 * candidates within r_err of each other become hits of the one with the
 * highest sigma; the list ends sorted by sigma.  Returns how many moved. */
int accel_sift_remove_duplicates(accel_sift *s, const accel_sift_params *p);
/* Candlist.remove_DM_problems (python/presto/sifting.py:829-919): fewer
 * than min_num_dms hits, the strongest hit at DM <= low_dm_cutoff, or no
 * two hits at neighbouring DMs of the list dms. */
int accel_sift_remove_dm_problems(accel_sift *s, int min_num_dms, const double *dms, int ndms,
                                  double low_dm_cutoff);
/* Candlist.remove_harmonics (python/presto/sifting.py:723-827) */
int accel_sift_remove_harmonics(accel_sift *s, const accel_sift_params *p);
int accel_sift_sort_sigma(accel_sift *s);
/* Candlist.extend: moves every candidate of `from` into s. */
int accel_sift_extend(accel_sift *s, accel_sift *from);
/* The number of candidates in a list, and one of them with its harmonic
 * powers and phases (numharm values each, or NULL) and up to maxhits hits
 * (triples); returns its hit count or -1. */
int accel_sift_count(const accel_sift *s, int list);
int accel_sift_get(const accel_sift *s, int list, int i, accel_sift_cand *c, double *harm_pows,
                   double *harm_phases, double *hits, int maxhits);
/* Candlist.to_file, write_cand_report and print_cand_summary
 * (python/presto/sifting.py:921-1071); path NULL writes to stdout. */
int accel_sift_write(const accel_sift *s, const char *path);
int accel_sift_write_report(const accel_sift *s, const char *path);
int accel_sift_write_summary(const accel_sift *s, const char *path);

const char *accel_last_error(void);


/* ---- the .cand consumer: prepfold -accelcand on a .dat ---- */

/* foldstats (include/presto.h:261-269), member for member. */
typedef struct {
    double numdata, data_avg, data_var, numprof, prof_avg, prof_var, redchi;
} accel_foldstats;

/* get_rzw_cand (src/get_candidates.c:27-37): record candnum (from 1) of a
 * .cand file of 88-byte fourierprops.  Returns 0, or -1 when the file or
 * the record is missing. */
int accel_read_rzw_cand(const char *candfile, int candnum, accel_fourierprops *out);
/* The start frequency and derivatives prepfold folds a candidate with
 * (src/prepfold_pipeline.c:759-767): the .cand holds the average r, z
 * over an observation of T seconds. */
void accel_cand_fold_params(const accel_fourierprops *c, double T, double *f, double *fd,
                            double *fdd);
/* prepfold's default profile length for a period of 1/f s sampled every
 * dt s (src/prepfold_pipeline.c:345-358): one bin per sample, at most 64. */
int accel_fold_proflen(double f, double dt);
/* Folds n samples (host memory) on the GPU as prepfold folds a .dat
 * (src/prepfold_pipeline.c:514-551, 1375-1594; src/fold.c:46-107,
 * 443-653): records of 1024 points, the number of records rounded down to
 * a multiple of npart, every point drizzled over the profile bins its phase
 * interval covers, a sub-integration's profile holding the full turns of
 * phase completed during it.  rawfolds: npart x proflen; parts: npart
 * statistics as fold() leaves them; total: the summed profile's (its
 * redchi = sum (prof - avg)^2 / (prof_var (proflen - 1)), prof_var =
 * data_var numdata / proflen, src/fold.c:627-640).  FP64 phases and sums;
 * the reference keeps phases in long double.  Returns the number of points
 * folded, or -1. */
long long accel_fold(const float *series, long long n, double dt, double f, double fd,
                     double fdd, int npart, int proflen, double *rawfolds,
                     accel_foldstats *parts, accel_foldstats *total, int device);

/* The best trial of prepfold's period / period-derivative search. */
typedef struct {
    int numtrials;              /* per axis: 2 npfact proflen + 1 */
    int ip, ipd, ipdd;          /* the best trial's period, p-dot, p-dotdot indices */
    double p, pd, pdd;          /* its topocentric period, p-dot, p-dotdot */
    accel_foldstats stats;      /* the combined profile's statistics */
} accel_fold_best;
/* prepfold's search of the folded sub-integrations for the best period
 * and period derivative (src/prepfold_pipeline.c:1633-1915, one subband,
 * p-dotdot search with searchpdd; combine_profs, src/fold.c:696-754) on the GPU:
 * numtrials = 2 npfact proflen + 1 periods and as many p-dots around the
 * fold's f, fd (fdd kept), spaced so that the last part moves by pstep
 * (periods) or pdstep (p-dots) bins; every trial delays part ii by its
 * phase drift at parttimes[ii] (s from the fold's start), rounds the delay
 * to whole bins, sums the parts and takes the reduced chi-squared.  T is
 * the folded span in seconds.  ppdot (numtrials x numtrials floats,
 * [ipd][ip]) and bestprof (proflen) may be NULL.  With nosearch the centre
 * trial (the fold's own parameters) is returned (and numtrials is 201
 * when proflen > 100, as the reference sizes its plane).  With searchpdd
 * (prepfold -searchpdd) as many p-dotdots are searched as well, the
 * outermost loop, and ppdot holds numtrials^3 floats [ipdd][ipd][ip] (the
 * reference keeps no plane then).  Returns 0 or -1. */
int accel_fold_search(const double *rawfolds, const accel_foldstats *parts, int npart,
                      int proflen, double f, double fd, double fdd, double T,
                      const double *parttimes, int npfact, int pstep, int pdstep, int nosearch,
                      int searchpdd, float *ppdot, double *bestprof, accel_fold_best *best,
                      int device);

#ifdef __cplusplus
}
#endif
#endif
