/*
 * oracle.h -- CPU restatement of PRESTO's accelsearch search path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * path is compared against (tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py).  No product code links, loads or calls
 * it.  Every function cites the reference file:line whose behaviour it
 * restates (paths relative to the PRESTO source tree).
 *
 * Two arithmetic modes run through the same code:
 *   fp64 = 0  "FP32-faithful": FP32 data, FP32 FFTs with FP64-computed
 *             twiddles, FP32 powers -- the reference's own arithmetic
 *             (the FFT algorithm differs from FFTW's, which is itself
 *             not bit-stable run to run: src/fftcalls.c:130).
 *   fp64 = 1  everything in FP64: the plane-power oracle.
 */
#ifndef PRESTO_ORACLE_H
#define PRESTO_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float r, i; } o_fcomplex;
typedef struct { double r, i; } o_dcomplex;

/* ---- constants (include/accel.h:18-35, include/presto.h:17-107) ---- */
#define O_NUMBETWEEN 2
#define O_DR 0.5
#define O_RDR 2
#define O_DZ 2
#define O_RDZ 0.5
#define O_DW 20
#define O_RDW 0.05
#define O_CLOSEST_R 15.0
#define O_DBLCORRECT 1e-14
#define O_NUMFINTBINS 16
#define O_NUMLOCPOWAVG 20
#define O_DELTAAVGBINS 5
#define O_MAXRAMUSE 2000000000LL

/* ---- math: o_math.c ---- */
int o_fresnl(double x, double *s, double *c);
int o_r_resp_halfwidth(int highacc);
int o_z_resp_halfwidth(double z, int highacc);
int o_w_resp_halfwidth(double z, double w, int highacc);
void o_gen_r_response(double roffset, int numbetween, int numkern, o_dcomplex *out);
void o_gen_z_response(double roffset, int numbetween, double z, int numkern,
                      o_dcomplex *out);
/* out32: the reference's FP32 construction; out64: FP64 DFT of the same
 * FP32-sampled chirp.  Either may be NULL. */
void o_gen_w_response(double roffset, int numbetween, double z, double w,
                      int numkern, o_fcomplex *out32, o_dcomplex *out64);
long long o_next2_to_n(long long x);
int o_next_good_fftlen(int n);
int o_fftlen_from_kernwidth(int kernwidth);
double o_gauss_Q(double x);
double o_gauss_Qinv(double p);
double o_chisq_Q_even(double chi2, double dof);
double o_chisq_Qinv_even(double q, double dof);
double o_chi2_logp(double chi2, double dof);
double o_equivalent_gaussian_sigma(double logp);
double o_candidate_sigma(double power, int numsum, double numtrials);
double o_power_for_sigma(double sigma, int numsum, double numtrials);
float o_median(float *arr, int n);
float o_invsqrtf(float x);

/* ---- FFTs: o_fft.c ---- */
void o_cfft_f(o_fcomplex *data, long n, int isign);
double o_time_cfft_ns(long n, int reps);
void o_cfft_d(o_dcomplex *data, long n, int isign);
void o_realfft_f(float *data, long n, int isign);
long o_deredden(o_fcomplex *fft, long numbins, int startwidth, int endwidth,
                double endfreq, double T);

/* ---- the search: o_accel.c ---- */
typedef struct {
    /* inputs (command line + .inf) */
    double N, dt;
    long long numbins;
    int numharm, zmax, wmax, wmaxP;
    float sigma;
    double flo, fhi, rlo_arg, rhi_arg;
    int floP, fhiP, rloP, rhiP;
    long long lobin_arg;
    int inmemP;
    int force_standard; /* test hook: never choose the in-memory path */
    int dat_input;
    /* derived by o_create_accelobs */
    long long lobin, highestbin;
    int maxkernlen, corr_uselen, fftlen, numharmstages, numz, numw, numbetween, inmem;
    double T, rlo, rhi, dr, zlo, zhi, dz, wlo, whi, dw;
    float powcut[6];
    long long numindep[6];
    /* -photon: the photon count whose 1/sqrt normalises every block
     * (src/accel_utils.c:1080-1086); 0 selects block medians.  Set before
     * o_create_accelobs, which clears it when lobin > 0 (:1626-1628). */
    double nph;
    /* -locpow: 1 normalises every bin of a window by invsqrtf of its
     * double-tophat local power (src/accel_utils.c:1100-1115); 0 medians */
    int norm_type;
} o_obs;

/* corr_loc_pow (src/corr_routines.c:309-366) for one window's FP32 powers:
 * (2/n) floor((n/2)/20) times the sum of the powers 6..15 bins either
 * side, the first and last 15 values set to their neighbours'.  The
 * reference sums with an FP32 FFT; this sums exactly, then rounds. */
void o_corr_loc_pow(const float *powers, int n, float *loc);

int o_create_accelobs(o_obs *obs);

typedef struct {
    int z, w, fftlen, numgoodbins, numbetween, kern_half_width;
    o_fcomplex *data;   /* fp32 mode */
    o_dcomplex *ddata;  /* fp64 mode */
} o_kernel;

typedef struct {
    int numharm, harmnum, zmax, wmax, numkern_zdim, numkern_wdim, numkern;
    o_kernel *kern;     /* [numkern_wdim][numkern_zdim] */
    unsigned short *rinds, *zinds;
} o_shi;

o_shi **o_create_subharminfos(const o_obs *obs, int fp64);
void o_free_subharminfos(const o_obs *obs, o_shi **shis);

typedef struct {
    long long rlo;
    int zlo, wlo, numrs, numzs, numws;
    float *powers;      /* fp32 mode, [numws][numzs][numrs] */
    double *dpowers;    /* fp64 mode */
    unsigned short *rinds, *zinds;
} o_ffdot;

o_ffdot *o_subharm_fderivs_vol(int numharm, int harmnum, double fullrlo,
                               double fullrhi, o_shi *shi, const o_obs *obs,
                               const o_fcomplex *fft, int fp64);
void o_free_ffdot(o_ffdot *f);
void o_add_ffdotpows(o_ffdot *fund, const o_ffdot *sub, int numharm, int harmnum);
void o_fund_to_ffdotplane(const o_ffdot *ffd, const o_obs *obs, float *plane,
                          double *dplane);
void o_inmem_add_ffdotpows(o_ffdot *fund, const o_obs *obs, const float *plane,
                           const double *dplane, int numharm, int harmnum);
/* block-median normalisation factor of the window the reference uses */
double o_block_norm(const o_obs *obs, const o_fcomplex *fft, long long lobin,
                    int numdata);

typedef struct o_cand {
    float power, sigma;
    int numharm;
    double r, z, w;
    struct o_cand *next;
} o_cand;

o_cand *o_insert_new_accelcand(o_cand *list, float power, float sigma, int numharm,
                               double rr, double zz, double ww, int *added);
o_cand *o_sort_accelcands(o_cand *list);
o_cand *o_eliminate_harmonics(o_cand *list, int *numcands);
void o_free_cands(o_cand *list);

typedef struct {
    int block, stage, wi, zi, ri;
    float pow, sigma;
} o_hit;

typedef struct {
    o_hit *hits;
    long long nhits, caphits;
    o_cand *cands;
    int nblocks;
} o_result;

/* The block loop of src/accelsearch.c:161-228.  max_blocks < 0: all. */
int o_accelsearch(const o_obs *obs, const o_fcomplex *fft, int fp64, int max_blocks,
                  o_result *res);
int o_accelsearch_range(const o_obs *obs, const o_fcomplex *fft, int fp64, int first_block,
                        int max_blocks, o_result *res);
void o_free_result(o_result *res);
void o_set_threads(int n);
long long o_prepare_dat(float *series, long long n, double dt);
/* flat accessors for ctypes */
int o_result_ncands(const o_result *res);
void o_result_cands(const o_result *res, double *out7);  /* power,sigma,numharm,r,z,w,0 */

/* ---- refinement: o_refine.c ---- */
typedef struct {
    double pow, phs, dpow, dphs, d2pow, d2phs, locpow;
} o_rderivs;   /* include/presto.h:160-168 */
typedef struct {
    double r;
    float rerr;
    double z;
    float zerr;
    double w;
    float werr;
    float pow, powerr, sig, rawpow, phs, phserr, cen, cenerr, pur, purerr, locpow;
} o_fourierprops;   /* include/presto.h:170-188 */
typedef struct {
    double p, pd, pdd, f, fd, fdd;
    float perr, pderr, pdderr, ferr, fderr, fdderr;
} o_rzwerrs;   /* include/presto.h:233-246 */

void o_rz_interp(const o_fcomplex *data, long numdata, double r, double z, int khw,
                 float *ar, float *ai);
void o_rzw_interp(const o_fcomplex *data, long numdata, double r, double z, double w,
                  int khw, float *ar, float *ai);
double o_localpower3d(const o_fcomplex *data, long numdata, double r, double z, double w);
void o_derivs3d(const o_fcomplex *data, long numdata, double r, double z, double w,
                double locpow, o_rderivs *res);
int o_calc_props(const o_rderivs *d, double r, double z, double w, o_fourierprops *p);
int o_calc_rzwerrs(const o_fourierprops *p, double T, o_rzwerrs *e);
double o_nm_minimize(double (*f)(const double *x, void *par), void *par, double *x,
                     double step, double tol, int maxiter, int *iters);
double o_max_rz_arr(const o_fcomplex *data, long numdata, double rin, double zin,
                    double *rout, double *zout, o_rderivs *derivs);
long o_max_rz_arr_harmonics(const o_fcomplex *data, long numdata, int numharm, double rin,
                            double zin, double *rout, double *zout, o_rderivs *derivs,
                            double *powers);
void o_max_rzw_arr_harmonics(const o_fcomplex *data, long numdata, int numharm, double rin,
                             double zin, double win, double *rout, double *zout, double *wout,
                             o_rderivs *derivs, double *powers);
double o_max_rzw_arr(const o_fcomplex *data, long numdata, double rin, double zin, double win,
                     double *rout, double *zout, double *wout, o_rderivs *derivs);

#ifdef __cplusplus
}
#endif

/* prepfold's fold of a .dat for an ACCEL candidate (o_fold.c) */
typedef struct {
    double numdata, data_avg, data_var, numprof, prof_avg, prof_var, redchi;
} o_foldstats;
void o_cand_fold_params(double r, double z, double w, double T, double *f, double *fd,
                        double *fdd);
long long o_fold_dat(const float *series, long long n, double dt, double f, double fd,
                     double fdd, int npart, int proflen, double *rawfolds, o_foldstats *stats);
/* prepfold's p / p-dot (and with searchpdd p-dotdot) search over the
 * sub-integrations of one subband: ppdot[ipdd][ipd][ip] = reduced
 * chi-squared of each trial (numtrials = 2 npfact proflen + 1, returned;
 * one ipdd plane without searchpdd); best = {ip, ipd, redchi, p, pd, pdd,
 * ipdd} of the first maximum in (ipdd, ipd, ip) order. */
int o_fold_search(const double *profs, const o_foldstats *parts, int npart, int proflen,
                  double f, double fd, double fdd, double T, const double *parttimes,
                  int npfact, int pstep, int pdstep, int searchpdd, float *ppdot,
                  double *best);

#endif
