/*
 * agpu.h -- the thin C ABI between the C host (accel_host.c) and the HIP
 * kernels of accel_gpu.hip.  Plain pointers, sizes and opaque stream/event
 * handles; no HIP types cross it.  All functions return 0 on success and a
 * negative code on failure; agpu_last_error() describes the failure.
 */
#ifndef PRESTO_AGPU_H
#define PRESTO_AGPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One block-median normalisation window: bins [lobin, lobin+numdata) of
 * the spectrum (absolute bin numbers), zero outside [lobin_file, numbins). */
typedef struct {
    long long lobin;
    int numdata;
    int pad;
} agpu_window;

/* One threshold crossing of the harmonic-sum kernel. */
typedef struct {
    int block;   /* fundamental block of the reference's block loop */
    int stage;   /* log2(number of harmonics summed) */
    int wi, zi;  /* w and z indices on the fundamental grid */
    int ii;      /* half-bin index inside the block */
    float pow;   /* summed normalised power */
} agpu_hit;

/* Raw plane layout: tiles of AGPU_TILE half-bins, each tile holding every
 * (w,z) row contiguously, each tile-row followed by a halo of AGPU_HALO
 * half-bins that repeats the start of the same row in the next tile.
 * Element (row, hb) lives at
 *   ((hb / AGPU_TILE) * nrows + row) * AGPU_ROWLEN + hb % AGPU_TILE
 * (and, for hb % AGPU_TILE < AGPU_HALO, also in the previous tile's halo,
 * which the correlation writes with the element).
 * All rows of a tile share a few 2 MiB pages, which keeps the harmonic-sum
 * gathers (16 rows at a time) inside the translation caches, and the halo
 * lets every sub-harmonic span (at most 1088 half-bins) be read from one
 * contiguous tile-row. */
#define AGPU_TILE_LOG2 13
#define AGPU_TILE (1 << AGPU_TILE_LOG2)
#define AGPU_HALO 1088
#define AGPU_ROWLEN (AGPU_TILE + AGPU_HALO)
#define AGPU_PLANE_IDX(row, hb, nrows)                                                  \
    (((size_t) ((hb) >> AGPU_TILE_LOG2) * (size_t) (nrows) + (size_t) (row)) *         \
         (size_t) AGPU_ROWLEN + (size_t) ((hb) & (AGPU_TILE - 1)))

#define AGPU_MAX_FRACS 32
#define AGPU_MAX_STAGES 6

/* Harmonic-sum / threshold launch description. */
typedef struct {
    const float *plane;      /* raw plane, tiled (AGPU_PLANE_IDX), column = absolute half-bin */
    long long ntiles;        /* tiles allocated (one past the last searched half-bin) */
    int numz, numw;          /* numw >= 1 */
    long long rh0;           /* first fundamental half-bin (2 * the first block's plane rlo) */
    int rh_off;              /* standard path: sub-harmonic lookups are taken at
                              * fundamental half-bin + rh_off (0 or 1) */
    const double *bstart;    /* standard path, start r off the half-bin grid: each block's
                              * start r [nblocks]; the lookups are then evaluated in FP64
                              * per cell as the reference does */
    int corr_uselen;         /* half-bins per block */
    int nblocks;
    int nstages;
    int inmem;               /* 1: in-memory index rules and norms */
    float powcut[AGPU_MAX_STAGES];
    int nfrac;               /* sub-harmonic fractions, stage order */
    int frac_k[AGPU_MAX_FRACS], frac_log2h[AGPU_MAX_FRACS];
    int stage_nfrac[AGPU_MAX_STAGES];  /* fractions summed through stage s */
    const int *zmap;         /* [nfrac][numz]: plane z row of the sub-harmonic */
    const int *wmap;         /* [nfrac][numw] */
    const float *nsq;        /* standard: [nblocks][nfrac+1]; in-memory: [npre+nblocks] */
    const float *nsq_hb;     /* in-memory: norm^2 for every half-bin [0, hb_end) */
    int npre;                /* in-memory pre-population blocks (start at half-bin 16) */
    agpu_hit *hits;
    unsigned int *nhits;
    unsigned int cap;
    /* address tables (agpu_harmsum_row_table / agpu_harmsum_chunk_tables) */
    const unsigned *row_tab;       /* [numz numw][AGPU_HS_RT] */
    const long long *chunk_base;   /* [nchunks][AGPU_HS_CT] */
    const int *chunk_lo;           /* [nchunks][AGPU_HS_CT] */
} agpu_harmsum_args;

/* Row table of the harmonic sum, per (w, z) row: [0] w index, [1] z index,
 * [AGPU_HS_RT_ROW] the row's offset in a tile (row x AGPU_ROWLEN floats),
 * [AGPU_HS_RT_SUB + f] the offset of fraction f's sub-harmonic row.  Chunk
 * tables, per chunk of 1024 fundamental half-bins: base[0] the plane offset
 * (floats) of the chunk's first half-bin in row 0's tile, base[1 + f] that
 * of fraction f's span start lo[1 + f] (its first half-bin). */
#define AGPU_HS_RT 36
#define AGPU_HS_RT_ROW 2
#define AGPU_HS_RT_SUB 3
#define AGPU_HS_CT 32
int agpu_harmsum_row_table(int numz, int numw, int nfrac, const int *zmap, const int *wmap,
                           unsigned *out);
/* mode: 0 standard, 1 in-memory, 2 standard with per-cell FP64 lookups;
 * rh_off as in agpu_harmsum_args */
int agpu_harmsum_chunk_tables(int mode, int rh_off, long long rh0, long long nchunks,
                              int nstages, int nrows, long long *base, int *lo);

int agpu_init(int device);
const char *agpu_last_error(void);
void agpu_set_error(const char *msg);
int agpu_device_name(char *buf, int len);

int agpu_malloc(void **p, size_t bytes);
/* One device block reserved up front; while it exists every agpu_malloc /
 * agpu_free is carved from / returned to it (no hipMalloc / hipFree). */
int agpu_arena_init(size_t bytes);
int agpu_arena_release(void);
size_t agpu_arena_used(void);
int agpu_free(void *p);
int agpu_h2d(void *dst, const void *src, size_t bytes, void *stream);
int agpu_d2h(void *dst, const void *src, size_t bytes, void *stream);
int agpu_memset(void *p, int v, size_t bytes, void *stream);
int agpu_stream_create(void **s);
int agpu_stream_destroy(void *s);
int agpu_stream_sync(void *s);
/* work queued on stream s after this waits for event e */
int agpu_stream_wait_event(void *s, void *e);
int agpu_event_create(void **e);
int agpu_event_destroy(void *e);
int agpu_event_record(void *e, void *stream);
int agpu_event_elapsed(float *ms, void *start, void *stop);

/* FP64 response templates for one segment length L (power of two):
 * T[t][f] = conj(FFT_L(placed kernel of (z[t], w[t])))[f] / L, rounded to
 * FP32.  |w| = 0: Fresnel z-response; otherwise the FP64 DFT of the
 * reference's FP32-sampled chirp.  zs/ws/numkern are host arrays of nt
 * entries. */
int agpu_gen_templates(const int *zs, const int *ws, const int *numkern, int nt, int L,
                       void *d_T, void *stream);
/* The same responses as taps, rounded to FP32 as the reference stores
 * them: R[t][k] for k < numkern[t] (centre at numkern[t]/2), zero up to the
 * row stride. */
int agpu_gen_response_taps(const int *zs, const int *ws, const int *numkern, int nt,
                           int stride, void *d_R, void *stream);

/* Exact block medians of |x|^2 (FP32 powers) and the reference's
 * normalisation: norm = 1/sqrt(median/ln 2) (FP64), nsq = (float)(norm^2).
 * maxn bounds every window's numdata (at most 15360). */
int agpu_block_norms(const void *d_spec, long long numbins, long long lobin_file,
                     const agpu_window *d_win, int nwin, int maxn, double *d_norm,
                     float *d_nsq, void *stream);

/* The same for one window given by value: d_norm[0], d_nsq[0]. */
int agpu_block_norm1(const void *d_spec, long long numbins, long long lobin_file,
                     long long lobin, int numdata, double *d_norm, float *d_nsq, void *stream);

/* ---- accel.h surface (surface_gpu.hip); planes are [numws][numzs][numrs] ---- */

/* Block plane from the raw plane: out = nsq[0] * raw(row (w+woff)*numz_plane
 * + z+zoff, half-bin hb0 + r). */
int agpu_sf_extract(const float *d_plane, int nrows, long long ntiles, int numz_plane, int woff,
                    int zoff, int numws, int numzs, int numrs, long long hb0, const float *d_nsq,
                    float *d_out, void *stream);

/* add_ffdotpows: the sub-harmonic's index tables are rebuilt on the device
 * from the values the reference builds them from. */
typedef struct {
    double hf;              /* harmnum / numharm */
    double sub_fullrlo;     /* fullrlo the sub-harmonic plane was made for */
    long long sub_rlo;
    int sub_zlo, sub_wlo;
    double obs_zlo;
    int fund_wlo;
    int fnumws, fnumzs, fnumrs;
    int snumws, snumzs, snumrs;
} agpu_sf_add_args;
int agpu_sf_add(float *d_fund, const float *d_sub, const agpu_sf_add_args *a, void *stream);

/* fund_to_ffdotplane / inmem_add_ffdotpows on a device full plane of
 * numz rows of rlen half-bins */
int agpu_sf_inmem_store(const float *d_fund, int numzs, int numrs, float *d_full,
                        long long rlen, long long col0, void *stream);
int agpu_sf_inmem_add(float *d_fund, int numzs, int numrs, const float *d_full, long long rlen,
                      int rlo, double hf, int zlo, void *stream);

/* Every element of a plane above powcut (unordered; the count may exceed
 * cap, then only cap entries are written). */
typedef struct {
    long long idx;
    float pow;
    int pad;
} agpu_sf_hit;
int agpu_sf_threshold(const float *d_pow, long long n, float powcut, agpu_sf_hit *d_hits,
                      unsigned *d_count, unsigned cap, void *stream);
int agpu_copy_d2d(void *dst, const void *src, size_t bytes, void *stream);
/* -locpow: the window's numdata bins from lobin, each scaled by the
 * inverse square root of its local power (surface_gpu.hip) */
int agpu_sf_locpow_norm(const void *d_spec, long long numbins, long long lobin_file,
                        long long lobin, int numdata, void *d_wdata, void *stream);
/* -locpow block plane [numzs][numrs] from a window's locally normalised
 * bins (numdata float2), z rows zlo + 2 i (surface_gpu.hip) */
int agpu_sf_locpow_plane(const void *d_wdata, int numdata, int binoffset, int zlo, int numzs,
                         int numrs, float *d_out, void *stream);
/* -locpow block plane [nrows][numrs] of a jerk window: row r correlates
 * with d_nk[r] FP32 taps at d_R + r*stride (surface_gpu.hip) */
int agpu_sf_locpow_plane_tab(const void *d_wdata, int numdata, int binoffset, int nrows,
                             int numrs, const void *d_R, const int *d_nk, int stride,
                             float *d_out, void *stream);

/* Fused overlap-save correlation: raw |corr|^2 plane rows for templates
 * [0, nt) (row = template), half-bins [0, hb_end), segment length L = 4096,
 * written in the tiled layout above, halos included.  Only segments [seg_lo, seg_hi)
 * (segment s: half-bins [s uselen, (s+1) uselen); seg_hi -1: to the end). */
int agpu_correlate(const void *d_spec, long long numbins, long long lobin_file,
                   const void *d_T, int nt, int L, int hw, int uselen, long long hb_end,
                   float *d_plane, long long ntiles, int groups, long long seg_lo,
                   long long seg_hi, void *stream);

/* Harmonic sums of fundamental chunks [c_lo, c_hi) (-1: to the end); chunk
 * c covers the 1024 half-bins from agpu_harmsum_chunk_hb(a, c). */
int agpu_harmsum(const agpu_harmsum_args *a, long long c_lo, long long c_hi, void *stream);
long long agpu_harmsum_nchunks(const agpu_harmsum_args *a);
long long agpu_harmsum_chunk_hb(const agpu_harmsum_args *a, long long c);

/* In-memory path: per-half-bin norm^2 of the block that last wrote it. */
int agpu_inmem_nsq_table(const float *d_nsq, int npre, int nblocks, long long rh0, int cu,
                         long long hb_end, float *d_out, void *stream);

/* De-reddening of a device spectrum in place (src/misc_utils.c:912-1067):
 * d_blk = the reference's median blocks {first bin, length} tiling
 * [1, numbins), maxn = the longest; d_mean receives nblk floats. */
int agpu_deredden(void *d_spec, long long numbins, const agpu_window *d_blk, int nblk,
                  int maxn, float *d_mean, void *stream);

/* Complex values of work agpu_realfft and agpu_realfft_inverse need for
 * an n-sample series. */
long long agpu_realfft_work(long long n);
/* Forward real FFT of n (even) FP32 samples in PRESTO's packing
 * into d_out (n/2 complex); d_work holds agpu_realfft_work(n) complex
 * values. */
int agpu_realfft(const float *d_x, long long n, void *d_out, void *d_work, void *stream);
/* Its inverse (realfft with isign = +1, scaled by 2/n): n/2 complex bins in
 * PRESTO's packing to n samples; d_X may be d_out; the same d_work size. */
int agpu_realfft_inverse(const void *d_X, long long n, float *d_out, void *d_work, void *stream);

/* Candidate refinement (harmonic polishing) on the device: max_rz_arr_harmonics
 * with one workgroup per harmonic of a candidate, max_rzw_arr_harmonics as a
 * sequence of kernels over many workgroups.
 * spec: the resident spectrum (numdata complex FP32 bins, bin 0 = lobin);
 * rin = candidate r - lobin; per candidate up to AGPU_REFINE_MAXH
 * harmonics; derivs[c][h][7] = pow, phs, dpow, dphs, d2pow, d2phs, locpow;
 * powers[c][h]; all output arrays are device memory. */
#define AGPU_REFINE_MAXH 32
typedef struct {
    const void *spec;
    long long numdata;
    int ncand;
    const int *numharm;
    const double *rin, *zin, *win;
    double *rout, *zout, *wout;
    double *derivs;
    double *powers;
    int *nevals;
    /* agpu_refine_rz_scratch(ncand) bytes (the candidates' group records)
     * or, for jerk refinement, agpu_refine_rzw_scratch(ncand, npts_cap) */
    void *scratch;
    int npts_cap;
    /* 1: one harmonic per candidate without polishing (-noharmpolish):
     * max_rz_arr (src/maximize_rz.c:103-136: the raw power, not over the
     * local power) or max_rzw_arr (src/maximize_rzw.c:3-55: the 9 x 9 x 24
     * grid of zin +- DZ, win +- DW, raw power, derivatives with the start's
     * local power, powers = the grid maximum); numharm must be 1 */
    int raw;
    /* the largest numharm of the candidates (workgroups per candidate) */
    int maxharm;
} agpu_refine_args;

/* max_rz_arr_harmonics per candidate, one workgroup per harmonic (a
 * cooperative launch); scratch must hold agpu_refine_rz_scratch(ncand)
 * bytes.  A candidate whose workgroups could not keep in step gets
 * rout = NaN, nevals = -1. */
size_t agpu_refine_rz_scratch(int ncand);
int agpu_refine_rz(const agpu_refine_args *args, void *stream);
/* The first attempt's barrier wait of agpu_refine_rz, 2^log2_polls polls
 * (0..24, default 24); candidates whose wait runs out are run again with
 * the full bound.  Lowered only to test that retry. */
void agpu_refine_set_first_wait(int log2_polls);
/* max_rzw_arr_harmonics per candidate, spread over many workgroups per
 * candidate; maxharm: the largest numharm of the candidates.  scratch must
 * hold agpu_refine_rzw_scratch(ncand, npts_cap) bytes.  A candidate whose
 * chirp length exceeds npts_cap gets rout = NaN, nevals = -1. */
size_t agpu_refine_rzw_scratch(int ncand, int npts_cap);
int agpu_refine_rzw(const agpu_refine_args *args, void *stream);

/* Subband de-dispersion (dedisp_gpu.hip).  raw: [nspec][numchan] FP32,
 * ascending frequency; ds: [nsub][nds]; out: [numdms][numout].  maxdelay /
 * maxoff bound idispdt / offsets, which are >= 0. */
/* the spectra the subbands read: the length of rawT's rows */
long long agpu_dedisp_ncols(long long nds, int downsamp, int maxdelay);
/* 1 when the subband sums need the channel-major copy d_rawT (a
 * subband wider than 64 channels, or an LDS tile of 256 output samples
 * plus maxdelay rows that does not fit); 0: they read d_raw through LDS
 * tiles and d_rawT may be NULL */
int agpu_dedisp_needs_transpose(int numchan, int nsub, int downsamp, int maxdelay);
/* the subband sums; through d_rawT [numchan][ncols] when the transpose is
 * needed */
int agpu_dedisp_subbands(const float *d_raw, long long nspec, int numchan, const int *d_idispdt,
                         int maxdelay, int nsub, int downsamp, long long nds, float *d_rawT,
                         float *d_ds, void *stream);
/* span: the largest spread of one subband's offsets over AGPU_DD_GROUP
 * consecutive DMs (groups from DM 0); it selects the LDS kernel */
#define AGPU_DD_GROUP 16
int agpu_dedisp_dms(const float *d_ds, long long nds, int nsub, const int *d_offsets,
                    int maxoff, int span, int numdms, long long nwrite, long long numout,
                    float *d_out, void *stream);
int agpu_dedisp_fill(float *d_out, int numdms, long long numout, long long from, float value,
                     void *stream);


/* prepfold's fold of a time series for one candidate (fold_gpu.hip):
 * records of 1024 points, npart sub-integrations of rpp records each;
 * part_end[p]: the phase added since the fold start at the end of part p;
 * rawfolds: npart x proflen FP64, zeroed by the caller; recsums: per
 * record the sum and the sum of squares of its points. */
typedef struct {
    const float *series;
    double dt, f, fd, fdd, phi_start;
    int npart, rpp, proflen;
    const double *part_end;
    double *rawfolds;
    double *recsums;
} agpu_fold_args;
int agpu_fold(const agpu_fold_args *args, void *stream);
/* prepfold's p / p-dot search over one subband's sub-integrations:
 * redchi[ipd][ip] = reduced chi-squared of trial (ip, ipd) of numtrials x
 * numtrials; profs npart x proflen, fdots and parttimes are device
 * arrays; prof_avg, prof_var the parts' sums. */
typedef struct {
    const double *profs, *fdots, *fdotdots, *parttimes;   /* fdotdots: NULL, no p-dotdot */
    double prof_avg, prof_var;
    int npart, proflen, numtrials, pstep;
    int ipdd;             /* the p-dotdot trial of this launch (with fdotdots) */
    double *redchi;       /* numtrials x numtrials, [ipd][ip], of trial ipdd */
} agpu_fold_search_args;
int agpu_fold_search(const agpu_fold_search_args *args, void *stream);

#ifdef __cplusplus
}
#endif
#endif
