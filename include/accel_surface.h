/*
 * accel_surface.h -- the accel.h surface of PRESTO's accelsearch
 * (include/accel.h:37-164), implemented on the MI355X path by
 * libpresto_accel.so.
 *
 * Every function below keeps the name, the argument meaning and the
 * ownership rules of the reference function it replaces, so that the
 * reference's driver (src/accelsearch.c:82-313), compiled against the
 * reference's own accel.h, links against this library in place of
 * src/accel_utils.c.  The structs are declared here with the same member
 * order and types as the reference's, hence the same x86-64 layout
 * (accel_surface_layout() reports it for the layout test).
 *
 * What differs from src/accel_utils.c, by design:
 *   - Planes live in device memory.  subharm_fderivs_vol() returns an
 *     ffdotpows whose powers are resident on the GPU (powers == NULL);
 *     add_ffdotpows(), fund_to_ffdotplane(), inmem_add_ffdotpows() and
 *     search_ffdotpows() work on them there.  accel_ffdotpows_sync()
 *     copies a plane to the host in the reference's float*** layout.
 *   - obs->ffdotplane (in-memory search) is device-resident too and stays
 *     NULL on the host.
 *   - subharminfo kernels are the device template bank of the search: each
 *     kernel's data is NULL; z, w, fftlen, numgoodbins, numbetween and
 *     kern_half_width are filled as the reference fills them.
 *   - The spectrum is uploaded to the GPU by create_subharminfos(), after
 *     the caller's zapping (src/accelsearch.c:86-111).
 *   - -photon normalises every block by 1/sqrt(nph), as the reference
 *     does.  With -locpow each subharm_fderivs_vol() normalises its window
 *     bin by bin on the GPU and correlates it directly with every row's
 *     FP32 response taps (z rows and, with -wmax, the jerk rows).
 *   - Errors print the reference's message and call exit() with its code,
 *     as the reference does; accel_surface_set_exit(0) makes them return
 *     instead, with accel_last_error() set and accel_surface_failed() true.
 *
 * GSList nodes are glib's { data, next } pair.  When the process has glib
 * loaded (the reference's accelsearch links it), nodes are allocated and
 * freed with glib's own g_slist_alloc() / g_slist_free_1(), so the caller's
 * g_slist_free() releases them; otherwise with calloc()/free(), and
 * accel_gslist_free() releases a list.
 */
#ifndef PRESTO_ACCEL_SURFACE_H
#define PRESTO_ACCEL_SURFACE_H

#include <stdio.h>
#include "presto_accel.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef accel_fcomplex fcomplex;          /* include/ransomfft.h:13-18 */
typedef accel_rderivs rderivs;            /* include/presto.h:160-168 */
typedef accel_fourierprops fourierprops;  /* include/presto.h:170-188 */

/* glib's singly linked list node (glib/gslist.h) */
typedef struct _GSList GSList;
struct _GSList {
    void *data;
    GSList *next;
};

#define MAXNUMONOFF 40

/* include/makeinf.h:23-56 */
typedef struct INFODATA {
    double ra_s, dec_s, N, dt, fov, mjd_f, dm, freq, freqband, chan_wid;
    double wavelen, waveband, energy, energyband;
    double onoff[MAXNUMONOFF * 2];
    int num_chan, mjd_i, ra_h, ra_m, dec_d, dec_m, bary, numonoff;
    char notes[500];
    char name[200];
    char object[100];
    char instrument[100];
    char observer[100];
    char analyzer[100];
    char telescope[40];
    char band[40];
    char filt[7];
} infodata;

/* include/accelsearch_cmd.h:12-80 (clig/accelsearch_cmd.cli:13-48) */
typedef struct s_Cmdline {
    int ncpusP, ncpus, ncpusC;
    int lobinP, lobin, lobinC;
    int numharmP, numharm, numharmC;
    int zmaxP, zmax, zmaxC;
    int wmaxP, wmax, wmaxC;
    int sigmaP;
    float sigma;
    int sigmaC;
    int rloP;
    double rlo;
    int rloC;
    int rhiP;
    double rhi;
    int rhiC;
    int floP;
    double flo;
    int floC;
    int fhiP;
    double fhi;
    int fhiC;
    int inmemP;
    int photonP;
    int medianP;
    int locpowP;
    int zaplistP;
    char *zaplist;
    int zaplistC;
    int baryvP;
    double baryv;
    int baryvC;
    int otheroptP;
    int noharmpolishP;
    int noharmremoveP;
    int argc;
    char **argv;
    char *full_cmd_line;
} Cmdline;

/* include/accel.h:37-81 */
typedef struct accelobs {
    long long N, numbins, lobin, highestbin;
    int maxkernlen, corr_uselen, fftlen, numharmstages, numz, numw, numbetween, numzap;
    int dat_input, mmap_file, inmem, norm_type;
    double dt, T, rlo, rhi, dr, zlo, zhi, dz, wlo, whi, dw, baryv;
    float nph, sigma;
    float *powcut;
    float *ffdotplane;
    double *lobins, *hibins;
    long long *numindep;
    FILE *fftfile, *workfile;
    fcomplex *fft;
    char *rootfilenm, *candnm, *accelnm, *workfilenm;
    int use_harmonic_polishing;
} accelobs;

/* include/accel.h:83-95 */
typedef struct accelcand {
    float power, sigma;
    int numharm;
    double r, z, w;
    double *pows, *hirs, *hizs, *hiws;
    rderivs *derivs;
} accelcand;

/* include/accel.h:97-105 */
typedef struct kernel {
    int z, w, fftlen, numgoodbins, numbetween, kern_half_width;
    fcomplex *data;
} kernel;

/* include/accel.h:107-118 */
typedef struct subharminfo {
    int numharm, harmnum, zmax, wmax, numkern_zdim, numkern_wdim, numkern;
    kernel **kern;
    unsigned short *rinds, *zinds;
} subharminfo;

/* include/accel.h:120-130 */
typedef struct ffdotpows {
    long long rlo;
    int zlo, wlo, numrs, numzs, numws;
    float ***powers;
    unsigned short *rinds, *zinds;
} ffdotpows;

/* ---- include/accel.h:134-164, each replacing the src/accel_utils.c function
 * of the same name (line numbers of its definition there) ---- */
subharminfo **create_subharminfos(accelobs *obs);                        /* :240 */
void free_subharminfos(accelobs *obs, subharminfo **shis);              /* :308 */
void create_accelobs(accelobs *obs, infodata *idata, Cmdline *cmd, int usemmap); /* :1455 */
GSList *sort_accelcands(GSList *list);                                  /* :380 */
GSList *eliminate_harmonics(GSList *cands, int *numcands);              /* :472 */
void optimize_accelcand(accelcand *cand, accelobs *obs);                /* :552 */
void output_fundamentals(fourierprops *props, GSList *list, accelobs *obs,
                         infodata *idata);                               /* :666 */
void output_harmonics(GSList *list, accelobs *obs, infodata *idata);    /* :844 */
void free_accelcand(void *data, void *user_data);                       /* :350 */
void print_accelcand(void *data, void *user_data);                      /* :982 */
fcomplex *get_fourier_amplitudes(long long lobin, int numbins, accelobs *obs); /* :993 */
ffdotpows *subharm_fderivs_vol(int numharm, int harmnum, double fullrlo, double fullrhi,
                               subharminfo *shi, accelobs *obs);         /* :1024 */
ffdotpows *copy_ffdotpows(ffdotpows *orig);                             /* :1200 */
void fund_to_ffdotplane(ffdotpows *ffd, accelobs *obs);                 /* :1219 */
void inmem_add_ffdotpows(ffdotpows *fundamental, accelobs *obs, int numharm,
                         int harmnum);                                   /* :1312 */
void fund_to_ffdotplane_trans(ffdotpows *ffd, accelobs *obs);           /* :1236 */
void inmem_add_ffdotpows_trans(ffdotpows *fundamental, accelobs *obs, int numharm,
                               int harmnum);                             /* :1364 */
void free_ffdotpows(ffdotpows *ffd);                                    /* :1253 */
void add_ffdotpows_ptrs(ffdotpows *fundamental, ffdotpows *subharmonic, int numharm,
                        int harmnum);                                    /* :1284 */
void add_ffdotpows(ffdotpows *fundamental, ffdotpows *subharmonic, int numharm,
                   int harmnum);                                         /* :1261 */
GSList *search_ffdotpows(ffdotpows *ffdot, int numharm, accelobs *obs,
                         GSList *cands);                                 /* :1412 */
void free_accelobs(accelobs *obs);                                      /* :1808 */

/* ---- additions of this library (no counterpart in the reference) ---- */

/* 1 (default): a failing surface call prints the reference's message and
 * exits with its code.  0: it returns (NULL or early), sets
 * accel_last_error() and accel_surface_failed().  Returns the old value. */
int accel_surface_set_exit(int on);
/* Nonzero if a surface call failed since the last accel_surface_clear(). */
int accel_surface_failed(void);
void accel_surface_clear(void);
/* HIP device the surface runs on (default 0); set before create_accelobs. */
int accel_surface_set_device(int device);
/* Copies a device-resident plane to the host in the reference's layout
 * (powers[w][z][r], one contiguous block, as gen_f3Darr makes it); the
 * host copy is freed by free_ffdotpows().  0 or a negative code. */
int accel_ffdotpows_sync(ffdotpows *ffd);
/* readinf / writeinf (src/ioinf.c:121-350): the <root>.inf of a data set
 * into / from an infodata, in the reference's line layout.  0 or -1 with
 * accel_last_error() set. */
int accel_readinf(infodata *d, const char *root);
int accel_writeinf(const infodata *d);
/* Frees the nodes (not the data) of a list built by this library. */
void accel_gslist_free(GSList *list);
/* Layout report for the structs above: for struct `name`, writes its size
 * and each member's offset, in declaration order, into out[0..n); returns
 * the number of values written or -1 for an unknown name. */
int accel_surface_layout(const char *name, long long *out, int n);

#ifdef __cplusplus
}
#endif
#endif
