/*
 * accel_surface.c -- the accel.h surface (include/accel.h:134-164) on the
 * MI355X search path.  Each function keeps the name, the arguments and the
 * ownership rules of its src/accel_utils.c namesake, so the reference's
 * driver (src/accelsearch.c:82-313) runs unchanged on top of it.
 *
 * How the block loop maps onto the device:
 *   create_accelobs     geometry (accel_geom_create), input files, a search
 *                       handle; a .dat series is FFT'd and de-reddened on
 *                       the GPU and copied to obs->fft for the caller's
 *                       zapping.
 *   create_subharminfos the subharminfo tables; uploads obs->fft and builds
 *                       the raw |corr|^2 plane of the whole spectrum once.
 *   subharm_fderivs_vol the block's normalisation (exact median on the GPU)
 *                       and its plane, cut from the raw plane and scaled by
 *                       norm^2 -- equal to the reference's correlation of
 *                       the normalised block, since correlation is linear.
 *   add_ffdotpows, inmem_add_ffdotpows, fund_to_ffdotplane
 *                       device gathers with the reference's index rules.
 *   search_ffdotpows    device threshold scan; the crossings are replayed on
 *                       the host in the reference's (w, z, r) loop order
 *                       through insert_new_accelcand.
 *   optimize_accelcand  the FP64 harmonic-polished refinement kernels.
 *   output_*            the ACCEL tables (accel_output.c).
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <fcntl.h>
#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include "accel_surface.h"
#include "accel_rules.h"
#include "accel_internal.h"
#include "agpu.h"

#define ACCEL_PADDING 2000   /* include/accel.h:30 */
#define SF_HIT_CAP0 (1u << 16)

/* ---------------------------------------------------------------------- */
/* errors                                                                  */

static int g_exit = 1;
static int g_failed = 0;
static int g_device = 0;

int accel_surface_set_exit(int on)
{
    int old = g_exit;
    g_exit = on ? 1 : 0;
    return old;
}

int accel_surface_failed(void) { return g_failed; }
void accel_surface_clear(void) { g_failed = 0; }

int accel_surface_set_device(int device)
{
    int old = g_device;
    g_device = device;
    return old;
}

/* The reference prints its message on stdout and calls exit(code); so does
 * this, unless accel_surface_set_exit(0) asked for a return. */
static void sf_fail(int code, const char *fmt, ...)
{
    char m[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(m, sizeof(m), fmt, ap);
    va_end(ap);
    accel_set_error("%s", m);
    g_failed = 1;
    if (g_exit) {
        printf("\n%s\n\n", m);
        fflush(stdout);
        exit(code);
    }
}

/* ---------------------------------------------------------------------- */
/* glib-compatible list nodes                                              */

typedef GSList *(*slist_alloc_fn)(void);
typedef void (*slist_free1_fn)(GSList *);
static slist_alloc_fn g_alloc_fn;
static slist_free1_fn g_free1_fn;
static pthread_once_t g_slist_once = PTHREAD_ONCE_INIT;

static void slist_resolve(void)
{
    void *a = dlsym(RTLD_DEFAULT, "g_slist_alloc");
    void *f = dlsym(RTLD_DEFAULT, "g_slist_free_1");
    if (a && f) {
        g_alloc_fn = (slist_alloc_fn) a;
        g_free1_fn = (slist_free1_fn) f;
    }
}

static GSList *node_new(void *data)
{
    pthread_once(&g_slist_once, slist_resolve);
    GSList *n = g_alloc_fn ? g_alloc_fn() : (GSList *) calloc(1, sizeof(GSList));
    if (!n) {
        sf_fail(-1, "Error allocating a candidate list node");
        return NULL;
    }
    n->data = data;
    n->next = NULL;
    return n;
}

static void node_free(GSList *n)
{
    pthread_once(&g_slist_once, slist_resolve);
    if (g_free1_fn)
        g_free1_fn(n);
    else
        free(n);
}

void accel_gslist_free(GSList *list)
{
    while (list) {
        GSList *next = list->next;
        node_free(list);
        list = next;
    }
}

static int slist_length(const GSList *l)
{
    int n = 0;
    for (; l; l = l->next)
        n++;
    return n;
}

/* ---------------------------------------------------------------------- */
/* index rules: src/accel_utils.c:36-102, src/responses.c:27-84            */

static int index_r(double r, double lor) { return (int) ((r - lor) * ACCEL_RDR + DBLCORRECT); }
static int index_z(double z, double loz) { return (int) ((z - loz) * ACCEL_RDZ + DBLCORRECT); }
static int index_w(double w, double low) { return (int) ((w - low) * ACCEL_RDW + DBLCORRECT); }

static int twon_to_index(int n)
{
    int x = 0;
    while (n > 1) {
        n >>= 1;
        x++;
    }
    return x;
}

/* ---------------------------------------------------------------------- */
/* per-accelobs device state                                               */

typedef struct sf_buf {
    float *d;
    size_t bytes;
    struct sf_buf *next;
} sf_buf;

#define SF_MAXTAPS 64

typedef struct sf_state {
    const accelobs *obs;
    accel_search *s;
    accel_dev_view v;
    int plane_ready;
    float *d_full;            /* in-memory full plane [numz][rlen] */
    long long rlen;
    double *d_norm;
    float *d_nsq;
    void *d_wdata;            /* -locpow: the normalised window */
    size_t wcap;
    struct {                  /* -locpow with w rows: FP32 taps per harmonic */
        int numharm, harmnum, stride;
        void *d_R;
        int *d_nk;
    } taps[SF_MAXTAPS];
    int ntaps;
    agpu_sf_hit *d_hits;
    unsigned *d_count;
    unsigned cap;
    sf_buf *pool;             /* free plane buffers */
    int refs;                 /* live ffdotpows, + 1 while the accelobs lives */
    float *series_base;       /* .dat input: the padded buffer obs->fft points into */
    size_t map_bytes;         /* .fft input: mmap length */
    struct sf_state *next;
} sf_state;

static sf_state *g_states = NULL;
static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;

static sf_state *state_of(const accelobs *obs)
{
    pthread_mutex_lock(&g_mu);
    sf_state *s = g_states;
    while (s && s->obs != obs)
        s = s->next;
    pthread_mutex_unlock(&g_mu);
    return s;
}

static void state_release(sf_state *st)
{
    if (--st->refs > 0)
        return;
    for (sf_buf *b = st->pool; b;) {
        sf_buf *n = b->next;
        agpu_free(b->d);
        free(b);
        b = n;
    }
    if (st->d_full)
        agpu_free(st->d_full);
    if (st->d_norm)
        agpu_free(st->d_norm);
    if (st->d_nsq)
        agpu_free(st->d_nsq);
    if (st->d_wdata)
        agpu_free(st->d_wdata);
    for (int i = 0; i < st->ntaps; i++) {
        agpu_free(st->taps[i].d_R);
        agpu_free(st->taps[i].d_nk);
    }
    if (st->d_hits)
        agpu_free(st->d_hits);
    if (st->d_count)
        agpu_free(st->d_count);
    if (st->s)
        accel_search_destroy(st->s);
    free(st);
}

static float *pool_get(sf_state *st, size_t bytes)
{
    sf_buf **pp = &st->pool;
    for (; *pp; pp = &(*pp)->next) {
        if ((*pp)->bytes == bytes) {
            sf_buf *b = *pp;
            float *d = b->d;
            *pp = b->next;
            free(b);
            return d;
        }
    }
    void *d = NULL;
    if (agpu_malloc(&d, bytes ? bytes : 4))
        return NULL;
    return (float *) d;
}

static void pool_put(sf_state *st, float *d, size_t bytes)
{
    sf_buf *b = (sf_buf *) malloc(sizeof(sf_buf));
    if (!b) {
        agpu_free(d);
        return;
    }
    b->d = d;
    b->bytes = bytes;
    b->next = st->pool;
    st->pool = b;
}

/* An ffdotpows as handed out: the reference's struct first, then where
 * its powers live on the device and what its index tables derive from. */
typedef struct {
    ffdotpows pub;
    sf_state *st;
    float *d_pow;
    size_t bytes;
    double fullrlo;
    float *h_block;           /* accel_ffdotpows_sync(): powers[0][0] */
    float **h_rows;           /* powers[0] */
} sf_ffd;

static sf_ffd *sf_of(ffdotpows *f) { return (sf_ffd *) f; }

static void drop_host_copy(sf_ffd *f)
{
    free(f->h_block);
    free(f->h_rows);
    free(f->pub.powers);
    f->h_block = NULL;
    f->h_rows = NULL;
    f->pub.powers = NULL;
}

static sf_ffd *ffd_new(sf_state *st, long long rlo, int zlo, int wlo, int numrs, int numzs,
                       int numws)
{
    sf_ffd *f = (sf_ffd *) calloc(1, sizeof(sf_ffd));
    if (!f)
        return NULL;
    f->pub.rlo = rlo;
    f->pub.zlo = zlo;
    f->pub.wlo = wlo;
    f->pub.numrs = numrs;
    f->pub.numzs = numzs;
    f->pub.numws = numws;
    f->st = st;
    f->bytes = sizeof(float) * (size_t) numrs * numzs * numws;
    f->d_pow = pool_get(st, f->bytes);
    if (!f->d_pow) {
        free(f);
        return NULL;
    }
    st->refs++;
    return f;
}

/* ---------------------------------------------------------------------- */
/* .inf reading: src/ioinf.c:14-350                                        */

static char *trim(char *s)
{
    while (*s == ' ' || *s == '\t')
        s++;
    size_t n = strlen(s);
    while (n && (s[n - 1] == ' ' || s[n - 1] == '\t' || s[n - 1] == '\n' || s[n - 1] == '\r'))
        s[--n] = '\0';
    return s;
}

/* The value after the key/value '=' of the next line (column 40 for a
 * standard line, else the last '='); "Unknown" if empty.  -1 at end of
 * file or on a blank line, as the reference's read_inf_line_valstr. */
static int inf_value(FILE *f, char *out, size_t outlen)
{
    char line[250];
    if (!fgets(line, sizeof(line), f) || line[0] == '\n' || !line[0])
        return -1;
    char *v;
    size_t n = strlen(line);
    if (n >= 40 && line[40] == '=') {
        v = line + 41;
    } else {
        char *e = strrchr(line, '=');
        if (!e)
            return -1;
        v = e + 1;
    }
    v = trim(v);
    snprintf(out, outlen, "%s", *v ? v : "Unknown");
    return 0;
}

/* ra_dec_from_string (src/misc_utils.c:715-737): hh:mm:ss.ssss; a '-'
 * before a zero hour/degree field is carried by minutes and seconds */
static void ra_dec(const char *s, int *h, int *m, double *sec)
{
    while (*s == ' ' || *s == '\t')
        s++;
    *h = 0;
    *m = 0;
    *sec = 0.0;
    sscanf(s, "%d:%d:%lf", h, m, sec);
    if (s[0] == '-' && *h == 0) {
        *m = -*m;
        *sec = -*sec;
    }
}

int accel_readinf(infodata *d, const char *root)
{
    char path[4200], t[256];
    snprintf(path, sizeof(path), "%s.inf", root);
    FILE *f = fopen(path, "r");
    if (!f) {
        accel_set_error("Error opening file '%s'", path);
        return -1;
    }
    memset(d, 0, sizeof(*d));
#define RD(dst)                                                                 \
    do {                                                                        \
        if (inf_value(f, dst, sizeof(dst))) {                                   \
            fclose(f);                                                          \
            accel_set_error("Error:  end-of-file or blank line reading '%s'",   \
                            path);                                              \
            return -1;                                                          \
        }                                                                       \
    } while (0)
    RD(d->name);
    RD(d->telescope);
    const int fake = strcmp(d->telescope, "None (Artificial Data Set)") == 0;
    if (!fake) {
        RD(d->instrument);
        RD(d->object);
        RD(t);
        ra_dec(t, &d->ra_h, &d->ra_m, &d->ra_s);
        RD(t);
        ra_dec(t, &d->dec_d, &d->dec_m, &d->dec_s);
        RD(d->observer);
        RD(t);
        char frac[128] = "0";
        if (sscanf(t, "%d.%120s", &d->mjd_i, frac + 1) == 2) {
            frac[0] = '.';
            char z[130];
            snprintf(z, sizeof(z), "0%s", frac);
            d->mjd_f = strtod(z, NULL);
        }
        RD(t);
        d->bary = (int) strtol(t, NULL, 10);
    } else {
        d->mjd_i = -1;
        strcpy(d->object, "fake pulsar");
    }
    RD(t);
    d->N = strtod(t, NULL);
    RD(t);
    d->dt = strtod(t, NULL);
    RD(t);
    d->numonoff = (int) strtol(t, NULL, 10);
    if (d->numonoff) {
        int ii = 0;
        do {
            RD(t);
            if (sscanf(t, "%lf %*[ ,] %lf", &d->onoff[ii], &d->onoff[ii + 1]) != 2) {
                fclose(f);
                accel_set_error("Error:  can not parse on-off pair (%d) in readinf()", ii / 2);
                return -1;
            }
            ii += 2;
        } while (d->onoff[ii - 1] < d->N - 1 && ii < 2 * MAXNUMONOFF);
        d->numonoff = ii / 2;
    } else {
        d->numonoff = 1;
        d->onoff[0] = 0;
        d->onoff[1] = d->N - 1;
    }
    if (!fake) {
        RD(d->band);
        if (strcmp(d->band, "Radio") == 0) {
            RD(t);
            d->fov = strtod(t, NULL);
            RD(t);
            d->dm = strtod(t, NULL);
            RD(t);
            d->freq = strtod(t, NULL);
            RD(t);
            d->freqband = strtod(t, NULL);
            RD(t);
            d->num_chan = (int) strtol(t, NULL, 10);
            RD(t);
            d->chan_wid = strtod(t, NULL);
        } else if (strcmp(d->band, "X-ray") == 0 || strcmp(d->band, "Gamma") == 0) {
            RD(t);
            d->fov = strtod(t, NULL);
            RD(t);
            d->energy = strtod(t, NULL);
            RD(t);
            d->energyband = strtod(t, NULL);
        } else {
            RD(d->filt);
            RD(t);
            d->fov = strtod(t, NULL);
            RD(t);
            d->wavelen = strtod(t, NULL);
            RD(t);
            d->waveband = strtod(t, NULL);
        }
    }
    RD(d->analyzer);
#undef RD
    /* the notes: everything after the "Any additional notes:" line */
    char line[100];
    size_t nl = 0;
    if (fgets(line, sizeof(line), f)) {
        while (fgets(line, sizeof(line), f)) {
            size_t k = strlen(line);
            if (nl + k > 500)
                break;
            if (nl == 0) {
                const char *p = line;
                while (*p == ' ' || *p == '\t')
                    p++;
                k = strlen(p);
                memcpy(d->notes, p, k + 1);
            } else {
                memcpy(d->notes + nl, line, k + 1);
            }
            nl += k;
        }
    }
    fclose(f);
    return 0;
}

/* ra_dec_to_string (src/misc_utils.c:700-712) */
static void ra_dec_str(char *out, size_t n, int h, int m, double s)
{
    int off = 0;
    if (h == 0 && (m < 0 || s < 0.0)) {
        out[0] = '-';
        off = 1;
    }
    snprintf(out + off, n - off, "%.2d:%.2d:%07.4f", h, abs(m), fabs(s));
}

static const char *or_unknown(const char *s, char *buf, size_t n)
{
    snprintf(buf, n, "%s", s);
    return *trim(buf) ? s : "Unknown";
}

int accel_writeinf(const infodata *d)
{
    char path[4200], t[128], u[512];
    snprintf(path, sizeof(path), "%s.inf", d->name);
    FILE *f = fopen(path, "w");
    if (!f) {
        accel_set_error("Error opening file '%s' for writing", path);
        return -1;
    }
    const int fake = strcmp(d->telescope, "None (Artificial Data Set)") == 0;
    fprintf(f, " Data file name without suffix          =  %s\n", d->name);
    fprintf(f, " Telescope used                         =  %s\n",
            or_unknown(d->telescope, u, sizeof(u)));
    if (!fake) {
        fprintf(f, " Instrument used                        =  %s\n",
                or_unknown(d->instrument, u, sizeof(u)));
        fprintf(f, " Object being observed                  =  %s\n",
                or_unknown(d->object, u, sizeof(u)));
        ra_dec_str(t, sizeof(t), d->ra_h, d->ra_m, d->ra_s);
        fprintf(f, " J2000 Right Ascension (hh:mm:ss.ssss)  =  %s\n", t);
        ra_dec_str(t, sizeof(t), d->dec_d, d->dec_m, d->dec_s);
        fprintf(f, " J2000 Declination     (dd:mm:ss.ssss)  =  %s\n", t);
        fprintf(f, " Data observed by                       =  %s\n",
                or_unknown(d->observer, u, sizeof(u)));
        char fr[64];
        snprintf(fr, sizeof(fr), "%.15f", d->mjd_f);
        const char *dot = strchr(fr, '.');
        fprintf(f, " Epoch of observation (MJD)             =  %d.%s\n", d->mjd_i,
                dot ? dot + 1 : "0");
        fprintf(f, " Barycentered?           (1 yes, 0 no)  =  %d\n", d->bary);
    }
    fprintf(f, " Number of bins in the time series      =  %-11.0f\n", d->N);
    fprintf(f, " Width of each time series bin (sec)    =  %.15g\n", d->dt);
    fprintf(f, " Any breaks in the data? (1 yes, 0 no)  =  %d\n", d->numonoff > 1 ? 1 : 0);
    if (d->numonoff > 1)
        for (int i = 0; i < d->numonoff; i++)
            fprintf(f, " On/Off bin pair #%3d                   =  %-11.0f, %-11.0f\n", i + 1,
                    d->onoff[2 * i], d->onoff[2 * i + 1]);
    if (!fake) {
        fprintf(f, " Type of observation (EM band)          =  %s\n", d->band);
        if (strcmp(d->band, "Radio") == 0) {
            fprintf(f, " Beam diameter (arcsec)                 =  %.0f\n", d->fov);
            fprintf(f, " Dispersion measure (cm-3 pc)           =  %.12g\n", d->dm);
            fprintf(f, " Central freq of low channel (MHz)      =  %.12g\n", d->freq);
            fprintf(f, " Total bandwidth (MHz)                  =  %.12g\n", d->freqband);
            fprintf(f, " Number of channels                     =  %d\n", d->num_chan);
            fprintf(f, " Channel bandwidth (MHz)                =  %.12g\n", d->chan_wid);
        } else if (strcmp(d->band, "X-ray") == 0 || strcmp(d->band, "Gamma") == 0) {
            fprintf(f, " Field-of-view diameter (arcsec)        =  %.2f\n", d->fov);
            fprintf(f, " Central energy (kev)                   =  %.1f\n", d->energy);
            fprintf(f, " Energy bandpass (kev)                  =  %.1f\n", d->energyband);
        } else {
            fprintf(f, " Photometric filter used                =  %s\n",
                    or_unknown(d->filt, u, sizeof(u)));
            fprintf(f, " Field-of-view diameter (arcsec)        =  %.2f\n", d->fov);
            fprintf(f, " Central wavelength (nm)                =  %.1f\n", d->wavelen);
            fprintf(f, " Bandpass (nm)                          =  %.1f\n", d->waveband);
        }
    }
    fprintf(f, " Data analyzed by                       =  %s\n",
            or_unknown(d->analyzer, u, sizeof(u)));
    fprintf(f, " Any additional notes:\n    %s\n\n", d->notes);
    if (fclose(f)) {
        accel_set_error("Error writing '%s'", path);
        return -1;
    }
    return 0;
}

/* ---------------------------------------------------------------------- */
/* create_accelobs / free_accelobs                                         */

static char *dupstr(const char *s)
{
    size_t n = strlen(s) + 1;
    char *d = (char *) malloc(n);
    if (d)
        memcpy(d, s, n);
    return d;
}

/* split_root_suffix: root = everything before the last '.', suffix after */
static int split_root_suffix(const char *in, char **root, char **suffix)
{
    const char *dot = strrchr(in, '.');
    if (!dot) {
        *root = dupstr(in);
        *suffix = NULL;
        return 0;
    }
    *root = (char *) malloc((size_t) (dot - in) + 1);
    memcpy(*root, in, (size_t) (dot - in));
    (*root)[dot - in] = '\0';
    *suffix = dupstr(dot + 1);
    return 1;
}

static void *read_all(const char *path, size_t elsize, long long *count)
{
    FILE *f = fopen(path, "rb");
    if (!f)
        return NULL;
    fseek(f, 0, SEEK_END);
    long long bytes = ftell(f);
    fseek(f, 0, SEEK_SET);
    *count = bytes / (long long) elsize;
    void *buf = malloc((size_t) (*count ? *count : 1) * elsize);
    if (buf && *count && fread(buf, elsize, (size_t) *count, f) != (size_t) *count) {
        free(buf);
        buf = NULL;
    }
    fclose(f);
    return buf;
}

void create_accelobs(accelobs *obs, infodata *idata, Cmdline *cmd, int usemmap)
{
    int input_shorts = 0;
    (void) usemmap;   /* .fft input is always memory-mapped (src/accel_utils.c:1568-1586) */
    memset(obs, 0, sizeof(*obs));
    {
        char *suffix = NULL;
        const char *in = (cmd->argc > 0 && cmd->argv) ? cmd->argv[0] : "";
        int has = split_root_suffix(in, &obs->rootfilenm, &suffix);
        if (!has || (strcmp(suffix, "fft") && strcmp(suffix, "dat") && strcmp(suffix, "sdat"))) {
            free(suffix);
            sf_fail(0, "Input file ('%s') must be an '.fft' or '.[s]dat' file!", in);
            return;
        }
        obs->dat_input = strcmp(suffix, "fft") != 0;
        input_shorts = strcmp(suffix, "sdat") == 0;
        free(suffix);
    }
    obs->use_harmonic_polishing = !cmd->noharmpolishP;   /* src/accel_utils.c:1490-1493 */
    if (accel_readinf(idata, obs->rootfilenm)) {
        sf_fail(-1, "%s", accel_last_error());
        return;
    }
    if (strlen(trim(idata->object)) > 0)
        printf("Analyzing %s data from '%s'.\n\n", idata->object, cmd->argv[0]);
    else
        printf("Analyzing data from '%s'.\n\n", cmd->argv[0]);

    sf_state *st = (sf_state *) calloc(1, sizeof(sf_state));
    if (!st) {
        sf_fail(-1, "Error allocating the search state");
        return;
    }
    st->obs = obs;
    st->refs = 1;

    float *series = NULL;
    long long nseries = 0;
    if (obs->dat_input) {
        printf("Reading and FFTing the time series...");
        fflush(stdout);
        long long n = 0;
        void *raw = read_all(cmd->argv[0], input_shorts ? sizeof(short) : sizeof(float), &n);
        if (!raw) {
            free(st);
            sf_fail(-1, "Error reading '%s'", cmd->argv[0]);
            return;
        }
        if (n > 67108864) {
            free(raw);
            free(st);
            sf_fail(0, "The input time series is too large.  Use 'realfft' first.");
            return;
        }
        /* the padded buffer of src/accel_utils.c:1527-1550 */
        st->series_base = (float *) calloc((size_t) n + 2 * ACCEL_PADDING, sizeof(float));
        if (!st->series_base) {
            free(raw);
            free(st);
            sf_fail(-1, "Error allocating the time series");
            return;
        }
        series = st->series_base + ACCEL_PADDING;
        for (long long i = 0; i < n; i++)
            series[i] = input_shorts ? (float) ((short *) raw)[i] : ((float *) raw)[i];
        free(raw);
        nseries = n;
        obs->fftfile = NULL;
        obs->fft = (fcomplex *) series;
        obs->numbins = n / 2;
    } else {
        int fd = open(cmd->argv[0], O_RDONLY);
        struct stat sb;
        if (fd < 0 || fstat(fd, &sb)) {
            if (fd >= 0)
                close(fd);
            free(st);
            sf_fail(-1, "Error opening '%s'", cmd->argv[0]);
            return;
        }
        obs->numbins = (long long) sb.st_size / (long long) sizeof(fcomplex);
        st->map_bytes = (size_t) obs->numbins * sizeof(fcomplex);
        void *m = st->map_bytes ? mmap(NULL, st->map_bytes, PROT_READ, MAP_SHARED, fd, 0)
                                : MAP_FAILED;
        if (m == MAP_FAILED) {
            close(fd);
            free(st);
            sf_fail(-1, "Error in mmap() of '%s'", cmd->argv[0]);
            return;
        }
        obs->mmap_file = fd;
        obs->fft = (fcomplex *) m;
    }

    /* src/accel_utils.c:1590-1664 */
    if (cmd->zmax % ACCEL_DZ)
        cmd->zmax = (cmd->zmax / ACCEL_DZ + 1) * ACCEL_DZ;
    if (cmd->wmaxP && cmd->wmax % ACCEL_DW)
        cmd->wmax = (cmd->wmax / ACCEL_DW + 1) * ACCEL_DW;
    obs->N = (long long) idata->N;
    obs->norm_type = 0;
    if (cmd->photonP) {
        /* bin 0's real part: the mapped .fft's, or 1 for a .dat series,
         * whose de-reddening sets bin 0 to (1, 0) (src/misc_utils.c:943-944)
         * and which then keeps it (src/accel_utils.c:1599-1605) */
        obs->nph = obs->dat_input ? 1.0 : (obs->numbins > 0 ? obs->fft[0].r : 0.0);
        printf("Normalizing powers using %.0f photons.\n\n", obs->nph);
    } else if (cmd->locpowP) {   /* src/accel_utils.c:1611-1613 */
        obs->nph = 0.0;
        obs->norm_type = 1;
        printf("Normalizing powers using local-power determination.\n\n");
    } else {
        obs->nph = 0.0;
        printf("Normalizing powers using median-blocks%s.\n\n", cmd->medianP ? "" : " (default)");
    }

    accel_params p;
    accel_params_default(&p);
    p.N = idata->N;
    p.dt = idata->dt;
    p.numbins = obs->numbins;
    p.lobin = cmd->lobin;
    p.numharm = cmd->numharm;
    p.zmax = cmd->zmax;
    p.wmax = cmd->wmax;
    p.wmaxP = cmd->wmaxP;
    p.sigma = cmd->sigma;
    p.flo = cmd->flo;
    p.fhi = cmd->fhi;
    p.rlo = cmd->rlo;
    p.rhi = cmd->rhi;
    p.floP = cmd->floP;
    p.fhiP = cmd->fhiP;
    p.rloP = cmd->rloP;
    p.rhiP = cmd->rhiP;
    p.inmem = cmd->inmemP;
    p.dat_input = obs->dat_input;
    p.noharmremove = cmd->noharmremoveP;
    p.nph = obs->nph;
    p.noharmpolish = !obs->use_harmonic_polishing;
    accel_geom g;
    if (accel_geom_create(&p, &g)) {
        char m[1100];
        snprintf(m, sizeof(m), "%s", accel_last_error());
        if (obs->dat_input)
            free(st->series_base);
        else {
            munmap(obs->fft, st->map_bytes);
            close(obs->mmap_file);
        }
        free(st);
        sf_fail(1, "%s.  Exiting.", m);
        return;
    }
    obs->lobin = g.lobin;
    obs->nph = g.nph;   /* 0 when lobin > 0, src/accel_utils.c:1626-1628 */
    obs->numharmstages = g.numharmstages;
    obs->dz = g.dz;
    obs->numz = g.numz;
    obs->whi = g.whi;
    obs->wlo = g.wlo;
    obs->dw = g.dw;
    obs->numw = g.numw;
    if (cmd->wmaxP)
        printf("Jerk search enabled with maximum fdotdot wmax = %d\n", cmd->wmax);

    /* output names, src/accel_utils.c:1666-1684 */
    size_t rl = strlen(obs->rootfilenm) + 45;
    obs->candnm = (char *) calloc(rl, 1);
    obs->accelnm = (char *) calloc(rl, 1);
    obs->workfilenm = (char *) calloc(rl, 1);
    if (obs->numw) {
        snprintf(obs->candnm, rl, "%s_ACCEL_%d_JERK_%d.cand", obs->rootfilenm, cmd->zmax,
                 cmd->wmax);
        snprintf(obs->accelnm, rl, "%s_ACCEL_%d_JERK_%d", obs->rootfilenm, cmd->zmax,
                 cmd->wmax);
        snprintf(obs->workfilenm, rl, "%s_ACCEL_%d_JERK_%d.txtcand", obs->rootfilenm,
                 cmd->zmax, cmd->wmax);
    } else {
        snprintf(obs->candnm, rl, "%s_ACCEL_%d.cand", obs->rootfilenm, cmd->zmax);
        snprintf(obs->accelnm, rl, "%s_ACCEL_%d", obs->rootfilenm, cmd->zmax);
        snprintf(obs->workfilenm, rl, "%s_ACCEL_%d.txtcand", obs->rootfilenm, cmd->zmax);
    }
    if (!obs->dat_input) {
        obs->workfile = fopen(obs->workfilenm, "w");
        if (!obs->workfile) {
            sf_fail(-1, "Error opening '%s'", obs->workfilenm);
            return;
        }
    }
    obs->numbetween = g.numbetween;
    obs->dt = g.dt;
    obs->T = g.T;
    obs->rlo = g.rlo;
    obs->highestbin = g.highestbin;
    obs->rhi = g.rhi;
    obs->dr = g.dr;
    obs->zhi = g.zhi;
    obs->zlo = g.zlo;
    obs->sigma = g.sigma;
    obs->powcut = (float *) malloc(sizeof(float) * obs->numharmstages);
    obs->numindep = (long long *) malloc(sizeof(long long) * obs->numharmstages);
    for (int i = 0; i < obs->numharmstages; i++) {
        obs->powcut[i] = g.powcut[i];
        obs->numindep[i] = g.numindep[i];
    }
    obs->numzap = 0;
    obs->maxkernlen = g.maxkernlen;
    obs->fftlen = g.fftlen;
    obs->corr_uselen = g.corr_uselen;
    obs->inmem = g.inmem;
    obs->ffdotplane = NULL;   /* device-resident */
    {
        long long memuse = (long long) sizeof(float) * (obs->highestbin + obs->corr_uselen) *
                           obs->numbetween * obs->numz * (obs->numw ? obs->numw : 1);
        printf("Full f-fdot%s would need %.2f GB: ", obs->numw ? "-fdotdot volume" : " plane",
               (float) memuse / (double) (1LL << 30));
        printf("using %s accelsearch.\n\n", obs->inmem ? "in-memory" : "standard");
    }

    /* the search handle; a time series is FFT'd and de-reddened on the
     * GPU (src/accel_utils.c:1554-1565) and copied back for the caller */
    p.rloP = 1;
    p.rlo = obs->rlo;
    p.floP = 0;
    p.rhiP = 1;
    p.rhi = (double) obs->highestbin;
    p.fhiP = 0;
    p.inmem = obs->inmem;
    p.force_standard = !obs->inmem;
    st->s = accel_search_create(&p, g_device);
    if (!st->s) {
        sf_fail(-1, "%s", accel_last_error());
        return;
    }
    const accel_geom *sg = accel_search_geom(st->s);
    if (sg->corr_uselen != obs->corr_uselen || sg->nblocks != g.nblocks ||
        sg->inmem != obs->inmem || sg->rlo != obs->rlo || sg->highestbin != obs->highestbin) {
        sf_fail(-1, "internal: search handle geometry differs from the accelobs");
        return;
    }
    if (obs->dat_input) {
        if (accel_search_set_series(st->s, series, nseries) ||
            accel_search_spectrum(st->s, obs->fft, obs->numbins)) {
            sf_fail(-1, "%s", accel_last_error());
            return;
        }
        printf("done.\n");
        printf("Removing red-noise...done.\n\n");
    }
    accel_search_dev_view(st->s, &st->v);
    if (agpu_malloc((void **) &st->d_norm, sizeof(double)) ||
        agpu_malloc((void **) &st->d_nsq, sizeof(float)) ||
        agpu_malloc((void **) &st->d_count, sizeof(unsigned)) ||
        agpu_malloc((void **) &st->d_hits, sizeof(agpu_sf_hit) * SF_HIT_CAP0)) {
        sf_fail(-1, "%s", agpu_last_error());
        return;
    }
    st->cap = SF_HIT_CAP0;
    pthread_mutex_lock(&g_mu);
    st->next = g_states;
    g_states = st;
    pthread_mutex_unlock(&g_mu);
}

void free_accelobs(accelobs *obs)
{
    sf_state *st = state_of(obs);
    if (st) {
        pthread_mutex_lock(&g_mu);
        sf_state **pp = &g_states;
        while (*pp && *pp != st)
            pp = &(*pp)->next;
        if (*pp)
            *pp = st->next;
        pthread_mutex_unlock(&g_mu);
        st->obs = NULL;
    }
    if (obs->mmap_file) {
        if (st && st->map_bytes)
            munmap(obs->fft, st->map_bytes);
        close(obs->mmap_file);
    } else if (obs->dat_input) {
        if (st)
            free(st->series_base);
    } else if (obs->fftfile) {
        fclose(obs->fftfile);
    }
    free(obs->powcut);
    free(obs->numindep);
    free(obs->rootfilenm);
    free(obs->candnm);
    free(obs->accelnm);
    free(obs->workfilenm);
    if (obs->numzap) {
        free(obs->lobins);
        free(obs->hibins);
    }
    obs->fft = NULL;
    if (st)
        state_release(st);
}

/* ---------------------------------------------------------------------- */
/* subharminfos                                                            */

static int init_subharminfo(int numharm, int harmnum, int zmax, int wmax, subharminfo *shi,
                            const accelobs *obs)
{
    const double hf = (double) harmnum / (double) numharm;
    memset(shi, 0, sizeof(*shi));
    shi->numharm = numharm;
    shi->harmnum = harmnum;
    shi->zmax = ar_required_z(hf, zmax);
    shi->wmax = ar_required_w(hf, wmax);
    if (numharm > 1) {
        shi->rinds = (unsigned short *) malloc(sizeof(unsigned short) * obs->corr_uselen);
        shi->zinds = (unsigned short *) malloc(sizeof(unsigned short) * obs->corr_uselen);
    }
    const int fftlen = (numharm == 1 && harmnum == 1)
                           ? obs->fftlen
                           : ar_calc_fftlen(numharm, harmnum, zmax, wmax, obs->corr_uselen);
    shi->numkern_zdim = (shi->zmax / ACCEL_DZ) * 2 + 1;
    shi->numkern_wdim = (shi->wmax / ACCEL_DW) * 2 + 1;
    shi->numkern = shi->numkern_zdim * shi->numkern_wdim;
    shi->kern = (kernel **) malloc(sizeof(kernel *) * shi->numkern_wdim);
    if (!shi->kern)
        return -1;
    shi->kern[0] = (kernel *) calloc((size_t) shi->numkern, sizeof(kernel));
    if (!shi->kern[0])
        return -1;
    for (int i = 1; i < shi->numkern_wdim; i++)
        shi->kern[i] = shi->kern[i - 1] + shi->numkern_zdim;
    /* init_kernel, src/accel_utils.c:153-169, without the FFT'd data: the
     * templates live in the search's device bank */
    for (int i = 0; i < shi->numkern_wdim; i++) {
        for (int j = 0; j < shi->numkern_zdim; j++) {
            kernel *k = &shi->kern[i][j];
            k->z = -shi->zmax + j * ACCEL_DZ;
            k->w = -shi->wmax + i * ACCEL_DW;
            k->fftlen = fftlen;
            k->numbetween = ACCEL_NUMBETWEEN;
            k->kern_half_width = ar_w_halfwidth((double) k->z, (double) k->w);
            k->numgoodbins = k->fftlen - 2 * k->numbetween * k->kern_half_width;
            k->data = NULL;
        }
    }
    return 0;
}

static void print_shi(const subharminfo *shi, const accelobs *obs)
{
    const int fftlen = shi->kern[0][0].fftlen;
    if (obs->numw)
        printf("  Harm %2d/%-2d: %5d kernels, %4d < z < %-4d and %5d < w < %-5d (%5d pt FFTs)\n",
               shi->harmnum, shi->numharm, shi->numkern, -shi->zmax, shi->zmax, -shi->wmax,
               shi->wmax, fftlen);
    else
        printf("  Harm %2d/%-2d: %5d kernels, %4d < z < %-4d (%d pt FFTs)\n", shi->harmnum,
               shi->numharm, shi->numkern, -shi->zmax, shi->zmax, fftlen);
}

subharminfo **create_subharminfos(accelobs *obs)
{
    sf_state *st = state_of(obs);
    if (!st) {
        sf_fail(-1, "create_subharminfos: the accelobs was not made by create_accelobs");
        return NULL;
    }
    subharminfo **shis = (subharminfo **) calloc((size_t) obs->numharmstages,
                                                 sizeof(subharminfo *));
    shis[0] = (subharminfo *) calloc(2, sizeof(subharminfo));
    if (init_subharminfo(1, 1, (int) obs->zhi, (int) obs->whi, &shis[0][0], obs)) {
        sf_fail(-1, "Error allocating the subharminfos");
        return NULL;
    }
    print_shi(&shis[0][0], obs);
    if (!obs->inmem) {
        for (int ii = 1; ii < obs->numharmstages; ii++) {
            const int harmtosum = 1 << ii;
            shis[ii] = (subharminfo *) calloc((size_t) harmtosum, sizeof(subharminfo));
            for (int jj = 1; jj < harmtosum; jj += 2) {
                if (init_subharminfo(harmtosum, jj, (int) obs->zhi, (int) obs->whi,
                                     &shis[ii][jj - 1], obs)) {
                    sf_fail(-1, "Error allocating the subharminfos");
                    return NULL;
                }
                print_shi(&shis[ii][jj - 1], obs);
            }
        }
    }
    /* the spectrum as the caller left it (after zapping), then the raw
     * plane of every template over the whole spectrum */
    if (accel_search_set_spectrum(st->s, obs->fft) || accel_search_plane_only(st->s)) {
        sf_fail(-1, "%s", accel_last_error());
        return shis;
    }
    accel_search_dev_view(st->s, &st->v);
    st->plane_ready = 1;
    printf("Total RAM used by correlation kernels:  %.3f GB (device)\n",
           (double) st->v.nrows * 4096 * sizeof(fcomplex) / (double) (1 << 30));
    return shis;
}

static void free_subharminfo(subharminfo *shi)
{
    if (shi->numharm > 1) {
        free(shi->rinds);
        free(shi->zinds);
    }
    if (shi->kern) {
        free(shi->kern[0]);
        free(shi->kern);
    }
}

void free_subharminfos(accelobs *obs, subharminfo **shis)
{
    if (!shis)
        return;
    if (!obs->inmem) {
        for (int ii = 1; ii < obs->numharmstages; ii++) {
            const int harmtosum = 1 << ii;
            if (!shis[ii])
                continue;
            for (int jj = 1; jj < harmtosum; jj += 2)
                free_subharminfo(&shis[ii][jj - 1]);
            free(shis[ii]);
        }
    }
    free_subharminfo(&shis[0][0]);
    free(shis[0]);
    free(shis);
}

/* ---------------------------------------------------------------------- */
/* planes                                                                  */

/* The FP32 response taps of every (w, z) row of one harmonic's kernels
 * (gen_w_response / gen_z_response of init_kernel, src/accel_utils.c:153-169),
 * made on the device once per harmonic and kept with the search: the
 * -locpow planes of jerk searches correlate each window with them
 * directly.  Returns the cache slot or -1. */
static int locpow_taps(sf_state *st, const subharminfo *shi)
{
    for (int i = 0; i < st->ntaps; i++)
        if (st->taps[i].numharm == shi->numharm && st->taps[i].harmnum == shi->harmnum)
            return i;
    if (st->ntaps == SF_MAXTAPS) {
        agpu_set_error("too many harmonics for the -locpow tap tables");
        return -1;
    }
    const int nt = shi->numkern;
    int *zs = (int *) malloc(sizeof(int) * nt), *ws = (int *) malloc(sizeof(int) * nt);
    int *nk = (int *) malloc(sizeof(int) * nt);
    int stride = 64, rc = -1;
    void *d_R = NULL;
    int *d_nk = NULL;
    if (!zs || !ws || !nk) {
        agpu_set_error("out of host memory");
        goto out;
    }
    for (int i = 0; i < shi->numkern_wdim; i++)
        for (int j = 0; j < shi->numkern_zdim; j++) {
            const kernel *k = &shi->kern[i][j];
            const int t = i * shi->numkern_zdim + j;
            zs[t] = k->z;
            ws[t] = k->w;
            nk[t] = 2 * ACCEL_NUMBETWEEN * k->kern_half_width;
            while (stride < nk[t])
                stride += 64;
        }
    if (agpu_malloc(&d_R, sizeof(fcomplex) * (size_t) nt * stride) ||
        agpu_malloc((void **) &d_nk, sizeof(int) * (size_t) nt) ||
        agpu_h2d(d_nk, nk, sizeof(int) * (size_t) nt, st->v.stream) ||
        agpu_gen_response_taps(zs, ws, nk, nt, stride, d_R, st->v.stream))
        goto out;
    st->taps[st->ntaps].numharm = shi->numharm;
    st->taps[st->ntaps].harmnum = shi->harmnum;
    st->taps[st->ntaps].stride = stride;
    st->taps[st->ntaps].d_R = d_R;
    st->taps[st->ntaps].d_nk = d_nk;
    d_R = NULL;
    d_nk = NULL;
    rc = st->ntaps++;
out:
    agpu_free(d_R);
    agpu_free(d_nk);
    free(zs);
    free(ws);
    free(nk);
    return rc;
}

fcomplex *get_fourier_amplitudes(long long lobin, int numbins, accelobs *obs)
{
    fcomplex *out = (fcomplex *) calloc((size_t) (numbins > 0 ? numbins : 1), sizeof(fcomplex));
    if (!out || !obs->fft)
        return out;
    for (int i = 0; i < numbins; i++) {
        long long b = lobin - obs->lobin + i;
        if (b >= 0 && b < obs->numbins)
            out[i] = obs->fft[b];
    }
    return out;
}

ffdotpows *subharm_fderivs_vol(int numharm, int harmnum, double fullrlo, double fullrhi,
                               subharminfo *shi, accelobs *obs)
{
    sf_state *st = state_of(obs);
    if (!st || !st->plane_ready) {
        sf_fail(-1, "subharm_fderivs_vol: call create_subharminfos first");
        return NULL;
    }
    const double hf = (double) harmnum / (double) numharm;
    const double drlo = ar_required_r(hf, fullrlo), drhi = ar_required_r(hf, fullrhi);
    const long long rlo = (long long) floor(drlo);
    const int zlo = ar_required_z(hf, obs->zlo), wlo = ar_required_w(hf, obs->wlo);
    /* the lookup tables, src/accel_utils.c:1042-1053 */
    if (numharm > 1 && !obs->inmem) {
        for (int ii = 0; ii < obs->corr_uselen; ii++) {
            double rr = fullrlo + ii * ACCEL_DR;
            shi->rinds[ii] = (unsigned short) index_r(ar_required_r(hf, rr), (double) rlo);
        }
        for (int ii = 0; ii < obs->numz; ii++) {
            double zz = obs->zlo + ii * ACCEL_DZ;
            shi->zinds[ii] = (unsigned short) index_z((double) ar_required_z(hf, zz), (double) zlo);
        }
    }
    int numrs = (int) ((ceil(drhi) - floor(drlo)) * ACCEL_RDR + DBLCORRECT) + 1;
    if (numharm == 1 && harmnum == 1)
        numrs = obs->corr_uselen;
    else if (numrs % ACCEL_RDR)
        numrs = (numrs / ACCEL_RDR + 1) * ACCEL_RDR;
    const int numzs = shi->numkern_zdim, numws = shi->numkern_wdim;
    const int binoffset = shi->kern[0][0].kern_half_width;
    const int fftlen = shi->kern[0][0].fftlen;
    const long long lobin = rlo - binoffset;
    const int numdata = fftlen / ACCEL_NUMBETWEEN;

    /* the block's rows on the search grid */
    const int zoff = index_z((double) zlo, obs->zlo);
    const int woff = obs->numw ? index_w((double) wlo, obs->wlo) : 0;
    if (zoff < 0 || zoff + numzs > obs->numz || woff < 0 ||
        woff + numws > (obs->numw ? obs->numw : 1)) {
        sf_fail(-1, "subharm_fderivs_vol: z/w range outside the search grid");
        return NULL;
    }
    const long long hb0 = 2 * rlo;
    if (hb0 < 0 || hb0 + numrs > st->v.hb_end) {
        sf_fail(-1, "subharm_fderivs_vol: r = %.1f..%.1f outside the searched range", drlo,
                drhi);
        return NULL;
    }
    sf_ffd *f = ffd_new(st, rlo, zlo, wlo, numrs, numzs, numws);
    if (!f) {
        sf_fail(-1, "subharm_fderivs_vol: %s", agpu_last_error());
        return NULL;
    }
    f->fullrlo = fullrlo;
    f->pub.rinds = shi->rinds;
    f->pub.zinds = shi->zinds;
    int nrc;
    if (st->obs->nph <= 0.0 && st->obs->norm_type == 1) {
        /* -locpow (src/accel_utils.c:1100-1188): each bin has its own
         * factor, so the plane is correlated from the normalised window */
        const size_t wb = sizeof(fcomplex) * (size_t) numdata;
        if (wb > st->wcap) {
            if (st->d_wdata)
                agpu_free(st->d_wdata);
            st->d_wdata = NULL;
            st->wcap = 0;
            if (agpu_malloc(&st->d_wdata, wb)) {
                sf_fail(-1, "subharm_fderivs_vol: %s", agpu_last_error());
                return &f->pub;
            }
            st->wcap = wb;
        }
        int bad = agpu_sf_locpow_norm(st->v.d_spec, st->v.numbins, st->v.lobin, lobin, numdata,
                                      st->d_wdata, st->v.stream);
        if (!bad && numws == 1) {
            bad = agpu_sf_locpow_plane(st->d_wdata, numdata, binoffset, zlo, numzs, numrs,
                                       f->d_pow, st->v.stream);
        } else if (!bad) {
            const int t = locpow_taps(st, shi);
            bad = t < 0 ||
                  agpu_sf_locpow_plane_tab(st->d_wdata, numdata, binoffset, numzs * numws, numrs,
                                           st->taps[t].d_R, st->taps[t].d_nk, st->taps[t].stride,
                                           f->d_pow, st->v.stream);
        }
        if (bad)
            sf_fail(-1, "subharm_fderivs_vol: %s", agpu_last_error());
        return &f->pub;
    }
    if (st->obs->nph > 0.0) {   /* -photon, src/accel_utils.c:1080-1086 */
        const double nv = 1.0 / sqrt(st->obs->nph);
        const float nsq = (float) (nv * nv);
        nrc = agpu_h2d(st->d_norm, &nv, sizeof(nv), st->v.stream) ||
              agpu_h2d(st->d_nsq, &nsq, sizeof(nsq), st->v.stream);
    } else {
        nrc = agpu_block_norm1(st->v.d_spec, st->v.numbins, st->v.lobin, lobin, numdata,
                               st->d_norm, st->d_nsq, st->v.stream);
    }
    if (nrc ||
        agpu_sf_extract(st->v.d_plane, st->v.nrows, st->v.ntiles, st->v.numz, woff, zoff, numws,
                        numzs, numrs, hb0, st->d_nsq, f->d_pow, st->v.stream)) {
        sf_fail(-1, "subharm_fderivs_vol: %s", agpu_last_error());
        return &f->pub;
    }
    return &f->pub;
}

ffdotpows *copy_ffdotpows(ffdotpows *orig)
{
    sf_ffd *o = sf_of(orig);
    sf_ffd *f = ffd_new(o->st, orig->rlo, orig->zlo, orig->wlo, orig->numrs, orig->numzs,
                        orig->numws);
    if (!f) {
        sf_fail(-1, "copy_ffdotpows: %s", agpu_last_error());
        return NULL;
    }
    f->fullrlo = o->fullrlo;
    if (agpu_copy_d2d(f->d_pow, o->d_pow, f->bytes, o->st->v.stream)) {
        sf_fail(-1, "copy_ffdotpows: %s", agpu_last_error());
        return &f->pub;
    }
    return &f->pub;
}

void free_ffdotpows(ffdotpows *ffd)
{
    if (!ffd)
        return;
    sf_ffd *f = sf_of(ffd);
    sf_state *st = f->st;
    drop_host_copy(f);
    /* the stream may still use the buffer: later work on the same stream
     * is ordered after it, so the pool hands it out only to that work */
    pool_put(st, f->d_pow, f->bytes);
    free(f);
    state_release(st);
}

int accel_ffdotpows_sync(ffdotpows *ffd)
{
    sf_ffd *f = sf_of(ffd);
    const int nw = ffd->numws, nz = ffd->numzs;
    drop_host_copy(f);
    f->h_block = (float *) malloc(f->bytes ? f->bytes : 4);
    f->h_rows = (float **) malloc(sizeof(float *) * (size_t) nw * nz);
    float ***pl = (float ***) malloc(sizeof(float **) * (size_t) nw);
    if (!f->h_block || !f->h_rows || !pl) {
        free(pl);
        drop_host_copy(f);
        accel_set_error("accel_ffdotpows_sync: out of host memory");
        return -1;
    }
    for (int i = 0; i < nw; i++) {
        pl[i] = f->h_rows + (size_t) i * nz;
        for (int j = 0; j < nz; j++)
            pl[i][j] = f->h_block + ((size_t) i * nz + j) * ffd->numrs;
    }
    ffd->powers = pl;
    if (agpu_d2h(f->h_block, f->d_pow, f->bytes, f->st->v.stream)) {
        accel_set_error("accel_ffdotpows_sync: %s", agpu_last_error());
        return -1;
    }
    return 0;
}

/* add_ffdotpows, src/accel_utils.c:1261-1282 */
void add_ffdotpows(ffdotpows *fundamental, ffdotpows *subharmonic, int numharm, int harmnum)
{
    sf_ffd *fu = sf_of(fundamental), *su = sf_of(subharmonic);
    sf_state *st = fu->st;
    const accelobs *obs = st->obs;
    const double hf = (double) harmnum / (double) numharm;
    agpu_sf_add_args a;
    memset(&a, 0, sizeof(a));
    a.hf = hf;
    a.sub_fullrlo = su->fullrlo;
    a.sub_rlo = subharmonic->rlo;
    a.sub_zlo = subharmonic->zlo;
    a.sub_wlo = subharmonic->wlo;
    a.obs_zlo = obs ? obs->zlo : (double) fundamental->zlo;
    a.fund_wlo = fundamental->wlo;
    a.fnumws = fundamental->numws;
    a.fnumzs = fundamental->numzs;
    a.fnumrs = fundamental->numrs;
    a.snumws = subharmonic->numws;
    a.snumzs = subharmonic->numzs;
    a.snumrs = subharmonic->numrs;
    /* every index the device will form, checked at its extremes (the
     * rules are monotonic in r, z and w); inside [0, 65535] the tables'
     * unsigned short storage changes nothing */
    const int r0 = index_r(ar_required_r(hf, a.sub_fullrlo), (double) a.sub_rlo);
    const int r1 = index_r(ar_required_r(hf, a.sub_fullrlo + (a.fnumrs - 1) * ACCEL_DR),
                           (double) a.sub_rlo);
    const int z0 = index_z((double) ar_required_z(hf, a.obs_zlo), (double) a.sub_zlo);
    const int z1 = index_z((double) ar_required_z(hf, a.obs_zlo + (a.fnumzs - 1) * ACCEL_DZ),
                           (double) a.sub_zlo);
    const int w0 = index_w((double) ar_required_w(hf, (double) a.fund_wlo), (double) a.sub_wlo);
    const int w1 = index_w(
        (double) ar_required_w(hf, (double) (a.fund_wlo + (a.fnumws - 1) * ACCEL_DW)),
        (double) a.sub_wlo);
    if (r0 < 0 || r1 >= a.snumrs || z0 < 0 || z1 >= a.snumzs || w0 < 0 || w1 >= a.snumws) {
        sf_fail(-1, "add_ffdotpows: the sub-harmonic plane does not cover the fundamental");
        return;
    }
    drop_host_copy(fu);
    if (agpu_sf_add(fu->d_pow, su->d_pow, &a, st->v.stream))
        sf_fail(-1, "add_ffdotpows: %s", agpu_last_error());
}

/* add_ffdotpows_ptrs (src/accel_utils.c:1284-1309) forms the same sums */
void add_ffdotpows_ptrs(ffdotpows *fundamental, ffdotpows *subharmonic, int numharm, int harmnum)
{
    add_ffdotpows(fundamental, subharmonic, numharm, harmnum);
}

static int ensure_full(sf_state *st, const accelobs *obs)
{
    if (st->d_full)
        return 0;
    st->rlen = (obs->highestbin + obs->corr_uselen) * ACCEL_RDR;
    const size_t bytes = sizeof(float) * (size_t) st->rlen * obs->numz;
    if (agpu_malloc((void **) &st->d_full, bytes) ||
        agpu_memset(st->d_full, 0, bytes, st->v.stream)) {
        st->d_full = NULL;
        return -1;
    }
    return 0;
}

/* fund_to_ffdotplane, src/accel_utils.c:1219-1234 */
void fund_to_ffdotplane(ffdotpows *ffd, accelobs *obs)
{
    sf_ffd *f = sf_of(ffd);
    sf_state *st = f->st;
    if (!obs->inmem || ffd->numws != 1 || ffd->numzs != obs->numz) {
        sf_fail(-1, "fund_to_ffdotplane: needs an in-memory search and a fundamental plane");
        return;
    }
    if (ensure_full(st, obs)) {
        sf_fail(-1, "fund_to_ffdotplane: %s", agpu_last_error());
        return;
    }
    const long long col0 = ffd->rlo * ACCEL_RDR;
    if (col0 < 0 || col0 + ffd->numrs > st->rlen) {
        sf_fail(-1, "fund_to_ffdotplane: block outside the full plane");
        return;
    }
    if (agpu_sf_inmem_store(f->d_pow, ffd->numzs, ffd->numrs, st->d_full, st->rlen, col0,
                            st->v.stream))
        sf_fail(-1, "fund_to_ffdotplane: %s", agpu_last_error());
}

/* the transposed layout of src/accel_utils.c:1236-1250 is a host-memory
 * choice; on the device both forms keep one layout and the same values */
void fund_to_ffdotplane_trans(ffdotpows *ffd, accelobs *obs) { fund_to_ffdotplane(ffd, obs); }

/* inmem_add_ffdotpows, src/accel_utils.c:1312-1362 */
void inmem_add_ffdotpows(ffdotpows *fundamental, accelobs *obs, int numharm, int harmnum)
{
    sf_ffd *f = sf_of(fundamental);
    sf_state *st = f->st;
    const double hf = (double) harmnum / (double) numharm;
    const int rlo = (int) fundamental->rlo;
    const int numrs = fundamental->numrs, numzs = fundamental->numzs, zlo = fundamental->zlo;
    if (!st->d_full) {
        sf_fail(-1, "inmem_add_ffdotpows: the full plane is empty (fund_to_ffdotplane first)");
        return;
    }
    const int rmax = (int) ((ACCEL_RDR * rlo + numrs - 1) * hf + 0.5);
    const int rmin = (int) ((ACCEL_RDR * rlo) * hf + 0.5);
    const int z0 = index_z((double) ar_required_z(hf, (double) zlo), (double) zlo);
    const int z1 = index_z((double) ar_required_z(hf, (double) (zlo + (numzs - 1) * ACCEL_DZ)),
                           (double) zlo);
    if (rmin < 0 || rmax >= st->rlen || z0 < 0 || z1 >= obs->numz || fundamental->numws != 1) {
        sf_fail(-1, "inmem_add_ffdotpows: indices outside the full plane");
        return;
    }
    drop_host_copy(f);
    if (agpu_sf_inmem_add(f->d_pow, numzs, numrs, st->d_full, st->rlen, rlo, hf, zlo,
                          st->v.stream))
        sf_fail(-1, "inmem_add_ffdotpows: %s", agpu_last_error());
}

void inmem_add_ffdotpows_trans(ffdotpows *fundamental, accelobs *obs, int numharm, int harmnum)
{
    inmem_add_ffdotpows(fundamental, obs, numharm, harmnum);
}

/* ---------------------------------------------------------------------- */
/* candidates                                                              */

static accelcand *create_accelcand(float power, float sigma, int numharm, double r, double z,
                                   double w)
{
    accelcand *c = (accelcand *) calloc(1, sizeof(accelcand));
    if (!c)
        return NULL;
    c->power = power;
    c->sigma = sigma;
    c->numharm = numharm;
    c->r = r;
    c->z = z;
    c->w = w;
    return c;
}

void free_accelcand(void *data, void *user_data)
{
    (void) user_data;
    accelcand *c = (accelcand *) data;
    if (!c)
        return;
    if (c->pows) {
        free(c->pows);
        free(c->hirs);
        free(c->hizs);
        free(c->hiws);
        free(c->derivs);
    }
    free(c);
}

void print_accelcand(void *data, void *user_data)
{
    (void) user_data;
    const accelcand *c = (const accelcand *) data;
    printf("sigma: %-7.4f  pow: %-7.2f  harm: %-2d  r: %-14.4f  z: %-10.4f  w: %-10.2f\n",
           c->sigma, c->power, c->numharm, c->r, c->z, c->w);
}

static double cand_r(const GSList *n) { return ((const accelcand *) n->data)->r; }
static float cand_sigma(const GSList *n) { return ((const accelcand *) n->data)->sigma; }

static void replace_data(GSList *n, float power, float sigma, int numharm, double r, double z,
                         double w)
{
    free_accelcand(n->data, NULL);
    n->data = create_accelcand(power, sigma, numharm, r, z, w);
}

static GSList *node_new_cand(float power, float sigma, int numharm, double r, double z,
                            double w)
{
    return node_new(create_accelcand(power, sigma, numharm, r, z, w));
}

static void node_drop(GSList *n)
{
    free_accelcand(n->data, NULL);
    node_free(n);
}

/* insert_new_accelcand, src/accel_utils.c:387-469 */
ACCEL_DEFINE_INSERT(insert_new_accelcand, GSList, cand_r, cand_sigma, replace_data, node_new_cand,
                    node_drop)

static int cmp_hit_idx(const void *a, const void *b)
{
    const long long x = ((const agpu_sf_hit *) a)->idx, y = ((const agpu_sf_hit *) b)->idx;
    return (x > y) - (x < y);
}

/* search_ffdotpows, src/accel_utils.c:1412-1452 (the serial loop order:
 * w, then z, then r) */
GSList *search_ffdotpows(ffdotpows *ffdot, int numharm, accelobs *obs, GSList *cands)
{
    sf_ffd *f = sf_of(ffdot);
    sf_state *st = f->st;
    const int stage = twon_to_index(numharm);
    const float powcut = obs->powcut[stage];
    const long long numindep = obs->numindep[stage];
    const long long n = (long long) ffdot->numws * ffdot->numzs * ffdot->numrs;
    unsigned cnt = 0;
    for (int pass = 0; pass < 2; pass++) {
        if (agpu_memset(st->d_count, 0, sizeof(unsigned), st->v.stream) ||
            agpu_sf_threshold(f->d_pow, n, powcut, st->d_hits, st->d_count, st->cap,
                              st->v.stream) ||
            agpu_d2h(&cnt, st->d_count, sizeof(unsigned), st->v.stream)) {
            sf_fail(-1, "search_ffdotpows: %s", agpu_last_error());
            return cands;
        }
        if (cnt <= st->cap)
            break;
        agpu_free(st->d_hits);
        st->d_hits = NULL;
        st->cap = cnt + cnt / 4 + 1024;
        if (agpu_malloc((void **) &st->d_hits, sizeof(agpu_sf_hit) * st->cap)) {
            st->cap = 0;
            sf_fail(-1, "search_ffdotpows: %s", agpu_last_error());
            return cands;
        }
    }
    if (!cnt)
        return cands;
    agpu_sf_hit *h = (agpu_sf_hit *) malloc(sizeof(agpu_sf_hit) * cnt);
    float *pw = (float *) malloc(sizeof(float) * cnt);
    float *sg = (float *) malloc(sizeof(float) * cnt);
    if (!h || !pw || !sg ||
        agpu_d2h(h, st->d_hits, sizeof(agpu_sf_hit) * cnt, st->v.stream)) {
        free(h);
        free(pw);
        free(sg);
        sf_fail(-1, "search_ffdotpows: %s", agpu_last_error());
        return cands;
    }
    qsort(h, cnt, sizeof(agpu_sf_hit), cmp_hit_idx);
    for (unsigned i = 0; i < cnt; i++)
        pw[i] = h[i].pow;
    accel_sigmas(pw, cnt, numharm, (double) numindep, sg);
    const long long plane = (long long) ffdot->numzs * ffdot->numrs;
    for (unsigned i = 0; i < cnt; i++) {
        const long long ii = h[i].idx / plane, rem = h[i].idx - ii * plane;
        const long long jj = rem / ffdot->numrs, kk = rem - jj * ffdot->numrs;
        const float pow = h[i].pow, sig = sg[i];
        const double rr = (ffdot->rlo + kk * (double) ACCEL_DR) / (double) numharm;
        const double zz = (ffdot->zlo + jj * (double) ACCEL_DZ) / (double) numharm;
        const double ww = (ffdot->wlo + ii * (double) ACCEL_DW) / (double) numharm;
        int added = 0;
        cands = insert_new_accelcand(cands, pow, sig, numharm, rr, zz, ww, &added);
        if (added && !obs->dat_input && obs->workfile)
            fprintf(obs->workfile, "%-7.2f  %-7.4f  %-2d  %-14.4f  %-14.9f  %-10.4f %-10.4f\n",
                    pow, sig, numharm, rr, rr / obs->T, zz, ww);
    }
    free(h);
    free(pw);
    free(sg);
    return cands;
}

static int cmp_cand(const GSList *a, const GSList *b)
{
    const accelcand *x = (const accelcand *) a->data, *y = (const accelcand *) b->data;
    int r = (x->sigma < y->sigma) - (x->sigma > y->sigma);
    return r ? r : (x->power < y->power) - (x->power > y->power);
}

/* sort_accelcands, src/accel_utils.c:380-384: a stable merge sort, as
 * g_slist_sort */
GSList *sort_accelcands(GSList *list)
{
    if (!list || !list->next)
        return list;
    GSList *slow = list, *fast = list->next;
    while (fast && fast->next) {
        slow = slow->next;
        fast = fast->next->next;
    }
    GSList *b = slow->next;
    slow->next = NULL;
    GSList *a = sort_accelcands(list);
    b = sort_accelcands(b);
    GSList head, *t = &head;
    while (a && b) {
        if (cmp_cand(a, b) <= 0) {
            t->next = a;
            a = a->next;
        } else {
            t->next = b;
            b = b->next;
        }
        t = t->next;
    }
    t->next = a ? a : b;
    return head.next;
}

ACCEL_DEFINE_ELIMINATE(eliminate_list, GSList, cand_r, node_drop)

/* eliminate_harmonics, src/accel_utils.c:472-549 */
GSList *eliminate_harmonics(GSList *cands, int *numcands)
{
    const int before = *numcands;
    cands = eliminate_list(cands, numcands);
    if (before - *numcands)
        printf("Removed %d likely harmonically related candidates.\n", before - *numcands);
    return cands;
}

/* optimize_accelcand, src/accel_utils.c:552-626: harmonic polishing, or
 * (-noharmpolish) each harmonic refined on its own */
void optimize_accelcand(accelcand *cand, accelobs *obs)
{
    sf_state *st = state_of(obs);
    const int nh = cand->numharm;
    cand->pows = (double *) calloc((size_t) nh, sizeof(double));
    cand->hirs = (double *) calloc((size_t) nh, sizeof(double));
    cand->hizs = (double *) calloc((size_t) nh, sizeof(double));
    cand->hiws = (double *) calloc((size_t) nh, sizeof(double));
    cand->derivs = (rderivs *) calloc((size_t) nh, sizeof(rderivs));
    if (!st || nh < 1 || nh > ACCEL_MAXHARM) {
        sf_fail(-1, "optimize_accelcand: numharm %d outside 1..%d, or no search state", nh,
                ACCEL_MAXHARM);
        return;
    }
    if (!obs->use_harmonic_polishing) {
        /* src/accel_utils.c:582-622: max_rz_arr / max_rzw_arr per harmonic */
        int one[ACCEL_MAXHARM];
        double v[6][ACCEL_MAXHARM];
        double *vp = (double *) calloc((size_t) nh * ACCEL_MAXHARM, sizeof(double));
        rderivs *vd = (rderivs *) calloc((size_t) nh * ACCEL_MAXHARM, sizeof(rderivs));
        if (!vp || !vd) {
            free(vp);
            free(vd);
            sf_fail(-1, "optimize_accelcand: out of host memory");
            return;
        }
        for (int i = 0; i < nh; i++) {
            one[i] = 1;
            v[0][i] = cand->r * (i + 1) - (double) obs->lobin;
            v[1][i] = cand->z * (i + 1);
            v[2][i] = cand->w * (i + 1);
        }
        int rc = accel_search_refine_dev(st->s, 1, nh, one, v[0], v[1],
                                         obs->numw ? v[2] : NULL, v[3], v[4], v[5], vp, vd);
        for (int i = 0; i < nh && !rc; i++) {
            cand->pows[i] = vp[i * ACCEL_MAXHARM];
            cand->derivs[i] = vd[i * ACCEL_MAXHARM];
            cand->hirs[i] = v[3][i] + obs->lobin;
            cand->hizs[i] = v[4][i];
            cand->hiws[i] = obs->numw ? v[5][i] : 0.0;
        }
        free(vp);
        free(vd);
        if (rc) {
            sf_fail(-1, "optimize_accelcand: %s", accel_last_error());
            return;
        }
    } else {
        double pows[ACCEL_MAXHARM];
        rderivs der[ACCEL_MAXHARM];
        double r = 0.0, z = 0.0, w = 0.0;
        const double rin = cand->r - (double) obs->lobin, zin = cand->z, win = cand->w;
        if (accel_search_refine_dev(st->s, 0, 1, &nh, &rin, &zin, obs->numw ? &win : NULL, &r,
                                    &z, &w, pows, der)) {
            sf_fail(-1, "optimize_accelcand: %s", accel_last_error());
            return;
        }
        for (int i = 0; i < nh; i++) {
            cand->pows[i] = pows[i];
            cand->derivs[i] = der[i];
            cand->hirs[i] = (r + obs->lobin) * (i + 1);
            cand->hizs[i] = z * (i + 1);
            cand->hiws[i] = obs->numw ? w * (i + 1) : 0.0;
        }
    }
    cand->sigma = (float) accel_candidate_sigma(cand->power, nh,
                                                (double) obs->numindep[twon_to_index(nh)]);
}

/* ---------------------------------------------------------------------- */
/* output                                                                  */

static accel_out_cand *out_cands(GSList *list, int *n)
{
    *n = slist_length(list);
    accel_out_cand *oc = (accel_out_cand *) calloc((size_t) (*n ? *n : 1),
                                                   sizeof(accel_out_cand));
    int i = 0;
    for (GSList *l = list; l && oc; l = l->next, i++) {
        const accelcand *c = (const accelcand *) l->data;
        oc[i].power = c->power;
        oc[i].sigma = c->sigma;
        oc[i].numharm = c->numharm;
        oc[i].r = c->r;
        oc[i].z = c->z;
        oc[i].w = c->w;
        oc[i].derivs = c->derivs;
        oc[i].hirs = c->hirs;
        oc[i].hizs = c->hizs;
        oc[i].hiws = c->hiws;
    }
    return oc;
}

/* output_fundamentals, src/accel_utils.c:666-842 (the pulsar-catalogue
 * notes of real telescopes are not built) */
void output_fundamentals(fourierprops *props, GSList *list, accelobs *obs, infodata *idata)
{
    (void) idata;
    int n;
    accel_out_cand *oc = out_cands(list, &n);
    for (int i = 0; i < n; i++)
        if (!oc[i].derivs) {
            free(oc);
            sf_fail(-1, "output_fundamentals: candidate %d was not optimized", i + 1);
            return;
        }
    if (!obs->dat_input && obs->workfile)
        fclose(obs->workfile);
    obs->workfile = fopen(obs->accelnm, "w");
    if (!obs->workfile) {
        free(oc);
        sf_fail(-1, "Error opening '%s'", obs->accelnm);
        return;
    }
    accel_out_fundamentals(obs->workfile, oc, props, n, obs->T, obs->numw != 0);
    fflush(obs->workfile);
    free(oc);
}

/* output_harmonics, src/accel_utils.c:844-979; the .inf is appended as
 * the reference's "cat <root>.inf >> <accelnm>" does */
void output_harmonics(GSList *list, accelobs *obs, infodata *idata)
{
    (void) idata;
    int n;
    accel_out_cand *oc = out_cands(list, &n);
    if (!obs->workfile) {
        free(oc);
        sf_fail(-1, "output_harmonics: output_fundamentals first");
        return;
    }
    accel_out_harmonics(obs->workfile, oc, n, obs->T, obs->numw != 0);
    free(oc);
    char inf[4200];
    snprintf(inf, sizeof(inf), "%s.inf", obs->rootfilenm);
    accel_out_append_file(obs->workfile, inf);
    fclose(obs->workfile);
    obs->workfile = NULL;
}

/* ---------------------------------------------------------------------- */
/* layout report                                                           */

#define OFF(T, m) (long long) offsetof(T, m)

int accel_surface_layout(const char *name, long long *out, int n)
{
    long long v[64];
    int k = 0;
#define PUT(x) (v[k++] = (x))
    if (!strcmp(name, "accelobs")) {
        PUT(sizeof(accelobs));
        PUT(OFF(accelobs, N)); PUT(OFF(accelobs, numbins)); PUT(OFF(accelobs, lobin));
        PUT(OFF(accelobs, highestbin)); PUT(OFF(accelobs, maxkernlen));
        PUT(OFF(accelobs, corr_uselen)); PUT(OFF(accelobs, fftlen));
        PUT(OFF(accelobs, numharmstages)); PUT(OFF(accelobs, numz)); PUT(OFF(accelobs, numw));
        PUT(OFF(accelobs, numbetween)); PUT(OFF(accelobs, numzap)); PUT(OFF(accelobs, dat_input));
        PUT(OFF(accelobs, mmap_file)); PUT(OFF(accelobs, inmem)); PUT(OFF(accelobs, norm_type));
        PUT(OFF(accelobs, dt)); PUT(OFF(accelobs, T)); PUT(OFF(accelobs, rlo));
        PUT(OFF(accelobs, rhi)); PUT(OFF(accelobs, dr)); PUT(OFF(accelobs, zlo));
        PUT(OFF(accelobs, zhi)); PUT(OFF(accelobs, dz)); PUT(OFF(accelobs, wlo));
        PUT(OFF(accelobs, whi)); PUT(OFF(accelobs, dw)); PUT(OFF(accelobs, baryv));
        PUT(OFF(accelobs, nph)); PUT(OFF(accelobs, sigma)); PUT(OFF(accelobs, powcut));
        PUT(OFF(accelobs, ffdotplane)); PUT(OFF(accelobs, lobins)); PUT(OFF(accelobs, hibins));
        PUT(OFF(accelobs, numindep)); PUT(OFF(accelobs, fftfile)); PUT(OFF(accelobs, workfile));
        PUT(OFF(accelobs, fft)); PUT(OFF(accelobs, rootfilenm)); PUT(OFF(accelobs, candnm));
        PUT(OFF(accelobs, accelnm)); PUT(OFF(accelobs, workfilenm));
        PUT(OFF(accelobs, use_harmonic_polishing));
    } else if (!strcmp(name, "accelcand")) {
        PUT(sizeof(accelcand));
        PUT(OFF(accelcand, power)); PUT(OFF(accelcand, sigma)); PUT(OFF(accelcand, numharm));
        PUT(OFF(accelcand, r)); PUT(OFF(accelcand, z)); PUT(OFF(accelcand, w));
        PUT(OFF(accelcand, pows)); PUT(OFF(accelcand, hirs)); PUT(OFF(accelcand, hizs));
        PUT(OFF(accelcand, hiws)); PUT(OFF(accelcand, derivs));
    } else if (!strcmp(name, "kernel")) {
        PUT(sizeof(kernel));
        PUT(OFF(kernel, z)); PUT(OFF(kernel, w)); PUT(OFF(kernel, fftlen));
        PUT(OFF(kernel, numgoodbins)); PUT(OFF(kernel, numbetween));
        PUT(OFF(kernel, kern_half_width)); PUT(OFF(kernel, data));
    } else if (!strcmp(name, "subharminfo")) {
        PUT(sizeof(subharminfo));
        PUT(OFF(subharminfo, numharm)); PUT(OFF(subharminfo, harmnum));
        PUT(OFF(subharminfo, zmax)); PUT(OFF(subharminfo, wmax));
        PUT(OFF(subharminfo, numkern_zdim)); PUT(OFF(subharminfo, numkern_wdim));
        PUT(OFF(subharminfo, numkern)); PUT(OFF(subharminfo, kern));
        PUT(OFF(subharminfo, rinds)); PUT(OFF(subharminfo, zinds));
    } else if (!strcmp(name, "ffdotpows")) {
        PUT(sizeof(ffdotpows));
        PUT(OFF(ffdotpows, rlo)); PUT(OFF(ffdotpows, zlo)); PUT(OFF(ffdotpows, wlo));
        PUT(OFF(ffdotpows, numrs)); PUT(OFF(ffdotpows, numzs)); PUT(OFF(ffdotpows, numws));
        PUT(OFF(ffdotpows, powers)); PUT(OFF(ffdotpows, rinds)); PUT(OFF(ffdotpows, zinds));
    } else if (!strcmp(name, "infodata")) {
        PUT(sizeof(infodata));
        PUT(OFF(infodata, ra_s)); PUT(OFF(infodata, dec_s)); PUT(OFF(infodata, N));
        PUT(OFF(infodata, dt)); PUT(OFF(infodata, fov)); PUT(OFF(infodata, mjd_f));
        PUT(OFF(infodata, dm)); PUT(OFF(infodata, freq)); PUT(OFF(infodata, freqband));
        PUT(OFF(infodata, chan_wid)); PUT(OFF(infodata, wavelen)); PUT(OFF(infodata, waveband));
        PUT(OFF(infodata, energy)); PUT(OFF(infodata, energyband)); PUT(OFF(infodata, onoff));
        PUT(OFF(infodata, num_chan)); PUT(OFF(infodata, mjd_i)); PUT(OFF(infodata, ra_h));
        PUT(OFF(infodata, ra_m)); PUT(OFF(infodata, dec_d)); PUT(OFF(infodata, dec_m));
        PUT(OFF(infodata, bary)); PUT(OFF(infodata, numonoff)); PUT(OFF(infodata, notes));
        PUT(OFF(infodata, name)); PUT(OFF(infodata, object)); PUT(OFF(infodata, instrument));
        PUT(OFF(infodata, observer)); PUT(OFF(infodata, analyzer));
        PUT(OFF(infodata, telescope)); PUT(OFF(infodata, band)); PUT(OFF(infodata, filt));
    } else if (!strcmp(name, "Cmdline")) {
        PUT(sizeof(Cmdline));
        PUT(OFF(Cmdline, ncpusP)); PUT(OFF(Cmdline, ncpus)); PUT(OFF(Cmdline, ncpusC));
        PUT(OFF(Cmdline, lobinP)); PUT(OFF(Cmdline, lobin)); PUT(OFF(Cmdline, lobinC));
        PUT(OFF(Cmdline, numharmP)); PUT(OFF(Cmdline, numharm)); PUT(OFF(Cmdline, numharmC));
        PUT(OFF(Cmdline, zmaxP)); PUT(OFF(Cmdline, zmax)); PUT(OFF(Cmdline, zmaxC));
        PUT(OFF(Cmdline, wmaxP)); PUT(OFF(Cmdline, wmax)); PUT(OFF(Cmdline, wmaxC));
        PUT(OFF(Cmdline, sigmaP)); PUT(OFF(Cmdline, sigma)); PUT(OFF(Cmdline, sigmaC));
        PUT(OFF(Cmdline, rloP)); PUT(OFF(Cmdline, rlo)); PUT(OFF(Cmdline, rloC));
        PUT(OFF(Cmdline, rhiP)); PUT(OFF(Cmdline, rhi)); PUT(OFF(Cmdline, rhiC));
        PUT(OFF(Cmdline, floP)); PUT(OFF(Cmdline, flo)); PUT(OFF(Cmdline, floC));
        PUT(OFF(Cmdline, fhiP)); PUT(OFF(Cmdline, fhi)); PUT(OFF(Cmdline, fhiC));
        PUT(OFF(Cmdline, inmemP)); PUT(OFF(Cmdline, photonP)); PUT(OFF(Cmdline, medianP));
        PUT(OFF(Cmdline, locpowP)); PUT(OFF(Cmdline, zaplistP)); PUT(OFF(Cmdline, zaplist));
        PUT(OFF(Cmdline, zaplistC)); PUT(OFF(Cmdline, baryvP)); PUT(OFF(Cmdline, baryv));
        PUT(OFF(Cmdline, baryvC)); PUT(OFF(Cmdline, otheroptP));
        PUT(OFF(Cmdline, noharmpolishP)); PUT(OFF(Cmdline, noharmremoveP));
        PUT(OFF(Cmdline, argc)); PUT(OFF(Cmdline, argv)); PUT(OFF(Cmdline, full_cmd_line));
    } else if (!strcmp(name, "fourierprops")) {
        PUT(sizeof(fourierprops));
        PUT(OFF(fourierprops, r)); PUT(OFF(fourierprops, rerr)); PUT(OFF(fourierprops, z));
        PUT(OFF(fourierprops, zerr)); PUT(OFF(fourierprops, w)); PUT(OFF(fourierprops, werr));
        PUT(OFF(fourierprops, pow)); PUT(OFF(fourierprops, powerr)); PUT(OFF(fourierprops, sig));
        PUT(OFF(fourierprops, rawpow)); PUT(OFF(fourierprops, phs)); PUT(OFF(fourierprops, phserr));
        PUT(OFF(fourierprops, cen)); PUT(OFF(fourierprops, cenerr)); PUT(OFF(fourierprops, pur));
        PUT(OFF(fourierprops, purerr)); PUT(OFF(fourierprops, locpow));
    } else if (!strcmp(name, "rderivs")) {
        PUT(sizeof(rderivs));
        PUT(OFF(rderivs, pow)); PUT(OFF(rderivs, phs)); PUT(OFF(rderivs, dpow));
        PUT(OFF(rderivs, dphs)); PUT(OFF(rderivs, d2pow)); PUT(OFF(rderivs, d2phs));
        PUT(OFF(rderivs, locpow));
    } else {
        return -1;
    }
#undef PUT
    for (int i = 0; i < k && i < n; i++)
        out[i] = v[i];
    return k;
}
