/*
 * accel_host.c -- the C host of the MI355X accelsearch path.
 *
 * Computes the reference's geometry and index rules on the host, lays out
 * device state (templates, raw plane, normalisation windows, index maps,
 * hit buffer), launches the HIP kernels through agpu.h and replays the
 * threshold crossings through the reference's candidate logic.
 *
 * Device layout (one search):
 *   spectrum   float2[numbins]                       (bin lobin first)
 *   templates  float2[numw'][numz][L]                L = 4096, FP32, 1/L folded in
 *   raw plane  float, tiles of 2048 half-bins x every  column = absolute half-bin,
 *              (w,z) row (AGPU_PLANE_IDX)             un-normalised |corr|^2
 *   windows    {lobin, numdata}[nwin] -> norm (FP64), nsq = norm^2 (FP32)
 *   zmap/wmap  int[nfrac][numz] / int[nfrac][numw']  sub-harmonic rows
 *   hits       accel_hit[cap] + counter
 * Correlation is linear, so the reference's normalised plane of any block
 * and any harmonic fraction is norm^2 times the raw plane at the same
 * (r, z, w); the harmonic-sum kernel applies the norm of the window that
 * the reference would have used for that block and fraction.
 */
#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>
#include "presto_accel.h"
#include "accel_internal.h"
#include "accel_rules.h"
#include "agpu.h"

#define MAXRAMUSE 2000000000LL
#define CORR_L 4096
#define HIT_CAP0 (1u << 20)

static __thread char g_err[1024];   /* per host thread */

static void set_err(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

const char *accel_last_error(void) { return g_err; }

void accel_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

#define GPU(call)                                                                \
    do {                                                                         \
        if ((call) != 0) {                                                       \
            set_err("%s: %s", #call, agpu_last_error());                         \
            return -1;                                                           \
        }                                                                        \
    } while (0)

static double now_ms(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return 1e3 * (double) t.tv_sec + 1e-6 * (double) t.tv_nsec;
}

/* ---- index rules (accel_rules.h) ---- */
/* First plane bin of the block that starts at startr: the fundamental's
 * ffdot->rlo = floor(calc_required_r(1, startr)), src/accel_utils.c:1035-1037
 * (startr itself when it is whole). */
static long long fund_rlo(double startr) { return (long long) floor(ar_required_r(1.0, startr)); }

static int fftlen_from_kernwidth(int kw)
{
    static const int lim[9] = {6, 52, 67, 378, 664, 1672, 3015, 3554, 6000};
    static const int len[9] = {128, 256, 512, 1024, 2048, 4096, 10240, 15360, 25600};
    for (int i = 0; i < 9; i++)
        if (kw < lim[i])
            return len[i];
    return (int) ar_next2((long long) kw * 5);
}


void accel_params_default(accel_params *p)
{
    memset(p, 0, sizeof(*p));
    p->numharm = 8;
    p->zmax = 200;
    p->sigma = 2.0f;
    p->flo = 1.0;
    p->fhi = 10000.0;
    p->floP = 1;
    p->fhiP = 1;
}

static int log2i(int n)
{
    int x = 0;
    while (n > 1) {
        n >>= 1;
        x++;
    }
    return x;
}

/* src/accel_utils.c:1596-1804 */
int accel_geom_create(const accel_params *p, accel_geom *g)
{
    memset(g, 0, sizeof(*g));
    int zmax = p->zmax, wmax = p->wmax;
    if (zmax < 0) {
        set_err("zmax must be >= 0");
        return -1;
    }
    if (zmax % ACCEL_DZ)
        zmax = (zmax / ACCEL_DZ + 1) * ACCEL_DZ;
    g->N = (long long) p->N;
    g->numbins = p->numbins;
    g->lobin = p->lobin;
    g->nph = (p->lobin > 0 || !(p->nph > 0.0)) ? 0.0 : p->nph;
    if (g->lobin > 0 && g->lobin > g->numbins - 1) {
        set_err("'lobin' is greater than the total number of frequencies in the data set");
        return -1;
    }
    if (p->numharm != 1 && p->numharm != 2 && p->numharm != 4 && p->numharm != 8 &&
        p->numharm != 16 && p->numharm != 32) {
        set_err("'numharm' = %d must be a power-of-two", p->numharm);
        return -1;
    }
    g->numharmstages = log2i(p->numharm) + 1;
    g->dz = ACCEL_DZ;
    g->numz = (zmax / ACCEL_DZ) * 2 + 1;
    if (p->wmaxP) {
        if (wmax % ACCEL_DW)
            wmax = (wmax / ACCEL_DW + 1) * ACCEL_DW;
        g->whi = wmax;
        g->wlo = -wmax;
        g->dw = ACCEL_DW;
        g->numw = (wmax / ACCEL_DW) * 2 + 1;
        if (wmax == 0)
            g->numw = 0;
    }
    g->numbetween = ACCEL_NUMBETWEEN;
    g->dt = p->dt;
    g->T = p->dt * p->N;
    if (p->floP) {
        g->rlo = floor(p->flo * g->T);
    } else {
        g->rlo = p->rloP ? p->rlo : 1.0;
    }
    if (g->rlo < g->lobin)
        g->rlo = (double) g->lobin;
    if (g->rlo > g->numbins - 1) {
        set_err("low frequency to search is above the highest available frequency");
        return -1;
    }
    g->highestbin = g->numbins - 1;
    g->rhi = (double) g->highestbin;
    if (p->fhiP || p->rhiP) {
        g->highestbin = p->fhiP ? (long long) ceil(p->fhi * g->T) : (long long) p->rhi;
        if (g->highestbin > g->numbins - 1)
            g->highestbin = g->numbins - 1;
        g->rhi = (double) g->highestbin;
        if (g->highestbin < g->rlo) {
            set_err("high frequency to search is below the low frequency");
            return -1;
        }
    }
    g->dr = ACCEL_DR;
    g->zhi = zmax;
    g->zlo = -zmax;
    g->sigma = p->sigma;
    for (int ii = 0; ii < g->numharmstages; ii++) {
        if (g->numz == 1 && g->numw == 0)
            g->numindep[ii] = (long long) ((g->rhi - g->rlo) / (1 << ii));
        else if (g->numz > 1 && g->numw == 0)
            g->numindep[ii] = (long long) ((g->rhi - g->rlo) * (g->numz + 1) *
                                           (g->dz / 6.95) / (1 << ii));
        else
            g->numindep[ii] = (long long) ((g->rhi - g->rlo) * (g->numz + 1) *
                                           (g->dz / 6.95) * (g->numw + 1) *
                                           (g->dw / 44.2) / (1 << ii));
        g->powcut[ii] = (float) accel_power_for_sigma(g->sigma, 1 << ii,
                                                      (double) g->numindep[ii]);
    }
    g->maxkernlen = 2 * ACCEL_NUMBETWEEN * ar_w_halfwidth(g->zhi, g->whi);
    g->fftlen = fftlen_from_kernwidth(g->maxkernlen);
    if (g->fftlen < 2048)
        g->fftlen = 2048;
    g->corr_uselen = g->fftlen - g->maxkernlen;
    if (g->corr_uselen % ACCEL_RDR)
        g->corr_uselen = g->corr_uselen / ACCEL_RDR * ACCEL_RDR;
    long long memuse = 4LL * (g->highestbin + g->corr_uselen) * g->numbetween * g->numz *
                       (p->wmaxP ? g->numw : 1);
    g->inmem = (!p->wmaxP && (memuse < MAXRAMUSE || p->inmem)) ? 1 : 0;
    if (p->force_standard)
        g->inmem = 0;
    /* src/accelsearch.c:186-227 */
    int rstep = (int) (g->corr_uselen * ACCEL_DR);
    double startr = g->rlo;
    int nb = 0;
    while (startr + rstep < g->highestbin) {
        startr += rstep;
        nb++;
    }
    g->nblocks = nb;
    return 0;
}

/* ---- candidate list: src/accel_utils.c:330-549 ---- */
typedef struct cnode {
    accel_cand c;
    struct cnode *next;
} cnode;

static cnode *cnew(float power, float sigma, int numharm, double r, double z, double w)
{
    cnode *n = (cnode *) calloc(1, sizeof(cnode));
    n->c.power = power;
    n->c.sigma = sigma;
    n->c.numharm = numharm;
    n->c.r = r;
    n->c.z = z;
    n->c.w = w;
    return n;
}

static void cset(cnode *n, float power, float sigma, int numharm, double r, double z, double w)
{
    n->c.power = power;
    n->c.sigma = sigma;
    n->c.numharm = numharm;
    n->c.r = r;
    n->c.z = z;
    n->c.w = w;
}

#define CN_R(n) ((n)->c.r)
#define CN_SIGMA(n) ((n)->c.sigma)
/* insert_new_accelcand, src/accel_utils.c:387-469 */
ACCEL_DEFINE_INSERT(insert_cand, cnode, CN_R, CN_SIGMA, cset, cnew, free)

/* compare_accelcand_sigma, src/accel_utils.c:364-377 */
static int cmp_sigma(const accel_cand *a, const accel_cand *b)
{
    int r = (a->sigma < b->sigma) - (a->sigma > b->sigma);
    return r ? r : (a->power < b->power) - (a->power > b->power);
}

/* sort_accelcands: a stable merge sort like g_slist_sort */
static cnode *sort_cands(cnode *l)
{
    if (!l || !l->next)
        return l;
    cnode *slow = l, *fast = l->next;
    while (fast && fast->next) {
        slow = slow->next;
        fast = fast->next->next;
    }
    cnode *b = slow->next;
    slow->next = NULL;
    cnode *a = sort_cands(l);
    b = sort_cands(b);
    cnode head, *t = &head;
    while (a && b) {
        if (cmp_sigma(&a->c, &b->c) <= 0) {
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

/* eliminate_harmonics, src/accel_utils.c:472-549 */
ACCEL_DEFINE_ELIMINATE(eliminate_harmonics, cnode, CN_R, free)

static void free_cands(cnode *l)
{
    while (l) {
        cnode *n = l->next;
        free(l);
        l = n;
    }
}

/* ---- the search ---- */
typedef struct refine_ws refine_ws;
static void ws_free(refine_ws *w);
struct accel_search {
    accel_params p;
    accel_geom g;
    int numw1;            /* max(numw, 1) */
    int nt;               /* templates = numw1 * numz */
    int hw;               /* largest kernel half width (bins) */
    int L, uselen, groups;
    long long nseg;
    long long hb_end, ntiles;
    int nfrac;
    int npre;             /* in-memory pre-population blocks */
    int nwin, maxn;
    void *stream;
    void *stream2;                    /* the block norms, beside the correlation */
    void *ev[5];
    void *ev_norms;
    void *d_spec;
    int own_spec;
    void *d_T;
    float *d_plane;
    agpu_window *d_win;
    double *d_norm;
    float *d_nsq;
    float *d_nsq_hb;      /* in-memory path: per-half-bin norm^2 */
    int *d_zmap, *d_wmap;
    unsigned *d_rowtab;   /* harmonic-sum address tables (agpu.h) */
    long long *d_cbase;
    int *d_clo;
    agpu_hit *d_hits;
    unsigned *d_nhits;
    unsigned cap;
    /* block start r, accumulated as src/accelsearch.c:189-224 does; a
     * fractional -rlo makes the sub-harmonic lookups depend on it */
    double *bstart;
    double *d_bstart;     /* device copy, only for a start r off the half-bin grid */
    int rh_off;           /* 2 startr - 2 fund_rlo(startr) when that is 0 or 1 */
    FILE *workfile;       /* .txtcand of an .fft input, or NULL */
    /* results */
    accel_hit *hits;
    long long nhits;
    cnode *cands;
    int ncands;
    float ms[5];
    double hms[4];                    /* host: launch, wait, copy, replay */
    /* refinement of the current list (accel_search_optimize) */
    int nopt;
    int *opt_nh;
    double *opt_r, *opt_z, *opt_w;    /* fundamental optimum; r includes lobin */
    double *opt_h;                    /* [nopt][3][ACCEL_MAXHARM]: every harmonic's
                                       * r (with lobin), z, w */
    double *opt_pows;                 /* [nopt][ACCEL_MAXHARM] */
    accel_rderivs *opt_derivs;        /* [nopt][ACCEL_MAXHARM] */
    refine_ws *ws;                    /* refinement device buffers */
    /* time-series input on the device (accel_search_set_series) */
    void *d_fftwork;
    agpu_window *d_dblk;
    float *d_dmean;
    int ndblk, dmaxn;
};

const accel_geom *accel_search_geom(const accel_search *s) { return &s->g; }

static int build_tables(accel_search *s, agpu_window **win_out, int **zmap_out,
                        int **wmap_out)
{
    const accel_geom *g = &s->g;
    const int numz = g->numz, numw1 = s->numw1;
    const int rstep = g->corr_uselen / 2;
    const int zmax = (int) g->zhi, wmax = (int) g->whi;
    s->nfrac = (1 << (g->numharmstages - 1)) - 1;
    int *zmap = (int *) calloc((size_t) (s->nfrac > 0 ? s->nfrac : 1) * numz, sizeof(int));
    int *wmap = (int *) calloc((size_t) (s->nfrac > 0 ? s->nfrac : 1) * numw1, sizeof(int));
    int f = 0;
    for (int st = 1; st < g->numharmstages; st++) {
        int H = 1 << st;
        for (int k = 1; k < H; k += 2, f++) {
            double hf = (double) k / (double) H;
            for (int zi = 0; zi < numz; zi++)
                zmap[f * numz + zi] = (ar_required_z(hf, g->zlo + zi * ACCEL_DZ) + zmax) / ACCEL_DZ;
            for (int wi = 0; wi < numw1; wi++)
                wmap[f * numw1 + wi] =
                    g->numw ? (ar_required_w(hf, g->wlo + wi * ACCEL_DW) + wmax) / ACCEL_DW : 0;
        }
    }
    /* normalisation windows */
    agpu_window *win;
    int fund_hw = ar_w_halfwidth(-g->zhi, -g->whi);
    int fund_nd = g->fftlen / ACCEL_NUMBETWEEN;
    s->maxn = fund_nd;
    if (g->inmem) {
        s->npre = 0;
        for (double sr = 8; sr < g->rlo; sr += rstep)
            s->npre++;
        s->nwin = s->npre + g->nblocks;
        win = (agpu_window *) calloc(s->nwin > 0 ? s->nwin : 1, sizeof(agpu_window));
        int i = 0;
        for (double sr = 8; sr < g->rlo; sr += rstep, i++) {
            win[i].lobin = (long long) floor(sr) - fund_hw;
            win[i].numdata = fund_nd;
        }
        for (int b = 0; b < g->nblocks; b++, i++) {
            win[i].lobin = fund_rlo(s->bstart[b]) - fund_hw;
            win[i].numdata = fund_nd;
        }
    } else {
        s->npre = 0;
        s->nwin = g->nblocks * (s->nfrac + 1);
        win = (agpu_window *) calloc(s->nwin > 0 ? s->nwin : 1, sizeof(agpu_window));
        int hwf[64], ndf[64];
        f = 0;
        for (int st = 1; st < g->numharmstages; st++) {
            int H = 1 << st;
            for (int k = 1; k < H; k += 2, f++) {
                double hf = (double) k / (double) H;
                int zmax_h = ar_required_z(hf, zmax), wmax_h = ar_required_w(hf, wmax);
                hwf[f] = ar_w_halfwidth(-zmax_h, -wmax_h);
                ndf[f] = ar_calc_fftlen(H, k, zmax, wmax, g->corr_uselen) / ACCEL_NUMBETWEEN;
                if (ndf[f] > s->maxn)
                    s->maxn = ndf[f];
            }
        }
        for (int b = 0; b < g->nblocks; b++) {
            double startr = s->bstart[b];
            agpu_window *wb = win + (size_t) b * (s->nfrac + 1);
            wb[0].lobin = fund_rlo(startr) - fund_hw;
            wb[0].numdata = fund_nd;
            f = 0;
            for (int st = 1; st < g->numharmstages; st++) {
                int H = 1 << st;
                for (int k = 1; k < H; k += 2, f++) {
                    double hf = (double) k / (double) H;
                    long long rlo_h = (long long) floor(ar_required_r(hf, startr));
                    wb[1 + f].lobin = rlo_h - hwf[f];
                    wb[1 + f].numdata = ndf[f];
                }
            }
        }
    }
    *win_out = win;
    *zmap_out = zmap;
    *wmap_out = wmap;
    return 0;
}

/* Template banks depend only on the (w, z) grid: they are generated once
 * per grid and shared by every handle (refcounted, kept while unused until
 * accel_template_cache_clear). */
typedef struct tbank {
    int numz, numw, zlo, wlo, L, refs;
    void *d_T;
    struct tbank *next;
} tbank;
static tbank *g_banks = NULL;
static pthread_mutex_t g_bank_mu = PTHREAD_MUTEX_INITIALIZER;

static int gen_templates(accel_search *s);

static int bank_get(accel_search *s)
{
    const accel_geom *g = &s->g;
    pthread_mutex_lock(&g_bank_mu);
    for (tbank *b = g_banks; b; b = b->next)
        if (b->numz == g->numz && b->numw == g->numw && b->zlo == (int) g->zlo &&
            b->wlo == (int) g->wlo && b->L == s->L) {
            b->refs++;
            s->d_T = b->d_T;
            pthread_mutex_unlock(&g_bank_mu);
            return 0;
        }
    int rc = -1;
    if (agpu_malloc(&s->d_T, sizeof(accel_fcomplex) * (size_t) s->nt * s->L)) {
        set_err("template bank: %s", agpu_last_error());
        s->d_T = NULL;
    } else if (gen_templates(s)) {
        agpu_free(s->d_T);
        s->d_T = NULL;
    } else {
        tbank *b = (tbank *) calloc(1, sizeof(tbank));
        b->numz = g->numz;
        b->numw = g->numw;
        b->zlo = (int) g->zlo;
        b->wlo = (int) g->wlo;
        b->L = s->L;
        b->refs = 1;
        b->d_T = s->d_T;
        b->next = g_banks;
        g_banks = b;
        rc = 0;
    }
    pthread_mutex_unlock(&g_bank_mu);
    return rc;
}

static void bank_put(void *d_T)
{
    if (!d_T)
        return;
    pthread_mutex_lock(&g_bank_mu);
    for (tbank *b = g_banks; b; b = b->next)
        if (b->d_T == d_T) {
            b->refs--;
            break;
        }
    pthread_mutex_unlock(&g_bank_mu);
}

void accel_template_cache_clear(void)
{
    pthread_mutex_lock(&g_bank_mu);
    tbank **pp = &g_banks;
    while (*pp) {
        tbank *b = *pp;
        if (b->refs <= 0) {
            agpu_free(b->d_T);
            *pp = b->next;
            free(b);
        } else {
            pp = &b->next;
        }
    }
    pthread_mutex_unlock(&g_bank_mu);
}

/* Template bank of the fundamental: (w, z) row-major, z from -zmax and w
 * from -wmax (src/accel_utils.c:223-237). */
static int gen_templates(accel_search *s)
{
    const accel_geom *g = &s->g;
    int *zs = (int *) malloc(sizeof(int) * s->nt);
    int *ws = (int *) malloc(sizeof(int) * s->nt);
    int *nk = (int *) malloc(sizeof(int) * s->nt);
    for (int wi = 0; wi < s->numw1; wi++) {
        for (int zi = 0; zi < g->numz; zi++) {
            int t = wi * g->numz + zi;
            zs[t] = (int) g->zlo + zi * ACCEL_DZ;
            ws[t] = g->numw ? (int) g->wlo + wi * ACCEL_DW : 0;
            nk[t] = 2 * ACCEL_NUMBETWEEN * ar_w_halfwidth(zs[t], ws[t]);
        }
    }
    int rc = agpu_gen_templates(zs, ws, nk, s->nt, s->L, s->d_T, s->stream);
    free(zs);
    free(ws);
    free(nk);
    if (rc) {
        set_err("template generation: %s", agpu_last_error());
        return -1;
    }
    return 0;
}

/* Start r of every block (the reference's accumulated startr) and how the
 * standard path's sub-harmonic lookups relate to the fundamental's plane
 * bins.  The reference computes them from startr + ii/2
 * (src/accel_utils.c:1042-1047) while the fundamental's cell ii sits at
 * fund_rlo(startr) + ii/2: the two differ by rh_off half-bins when startr
 * lies on the half-bin grid, by a fraction otherwise (then the kernel
 * evaluates the reference's FP64 expression per cell from d_bstart). */
static int block_starts(accel_search *s)
{
    const accel_geom *g = &s->g;
    const int rstep = g->corr_uselen / 2;
    s->bstart = (double *) malloc(sizeof(double) * (size_t) (g->nblocks > 0 ? g->nblocks : 1));
    if (!s->bstart) {
        set_err("out of host memory for %d block starts", g->nblocks);
        return -1;
    }
    double startr = g->rlo;
    int ongrid = 1;
    const long long r0 = fund_rlo(g->rlo);
    s->rh_off = 0;
    for (int b = 0; b < g->nblocks; b++) {
        s->bstart[b] = startr;
        if (fund_rlo(startr) != r0 + (long long) b * rstep) {
            set_err("block %d of -rlo %.17g does not start %d bins after the previous one",
                    b, g->rlo, rstep);
            return -1;
        }
        const double d = 2.0 * startr - 2.0 * (double) fund_rlo(startr);
        if ((d != 0.0 && d != 1.0) || (b > 0 && (int) d != s->rh_off))
            ongrid = 0;
        else
            s->rh_off = (int) d;
        startr += rstep;
    }
    if (!ongrid) {
        s->rh_off = 0;
        if (!g->inmem && g->numharmstages > 1) {
            GPU(agpu_malloc((void **) &s->d_bstart, sizeof(double) * (size_t) g->nblocks));
            GPU(agpu_h2d(s->d_bstart, s->bstart, sizeof(double) * (size_t) g->nblocks,
                         s->stream));
        }
    }
    return 0;
}

/* -photon: every window's norm is 1/sqrt(nph) (src/accel_utils.c:1080-1086),
 * the same FP64 value and FP32 square that the median path produces */
static int photon_norms(accel_search *s)
{
    const int n = s->nwin > 0 ? s->nwin : 1;
    double *norm = (double *) malloc(sizeof(double) * (size_t) n);
    float *nsq = (float *) malloc(sizeof(float) * (size_t) n);
    if (!norm || !nsq) {
        free(norm);
        free(nsq);
        set_err("out of host memory for %d window norms", n);
        return -1;
    }
    const double nv = 1.0 / sqrt(s->g.nph);
    for (int i = 0; i < n; i++) {
        norm[i] = nv;
        nsq[i] = (float) (nv * nv);
    }
    int rc = agpu_h2d(s->d_norm, norm, sizeof(double) * (size_t) n, s->stream) ||
             agpu_h2d(s->d_nsq, nsq, sizeof(float) * (size_t) n, s->stream) ||
             agpu_stream_sync(s->stream);
    free(norm);
    free(nsq);
    if (rc) {
        set_err("photon norms: %s", agpu_last_error());
        return -1;
    }
    return 0;
}

static int search_init(accel_search *s, int device, int nstreams)
{
    accel_geom *g = &s->g;
    GPU(agpu_init(device));
    GPU(agpu_stream_create(&s->stream));
    if (nstreams > 1)
        GPU(agpu_stream_create(&s->stream2));
    else
        s->stream2 = s->stream;
    for (int i = 0; i < 5; i++)
        GPU(agpu_event_create(&s->ev[i]));
    GPU(agpu_event_create(&s->ev_norms));
    s->numw1 = g->numw ? g->numw : 1;
    s->nt = s->numw1 * g->numz;
    s->hw = g->maxkernlen / (2 * ACCEL_NUMBETWEEN);
    s->L = CORR_L;
    if (4 * s->hw + 64 > s->L) {
        set_err("kernel width %d too large for %d-point segments", 4 * s->hw, s->L);
        return -1;
    }
    s->uselen = (s->L - 4 * s->hw) & ~1;
    if (block_starts(s))
        return -1;
    s->hb_end = 2 * fund_rlo(g->rlo) + (long long) g->nblocks * g->corr_uselen;
    /* one tile of padding past the last half-bin: the tail chunk of the
     * harmonic sum reads it instead of clamping indices */
    s->ntiles = (s->hb_end + AGPU_TILE - 1) / AGPU_TILE + 1;
    /* template groups of about 26 (the forward FFT of a segment is shared
     * by a group), a multiple of 8 so that each XCD keeps one residue
     * class of templates in its L2 */
    s->groups = 8 * (s->nt + 103 >= 208 ? (s->nt + 103) / 208 : 1);
    s->nseg = (s->hb_end + s->uselen - 1) / s->uselen;
    agpu_window *win = NULL;
    int *zmap = NULL, *wmap = NULL;
    build_tables(s, &win, &zmap, &wmap);
    size_t plane_bytes = sizeof(float) * (size_t) s->nt * (size_t) s->ntiles * AGPU_ROWLEN;
    GPU(agpu_malloc(&s->d_spec, sizeof(accel_fcomplex) * (size_t) (g->numbins > 0 ? g->numbins : 1)));
    s->own_spec = 1;
    GPU(agpu_malloc((void **) &s->d_plane, plane_bytes > 0 ? plane_bytes : 4));
    GPU(agpu_malloc((void **) &s->d_win, sizeof(agpu_window) * (size_t) (s->nwin > 0 ? s->nwin : 1)));
    GPU(agpu_malloc((void **) &s->d_norm, sizeof(double) * (size_t) (s->nwin > 0 ? s->nwin : 1)));
    GPU(agpu_malloc((void **) &s->d_nsq, sizeof(float) * (size_t) (s->nwin > 0 ? s->nwin : 1)));
    if (g->inmem)
        GPU(agpu_malloc((void **) &s->d_nsq_hb, sizeof(float) * (size_t) s->ntiles * AGPU_TILE));
    int nf = s->nfrac > 0 ? s->nfrac : 1;
    GPU(agpu_malloc((void **) &s->d_zmap, sizeof(int) * (size_t) nf * g->numz));
    GPU(agpu_malloc((void **) &s->d_wmap, sizeof(int) * (size_t) nf * s->numw1));
    s->cap = HIT_CAP0;
    GPU(agpu_malloc((void **) &s->d_hits, sizeof(agpu_hit) * s->cap));
    GPU(agpu_malloc((void **) &s->d_nhits, sizeof(unsigned)));
    GPU(agpu_h2d(s->d_win, win, sizeof(agpu_window) * (size_t) (s->nwin > 0 ? s->nwin : 1), s->stream));
    if (g->nph > 0.0 && photon_norms(s))
        return -1;
    GPU(agpu_h2d(s->d_zmap, zmap, sizeof(int) * (size_t) nf * g->numz, s->stream));
    GPU(agpu_h2d(s->d_wmap, wmap, sizeof(int) * (size_t) nf * s->numw1, s->stream));
    free(win);
    /* harmonic-sum address tables: per (w, z) row and per chunk */
    const long long rows = (long long) s->nt;
    const long long ncell = (long long) g->nblocks * g->corr_uselen;
    const long long rh0 = 2 * fund_rlo(g->rlo);
    const long long nchunks = ncell > 0 ? (rh0 + ncell - 1) / 1024 - rh0 / 1024 + 1 : 1;
    const int mode = g->inmem ? 1 : (s->d_bstart ? 2 : 0);
    unsigned *rowtab = (unsigned *) malloc(sizeof(unsigned) * AGPU_HS_RT * (size_t) rows);
    long long *cbase = (long long *) malloc(sizeof(long long) * AGPU_HS_CT * (size_t) nchunks);
    int *clo = (int *) malloc(sizeof(int) * AGPU_HS_CT * (size_t) nchunks);
    int rc = agpu_harmsum_row_table(g->numz, s->numw1, s->nfrac, zmap, wmap, rowtab) ||
             agpu_harmsum_chunk_tables(mode, g->inmem ? 0 : s->rh_off, rh0, nchunks,
                                       g->numharmstages, (int) rows, cbase, clo) ||
             agpu_malloc((void **) &s->d_rowtab, sizeof(unsigned) * AGPU_HS_RT * (size_t) rows) ||
             agpu_malloc((void **) &s->d_cbase,
                         sizeof(long long) * AGPU_HS_CT * (size_t) nchunks) ||
             agpu_malloc((void **) &s->d_clo, sizeof(int) * AGPU_HS_CT * (size_t) nchunks) ||
             agpu_h2d(s->d_rowtab, rowtab, sizeof(unsigned) * AGPU_HS_RT * (size_t) rows,
                      s->stream) ||
             agpu_h2d(s->d_cbase, cbase, sizeof(long long) * AGPU_HS_CT * (size_t) nchunks,
                      s->stream) ||
             agpu_h2d(s->d_clo, clo, sizeof(int) * AGPU_HS_CT * (size_t) nchunks, s->stream);
    free(rowtab);
    free(cbase);
    free(clo);
    free(zmap);
    free(wmap);
    if (rc) {
        set_err("harmonic-sum tables: %s", agpu_last_error());
        return -1;
    }
    if (bank_get(s))
        return -1;
    return 0;
}

accel_search *accel_search_create(const accel_params *p, int device)
{
    return accel_search_create_streams(p, device, 2);
}

accel_search *accel_search_create_streams(const accel_params *p, int device, int nstreams)
{
    accel_search *s = (accel_search *) calloc(1, sizeof(accel_search));
    s->p = *p;
    if (accel_geom_create(p, &s->g) || search_init(s, device, nstreams)) {
        char keep[sizeof(g_err)];
        memcpy(keep, g_err, sizeof(keep));
        accel_search_destroy(s);
        memcpy(g_err, keep, sizeof(keep));
        return NULL;
    }
    return s;
}

void accel_search_destroy(accel_search *s)
{
    if (!s)
        return;
    if (s->stream)
        agpu_stream_sync(s->stream);
    if (s->stream2 && s->stream2 != s->stream)
        agpu_stream_sync(s->stream2);
    if (s->own_spec)
        agpu_free(s->d_spec);
    bank_put(s->d_T);
    agpu_free(s->d_plane);
    agpu_free(s->d_win);
    agpu_free(s->d_norm);
    agpu_free(s->d_nsq);
    agpu_free(s->d_nsq_hb);
    agpu_free(s->d_zmap);
    agpu_free(s->d_wmap);
    agpu_free(s->d_rowtab);
    agpu_free(s->d_cbase);
    agpu_free(s->d_clo);
    agpu_free(s->d_hits);
    agpu_free(s->d_nhits);
    agpu_free(s->d_bstart);
    free(s->bstart);
    if (s->workfile)
        fclose(s->workfile);
    for (int i = 0; i < 5; i++)
        if (s->ev[i])
            agpu_event_destroy(s->ev[i]);
    if (s->ev_norms)
        agpu_event_destroy(s->ev_norms);
    if (s->stream)
        agpu_stream_destroy(s->stream);
    if (s->stream2 && s->stream2 != s->stream)
        agpu_stream_destroy(s->stream2);
    free(s->hits);
    free_cands(s->cands);
    free(s->opt_nh);
    free(s->opt_r);
    free(s->opt_z);
    free(s->opt_w);
    free(s->opt_pows);
    free(s->opt_derivs);
    free(s->opt_h);
    if (s->ws) {
        ws_free(s->ws);
        free(s->ws);
    }
    agpu_free(s->d_fftwork);
    agpu_free(s->d_dblk);
    agpu_free(s->d_dmean);
    free(s);
}

int accel_search_set_spectrum(accel_search *s, const accel_fcomplex *fft)
{
    if (!s->own_spec) {
        GPU(agpu_malloc(&s->d_spec, sizeof(accel_fcomplex) * (size_t) s->g.numbins));
        s->own_spec = 1;
    }
    GPU(agpu_h2d(s->d_spec, fft, sizeof(accel_fcomplex) * (size_t) s->g.numbins, s->stream));
    return 0;
}

int accel_search_set_device_spectrum(accel_search *s, const void *d_fft)
{
    if (s->own_spec)
        agpu_free(s->d_spec);
    s->own_spec = 0;
    s->d_spec = (void *) d_fft;
    return 0;
}

static void fill_harmsum_args(const accel_search *s, agpu_harmsum_args *a)
{
    const accel_geom *g = &s->g;
    memset(a, 0, sizeof(*a));
    a->plane = s->d_plane;
    a->ntiles = s->ntiles;
    a->numz = g->numz;
    a->numw = s->numw1;
    a->rh0 = 2 * fund_rlo(g->rlo);
    a->rh_off = g->inmem ? 0 : s->rh_off;
    a->bstart = g->inmem ? NULL : s->d_bstart;
    a->corr_uselen = g->corr_uselen;
    a->nblocks = g->nblocks;
    a->nstages = g->numharmstages;
    a->inmem = g->inmem;
    for (int i = 0; i < g->numharmstages; i++)
        a->powcut[i] = g->powcut[i];
    a->nfrac = s->nfrac;
    a->zmap = s->d_zmap;
    a->wmap = s->d_wmap;
    a->row_tab = s->d_rowtab;
    a->chunk_base = s->d_cbase;
    a->chunk_lo = s->d_clo;
    a->nsq = s->d_nsq;
    a->nsq_hb = s->d_nsq_hb;
    a->npre = s->npre;
    a->hits = s->d_hits;
    a->nhits = s->d_nhits;
    a->cap = s->cap;
}

int accel_search_launch(accel_search *s)
{
    const accel_geom *g = &s->g;
    agpu_harmsum_args a;
    if (!s->d_hits) {
        set_err("search handle has no hit buffer (a regrowth failed); destroy it");
        return -1;
    }
    const double t0 = now_ms();
    GPU(agpu_memset(s->d_nhits, 0, sizeof(unsigned), s->stream));
    GPU(agpu_event_record(s->ev[0], s->stream));
    /* the block norms (latency-bound, few workgroups) run on a second
     * stream beside the correlation (VALU-bound); the harmonic sum waits
     * for both */
    const int two = s->stream2 != s->stream;
    if (two)
        GPU(agpu_stream_wait_event(s->stream2, s->ev[0]));
    if (!(g->nph > 0.0))   /* -photon: the constant norms set at creation */
        GPU(agpu_block_norms(s->d_spec, g->numbins, g->lobin, s->d_win, s->nwin, s->maxn,
                             s->d_norm, s->d_nsq, s->stream2));
    if (g->inmem)
        GPU(agpu_inmem_nsq_table(s->d_nsq, s->npre, g->nblocks, 2 * fund_rlo(g->rlo),
                                 g->corr_uselen, s->hb_end, s->d_nsq_hb, s->stream2));
    GPU(agpu_event_record(s->ev_norms, s->stream2));
    GPU(agpu_event_record(s->ev[1], s->stream));
    GPU(agpu_correlate(s->d_spec, g->numbins, g->lobin, s->d_T, s->nt, s->L, s->hw, s->uselen,
                       s->hb_end, s->d_plane, s->ntiles, s->groups, 0, -1, s->stream));
    GPU(agpu_event_record(s->ev[2], s->stream));
    /* the tile halos are written by the correlation itself */
    GPU(agpu_event_record(s->ev[3], s->stream));
    if (two)
        GPU(agpu_stream_wait_event(s->stream, s->ev_norms));
    fill_harmsum_args(s, &a);
    GPU(agpu_harmsum(&a, 0, -1, s->stream));
    GPU(agpu_event_record(s->ev[4], s->stream));
    s->hms[0] = now_ms() - t0;
    return 0;
}

static int cmp_hit(const void *pa, const void *pb)
{
    const accel_hit *a = (const accel_hit *) pa, *b = (const accel_hit *) pb;
    if (a->block != b->block) return a->block < b->block ? -1 : 1;
    if (a->stage != b->stage) return a->stage < b->stage ? -1 : 1;
    if (a->wi != b->wi) return a->wi < b->wi ? -1 : 1;
    if (a->zi != b->zi) return a->zi < b->zi ? -1 : 1;
    if (a->ii != b->ii) return a->ii < b->ii ? -1 : 1;
    return 0;
}

static void clear_opt(accel_search *s);

static int bits_for(unsigned v)
{
    int b = 0;
    while (v) {
        b++;
        v >>= 1;
    }
    return b;
}

/* Puts the hits into the reference's search order.  Every (block, stage, w,
 * z, r) appears at most once, so the order is total: the fields are packed
 * into one 64-bit key (widths from the largest value of each field) and
 * sorted by an LSD radix sort of 11-bit digits.  Falls back to qsort if a
 * field is negative or the key would not fit. */
static int sort_hits(accel_hit *h, long long n)
{
    if (n < 2)
        return 0;
    unsigned mx[5] = {0, 0, 0, 0, 0};
    for (long long i = 0; i < n; i++) {
        const int f[5] = {h[i].block, h[i].stage, h[i].wi, h[i].zi, h[i].ii};
        for (int k = 0; k < 5; k++) {
            if (f[k] < 0) {
                qsort(h, (size_t) n, sizeof(accel_hit), cmp_hit);
                return 0;
            }
            if ((unsigned) f[k] > mx[k])
                mx[k] = (unsigned) f[k];
        }
    }
    int nb[5], tot = 0;
    for (int k = 0; k < 5; k++)
        tot += nb[k] = bits_for(mx[k]);
    if (tot > 64 || n > 0xffffffffLL) {
        qsort(h, (size_t) n, sizeof(accel_hit), cmp_hit);
        return 0;
    }
    unsigned long long *key = (unsigned long long *) malloc(sizeof(unsigned long long) * 2 * n);
    unsigned *idx = (unsigned *) malloc(sizeof(unsigned) * 2 * n);
    accel_hit *tmp = (accel_hit *) malloc(sizeof(accel_hit) * n);
    if (!key || !idx || !tmp) {
        free(key);
        free(idx);
        free(tmp);
        set_err("out of host memory sorting %lld hits", n);
        return -1;
    }
    for (long long i = 0; i < n; i++) {
        const unsigned f[5] = {(unsigned) h[i].block, (unsigned) h[i].stage, (unsigned) h[i].wi,
                               (unsigned) h[i].zi, (unsigned) h[i].ii};
        unsigned long long kv = 0;
        for (int k = 0; k < 5; k++)
            kv = (nb[k] ? (kv << nb[k]) : kv) | f[k];
        key[i] = kv;
        idx[i] = (unsigned) i;
    }
    unsigned long long *ka = key, *kb = key + n;
    unsigned *ia = idx, *ib = idx + n;
    enum { DIG = 11, NBK = 1 << DIG };
    static __thread long long cnt[NBK];
    for (int sh = 0; sh < tot; sh += DIG) {
        memset(cnt, 0, sizeof(cnt));
        for (long long i = 0; i < n; i++)
            cnt[(ka[i] >> sh) & (NBK - 1)]++;
        long long run = 0;
        for (int d = 0; d < NBK; d++) {
            long long c = cnt[d];
            cnt[d] = run;
            run += c;
        }
        for (long long i = 0; i < n; i++) {
            long long o = cnt[(ka[i] >> sh) & (NBK - 1)]++;
            kb[o] = ka[i];
            ib[o] = ia[i];
        }
        unsigned long long *tk = ka;
        ka = kb;
        kb = tk;
        unsigned *ti = ia;
        ia = ib;
        ib = ti;
    }
    for (long long i = 0; i < n; i++)
        tmp[i] = h[ia[i]];
    memcpy(h, tmp, sizeof(accel_hit) * n);
    free(key);
    free(idx);
    free(tmp);
    return 0;
}

/* Sigma of every hit (src/accel_utils.c:1434-1437), spread over host
 * threads when there are many: each value is a pure function of its hit,
 * so the result does not depend on the split. */
typedef struct {
    const accel_hit *h;
    const accel_geom *g;
    float *sig;
    long long lo, hi;
} sig_job;

static void *sig_worker(void *arg)
{
    sig_job *j = (sig_job *) arg;
    for (long long i = j->lo; i < j->hi; i++) {
        const accel_hit *h = &j->h[i];
        j->sig[i] = (float) accel_candidate_sigma(h->pow, 1 << h->stage,
                                                  (double) j->g->numindep[h->stage]);
    }
    return NULL;
}

#define SIG_MAX_THREADS 16
#define SIG_PER_THREAD 128   /* ~1 us per sigma: threads pay off above a few hundred hits */

static void hit_sigmas(const accel_hit *h, long long n, const accel_geom *g, float *sig)
{
    long np = sysconf(_SC_NPROCESSORS_ONLN);
    long long nt = n / SIG_PER_THREAD;
    if (nt > np)
        nt = np;
    if (nt > SIG_MAX_THREADS)
        nt = SIG_MAX_THREADS;
    sig_job jb[SIG_MAX_THREADS];
    pthread_t th[SIG_MAX_THREADS];
    int started[SIG_MAX_THREADS] = {0};
    if (nt < 2) {
        sig_job one = {h, g, sig, 0, n};
        sig_worker(&one);
        return;
    }
    for (int t = 0; t < nt; t++) {
        jb[t] = (sig_job){h, g, sig, n * t / nt, n * (t + 1) / nt};
        if (t > 0)
            started[t] = pthread_create(&th[t], NULL, sig_worker, &jb[t]) == 0;
    }
    sig_worker(&jb[0]);
    for (int t = 1; t < nt; t++) {
        if (started[t])
            pthread_join(th[t], NULL);
        else
            sig_worker(&jb[t]);
    }
}

typedef struct {
    const float *pow;
    float *sig;
    int numharm;
    double numindep;
    long long lo, hi;
} psig_job;

static void *psig_worker(void *arg)
{
    psig_job *j = (psig_job *) arg;
    for (long long i = j->lo; i < j->hi; i++)
        j->sig[i] = (float) accel_candidate_sigma(j->pow[i], j->numharm, j->numindep);
    return NULL;
}

void accel_sigmas(const float *pow, long long n, int numharm, double numindep, float *sig)
{
    long np = sysconf(_SC_NPROCESSORS_ONLN);
    long long nt = n / SIG_PER_THREAD;
    if (nt > np)
        nt = np;
    if (nt > SIG_MAX_THREADS)
        nt = SIG_MAX_THREADS;
    psig_job jb[SIG_MAX_THREADS];
    pthread_t th[SIG_MAX_THREADS];
    int started[SIG_MAX_THREADS] = {0};
    if (nt < 2) {
        psig_job one = {pow, sig, numharm, numindep, 0, n};
        psig_worker(&one);
        return;
    }
    for (int t = 0; t < nt; t++) {
        jb[t] = (psig_job){pow, sig, numharm, numindep, n * t / nt, n * (t + 1) / nt};
        if (t > 0)
            started[t] = pthread_create(&th[t], NULL, psig_worker, &jb[t]) == 0;
    }
    psig_worker(&jb[0]);
    for (int t = 1; t < nt; t++) {
        if (started[t])
            pthread_join(th[t], NULL);
        else
            psig_worker(&jb[t]);
    }
}

/* Replays the hits in the reference's order: block, then stage, then
 * w, z, r (src/accelsearch.c:189-227, src/accel_utils.c:1419-1450). */
static int replay(accel_search *s)
{
    const accel_geom *g = &s->g;
    if (sort_hits(s->hits, s->nhits))
        return -1;
    float *sigs = (float *) malloc(sizeof(float) * (s->nhits ? s->nhits : 1));
    if (!sigs) {
        set_err("out of host memory for %lld hit sigmas", s->nhits);
        return -1;
    }
    hit_sigmas(s->hits, s->nhits, g, sigs);
    clear_opt(s);
    free_cands(s->cands);
    s->cands = NULL;
    for (long long i = 0; i < s->nhits; i++) {
        const accel_hit *h = &s->hits[i];
        int numharm = 1 << h->stage;
        float sig = sigs[i];
        double rr = (fund_rlo(s->bstart[h->block]) + h->ii * (double) ACCEL_DR) /
                    (double) numharm;
        double zz = (g->zlo + h->zi * (double) ACCEL_DZ) / (double) numharm;
        double ww = (g->wlo + h->wi * (double) ACCEL_DW) / (double) numharm;
        int added;
        s->cands = insert_cand(s->cands, h->pow, sig, numharm, rr, zz, ww, &added);
        if (added && s->workfile)   /* src/accel_utils.c:1443-1447 */
            fprintf(s->workfile, "%-7.2f  %-7.4f  %-2d  %-14.4f  %-14.9f  %-10.4f %-10.4f\n",
                    h->pow, sig, numharm, rr, rr / g->T, zz, ww);
    }
    free(sigs);
    int n = 0;
    for (cnode *c = s->cands; c; c = c->next)
        n++;
    if (n) {
        s->cands = sort_cands(s->cands);
        /* src/accelsearch.c:239-241 */
        if (s->p.numharm > 1 && !s->p.noharmremove)
            s->cands = eliminate_harmonics(s->cands, &n);
    }
    s->ncands = n;
    return 0;
}

int accel_search_collect(accel_search *s)
{
    unsigned nh = 0;
    double t0 = now_ms();
    GPU(agpu_d2h(&nh, s->d_nhits, sizeof(unsigned), s->stream));
    if (nh > s->cap) {
        /* too many crossings for the buffer: grow it and sum again.  The
         * handle never keeps a freed buffer: on a failed allocation it has
         * no hit buffer (cap 0) and refuses to launch until rebuilt. */
        GPU(agpu_free(s->d_hits));
        s->d_hits = NULL;
        s->cap = 0;
        const unsigned cap = nh + nh / 4 + 1024;
        void *d_new = NULL;
        GPU(agpu_malloc(&d_new, sizeof(agpu_hit) * cap));
        s->d_hits = (agpu_hit *) d_new;
        s->cap = cap;
        agpu_harmsum_args a;
        GPU(agpu_memset(s->d_nhits, 0, sizeof(unsigned), s->stream));
        fill_harmsum_args(s, &a);
        GPU(agpu_harmsum(&a, 0, -1, s->stream));
        GPU(agpu_d2h(&nh, s->d_nhits, sizeof(unsigned), s->stream));
        if (nh > s->cap) {
            set_err("hit buffer overflow after regrowth (%u > %u)", nh, s->cap);
            return -1;
        }
    }
    GPU(agpu_event_elapsed(&s->ms[0], s->ev[0], s->ev_norms));
    for (int i = 1; i < 4; i++)
        GPU(agpu_event_elapsed(&s->ms[i], s->ev[i], s->ev[i + 1]));
    GPU(agpu_event_elapsed(&s->ms[4], s->ev[0], s->ev[4]));
    double t1 = now_ms();
    s->hms[1] = t1 - t0;
    free(s->hits);
    s->nhits = nh;
    s->hits = (accel_hit *) malloc(sizeof(accel_hit) * (nh ? nh : 1));
    if (nh)
        GPU(agpu_d2h(s->hits, s->d_hits, sizeof(agpu_hit) * nh, s->stream));
    t0 = now_ms();
    s->hms[2] = t0 - t1;
    int rc = replay(s);
    s->hms[3] = now_ms() - t0;
    return rc;
}

static void clear_opt(accel_search *s)
{
    free(s->opt_nh);
    free(s->opt_r);
    free(s->opt_z);
    free(s->opt_w);
    free(s->opt_pows);
    free(s->opt_derivs);
    free(s->opt_h);
    s->opt_h = NULL;
    s->opt_nh = NULL;
    s->opt_r = s->opt_z = s->opt_w = s->opt_pows = NULL;
    s->opt_derivs = NULL;
    s->nopt = 0;
}

int accel_search_run(accel_search *s)
{
    if (accel_search_launch(s))
        return -1;
    return accel_search_collect(s);
}

int accel_search_ncands(const accel_search *s) { return s->ncands; }

int accel_search_cands(const accel_search *s, accel_cand *out, int n)
{
    int i = 0;
    for (cnode *c = s->cands; c && i < n; c = c->next, i++)
        out[i] = c->c;
    return i;
}

long long accel_search_nhits(const accel_search *s) { return s->nhits; }

int accel_search_hits(const accel_search *s, accel_hit *out, long long n)
{
    long long m = n < s->nhits ? n : s->nhits;
    memcpy(out, s->hits, sizeof(accel_hit) * (size_t) m);
    return (int) m;
}

int accel_search_plane_row(const accel_search *s, int wi, int zi, long long hb0, int n,
                           float *out)
{
    if (wi < 0 || wi >= s->numw1 || zi < 0 || zi >= s->g.numz || hb0 < 0 || n < 0 ||
        hb0 + n > s->ntiles * AGPU_TILE) {
        set_err("plane_row: out of range");
        return -1;
    }
    const int row = wi * s->g.numz + zi;
    long long hb = hb0;
    while (hb < hb0 + n) {
        long long len = AGPU_TILE - (hb & (AGPU_TILE - 1));
        if (hb + len > hb0 + n)
            len = hb0 + n - hb;
        GPU(agpu_d2h(out + (hb - hb0), s->d_plane + AGPU_PLANE_IDX(row, hb, s->nt),
                     sizeof(float) * (size_t) len, s->stream));
        hb += len;
    }
    return 0;
}

int accel_search_nwindows(const accel_search *s) { return s->nwin; }

int accel_search_windows(const accel_search *s, long long *lobin, int *numdata, int n)
{
    if (n > s->nwin)
        n = s->nwin;
    agpu_window *w = (agpu_window *) malloc(sizeof(agpu_window) * (n > 0 ? n : 1));
    if (agpu_d2h(w, s->d_win, sizeof(agpu_window) * (size_t) n, s->stream)) {
        free(w);
        set_err("windows: %s", agpu_last_error());
        return -1;
    }
    for (int i = 0; i < n; i++) {
        lobin[i] = w[i].lobin;
        numdata[i] = w[i].numdata;
    }
    free(w);
    return n;
}

int accel_search_norms(const accel_search *s, double *out, int n)
{
    if (n > s->nwin)
        n = s->nwin;
    GPU(agpu_d2h(out, s->d_norm, sizeof(double) * (size_t) n, s->stream));
    return n;
}

int accel_search_timings(const accel_search *s, float *ms5)
{
    memcpy(ms5, s->ms, sizeof(s->ms));
    return 0;
}

int accel_search_host_timings(const accel_search *s, double *ms4)
{
    memcpy(ms4, s->hms, sizeof(s->hms));
    return 0;
}

int accel_search_corr_shape(const accel_search *s, int *L, int *uselen, int *groups,
                            long long *nseg)
{
    *L = s->L;
    *uselen = s->uselen;
    *groups = s->groups;
    *nseg = s->nseg;
    return 0;
}

int accel_search_tune(accel_search *s, int ncand, int *groups, float *ms, int *best)
{
    if (!s || ncand <= 0 || !groups || !ms || !best) {
        set_err("accel_search_tune: bad arguments");
        return -1;
    }
    /* group counts (multiples of 8) whose group sizes lie in [4, 64],
     * ordered by distance from the default */
    int cand[64], nc = 0;
    for (int G = 8; G <= 8 * ((s->nt + 31) / 32) && nc < 64; G += 8) {
        const int tpg = (s->nt + G - 1) / G;
        if (tpg > 64 || tpg < 4 || (long long) G * s->nseg >= (1LL << 24))
            continue;
        if (nc && (s->nt + cand[nc - 1] - 1) / cand[nc - 1] == tpg)
            continue;   /* same group size as a smaller count */
        cand[nc++] = G;
    }
    if (nc == 0)
        cand[nc++] = s->groups;
    for (int i = 1; i < nc; i++)   /* insertion sort by |G - default| */
        for (int k = i; k > 0 && abs(cand[k] - s->groups) < abs(cand[k - 1] - s->groups); k--) {
            const int t = cand[k];
            cand[k] = cand[k - 1];
            cand[k - 1] = t;
        }
    if (nc > ncand)
        nc = ncand;
    /* the whole spectrum (a sample of segments favours small groups: its
     * last round of workgroups weighs more), one warm-up launch, then
     * three rounds over the candidates, best of each */
    const accel_geom *g = &s->g;
    for (int i = 0; i < nc; i++)
        ms[i] = 0.0f;
    for (int rep = -1; rep < 3; rep++) {
        for (int i = 0; i < nc; i++) {
            float t = 0.0f;
            GPU(agpu_event_record(s->ev[1], s->stream));
            GPU(agpu_correlate(s->d_spec, g->numbins, g->lobin, s->d_T, s->nt, s->L, s->hw,
                               s->uselen, s->hb_end, s->d_plane, s->ntiles, cand[i], 0, -1,
                               s->stream));
            GPU(agpu_event_record(s->ev[2], s->stream));
            GPU(agpu_stream_sync(s->stream));
            GPU(agpu_event_elapsed(&t, s->ev[1], s->ev[2]));
            if (rep < 0)
                break;   /* warm-up */
            if (rep == 0 || t < ms[i])
                ms[i] = t;
        }
    }
    int kept = cand[0];
    float kept_ms = ms[0];
    for (int i = 0; i < nc; i++) {
        groups[i] = cand[i];
        if (ms[i] < kept_ms) {
            kept = cand[i];
            kept_ms = ms[i];
        }
    }
    s->groups = kept;
    *best = kept;
    return nc;
}

long long accel_search_plane_cells(const accel_search *s)
{
    return (long long) s->nt * s->hb_end;
}

/* ---- refinement: optimize_accelcand, src/accel_utils.c:552-626 ---- */

/* Largest chirp length max_rzw_arr_harmonics can need for a start
 * (zin, win) with numharm harmonics: every template of harmonic n has
 * |z| <= |zin| n + 4 and |w| <= |win| n + 40, so its half width is at most
 * 1.1 (13/12 |w| + 3/2 |z|) + 16 (w_resp_halfwidth, src/responses.c:48-84)
 * and its chirp next2_to_n(6 hw + 200 + 2 hw_vol) points (src/responses.c:381). */
static int rzw_npts_bound(int numharm, double zin, double win)
{
    double zb = fabs(zin) * numharm + 4.0, wb = fabs(win) * numharm + 40.0;
    long long hw = (long long) (1.1 * (13.0 / 12.0 * wb + 1.5 * zb)) + 17;
    return (int) ar_next2(8 * hw + 200);
}

/* Device buffers of the refinement, kept by their owner between calls:
 * freeing device memory waits for the whole device, which would stop a
 * refinement from overlapping the next search. */
#define REFINE_CHUNK 256
#define REFINE_RZW_BUDGET ((size_t) 3 << 30)
struct refine_ws {
    int ready;
    size_t scr_bytes;
    int *d_nh, *d_nev;
    double *d_rin, *d_zin, *d_win, *d_rout, *d_zout, *d_wout, *d_der, *d_pow;
    void *d_scr;
};

static void ws_free(refine_ws *w)
{
    if (!w->ready)
        return;
    agpu_free(w->d_nh);
    agpu_free(w->d_nev);
    agpu_free(w->d_rin);
    agpu_free(w->d_zin);
    agpu_free(w->d_win);
    agpu_free(w->d_rout);
    agpu_free(w->d_zout);
    agpu_free(w->d_wout);
    agpu_free(w->d_der);
    agpu_free(w->d_pow);
    if (w->d_scr)
        agpu_free(w->d_scr);
    memset(w, 0, sizeof(*w));
}

static int ws_reserve(refine_ws *w, size_t scr_bytes)
{
    const size_t nd = sizeof(double) * REFINE_CHUNK;
    const size_t nhx = (size_t) REFINE_CHUNK * ACCEL_MAXHARM;
    if (!w->ready) {
        if (agpu_malloc((void **) &w->d_nh, sizeof(int) * REFINE_CHUNK) ||
            agpu_malloc((void **) &w->d_nev, sizeof(int) * REFINE_CHUNK) ||
            agpu_malloc((void **) &w->d_rin, nd) || agpu_malloc((void **) &w->d_zin, nd) ||
            agpu_malloc((void **) &w->d_win, nd) || agpu_malloc((void **) &w->d_rout, nd) ||
            agpu_malloc((void **) &w->d_zout, nd) || agpu_malloc((void **) &w->d_wout, nd) ||
            agpu_malloc((void **) &w->d_der, sizeof(double) * 7 * nhx) ||
            agpu_malloc((void **) &w->d_pow, sizeof(double) * nhx)) {
            set_err("refinement: %s", agpu_last_error());
            w->ready = 1;
            ws_free(w);
            return -1;
        }
        w->ready = 1;
    }
    if (scr_bytes > w->scr_bytes) {
        if (w->d_scr)
            agpu_free(w->d_scr);
        w->d_scr = NULL;
        w->scr_bytes = 0;
        if (agpu_malloc(&w->d_scr, scr_bytes)) {
            set_err("refinement scratch: %s", agpu_last_error());
            return -1;
        }
        w->scr_bytes = scr_bytes;
    }
    return 0;
}

/* Runs max_rz_arr_harmonics (win == NULL) or max_rzw_arr_harmonics for n
 * starts on the device spectrum d_spec, in chunks of at most REFINE_CHUNK
 * candidates (one workgroup each), with the buffers of w. */
static int refine_device(refine_ws *w, const void *d_spec, long long numdata, int raw, int n,
                         const int *nh, const double *rin, const double *zin, const double *win,
                         double *rout, double *zout, double *wout, double *pows,
                         accel_rderivs *derivs, void *stream)
{
    int cap = 0;
    for (int i = 0; i < n; i++) {
        if (nh[i] < 1 || nh[i] > ACCEL_MAXHARM || (raw && nh[i] != 1)) {
            set_err("refinement: numharm %d out of range 1..%d", nh[i], raw ? 1 : ACCEL_MAXHARM);
            return -1;
        }
        if (win) {
            int b = rzw_npts_bound(nh[i], zin[i], win[i]);
            if (b > cap)
                cap = b;
        }
    }
    if (win && cap > (1 << 20)) {
        set_err("refinement: jerk chirp of %d points is too long", cap);
        return -1;
    }
    int chunk = n < REFINE_CHUNK ? n : REFINE_CHUNK;
    if (win) {
        /* the jerk refinement's scratch grows with the chirp bound: keep a
         * chunk within REFINE_RZW_BUDGET bytes */
        const size_t per = agpu_refine_rzw_scratch(1, cap);
        const int fit = (int) (REFINE_RZW_BUDGET / (per ? per : 1));
        if (chunk > fit)
            chunk = fit > 0 ? fit : 1;
    }
    if (ws_reserve(w, win ? agpu_refine_rzw_scratch(chunk, cap) : agpu_refine_rz_scratch(chunk)))
        return -1;
    int nev[REFINE_CHUNK];
    for (int c0 = 0; c0 < n; c0 += chunk) {
        const int m = (n - c0) < chunk ? n - c0 : chunk;
        const size_t md = sizeof(double) * (size_t) m, mhx = (size_t) m * ACCEL_MAXHARM;
        int maxh = 1;
        for (int i = 0; i < m; i++)
            maxh = nh[c0 + i] > maxh ? nh[c0 + i] : maxh;
        if (agpu_h2d(w->d_nh, nh + c0, sizeof(int) * m, stream) ||
            agpu_h2d(w->d_rin, rin + c0, md, stream) ||
            agpu_h2d(w->d_zin, zin + c0, md, stream) ||
            (win && agpu_h2d(w->d_win, win + c0, md, stream))) {
            set_err("refinement: %s", agpu_last_error());
            return -1;
        }
        agpu_refine_args a;
        memset(&a, 0, sizeof(a));
        a.spec = d_spec;
        a.numdata = numdata;
        a.ncand = m;
        a.numharm = w->d_nh;
        a.rin = w->d_rin;
        a.zin = w->d_zin;
        a.win = w->d_win;
        a.rout = w->d_rout;
        a.zout = w->d_zout;
        a.wout = w->d_wout;
        a.derivs = w->d_der;
        a.powers = w->d_pow;
        a.nevals = w->d_nev;
        a.scratch = w->d_scr;
        a.npts_cap = cap;
        a.raw = raw;
        a.maxharm = maxh;
        if (agpu_memset(w->d_der, 0, sizeof(double) * 7 * mhx, stream) ||
            agpu_memset(w->d_pow, 0, sizeof(double) * mhx, stream) ||
            (win ? agpu_refine_rzw(&a, stream) : agpu_refine_rz(&a, stream)) ||
            agpu_d2h(rout + c0, w->d_rout, md, stream) ||
            agpu_d2h(zout + c0, w->d_zout, md, stream) ||
            (wout && agpu_d2h(wout + c0, w->d_wout, md, stream)) ||
            agpu_d2h(nev, w->d_nev, sizeof(int) * m, stream) ||
            agpu_d2h(pows + (size_t) c0 * ACCEL_MAXHARM, w->d_pow, sizeof(double) * mhx,
                     stream) ||
            agpu_d2h(derivs + (size_t) c0 * ACCEL_MAXHARM, w->d_der, sizeof(double) * 7 * mhx,
                     stream)) {
            set_err("refinement: %s", agpu_last_error());
            return -1;
        }
        for (int i = 0; i < m; i++)
            if (nev[i] < 0) {
                set_err(win ? "refinement: candidate %d exceeded the chirp bound"
                            : "refinement: candidate %d: its workgroups' barrier wait ran out",
                        c0 + i);
                return -1;
            }
    }
    return 0;
}

/* -noharmpolish (src/accel_utils.c:582-622): harmonic k of each candidate
 * refined on its own from (k r - lobin, k z, k w), max_rz_arr /
 * max_rzw_arr, all harmonics of all candidates in one launch */
static int refine_unpolished(accel_search *s, int n, const int *nh, const double *rin,
                             const double *zin, const double *win)
{
    const accel_geom *g = &s->g;
    int nv = 0;
    for (int i = 0; i < n; i++)
        nv += nh[i];
    int *vnh = (int *) malloc(sizeof(int) * nv);
    double *v = (double *) malloc(sizeof(double) * 6 * (size_t) nv);
    double *vpow = (double *) calloc((size_t) nv * ACCEL_MAXHARM, sizeof(double));
    accel_rderivs *vder = (accel_rderivs *) calloc((size_t) nv * ACCEL_MAXHARM,
                                                   sizeof(accel_rderivs));
    if (!vnh || !v || !vpow || !vder) {
        free(vnh);
        free(v);
        free(vpow);
        free(vder);
        set_err("refinement: out of host memory");
        return -1;
    }
    double *vr = v, *vz = v + nv, *vw = v + 2 * nv, *vro = v + 3 * nv, *vzo = v + 4 * nv,
           *vwo = v + 5 * nv;
    int j = 0;
    for (int i = 0; i < n; i++) {
        const double r = rin[i] + (double) g->lobin;
        for (int k = 0; k < nh[i]; k++, j++) {
            vnh[j] = 1;
            vr[j] = r * (k + 1) - (double) g->lobin;
            vz[j] = zin[i] * (k + 1);
            vw[j] = win[i] * (k + 1);
        }
    }
    int rc = refine_device(s->ws, s->d_spec, g->numbins, 1, nv, vnh, vr, vz,
                           g->numw ? vw : NULL, vro, vzo, vwo, vpow, vder, s->stream);
    if (rc == 0) {
        j = 0;
        for (int i = 0; i < n; i++) {
            double *h = s->opt_h + (size_t) i * 3 * ACCEL_MAXHARM;
            for (int k = 0; k < nh[i]; k++, j++) {
                h[k] = vro[j] + (double) g->lobin;
                h[ACCEL_MAXHARM + k] = vzo[j];
                h[2 * ACCEL_MAXHARM + k] = g->numw ? vwo[j] : 0.0;
                s->opt_pows[(size_t) i * ACCEL_MAXHARM + k] = vpow[(size_t) j * ACCEL_MAXHARM];
                s->opt_derivs[(size_t) i * ACCEL_MAXHARM + k] = vder[(size_t) j * ACCEL_MAXHARM];
            }
            s->opt_r[i] = h[0];
            s->opt_z[i] = h[ACCEL_MAXHARM];
            s->opt_w[i] = h[2 * ACCEL_MAXHARM];
        }
    }
    free(vnh);
    free(v);
    free(vpow);
    free(vder);
    return rc;
}

void accel_refine_set_first_wait(int log2_polls) { agpu_refine_set_first_wait(log2_polls); }

int accel_search_optimize(accel_search *s)
{
    const accel_geom *g = &s->g;
    clear_opt(s);
    const int n = s->ncands;
    if (n == 0)
        return 0;
    int *nh = (int *) malloc(sizeof(int) * n);
    double *rin = (double *) malloc(sizeof(double) * n);
    double *zin = (double *) malloc(sizeof(double) * n);
    double *win = (double *) malloc(sizeof(double) * n);
    s->opt_nh = (int *) malloc(sizeof(int) * n);
    s->opt_r = (double *) malloc(sizeof(double) * n);
    s->opt_z = (double *) malloc(sizeof(double) * n);
    s->opt_w = (double *) calloc(n, sizeof(double));
    s->opt_pows = (double *) calloc((size_t) n * ACCEL_MAXHARM, sizeof(double));
    s->opt_derivs = (accel_rderivs *) calloc((size_t) n * ACCEL_MAXHARM, sizeof(accel_rderivs));
    int i = 0;
    for (cnode *c = s->cands; c; c = c->next, i++) {
        nh[i] = c->c.numharm;
        rin[i] = c->c.r - (double) g->lobin;
        zin[i] = c->c.z;
        win[i] = c->c.w;
    }
    /* src/accel_utils.c:563-575: max_rzw_arr_harmonics when numw, else
     * max_rz_arr_harmonics */
    if (!s->ws)
        s->ws = (refine_ws *) calloc(1, sizeof(refine_ws));
    s->opt_h = (double *) calloc((size_t) n * 3 * ACCEL_MAXHARM, sizeof(double));
    int rc = s->p.noharmpolish ? refine_unpolished(s, n, nh, rin, zin, win)
                               : refine_device(s->ws, s->d_spec, g->numbins, 0, n, nh, rin, zin,
                                               g->numw ? win : NULL, s->opt_r, s->opt_z,
                                               s->opt_w, s->opt_pows, s->opt_derivs, s->stream);
    if (rc == 0) {
        for (i = 0; i < n; i++) {
            s->opt_nh[i] = nh[i];
            if (s->p.noharmpolish)
                continue;
            /* src/accel_utils.c:577-581 */
            s->opt_r[i] += (double) g->lobin;
            double *h = s->opt_h + (size_t) i * 3 * ACCEL_MAXHARM;
            for (int k = 0; k < nh[i]; k++) {
                h[k] = s->opt_r[i] * (k + 1);
                h[ACCEL_MAXHARM + k] = s->opt_z[i] * (k + 1);
                h[2 * ACCEL_MAXHARM + k] = g->numw ? s->opt_w[i] * (k + 1) : 0.0;
            }
        }
        s->nopt = n;
    } else {
        clear_opt(s);
    }
    free(nh);
    free(rin);
    free(zin);
    free(win);
    return rc;
}

int accel_search_refine_dev(accel_search *s, int raw, int n, const int *nh, const double *rin,
                            const double *zin, const double *win, double *rout, double *zout,
                            double *wout, double *pows, accel_rderivs *derivs)
{
    if (!s->ws)
        s->ws = (refine_ws *) calloc(1, sizeof(refine_ws));
    if (!s->ws) {
        set_err("refinement: out of host memory");
        return -1;
    }
    return refine_device(s->ws, s->d_spec, s->g.numbins, raw, n, nh, rin, zin, win, rout, zout,
                         wout, pows, derivs, s->stream);
}

void accel_search_dev_view(const accel_search *s, accel_dev_view *v)
{
    v->d_spec = s->d_spec;
    v->numbins = s->g.numbins;
    v->lobin = s->g.lobin;
    v->d_plane = s->d_plane;
    v->ntiles = s->ntiles;
    v->hb_end = s->hb_end;
    v->nrows = s->nt;
    v->numz = s->g.numz;
    v->numw1 = s->numw1;
    v->stream = s->stream;
}

int accel_search_set_txtcand(accel_search *s, const char *path)
{
    if (s->workfile && fclose(s->workfile)) {
        s->workfile = NULL;
        set_err("error closing the .txtcand file");
        return -1;
    }
    s->workfile = NULL;
    if (!path)
        return 0;
    s->workfile = fopen(path, "w");
    if (!s->workfile) {
        set_err("cannot open '%s' for writing", path);
        return -1;
    }
    return 0;
}

int accel_search_plane_only(accel_search *s)
{
    const accel_geom *g = &s->g;
    GPU(agpu_correlate(s->d_spec, g->numbins, g->lobin, s->d_T, s->nt, s->L, s->hw, s->uselen,
                       s->hb_end, s->d_plane, s->ntiles, s->groups, 0, -1, s->stream));
    GPU(agpu_stream_sync(s->stream));
    return 0;
}

int accel_search_cand_harmonics(const accel_search *s, int i, double *hirs, double *hizs,
                                double *hiws, double *pows, accel_rderivs *derivs, int nmax)
{
    if (i < 0 || i >= s->nopt) {
        set_err("accel_search_cand_harmonics: candidate %d not refined", i);
        return -1;
    }
    int nh = s->opt_nh[i];
    int m = nh < nmax ? nh : nmax;
    for (int k = 0; k < m; k++) {
        const double *h = s->opt_h + (size_t) i * 3 * ACCEL_MAXHARM;
        if (hirs)
            hirs[k] = h[k];
        if (hizs)
            hizs[k] = h[ACCEL_MAXHARM + k];
        if (hiws)
            hiws[k] = h[2 * ACCEL_MAXHARM + k];
        if (pows)
            pows[k] = s->opt_pows[(size_t) i * ACCEL_MAXHARM + k];
        if (derivs)
            derivs[k] = s->opt_derivs[(size_t) i * ACCEL_MAXHARM + k];
    }
    return nh;
}

int accel_search_fund_props(const accel_search *s, accel_fourierprops *props, int n)
{
    if (s->nopt != s->ncands) {
        set_err("accel_search_fund_props: run accel_search_optimize first");
        return -1;
    }
    int i = 0;
    for (cnode *c = s->cands; c && i < n; c = c->next, i++) {
        /* src/accelsearch.c:262-275: props at the search's r, z, w with the
         * fundamental's derivatives, errors from the search grid */
        if (accel_calc_props(&s->opt_derivs[(size_t) i * ACCEL_MAXHARM], c->c.r, c->c.z,
                             c->c.w, &props[i]))
            return -1;
        props[i].rerr = (float) (ACCEL_DR) / c->c.numharm;
        props[i].zerr = (float) (ACCEL_DZ) / c->c.numharm;
        props[i].werr = (float) (ACCEL_DW) / c->c.numharm;
    }
    return i;
}

int accel_max_rz_harmonics(const accel_fcomplex *spec, long long numdata, int n,
                           const int *numharm, const double *rin, const double *zin,
                           double *rout, double *zout, double *pows, accel_rderivs *derivs,
                           int device)
{
    void *stream = NULL, *d_spec = NULL;
    int rc = -1;
    if (n <= 0 || numdata <= 0) {
        set_err("accel_max_rz_harmonics: empty input");
        return -1;
    }
    if (agpu_init(device) || agpu_stream_create(&stream) ||
        agpu_malloc(&d_spec, sizeof(accel_fcomplex) * (size_t) numdata) ||
        agpu_h2d(d_spec, spec, sizeof(accel_fcomplex) * (size_t) numdata, stream)) {
        set_err("accel_max_rz_harmonics: %s", agpu_last_error());
        goto done;
    }
    refine_ws w;
    memset(&w, 0, sizeof(w));
    rc = refine_device(&w, d_spec, numdata, 0, n, numharm, rin, zin, NULL, rout, zout, NULL, pows,
                       derivs, stream);
    ws_free(&w);
done:
    if (d_spec)
        agpu_free(d_spec);
    if (stream)
        agpu_stream_destroy(stream);
    return rc;
}

int accel_max_rzw_harmonics(const accel_fcomplex *spec, long long numdata, int n,
                            const int *numharm, const double *rin, const double *zin,
                            const double *win, double *rout, double *zout, double *wout,
                            double *pows, accel_rderivs *derivs, int device)
{
    void *stream = NULL, *d_spec = NULL;
    int rc = -1;
    if (n <= 0 || numdata <= 0) {
        set_err("accel_max_rzw_harmonics: empty input");
        return -1;
    }
    if (agpu_init(device) || agpu_stream_create(&stream) ||
        agpu_malloc(&d_spec, sizeof(accel_fcomplex) * (size_t) numdata) ||
        agpu_h2d(d_spec, spec, sizeof(accel_fcomplex) * (size_t) numdata, stream)) {
        set_err("accel_max_rzw_harmonics: %s", agpu_last_error());
        goto done;
    }
    refine_ws w;
    memset(&w, 0, sizeof(w));
    rc = refine_device(&w, d_spec, numdata, 0, n, numharm, rin, zin, win, rout, zout, wout, pows,
                       derivs, stream);
    ws_free(&w);
done:
    if (d_spec)
        agpu_free(d_spec);
    if (stream)
        agpu_stream_destroy(stream);
    return rc;
}

int accel_realfft_forward(const float *series, long long n, accel_fcomplex *out, int device)
{
    void *stream = NULL, *d_x = NULL, *d_out = NULL, *d_work = NULL;
    int rc = -1;
    if (n < 4 || (n & 1)) {
        set_err("accel_realfft_forward: length %lld is not even and >= 4", n);
        return -1;
    }
    if (agpu_init(device) || agpu_stream_create(&stream) ||
        agpu_malloc(&d_x, sizeof(float) * (size_t) n) ||
        agpu_malloc(&d_out, sizeof(accel_fcomplex) * (size_t) (n / 2)) ||
        agpu_malloc(&d_work, sizeof(accel_fcomplex) * (size_t) agpu_realfft_work(n)) ||
        agpu_h2d(d_x, series, sizeof(float) * (size_t) n, stream) ||
        agpu_realfft((const float *) d_x, n, d_out, d_work, stream) ||
        agpu_d2h(out, d_out, sizeof(accel_fcomplex) * (size_t) (n / 2), stream)) {
        set_err("accel_realfft_forward: %s", agpu_last_error());
        goto done;
    }
    rc = 0;
done:
    if (d_x)
        agpu_free(d_x);
    if (d_out)
        agpu_free(d_out);
    if (d_work)
        agpu_free(d_work);
    if (stream)
        agpu_stream_destroy(stream);
    return rc;
}

int accel_realfft_inverse(const accel_fcomplex *spec, long long n, float *out, int device)
{
    void *stream = NULL, *d_x = NULL, *d_work = NULL;
    int rc = -1;
    if (n < 4 || (n & 1)) {
        set_err("accel_realfft_inverse: length %lld is not even and >= 4", n);
        return -1;
    }
    if (agpu_init(device) || agpu_stream_create(&stream) ||
        agpu_malloc(&d_x, sizeof(float) * (size_t) n) ||
        agpu_malloc(&d_work, sizeof(accel_fcomplex) * (size_t) agpu_realfft_work(n)) ||
        agpu_h2d(d_x, spec, sizeof(float) * (size_t) n, stream) ||
        agpu_realfft_inverse(d_x, n, (float *) d_x, d_work, stream) ||
        agpu_d2h(out, d_x, sizeof(float) * (size_t) n, stream)) {
        set_err("accel_realfft_inverse: %s", agpu_last_error());
        goto done;
    }
    rc = 0;
done:
    if (d_x)
        agpu_free(d_x);
    if (d_work)
        agpu_free(d_work);
    if (stream)
        agpu_stream_destroy(stream);
    return rc;
}

long accel_deredden_gpu(accel_fcomplex *fft, long long numbins, int startwidth, int endwidth,
                        double endfreq, double T, int device)
{
    void *stream = NULL, *d_spec = NULL, *d_blk = NULL, *d_mean = NULL;
    long rc = -1;
    const int nb = accel_dered_blocks((long) numbins, startwidth, endwidth, endfreq, T, NULL,
                                      NULL, 0);
    if (nb <= 0) {
        set_err("accel_deredden_gpu: spectrum of %lld bins too short for a %d-bin window",
                numbins, startwidth);
        return -1;
    }
    long long *st = (long long *) malloc(sizeof(long long) * nb);
    int *ln = (int *) malloc(sizeof(int) * nb);
    agpu_window *w = (agpu_window *) calloc(nb, sizeof(agpu_window));
    if (!st || !ln || !w) {
        set_err("accel_deredden_gpu: out of host memory");
        goto done;
    }
    accel_dered_blocks((long) numbins, startwidth, endwidth, endfreq, T, st, ln, nb);
    int maxn = 0;
    for (int i = 0; i < nb; i++) {
        w[i].lobin = st[i];
        w[i].numdata = ln[i];
        if (ln[i] > maxn)
            maxn = ln[i];
    }
    if (agpu_init(device) || agpu_stream_create(&stream) ||
        agpu_malloc(&d_spec, sizeof(accel_fcomplex) * (size_t) numbins) ||
        agpu_malloc(&d_blk, sizeof(agpu_window) * nb) ||
        agpu_malloc(&d_mean, sizeof(float) * nb) ||
        agpu_h2d(d_blk, w, sizeof(agpu_window) * nb, stream) ||
        agpu_h2d(d_spec, fft, sizeof(accel_fcomplex) * (size_t) numbins, stream) ||
        agpu_deredden(d_spec, numbins, (const agpu_window *) d_blk, nb, maxn, (float *) d_mean,
                      stream) ||
        agpu_d2h(fft, d_spec, sizeof(accel_fcomplex) * (size_t) numbins, stream)) {
        set_err("accel_deredden_gpu: %s", agpu_last_error());
        goto done;
    }
    rc = (long) numbins;
done:
    free(st);
    free(ln);
    free(w);
    if (d_spec)
        agpu_free(d_spec);
    if (d_blk)
        agpu_free(d_blk);
    if (d_mean)
        agpu_free(d_mean);
    if (stream)
        agpu_stream_destroy(stream);
    return rc;
}

static int set_series(accel_search *s, const float *series, long long n, int on_device)
{
    const accel_geom *g = &s->g;
    if (n < 4 || (n & 1) || n / 2 != g->numbins || g->lobin != 0) {
        set_err("accel_search_set_series: %lld samples do not give the %lld-bin spectrum "
                "of this search (even length, lobin 0)", n, g->numbins);
        return -1;
    }
    if (!s->own_spec) {
        GPU(agpu_malloc(&s->d_spec, sizeof(accel_fcomplex) * (size_t) g->numbins));
        s->own_spec = 1;
    }
    if (!s->d_fftwork)
        GPU(agpu_malloc(&s->d_fftwork, sizeof(accel_fcomplex) * (size_t) agpu_realfft_work(n)));
    if (!s->d_dblk) {
        /* src/accel_utils.c:1561: deredden(fft, numbins, 6, 200, 0.5/dt, dt N) */
        const double T = g->dt * (double) n;
        int nb = accel_dered_blocks((long) g->numbins, 6, 200, 0.5 / g->dt, T, NULL, NULL, 0);
        if (nb <= 0) {
            set_err("accel_search_set_series: spectrum too short to de-redden");
            return -1;
        }
        long long *st = (long long *) malloc(sizeof(long long) * nb);
        int *ln = (int *) malloc(sizeof(int) * nb);
        accel_dered_blocks((long) g->numbins, 6, 200, 0.5 / g->dt, T, st, ln, nb);
        agpu_window *w = (agpu_window *) calloc(nb, sizeof(agpu_window));
        int maxn = 0;
        for (int i = 0; i < nb; i++) {
            w[i].lobin = st[i];
            w[i].numdata = ln[i];
            if (ln[i] > maxn)
                maxn = ln[i];
        }
        int rc = agpu_malloc((void **) &s->d_dblk, sizeof(agpu_window) * nb) ||
                 agpu_malloc((void **) &s->d_dmean, sizeof(float) * nb) ||
                 agpu_h2d(s->d_dblk, w, sizeof(agpu_window) * nb, s->stream);
        free(st);
        free(ln);
        free(w);
        if (rc) {
            set_err("accel_search_set_series: %s", agpu_last_error());
            return -1;
        }
        s->ndblk = nb;
        s->dmaxn = maxn;
    }
    /* realfft (src/accel_utils.c:1547), deredden (:1561), bin 0 := 1+1i
     * (:1621-1624; left as de-reddened under -photon), all on the device */
    static const accel_fcomplex one = {1.0f, 1.0f};
    if (on_device)
        GPU(agpu_copy_d2d(s->d_spec, series, sizeof(float) * (size_t) n, s->stream));
    else
        GPU(agpu_h2d(s->d_spec, series, sizeof(float) * (size_t) n, s->stream));
    GPU(agpu_realfft((const float *) s->d_spec, n, s->d_spec, s->d_fftwork, s->stream));
    GPU(agpu_deredden(s->d_spec, g->numbins, s->d_dblk, s->ndblk, s->dmaxn, s->d_dmean,
                      s->stream));
    if (!(s->p.nph > 0.0))
        GPU(agpu_h2d(s->d_spec, &one, sizeof(one), s->stream));
    return 0;
}

int accel_search_set_series(accel_search *s, const float *series, long long n)
{
    return set_series(s, series, n, 0);
}

int accel_search_set_device_series(accel_search *s, const float *d_series, long long n)
{
    return set_series(s, d_series, n, 1);
}

int accel_search_spectrum(const accel_search *s, accel_fcomplex *out, long long n)
{
    long long m = n < s->g.numbins ? n : s->g.numbins;
    GPU(agpu_d2h(out, s->d_spec, sizeof(accel_fcomplex) * (size_t) m, s->stream));
    return 0;
}

long long accel_search_footprint(const accel_params *p)
{
    accel_geom g;
    if (accel_geom_create(p, &g))
        return -1;
    const long long numw1 = g.numw ? g.numw : 1, nt = numw1 * g.numz;
    const long long hb_end = 2 * fund_rlo(g.rlo) + (long long) g.nblocks * g.corr_uselen;
    const long long ntiles = (hb_end + AGPU_TILE - 1) / AGPU_TILE + 1;
    const int nfrac = (1 << (g.numharmstages - 1)) - 1;
    long long bytes = 4LL * nt * ntiles * AGPU_ROWLEN;        /* raw plane */
    bytes += 8LL * g.numbins;                                 /* spectrum */
    bytes += 8LL * nt * CORR_L;                               /* templates */
    bytes += 24LL * (long long) (g.nblocks + 64) * (nfrac + 1); /* windows, norms */
    if (g.inmem)
        bytes += 4LL * ntiles * AGPU_TILE;                    /* per-half-bin norm^2 */
    bytes += 24LL * HIT_CAP0;                                 /* hit buffer */
    bytes += 8LL * agpu_realfft_work(2 * g.numbins);          /* FFT work of series input */
    bytes += 64LL << 20;                                      /* refinement, tables, slack */
    if (g.numw)   /* jerk refinement scratch for a full chunk of candidates */
        bytes += 16LL * 26 * REFINE_CHUNK * rzw_npts_bound(1, g.zhi, g.whi);
    return bytes;
}
