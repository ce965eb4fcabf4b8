/*
 * accel_dered.c -- spectrum preparation for time-series (.dat) input:
 * de-reddening by running block medians (src/misc_utils.c:912-1067, with
 * invsqrtf from :291-302) and the bin-0 convention of create_accelobs
 * (src/accel_utils.c:1621-1624).  Built with -ffp-contract=off so the FP32
 * expressions round as the reference's do.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "presto_accel.h"

/* k-th smallest of a[0..n) (partial sort in place) */
static float kth(float *a, int n, int k)
{
    int lo = 0, hi = n - 1;
    while (hi > lo) {
        float piv = a[lo + (hi - lo) / 2];
        int i = lo, j = hi;
        do {
            while (a[i] < piv) i++;
            while (piv < a[j]) j--;
            if (i <= j) {
                float t = a[i];
                a[i] = a[j];
                a[j] = t;
                i++;
                j--;
            }
        } while (i <= j);
        if (k <= j) hi = j;
        else if (k >= i) lo = i;
        else break;
    }
    return a[k];
}

static float median_f(float *a, int n)
{
    int lhs = (n - 1) / 2, rhs = n / 2;
    if (lhs == rhs)
        return kth(a, n, lhs);
    double x = kth(a, n, lhs), y = kth(a, n, rhs);
    return (float) (0.5 * (x + y));
}

static float fast_invsqrt(float x)
{
    union { float f; int i; } u;
    u.f = x;
    u.i = 0x5f3759df - (u.i >> 1);
    float y = u.f;
    return y * (1.5f - 0.5f * x * y * y);
}

static void block_powers(const accel_fcomplex *b, int n, float *p)
{
    for (int i = 0; i < n; i++) {
        float re = b[i].r, im = b[i].i;
        p[i] = re * re + im * im;
    }
}

long accel_deredden(accel_fcomplex *fft, long numbins, int startwidth, int endwidth,
                    double endfreq, double T)
{
    if (numbins < 1 + startwidth || startwidth < 1 || endwidth < startwidth)
        return -1;
    const float Tf = (float) T;
    accel_fcomplex *bufa = (accel_fcomplex *) malloc(sizeof(accel_fcomplex) * endwidth);
    accel_fcomplex *bufb = (accel_fcomplex *) malloc(sizeof(accel_fcomplex) * endwidth);
    accel_fcomplex *out = (accel_fcomplex *) malloc(sizeof(accel_fcomplex) * endwidth);
    float *pw = (float *) malloc(sizeof(float) * endwidth);
    accel_fcomplex *oldb = bufa, *newb = bufb;
    long rpos = 0, wpos = 0, binnum = 1;
    /* DC / Nyquist bin */
    fft[0].r = 1.0f;
    fft[0].i = 0.0f;
    rpos = wpos = 1;
    int bufflen = startwidth;
    int nold = bufflen;
    memcpy(oldb, fft + rpos, sizeof(accel_fcomplex) * nold);
    rpos += nold;
    int mid_old = nold / 2;
    block_powers(oldb, nold, pw);
    float mean_old = median_f(pw, nold) / log(2.0);
    float norm = fast_invsqrt(mean_old), dslope = 1.0f, mean_new;
    for (int i = 0; i < mid_old; i++) {
        fft[wpos + i].r = oldb[i].r * norm;
        fft[wpos + i].i = oldb[i].i * norm;
    }
    wpos += mid_old;
    binnum += nold;
    bufflen = startwidth * log(binnum);
    if (bufflen > endwidth)
        bufflen = endwidth;
    for (;;) {
        long avail = numbins - rpos;
        int nnew = (int) (avail < bufflen ? avail : bufflen);
        if (nnew <= 0)
            break;
        memcpy(newb, fft + rpos, sizeof(accel_fcomplex) * nnew);
        rpos += nnew;
        int mid_new = nnew / 2;
        block_powers(newb, nnew, pw);
        mean_new = median_f(pw, nnew) / log(2.0);
        dslope = (mean_new - mean_old) / (0.5 * (nold + nnew));
        int ii = 0;
        for (int ind = mid_old; ind < nold; ii++, ind++) {
            norm = fast_invsqrt(mean_old + dslope * ii);
            out[ii].r = oldb[ind].r * norm;
            out[ii].i = oldb[ind].i * norm;
        }
        for (int ind = 0; ind < mid_new; ii++, ind++) {
            norm = fast_invsqrt(mean_old + dslope * ii);
            out[ii].r = newb[ind].r * norm;
            out[ii].i = newb[ind].i * norm;
        }
        memcpy(fft + wpos, out, sizeof(accel_fcomplex) * ii);
        wpos += ii;
        binnum += nnew;
        if ((float) binnum / Tf < endfreq) {
            bufflen = startwidth * log(binnum);
            if (bufflen > endwidth)
                bufflen = endwidth;
        } else {
            bufflen = endwidth;
        }
        accel_fcomplex *t = oldb;
        oldb = newb;
        newb = t;
        nold = nnew;
        mean_old = mean_new;
        mid_old = mid_new;
    }
    int ii = 0;
    for (int ind = mid_old; ind < nold; ii++, ind++) {
        norm = fast_invsqrt(mean_old + dslope * ii);
        out[ii].r = oldb[ind].r * norm;
        out[ii].i = oldb[ind].i * norm;
    }
    memcpy(fft + wpos, out, sizeof(accel_fcomplex) * ii);
    wpos += ii;
    free(bufa);
    free(bufb);
    free(out);
    free(pw);
    return wpos;
}

/* After realfft of a .dat of N samples (dt each): de-redden with the
 * window growing to 200 bins at Nyquist, then bin 0 := 1 + 1i
 * (src/accel_utils.c:1561, 1621-1624). */
int accel_dered_blocks(long numbins, int startwidth, int endwidth, double endfreq, double T,
                       long long *start, int *len, int max)
{
    if (numbins < 1 + startwidth || startwidth < 1 || endwidth < startwidth)
        return -1;
    const float Tf = (float) T;
    long rpos = 1, binnum = 1;
    int nb = 0;
    int bufflen = startwidth;
    for (;;) {
        long avail = numbins - rpos;
        int n = (int) (avail < bufflen ? avail : bufflen);
        if (n <= 0)
            break;
        if (nb < max) {
            start[nb] = rpos;
            len[nb] = n;
        }
        nb++;
        rpos += n;
        binnum += n;
        /* the window growth of src/misc_utils.c:980-984 and :1036-1042 */
        if (nb == 1 || (float) binnum / Tf < endfreq) {
            bufflen = startwidth * log(binnum);
            if (bufflen > endwidth)
                bufflen = endwidth;
        } else {
            bufflen = endwidth;
        }
    }
    return nb;
}

long accel_prepare_dat_spectrum(accel_fcomplex *fft, long numbins, double N, double dt)
{
    long n = accel_deredden(fft, numbins, 6, 200, 0.5 / dt, dt * N);
    fft[0].r = 1.0f;
    fft[0].i = 1.0f;
    return n;
}

/* ------------------------------------------------------------------ */
/* Birdie zapping for in-memory (.dat) spectra: get_birdies             */
/* (src/birdzap.c:6-90) and zapbirds (src/zapping.c:58-160) as          */
/* accelsearch applies them (src/accelsearch.c:87-110).                 */
/* ------------------------------------------------------------------ */

/* MT19937 (Matsumoto & Nishimura 1998; the 2002 init_genrand seeding),
 * the generator GSL's gsl_rng_mt19937 implements. */
typedef struct {
    unsigned int mt[624];
    int mti;
} mt19937;

static void mt_seed(mt19937 *g, unsigned int s)
{
    if (s == 0)
        s = 4357;   /* GSL's substitute for a zero seed */
    g->mt[0] = s;
    for (int i = 1; i < 624; i++)
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (unsigned) i;
    g->mti = 624;
}

static unsigned int mt_next(mt19937 *g)
{
    if (g->mti >= 624) {
        for (int k = 0; k < 624; k++) {
            unsigned int y = (g->mt[k] & 0x80000000u) | (g->mt[(k + 1) % 624] & 0x7fffffffu);
            g->mt[k] = g->mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        g->mti = 0;
    }
    unsigned int y = g->mt[g->mti++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

int accel_mt19937_raw(unsigned int seed, unsigned int *out, int n)
{
    mt19937 g;
    mt_seed(&g, seed);
    for (int i = 0; i < n; i++)
        out[i] = mt_next(&g);
    return n;
}

typedef struct {
    double lo, hi;
} bird_t;

static int cmp_bird(const void *a, const void *b)
{
    double x = ((const bird_t *) a)->lo, y = ((const bird_t *) b)->lo;
    return (x > y) - (x < y);
}

int accel_read_birdies(const char *zaplist, double T, double baryv, double *lobins,
                       double *hibins, int maxbirds)
{
    FILE *f = fopen(zaplist, "r");
    if (!f)
        return -1;
    char line[200];
    int n = 0, cap = 64;
    bird_t *b = (bird_t *) malloc(sizeof(bird_t) * cap);
    while (fgets(line, sizeof(line), f)) {
        double freq, width;
        int isb = (line[0] == 'B');
        if (line[0] == '#')
            continue;
        if (sscanf(line + isb, "%lf %lf", &freq, &width) != 2)
            continue;
        if (n == cap) {
            cap *= 2;
            b = (bird_t *) realloc(b, sizeof(bird_t) * cap);
        }
        /* 'B' lines are barycentric already; others are Doppler-shifted */
        double scale = isb ? T : T * (1.0 + baryv);
        b[n].lo = (freq - 0.5 * width) * scale;
        b[n].hi = (freq + 0.5 * width) * scale;
        if (b[n].hi - b[n].lo < 1.0) {   /* at least one bin wide */
            double mid = 0.5 * (b[n].hi + b[n].lo);
            b[n].lo = mid - 0.5;
            b[n].hi = mid + 0.5;
        }
        n++;
    }
    fclose(f);
    qsort(b, n, sizeof(bird_t), cmp_bird);
    int m = n < maxbirds ? n : maxbirds;
    for (int i = 0; i < m; i++) {
        if (lobins)
            lobins[i] = b[i].lo;
        if (hibins)
            hibins[i] = b[i].hi;
    }
    free(b);
    return n;
}

#define ZAP_MEDIANBINS 200
#define ZAP_GUARDBINS 100

static float median_power(const accel_fcomplex *d, int n)
{
    float *p = (float *) malloc(sizeof(float) * n);
    for (int i = 0; i < n; i++)
        p[i] = d[i].r * d[i].r + d[i].i * d[i].i;
    float m = median_f(p, n);
    free(p);
    return m;
}

/* zapbirds on an in-memory spectrum (src/zapping.c:58-160): the bins
 * [floor(lo), ceil(hi)) become chi^2(2)-distributed noise at the local
 * mean power measured MEDIANBINS wide, GUARDBINS away on each side. */
static int zap_one(accel_fcomplex *fft, long long numbins, double lobin, double hibin,
                   mt19937 *g, int constamp)
{
    const double TWOPI = 6.28318530717958647692;
    long long ilo = (long long) floor(lobin), ihi = (long long) ceil(hibin);
    long long nz = ihi - ilo;
    if (nz < 1)
        return 0;
    long long lostart = ilo - ZAP_GUARDBINS - ZAP_MEDIANBINS, histart = ihi + ZAP_GUARDBINS;
    int uselo = lostart >= 1, usehi = histart + ZAP_MEDIANBINS <= numbins;
    if (!uselo && !usehi)
        return 0;
    if (ilo < 0 || ihi > numbins)
        return -1;
    float medpow = 0.0f;
    int sides = 0;
    if (uselo) {
        medpow += median_power(fft + lostart, ZAP_MEDIANBINS);
        sides++;
    }
    if (usehi) {
        medpow += median_power(fft + histart, ZAP_MEDIANBINS);
        sides++;
    }
    medpow /= sides;
    const float meanpow = medpow / -log(0.5);
    if (constamp) {
        /* the PRESTO 6.0.1 behaviour: sqrt(local mean power), zero phase */
        const float avgamp = sqrt(meanpow);
        for (long long i = 0; i < nz; i++) {
            fft[ilo + i].r = avgamp;
            fft[ilo + i].i = 0.0f;
        }
        return 0;
    }
    for (long long i = 0; i < nz; i++) {
        /* gsl_ran_exponential(rng, mu) = -mu log1p(-u), u = next / 2^32 */
        double u = mt_next(g) / 4294967296.0;
        double amp = sqrt(-meanpow * log1p(-u));
        double phase = TWOPI * (mt_next(g) / 4294967296.0);
        fft[ilo + i].r = (float) (amp * cos(phase));
        fft[ilo + i].i = (float) (amp * sin(phase));
    }
    return 0;
}

int accel_zap_birdies(accel_fcomplex *fft, long long numbins, double N, const double *lobins,
                      const double *hibins, int nbirds)
{
    mt19937 g;
    mt_seed(&g, 1);   /* ZAPSEED: the same zaplist always zaps alike */
    const double hibin = N / 2;
    for (int i = 0; i < nbirds; i++) {
        double lo = lobins[i], hi = hibins[i];
        if (lo >= hibin)
            break;
        if (hi >= hibin)
            hi = hibin - 1;
        if (zap_one(fft, numbins, lo, hi, &g, 0))
            return -1;
    }
    return 0;
}

int accel_zapbirds_fft(accel_fcomplex *fft, long long numbins, double N, const double *lobins,
                       const double *hibins, int nbirds, int constamp)
{
    mt19937 g;
    mt_seed(&g, 1);
    const double hibin = N / 2;
    const long long nb = (long long) hibin < numbins ? (long long) hibin : numbins;
    for (int i = 0; i < nbirds; i++) {
        double lo = lobins[i], hi = hibins[i];
        if (lo >= hibin - 1)
            break;
        if (hi >= hibin)
            hi = hibin - 1;
        if (zap_one(fft, nb, lo, hi, &g, constamp))
            return -1;
    }
    return 0;
}
