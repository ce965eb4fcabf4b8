/*
 * o_fft.c -- oracle FFTs: mixed-radix (2,3,4,5,generic) Stockham complex
 * FFT in FP32 and FP64, the reference's packed real FFT, and the running
 * block-median de-reddening.  Test infrastructure only (see oracle.h).
 *
 * Sign convention follows src/fftcalls.c: isign = -1 is the forward
 * transform exp(-2 pi i jk/n); transforms are unnormalised.
 */
#include <math.h>
#include <stdlib.h>
#include <time.h>
#include <string.h>
#include <pthread.h>
#include "oracle.h"

#define O_TWOPI 6.2831853071795864769252867665590057683943387987502

typedef struct {
    long n;
    int nf, fac[64];
    o_dcomplex *tw;   /* exp(-2 pi i k / n), k < n */
    o_fcomplex *twf;
} fftplan;

#define NPLANS 48
static fftplan *g_plans[NPLANS];
static int g_nplans = 0;
static pthread_mutex_t g_plan_lock = PTHREAD_MUTEX_INITIALIZER;

static fftplan *get_plan(long n)
{
    pthread_mutex_lock(&g_plan_lock);
    for (int i = 0; i < g_nplans; i++)
        if (g_plans[i]->n == n) {
            fftplan *p = g_plans[i];
            pthread_mutex_unlock(&g_plan_lock);
            return p;
        }
    fftplan *p = (fftplan *) calloc(1, sizeof(fftplan));
    p->n = n;
    long m = n;
    while (m % 4 == 0) { p->fac[p->nf++] = 4; m /= 4; }
    while (m % 2 == 0) { p->fac[p->nf++] = 2; m /= 2; }
    while (m % 3 == 0) { p->fac[p->nf++] = 3; m /= 3; }
    while (m % 5 == 0) { p->fac[p->nf++] = 5; m /= 5; }
    for (long f = 7; m > 1; f += 2)
        while (m % f == 0) { p->fac[p->nf++] = (int) f; m /= f; }
    p->tw = (o_dcomplex *) malloc(sizeof(o_dcomplex) * n);
    p->twf = (o_fcomplex *) malloc(sizeof(o_fcomplex) * n);
    for (long k = 0; k < n; k++) {
        double a = -O_TWOPI * (double) k / (double) n;
        p->tw[k].r = cos(a);
        p->tw[k].i = sin(a);
        p->twf[k].r = (float) p->tw[k].r;
        p->twf[k].i = (float) p->tw[k].i;
    }
    if (g_nplans == NPLANS) {          /* evict the oldest */
        free(g_plans[0]->tw);
        free(g_plans[0]->twf);
        free(g_plans[0]);
        memmove(g_plans, g_plans + 1, sizeof(fftplan *) * (NPLANS - 1));
        g_nplans--;
    }
    g_plans[g_nplans++] = p;
    pthread_mutex_unlock(&g_plan_lock);
    return p;
}

/* One Stockham pass: radix R, Ns = product of earlier radices.
 * y[(j/Ns)*Ns*R + j%Ns + r*Ns] = DFT_R over r of x[j + r*N/R]*w^(r*(j%Ns)).
 * j runs as (b, jm) = (j/Ns, j%Ns); the twiddle index jm*r*step stays
 * below n, so no reduction modulo n is needed. */
#define STOCKHAM_IMPL(NAME, T, TC, TWF)                                          \
static void NAME(TC *x, TC *y, long n, int isign, const fftplan *p)              \
{                                                                                \
    long Ns = 1;                                                                 \
    TC *in = x, *out = y;                                                        \
    const TC *tw = (const TC *) p->TWF;                                          \
    const T sgn = (isign < 0) ? (T) 1 : (T) -1;                                  \
    for (int f = 0; f < p->nf; f++) {                                            \
        int R = p->fac[f];                                                       \
        long NR = n / R, step = n / (Ns * R), nb = NR / Ns;                      \
        TC v[64], u[64];                                                         \
        for (long b = 0; b < nb; b++)                                            \
        for (long jm = 0; jm < Ns; jm++) {                                       \
            const long j = b * Ns + jm;                                          \
            for (int r = 0; r < R; r++) {                                        \
                TC a = in[j + r * NR];                                           \
                if (jm && r) {                                                   \
                    const long k = jm * r * step;                                \
                    T wr = tw[k].r, wi = sgn * tw[k].i;                          \
                    TC c;                                                        \
                    c.r = a.r * wr - a.i * wi;                                   \
                    c.i = a.r * wi + a.i * wr;                                   \
                    a = c;                                                       \
                }                                                                \
                v[r] = a;                                                        \
            }                                                                    \
            if (R == 2) {                                                        \
                u[0].r = v[0].r + v[1].r; u[0].i = v[0].i + v[1].i;              \
                u[1].r = v[0].r - v[1].r; u[1].i = v[0].i - v[1].i;              \
            } else if (R == 4) {                                                 \
                TC s0, s1, d0, d1;                                               \
                s0.r = v[0].r + v[2].r; s0.i = v[0].i + v[2].i;                  \
                d0.r = v[0].r - v[2].r; d0.i = v[0].i - v[2].i;                  \
                s1.r = v[1].r + v[3].r; s1.i = v[1].i + v[3].i;                  \
                d1.r = v[1].r - v[3].r; d1.i = v[1].i - v[3].i;                  \
                /* multiply d1 by -i*isign' : forward (isign<0) uses -i */       \
                TC e;                                                            \
                if (isign < 0) { e.r = d1.i; e.i = -d1.r; }                      \
                else { e.r = -d1.i; e.i = d1.r; }                                \
                u[0].r = s0.r + s1.r; u[0].i = s0.i + s1.i;                      \
                u[2].r = s0.r - s1.r; u[2].i = s0.i - s1.i;                      \
                u[1].r = d0.r + e.r;  u[1].i = d0.i + e.i;                       \
                u[3].r = d0.r - e.r;  u[3].i = d0.i - e.i;                       \
            } else {                                                             \
                /* generic O(R^2) DFT with exact root table from the plan */     \
                long base = n / R;                                               \
                for (int q = 0; q < R; q++) {                                    \
                    T sr = 0, si = 0;                                            \
                    for (int r = 0; r < R; r++) {                                \
                        long k = ((long) q * r % R) * base;                      \
                        T wr = tw[k].r, wi = sgn * tw[k].i;                      \
                        sr += v[r].r * wr - v[r].i * wi;                         \
                        si += v[r].r * wi + v[r].i * wr;                         \
                    }                                                            \
                    u[q].r = sr; u[q].i = si;                                    \
                }                                                                \
            }                                                                    \
            const long idxD = b * Ns * R + jm;                                   \
            for (int r = 0; r < R; r++)                                          \
                out[idxD + r * Ns] = u[r];                                       \
        }                                                                        \
        Ns *= R;                                                                 \
        TC *t = in; in = out; out = t;                                           \
    }                                                                            \
    if (in != x)                                                                 \
        memcpy(x, in, sizeof(TC) * n);                                           \
}

STOCKHAM_IMPL(stockham_f, float, o_fcomplex, twf)
STOCKHAM_IMPL(stockham_d, double, o_dcomplex, tw)

void o_cfft_f(o_fcomplex *data, long n, int isign)
{
    if (n <= 1)
        return;
    fftplan *p = get_plan(n);
    o_fcomplex *work = (o_fcomplex *) malloc(sizeof(o_fcomplex) * n);
    stockham_f(data, work, n, isign, p);
    free(work);
}

void o_cfft_d(o_dcomplex *data, long n, int isign)
{
    if (n <= 1)
        return;
    fftplan *p = get_plan(n);
    o_dcomplex *work = (o_dcomplex *) malloc(sizeof(o_dcomplex) * n);
    stockham_d(data, work, n, isign, p);
    free(work);
}

/* src/fastffts.c:163-243: packed real FFT (data[0] = (DC, Nyquist)) with
 * the FP64 trig recursion of the reference around an FP32 complex FFT. */
void o_realfft_f(float *idata, long n, int isign)
{
    long nby2 = n >> 1, nby4 = n >> 2, il, ih;
    double cc, h1r, h1i, h2r, h2i, h2rwr, h2iwr, h2rwi, h2iwi;
    double wr, wi, wpr, wpi, tmp1, theta;
    o_fcomplex *data = (o_fcomplex *) idata;

    if (isign == -1) {
        cc = -0.5;
        theta = -O_TWOPI / (double) n;
        o_cfft_f(data, nby2, -1);
    } else {
        cc = 0.5;
        theta = O_TWOPI / (double) n;
        if ((n + 2) % 4)
            data[nby4].i = -data[nby4].i;
    }
    wr = cos(theta);
    wi = sin(theta);
    tmp1 = sin(0.5 * theta);
    wpr = -2.0 * tmp1 * tmp1;
    wpi = wi;
    il = 1;
    ih = nby2 - il;
    for (; il <= nby4; il++, ih--) {
        h1r = 0.5 * (data[il].r + data[ih].r);
        h1i = 0.5 * (data[il].i - data[ih].i);
        h2r = -cc * (data[il].i + data[ih].i);
        h2i = cc * (data[il].r - data[ih].r);
        h2rwr = h2r * wr;
        h2rwi = h2r * wi;
        h2iwr = h2i * wr;
        h2iwi = h2i * wi;
        data[il].r = h1r + h2rwr - h2iwi;
        data[il].i = h1i + h2iwr + h2rwi;
        data[ih].r = h1r - h2rwr + h2iwi;
        data[ih].i = -h1i + h2iwr + h2rwi;
        tmp1 = wr;
        wr = tmp1 * wpr - wi * wpi + wr;
        wi = wi * wpr + tmp1 * wpi + wi;
    }
    if (isign == -1) {
        tmp1 = data[0].r;
        data[0].r = tmp1 + data[0].i;
        data[0].i = tmp1 - data[0].i;
    } else {
        if ((n + 2) % 4)
            data[nby4].i = -data[nby4].i;
        tmp1 = data[0].r;
        data[0].r = 0.5 * (tmp1 + data[0].i);
        data[0].i = 0.5 * (tmp1 - data[0].i);
        o_cfft_f(data, nby2, 1);
        tmp1 = 2.0 / (double) n;
        for (il = 0; il < n; il++)
            idata[il] *= tmp1;
    }
}

/* src/misc_utils.c:912-1067 (dered_engine over an in-core array) */
long o_deredden(o_fcomplex *fft, long numbins, int startwidth, int endwidth,
                double endfreq, double T)
{
    long binnum = 1, numwrote = 0, rpos = 0, wpos = 0;
    int bufflen, nblk_old, nblk_new, mid_old, mid_new, ii, ind;
    float mean_old, mean_new, dslope = 1.0f, norm, powr, powi;
    const float Tf = (float) T;
    o_fcomplex *inbuf1 = (o_fcomplex *) malloc(sizeof(o_fcomplex) * endwidth);
    o_fcomplex *inbuf2 = (o_fcomplex *) malloc(sizeof(o_fcomplex) * endwidth);
    o_fcomplex *outbuf = (o_fcomplex *) malloc(sizeof(o_fcomplex) * endwidth);
    float *powbuf = (float *) malloc(sizeof(float) * endwidth);
    o_fcomplex *oldbuf = inbuf1, *newbuf = inbuf2, *tmpbuf;

#define RD(buf, nb) ({ long _n = numbins - rpos; if (_n > (nb)) _n = (nb);          \
                       if (_n < 0) _n = 0;                                           \
                       memcpy((buf), fft + rpos, _n * sizeof(o_fcomplex));          \
                       rpos += _n; _n; })
#define WR(buf, nb) do { if ((nb) > 0) { memcpy(fft + wpos, (buf),                    \
                         (nb) * sizeof(o_fcomplex)); wpos += (nb); } } while (0)

    if (RD(oldbuf, 1) != 1)
        goto error;
    oldbuf[0].r = 1.0f;
    oldbuf[0].i = 0.0f;
    WR(oldbuf, 1);
    numwrote += 1;

    bufflen = startwidth;
    nblk_old = (int) RD(oldbuf, bufflen);
    if (nblk_old != bufflen)
        goto error;
    mid_old = nblk_old / 2;
    for (ii = 0; ii < nblk_old; ii++) {
        powr = oldbuf[ii].r;
        powi = oldbuf[ii].i;
        powbuf[ii] = powr * powr + powi * powi;
    }
    mean_old = o_median(powbuf, nblk_old) / log(2.0);
    norm = o_invsqrtf(mean_old);
    for (ii = 0; ii < mid_old; ii++) {
        outbuf[ii].r = oldbuf[ii].r * norm;
        outbuf[ii].i = oldbuf[ii].i * norm;
    }
    WR(outbuf, mid_old);
    numwrote += mid_old;

    binnum += nblk_old;
    bufflen = startwidth * log(binnum);
    if (bufflen > endwidth)
        bufflen = endwidth;

    while ((nblk_new = (int) RD(newbuf, bufflen))) {
        mid_new = nblk_new / 2;
        for (ii = 0; ii < nblk_new; ii++) {
            powr = newbuf[ii].r;
            powi = newbuf[ii].i;
            powbuf[ii] = powr * powr + powi * powi;
        }
        mean_new = o_median(powbuf, nblk_new) / log(2.0);
        dslope = (mean_new - mean_old) / (0.5 * (nblk_old + nblk_new));
        for (ii = 0, ind = mid_old; ind < nblk_old; ii++, ind++) {
            norm = o_invsqrtf(mean_old + dslope * ii);
            outbuf[ii].r = oldbuf[ind].r * norm;
            outbuf[ii].i = oldbuf[ind].i * norm;
        }
        for (ind = 0; ind < mid_new; ii++, ind++) {
            norm = o_invsqrtf(mean_old + dslope * ii);
            outbuf[ii].r = newbuf[ind].r * norm;
            outbuf[ii].i = newbuf[ind].i * norm;
        }
        WR(outbuf, ii);
        numwrote += ii;
        binnum += nblk_new;
        if ((float) binnum / Tf < endfreq) {
            bufflen = startwidth * log(binnum);
            if (bufflen > endwidth)
                bufflen = endwidth;
        } else {
            bufflen = endwidth;
        }
        tmpbuf = oldbuf;
        oldbuf = newbuf;
        newbuf = tmpbuf;
        nblk_old = nblk_new;
        mean_old = mean_new;
        mid_old = mid_new;
    }
    for (ii = 0, ind = mid_old; ind < nblk_old; ii++, ind++) {
        norm = o_invsqrtf(mean_old + dslope * ii);
        outbuf[ii].r = oldbuf[ind].r * norm;
        outbuf[ii].i = oldbuf[ind].i * norm;
    }
    WR(outbuf, nblk_old - mid_old);
    numwrote += nblk_old - mid_old;
    free(inbuf1); free(inbuf2); free(outbuf); free(powbuf);
    return numwrote;
error:
    free(inbuf1); free(inbuf2); free(outbuf); free(powbuf);
    return -1;
#undef RD
#undef WR
}

/* Nanoseconds per in-place forward transform of n complex floats, the
 * median of `reps` timed batches of 64 transforms on one thread: the
 * figure reported beside the CPU baseline, so that this FFT can be set
 * against FFTW's. */
double o_time_cfft_ns(long n, int reps)
{
    o_fcomplex *x = (o_fcomplex *) malloc(sizeof(o_fcomplex) * n);
    double *t = (double *) malloc(sizeof(double) * (reps > 0 ? reps : 1));
    for (long i = 0; i < n; i++) {
        x[i].r = (float) ((i * 7919) % 101) / 101.0f;
        x[i].i = (float) ((i * 104729) % 97) / 97.0f;
    }
    o_cfft_f(x, n, -1);   /* plan made before timing */
    for (int k = 0; k < reps; k++) {
        struct timespec a, b;
        clock_gettime(CLOCK_MONOTONIC, &a);
        for (int j = 0; j < 64; j++)
            o_cfft_f(x, n, (j & 1) ? 1 : -1);
        clock_gettime(CLOCK_MONOTONIC, &b);
        t[k] = ((b.tv_sec - a.tv_sec) * 1e9 + (b.tv_nsec - a.tv_nsec)) / 64.0;
    }
    for (int i = 1; i < reps; i++)   /* insertion sort for the median */
        for (int j = i; j > 0 && t[j - 1] > t[j]; j--) {
            double tmp = t[j];
            t[j] = t[j - 1];
            t[j - 1] = tmp;
        }
    double med = reps > 0 ? t[reps / 2] : 0.0;
    free(x);
    free(t);
    return med;
}
