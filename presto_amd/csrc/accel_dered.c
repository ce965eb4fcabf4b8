/*
 * accel_dered.c -- spectrum preparation for time-series (.dat) input:
 * de-reddening by running block medians (src/misc_utils.c:912-1067, with
 * invsqrtf from :291-302) and the bin-0 convention of create_accelobs
 * (src/accel_utils.c:1621-1624).  Built with -ffp-contract=off so the FP32
 * expressions round as the reference's do.
 */
#include <math.h>
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
long accel_prepare_dat_spectrum(accel_fcomplex *fft, long numbins, double N, double dt)
{
    long n = accel_deredden(fft, numbins, 6, 200, 0.5 / dt, dt * N);
    fft[0].r = 1.0f;
    fft[0].i = 1.0f;
    return n;
}
