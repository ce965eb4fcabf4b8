/*
 * accel_dedisp.c -- subband de-dispersion of a filterbank in HBM into DM
 * trials, as prepsubband -nobary -noclip computes them (no mask, no
 * ignored channels, no barycentring), for the accelsearch batch.
 *
 * The plan follows src/prepsubband.c:255-306 and 364-387: DMs lodm +
 * i dmstep, the average DM, blocksperread from the largest DM's delay
 * across the band, the channel delays inside each subband at the average
 * DM (subband_search_delays) and the subband offsets of each DM
 * (subband_delays), rounded with prepsubband's NEAREST_LONG.  The
 * reference reads blocks of spectra_per_subint spectra and writes output
 * blocks of blocksperread of them, one block behind its reads in each of
 * its two stages; with K raw blocks it writes
 * J = floor((K - 1) / blocksperread) - 1 output blocks, pads the rest of
 * numout with the running mean of the first DM's data, and its samples
 * are the continuous shifted sums the kernels compute (dedisp_gpu.hip).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "presto_accel.h"
#include "agpu.h"
#include "accel_internal.h"

static long nearest_long(double x) { return (long) (x < 0 ? ceil(x - 0.5) : floor(x + 0.5)); }

/* delay_from_dm, src/dispersion.c:34-43 */
static double delay_from_dm(double dm, double f)
{
    return f == 0.0 ? 0.0 : dm / (0.000241 * f * f);
}

/* dedisp_delays without a Doppler factor (voverc 0), src/dispersion.c:64-81 */
static void chan_delays(int n, double dm, double lofreq, double width, double *out)
{
    for (int i = 0; i < n; i++)
        out[i] = delay_from_dm(dm, (lofreq + i * width) * (1.0 + 0.0));
}

/* subband_delays, src/dispersion.c:101-118 */
static void sub_delays(int numchan, int nsub, double dm, double lofreq, double width, double *out)
{
    const int cps = numchan / nsub;
    const double subbw = width * cps;
    chan_delays(nsub, dm, lofreq + subbw - width, subbw, out);
}

int accel_dedisp_plan(const accel_dedisp_params *p, long long nspec, int *idispdt, int *offsets,
                      int *blocksperread, long long *nwritten)
{
    if (!p || p->numchan <= 0 || p->nsub <= 0 || p->numchan % p->nsub || p->numdms <= 0 ||
        p->downsamp < 1 || p->spectra_per_subint <= 0 ||
        p->spectra_per_subint % p->downsamp || !(p->dt > 0.0) || !(p->chanwidth > 0.0) ||
        !(p->lofreq > 0.0) || p->numout < 0 || nspec < 0) {
        accel_set_error("accel_dedisp_plan: channels must split into subbands, the block "
                        "length into downsampled samples; dt, lofreq and an ascending "
                        "channel width must be positive");
        return -1;
    }
    const int numchan = p->numchan, nsub = p->nsub, B = p->spectra_per_subint;
    double avgdm = 0.0;
    for (int i = 0; i < p->numdms; i++)
        avgdm += p->lodm + i * p->dmstep;
    avgdm /= p->numdms;
    const double maxdm = p->lodm + (p->numdms - 1) * p->dmstep;
    const double bw_ddelay = delay_from_dm(maxdm, p->lofreq) -
                             delay_from_dm(maxdm, p->lofreq + (numchan - 1) * p->chanwidth);
    const int bpr = ((int) (bw_ddelay / p->dt) / B + 1);
    const int worklen = B * bpr, dsworklen = worklen / p->downsamp;
    double *d = (double *) malloc(sizeof(double) * (size_t) numchan);
    double *sd = (double *) malloc(sizeof(double) * (size_t) nsub);
    if (!d || !sd) {
        free(d);
        free(sd);
        accel_set_error("accel_dedisp_plan: out of host memory");
        return -1;
    }
    /* subband_search_delays at the average DM, src/dispersion.c:121-161 */
    chan_delays(numchan, avgdm, p->lofreq, p->chanwidth, d);
    sub_delays(numchan, nsub, avgdm, p->lofreq, p->chanwidth, sd);
    const int cps = numchan / nsub;
    int bad = 0;
    for (int c = 0; c < numchan; c++) {
        const long v = nearest_long((d[c] - sd[c / cps]) / p->dt);
        if (idispdt)
            idispdt[c] = (int) v;
        if (c == 0 && v > B)
            bad = 1;   /* src/backend_common.c:785-789 */
        if (v < 0)
            bad = 1;
    }
    const double dsdt = p->downsamp * p->dt;
    for (int i = 0; i < p->numdms && !bad; i++) {
        sub_delays(numchan, nsub, p->lodm + i * p->dmstep, p->lofreq, p->chanwidth, sd);
        const double top = sd[nsub - 1];
        for (int s = 0; s < nsub; s++) {
            const long v = nearest_long((sd[s] - top) / dsdt);
            if (offsets)
                offsets[(size_t) i * nsub + s] = (int) v;
            if (v < 0 || v > dsworklen)   /* src/prepsubband.c:937-943 */
                bad = 1;
        }
    }
    free(d);
    free(sd);
    if (bad) {
        accel_set_error("accel_dedisp_plan: the dispersion across a subband exceeds a block of "
                        "%d spectra, or a DM offset exceeds a working block of %d samples",
                        B, dsworklen);
        return -1;
    }
    long long K = nspec / B, J = (K - 1) / bpr - 1;
    if (K < 1 || J < 0)
        J = 0;
    long long w = J * (long long) dsworklen;
    if (w > p->numout)
        w = p->numout;
    if (blocksperread)
        *blocksperread = bpr;
    if (nwritten)
        *nwritten = w;
    return 0;
}

int accel_dedisperse_dev(const accel_dedisp_params *p, const float *d_raw, long long nspec,
                         float *d_out, double *padval, float *ms, int device)
{
    int bpr;
    long long nw;
    if (accel_dedisp_plan(p, nspec, NULL, NULL, &bpr, &nw))
        return -1;
    const int numchan = p->numchan, nsub = p->nsub, nd = p->numdms;
    int *idd = (int *) malloc(sizeof(int) * (size_t) numchan);
    int *off = (int *) malloc(sizeof(int) * (size_t) nd * nsub);
    float *row0 = (float *) malloc(sizeof(float) * (size_t) (nw > 0 ? nw : 1));
    void *stream = NULL;
    int *d_idd = NULL, *d_off = NULL;
    float *d_ds = NULL, *d_rawT = NULL;
    void *ev[3] = {NULL, NULL, NULL};
    int rc = -1;
    if (!idd || !off || !row0) {
        accel_set_error("accel_dedisperse: out of host memory");
        goto done;
    }
    accel_dedisp_plan(p, nspec, idd, off, NULL, NULL);
    int maxd = 0, maxo = 0;
    for (int c = 0; c < numchan; c++)
        maxd = idd[c] > maxd ? idd[c] : maxd;
    for (int i = 0; i < nd * nsub; i++)
        maxo = off[i] > maxo ? off[i] : maxo;
    int span = 0;
    for (int d0 = 0; d0 < nd; d0 += AGPU_DD_GROUP)
        for (int s = 0; s < nsub; s++) {
            int lo = off[(size_t) d0 * nsub + s], hi = lo;
            for (int d = d0 + 1; d < nd && d < d0 + AGPU_DD_GROUP; d++) {
                const int v = off[(size_t) d * nsub + s];
                lo = v < lo ? v : lo;
                hi = v > hi ? v : hi;
            }
            span = hi - lo > span ? hi - lo : span;
        }
    /* downsampled subband samples the written outputs read */
    const long long nds = nw > 0 ? nw + maxo : 0;
    if (agpu_init(device) || agpu_stream_create(&stream) ||
        agpu_event_create(&ev[0]) || agpu_event_create(&ev[1]) || agpu_event_create(&ev[2]) ||
        agpu_malloc((void **) &d_idd, sizeof(int) * (size_t) numchan) ||
        agpu_malloc((void **) &d_off, sizeof(int) * (size_t) nd * nsub) ||
        agpu_malloc((void **) &d_ds, sizeof(float) * (size_t) nsub * (size_t) (nds > 0 ? nds : 1)) ||
        (agpu_dedisp_needs_transpose(numchan, nsub, p->downsamp, maxd) &&
         agpu_malloc((void **) &d_rawT,
                     sizeof(float) * (size_t) numchan *
                         (size_t) (nds > 0 ? agpu_dedisp_ncols(nds, p->downsamp, maxd) : 1))) ||
        agpu_h2d(d_idd, idd, sizeof(int) * (size_t) numchan, stream) ||
        agpu_h2d(d_off, off, sizeof(int) * (size_t) nd * nsub, stream) ||
        agpu_event_record(ev[0], stream) ||
        agpu_dedisp_subbands(d_raw, nspec, numchan, d_idd, maxd, nsub, p->downsamp, nds, d_rawT,
                             d_ds, stream) ||
        agpu_event_record(ev[1], stream) ||
        agpu_dedisp_dms(d_ds, nds, nsub, d_off, maxo, span, nd, nw, p->numout, d_out, stream) ||
        agpu_event_record(ev[2], stream) ||
        (nw > 0 && agpu_d2h(row0, d_out, sizeof(float) * (size_t) nw, stream)) ||
        (ms && (agpu_event_elapsed(&ms[0], ev[0], ev[1]) ||
                agpu_event_elapsed(&ms[1], ev[1], ev[2])))) {
        accel_set_error("accel_dedisperse: %s", agpu_last_error());
        goto done;
    }
    /* the padding: update_stats' running mean of the first DM's written
     * samples (include/misc_utils.h:146-158, src/prepsubband.c:429-431,
     * 667-718) */
    double avg = 0.0;
    for (long long i = 0; i < nw; i++) {
        const double x = row0[i], dev = x - avg;
        avg += dev / (i + 1.0);
    }
    if (agpu_dedisp_fill(d_out, nd, p->numout, nw, (float) avg, stream) ||
        agpu_stream_sync(stream)) {
        accel_set_error("accel_dedisperse: %s", agpu_last_error());
        goto done;
    }
    if (padval)
        *padval = avg;
    rc = 0;
done:
    if (stream)
        agpu_stream_sync(stream);
    for (int i = 0; i < 3; i++)
        if (ev[i])
            agpu_event_destroy(ev[i]);
    agpu_free(d_ds);
    agpu_free(d_rawT);
    agpu_free(d_off);
    agpu_free(d_idd);
    if (stream)
        agpu_stream_destroy(stream);
    free(idd);
    free(off);
    free(row0);
    return rc;
}

long long accel_dedisperse(const accel_dedisp_params *p, const float *raw, long long nspec,
                           float *out, double *padval, int device)
{
    long long nw;
    if (accel_dedisp_plan(p, nspec, NULL, NULL, NULL, &nw))
        return -1;
    float *d_raw = NULL, *d_out = NULL;
    void *stream = NULL;
    long long rc = -1;
    const size_t rawb = sizeof(float) * (size_t) nspec * (size_t) p->numchan;
    const size_t outb = sizeof(float) * (size_t) p->numdms * (size_t) p->numout;
    if (agpu_init(device) || agpu_stream_create(&stream) ||
        agpu_malloc((void **) &d_raw, rawb ? rawb : 4) ||
        agpu_malloc((void **) &d_out, outb ? outb : 4) ||
        (rawb && agpu_h2d(d_raw, raw, rawb, stream))) {
        accel_set_error("accel_dedisperse: %s", agpu_last_error());
        goto done;
    }
    if (accel_dedisperse_dev(p, d_raw, nspec, d_out, padval, NULL, device))
        goto done;
    if (outb && agpu_d2h(out, d_out, outb, stream)) {
        accel_set_error("accel_dedisperse: %s", agpu_last_error());
        goto done;
    }
    rc = nw;
done:
    agpu_free(d_out);
    agpu_free(d_raw);
    if (stream)
        agpu_stream_destroy(stream);
    return rc;
}

void *accel_device_alloc(size_t bytes, int device)
{
    void *p = NULL;
    if (agpu_init(device) || agpu_malloc(&p, bytes ? bytes : 4)) {
        accel_set_error("accel_device_alloc: %s", agpu_last_error());
        return NULL;
    }
    return p;
}

int accel_device_free(void *p)
{
    return agpu_free(p);
}

int accel_device_upload(void *d_dst, const void *src, size_t bytes)
{
    if (agpu_h2d(d_dst, src, bytes, NULL)) {
        accel_set_error("accel_device_upload: %s", agpu_last_error());
        return -1;
    }
    return 0;
}

int accel_device_download(void *dst, const void *d_src, size_t bytes)
{
    if (agpu_d2h(dst, d_src, bytes, NULL)) {
        accel_set_error("accel_device_download: %s", agpu_last_error());
        return -1;
    }
    return 0;
}
