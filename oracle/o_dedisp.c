/*
 * o_dedisp.c -- CPU restatement of PRESTO's subband de-dispersion
 * (prepsubband's topocentric path without masking, clipping or ignored
 * channels).  TEST INFRASTRUCTURE ONLY: the checker for the HIP
 * de-dispersion kernels; no product code uses it.
 *
 *   o_delay_from_dm, o_dedisp_delays          src/dispersion.c:34-81
 *   o_subband_delays, o_subband_search_delays src/dispersion.c:101-161
 *   o_dedisp_plan      the DM list, average DM, blocksperread, channel
 *                      and subband delays: src/prepsubband.c:255-306,
 *                      364-387 (NEAREST_LONG of prepsubband.c:23)
 *   o_prepsubband      the data path, block by block as the reference
 *                      runs it: prep_subbands' transpose + dedisp_subbands
 *                      (src/backend_common.c:659-760, read_subbands
 *                      :761-813, src/dispersion.c:164-198), get_data's
 *                      double buffers and downsampling
 *                      (src/prepsubband.c:918-1040), float_dedisp
 *                      (src/dispersion.c:201-224), the write loop
 *                      (:399-431) and the padding with the running mean
 *                      of the first DM's data (:667-718, update_stats of
 *                      include/misc_utils.h:146-158).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

static long nearest_long(double x) { return (long) (x < 0 ? ceil(x - 0.5) : floor(x + 0.5)); }

double o_delay_from_dm(double dm, double freq_emitted)
{
    if (freq_emitted == 0.0)
        return 0.0;
    return dm / (0.000241 * freq_emitted * freq_emitted);
}

void o_dedisp_delays(int numchan, double dm, double lofreq, double chanwidth, double voverc,
                     double *delays)
{
    for (int ii = 0; ii < numchan; ii++) {
        double freq = (lofreq + ii * chanwidth) * (1.0 + voverc);   /* doppler() */
        delays[ii] = o_delay_from_dm(dm, freq);
    }
}

void o_subband_delays(int numchan, int numsubbands, double dm, double lofreq, double chanwidth,
                      double voverc, double *delays)
{
    int chan_per_subband = numchan / numsubbands;
    double subbandwidth = chanwidth * chan_per_subband;
    double losub_hifreq = lofreq + subbandwidth - chanwidth;
    o_dedisp_delays(numsubbands, dm, losub_hifreq, subbandwidth, voverc, delays);
}

void o_subband_search_delays(int numchan, int numsubbands, double dm, double lofreq,
                             double chanwidth, double voverc, double *delays)
{
    int chan_per_subband = numchan / numsubbands;
    double *sub = (double *) malloc(sizeof(double) * numsubbands);
    o_subband_delays(numchan, numsubbands, dm, lofreq, chanwidth, voverc, sub);
    o_dedisp_delays(numchan, dm, lofreq, chanwidth, voverc, delays);
    for (int ii = 0; ii < numsubbands; ii++)
        for (int jj = 0; jj < chan_per_subband; jj++)
            delays[ii * chan_per_subband + jj] -= sub[ii];
    free(sub);
}

/* idispdt[numchan] (bins of dt), offsets[numdms][nsub] (bins of dt *
 * downsamp), *bpr = blocksperread; returns the average DM */
double o_dedisp_plan(int numchan, double dt, double lofreq, double chanwidth,
                     int spectra_per_subint, int nsub, int downsamp, int numdms, double lodm,
                     double dmstep, int *idispdt, int *offsets, int *bpr)
{
    double avgdm = 0.0, maxdm;
    double *dms = (double *) malloc(sizeof(double) * numdms);
    for (int ii = 0; ii < numdms; ii++) {
        dms[ii] = lodm + ii * dmstep;
        avgdm += dms[ii];
    }
    avgdm /= numdms;
    maxdm = dms[numdms - 1];
    const double dsdt = downsamp * dt;
    const double bw_ddelay = o_delay_from_dm(maxdm, lofreq) -
                             o_delay_from_dm(maxdm, lofreq + (numchan - 1) * chanwidth);
    *bpr = ((int) (bw_ddelay / dt) / spectra_per_subint + 1);
    double *d = (double *) malloc(sizeof(double) * (numchan > nsub ? numchan : nsub));
    o_subband_search_delays(numchan, nsub, avgdm, lofreq, chanwidth, 0.0, d);
    for (int ii = 0; ii < numchan; ii++)
        idispdt[ii] = (int) nearest_long(d[ii] / dt);
    for (int ii = 0; ii < numdms; ii++) {
        o_subband_delays(numchan, nsub, dms[ii], lofreq, chanwidth, 0.0, d);
        double dtmp = d[nsub - 1];
        for (int jj = 0; jj < nsub; jj++)
            offsets[ii * nsub + jj] = (int) nearest_long((d[jj] - dtmp) / dsdt);
    }
    free(d);
    free(dms);
    return avgdm;
}

/* dedisp_subbands: data and lastdata channel-major [numchan][numpts];
 * result subband-major [nsub][numpts] */
static void dedisp_subbands(const float *data, const float *lastdata, int numpts, int numchan,
                            const int *delays, int numsubbands, float *result)
{
    const int cps = numchan / numsubbands;
    for (long long ii = 0; ii < (long long) numpts * numsubbands; ii++)
        result[ii] = 0.0f;
    for (int ii = 0; ii < numchan; ii++) {
        const int subnum = ii / cps, dind = delays[ii];
        float *sub = result + (long long) subnum * numpts;
        const float *chan = lastdata + (long long) ii * numpts + dind;
        long long jj, kk;
        for (jj = 0; jj < numpts - dind; jj++)
            sub[jj] += chan[jj];
        chan = data + (long long) ii * numpts;
        for (jj = numpts - dind, kk = 0; jj < numpts; jj++, kk++)
            sub[jj] += chan[kk];
    }
}

/* float_dedisp: data / lastdata time-major [numpts][numchan] */
static void float_dedisp(const float *data, const float *lastdata, int numpts, int numchan,
                         const int *delays, float approx_mean, float *result)
{
    for (long long ii = 0; ii < numpts; ii++)
        result[ii] = -approx_mean;
    for (long long ii = 0; ii < numchan; ii++) {
        long long jj = ii + (long long) delays[ii] * numchan, kk;
        for (kk = 0; kk < numpts - delays[ii]; kk++, jj += numchan)
            result[kk] += lastdata[jj];
        jj = ii;
        for (; kk < numpts; kk++, jj += numchan)
            result[kk] += data[jj];
    }
}

typedef struct {
    /* input: raw spectra, time-major [nspec][numchan], ascending frequency */
    const float *raw;
    long long nspec, next;       /* spectra, next unread spectrum */
    int numchan, B, nsub;
    const int *idispdt;
    /* prep_subbands' buffers, channel-major */
    float *cur, *last;
    int primed;
} rawstream;

/* read_subbands (transpose 0): the subband block of the oldest raw block,
 * time-major [B][nsub] in fdata; returns B, or 0 when no raw block is left */
static int read_subbands(rawstream *r, float *fdata, float *tmp)
{
    for (;;) {
        if (r->next + r->B > r->nspec)
            return 0;
        /* the raw block, transposed to channel-major */
        const float *blk = r->raw + r->next * r->numchan;
        for (int t = 0; t < r->B; t++)
            for (int c = 0; c < r->numchan; c++)
                r->cur[(long long) c * r->B + t] = blk[(long long) t * r->numchan + c];
        r->next += r->B;
        if (!r->primed) {
            float *s = r->cur;
            r->cur = r->last;
            r->last = s;
            r->primed = 1;
            continue;
        }
        dedisp_subbands(r->cur, r->last, r->B, r->numchan, r->idispdt, r->nsub, tmp);
        float *s = r->cur;
        r->cur = r->last;
        r->last = s;
        for (int t = 0; t < r->B; t++)
            for (int k = 0; k < r->nsub; k++)
                fdata[(long long) t * r->nsub + k] = tmp[(long long) k * r->B + t];
        return r->B;
    }
}

/* Writes out[numdms][numout]; returns the data points written before the
 * padding (the rest is the first DM's running mean), or -1 on bad
 * geometry.  *padval gets that mean. */
long long o_prepsubband(const float *raw, long long nspec, int numchan, double dt,
                        double lofreq, double chanwidth, int spectra_per_subint, int nsub,
                        int downsamp, int numdms, double lodm, double dmstep, long long numout,
                        float *out, double *padval)
{
    const int B = spectra_per_subint;
    if (numchan % nsub || B % downsamp)
        return -1;
    int *idispdt = (int *) malloc(sizeof(int) * numchan);
    int *offsets = (int *) malloc(sizeof(int) * numdms * nsub);
    int bpr;
    o_dedisp_plan(numchan, dt, lofreq, chanwidth, B, nsub, downsamp, numdms, lodm, dmstep,
                  idispdt, offsets, &bpr);
    const int worklen = B * bpr, dsworklen = worklen / downsamp;
    /* the reference stops on the first (src/backend_common.c:785-789) and
     * would read past its buffers on the second (src/prepsubband.c:937-943) */
    int bad = idispdt[0] > B;
    for (int ii = 0; ii < numdms * nsub; ii++)
        bad |= offsets[ii] > dsworklen;
    if (bad) {
        free(idispdt);
        free(offsets);
        return -1;
    }
    rawstream r = {raw, nspec, 0, numchan, B, nsub, idispdt,
                   (float *) malloc(sizeof(float) * numchan * B),
                   (float *) malloc(sizeof(float) * numchan * B), 0};
    float *tmp = (float *) malloc(sizeof(float) * nsub * B);
    float *d1 = (float *) malloc(sizeof(float) * nsub * worklen);
    float *d2 = (float *) malloc(sizeof(float) * nsub * worklen);
    float *ds1 = downsamp > 1 ? (float *) malloc(sizeof(float) * nsub * dsworklen) : d1;
    float *ds2 = downsamp > 1 ? (float *) malloc(sizeof(float) * nsub * dsworklen) : d2;
    float *cur = d1, *last = d2, *curds = ds1, *lastds = ds2;
    float *outdata = (float *) malloc(sizeof(float) * numdms * dsworklen);
    int firsttime = 1;
    long long totwrote = 0, statnum = 0;
    double avg = 0.0;
    for (;;) {
        /* get_data */
        long long totnumread = 0;
        for (;;) {
            for (int ii = 0; ii < bpr; ii++) {
                int numread = read_subbands(&r, cur + (long long) ii * B * nsub, tmp);
                if (!firsttime)
                    totnumread += numread;
                if (numread != B)
                    memset(cur + (long long) ii * B * nsub, 0, sizeof(float) * B * nsub);
            }
            if (downsamp > 1) {
                for (int ii = 0; ii < dsworklen; ii++) {
                    const int dsoffset = ii * nsub, offset = dsoffset * downsamp;
                    for (int jj = 0; jj < nsub; jj++) {
                        int index = offset + jj;
                        float ftmp = 0.0f;
                        curds[dsoffset + jj] = 0.0f;
                        for (int kk = 0; kk < downsamp; kk++) {
                            ftmp += cur[index];
                            index += nsub;
                        }
                        curds[dsoffset + jj] += ftmp / downsamp;
                    }
                }
            }
            if (firsttime) {
                float *s = cur;
                cur = last;
                last = s;
                s = curds;
                curds = lastds;
                lastds = s;
                firsttime = 0;
            } else {
                break;
            }
        }
        for (int ii = 0; ii < numdms; ii++)
            float_dedisp(curds, lastds, dsworklen, nsub, offsets + ii * nsub, 0.0f,
                         outdata + (long long) ii * dsworklen);
        {
            float *s = cur;
            cur = last;
            last = s;
            s = curds;
            curds = lastds;
            lastds = s;
        }
        if (totnumread != worklen)
            break;
        /* the write loop */
        long long numtowrite = dsworklen;
        if (totwrote + numtowrite > numout)
            numtowrite = numout - totwrote;
        for (int ii = 0; ii < numdms; ii++)
            memcpy(out + (long long) ii * numout + totwrote, outdata + (long long) ii * dsworklen,
                   sizeof(float) * numtowrite);
        for (long long ii = 0; ii < numtowrite; ii++) {
            /* update_stats: avg only */
            double x = outdata[ii], dev = x - avg;
            avg += dev / ((statnum + ii) + 1.0);
        }
        statnum += numtowrite;
        totwrote += numtowrite;
        if (totwrote == numout)
            break;
    }
    for (int ii = 0; ii < numdms; ii++)
        for (long long jj = totwrote; jj < numout; jj++)
            out[(long long) ii * numout + jj] = (float) avg;
    *padval = avg;
    free(idispdt);
    free(offsets);
    free(r.cur);
    free(r.last);
    free(tmp);
    if (downsamp > 1) {
        free(ds1);
        free(ds2);
    }
    free(d1);
    free(d2);
    free(outdata);
    return totwrote;
}
