/*
 * o_accel.c -- oracle: accelsearch geometry, template banks, the f-fdot
 * plane, harmonic sums, thresholding, the candidate list and the block
 * loop of the driver.  Test infrastructure only (see oracle.h).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include <omp.h>

#define NEAREST_INT(x) ((int) ((x) < 0 ? (x) - 0.5 : (x) + 0.5))

static int log2i(int n)
{
    int x = 0;
    while (n > 1) {
        n >>= 1;
        x++;
    }
    return x;
}

/* src/accel_utils.c:54-102 */
static double calc_required_r(double harm_fract, double rfull)
{
    return rint(O_RDR * rfull * harm_fract) * O_DR;
}

static int calc_required_z(double harm_fract, double zfull)
{
    double x = O_RDZ * zfull * harm_fract;
    return NEAREST_INT(x) * O_DZ;
}

static int calc_required_w(double harm_fract, double wfull)
{
    double x = O_RDW * wfull * harm_fract;
    return NEAREST_INT(x) * O_DW;
}

static int index_from_r(double r, double lor)
{
    return (int) ((r - lor) * O_RDR + O_DBLCORRECT);
}

static int index_from_z(double z, double loz)
{
    return (int) ((z - loz) * O_RDZ + O_DBLCORRECT);
}

static int index_from_w(double w, double low)
{
    return (int) ((w - low) * O_RDW + O_DBLCORRECT);
}

/* src/accel_utils.c:1596-1804 (the numeric part of create_accelobs) */
int o_create_accelobs(o_obs *obs)
{
    if (obs->zmax % O_DZ)
        obs->zmax = (obs->zmax / O_DZ + 1) * O_DZ;
    obs->lobin = obs->lobin_arg;
    if (obs->lobin > 0)
        obs->nph = 0.0;
    if (obs->lobin > 0 && obs->lobin > obs->numbins - 1)
        return 1;
    if (obs->numharm != 1 && obs->numharm != 2 && obs->numharm != 4 &&
        obs->numharm != 8 && obs->numharm != 16 && obs->numharm != 32)
        return 2;
    obs->numharmstages = log2i(obs->numharm) + 1;
    obs->dz = O_DZ;
    obs->numz = (obs->zmax / O_DZ) * 2 + 1;
    if (obs->wmaxP) {
        if (obs->wmax % O_DW)
            obs->wmax = (obs->wmax / O_DW + 1) * O_DW;
        obs->whi = obs->wmax;
        obs->wlo = -obs->wmax;
        obs->dw = O_DW;
        obs->numw = (obs->wmax / O_DW) * 2 + 1;
        if (obs->wmax == 0)
            obs->numw = 0;
    } else {
        obs->whi = obs->wlo = obs->dw = 0.0;
        obs->numw = 0;
    }
    obs->numbetween = O_NUMBETWEEN;
    obs->T = obs->dt * obs->N;
    if (obs->floP) {
        obs->rlo = floor(obs->flo * obs->T);
        if (obs->rlo < obs->lobin)
            obs->rlo = obs->lobin;
        if (obs->rlo > obs->numbins - 1)
            return 3;
    } else {
        obs->rlo = obs->rloP ? obs->rlo_arg : 1.0;
        if (obs->rlo < obs->lobin)
            obs->rlo = obs->lobin;
        if (obs->rlo > obs->numbins - 1)
            return 3;
    }
    obs->highestbin = obs->numbins - 1;
    obs->rhi = obs->highestbin;
    if (obs->fhiP) {
        obs->highestbin = (long long) ceil(obs->fhi * obs->T);
        if (obs->highestbin > obs->numbins - 1)
            obs->highestbin = obs->numbins - 1;
        obs->rhi = obs->highestbin;
        if (obs->highestbin < obs->rlo)
            return 4;
    } else if (obs->rhiP) {
        obs->highestbin = (long long) obs->rhi_arg;
        if (obs->highestbin > obs->numbins - 1)
            obs->highestbin = obs->numbins - 1;
        obs->rhi = obs->highestbin;
        if (obs->highestbin < obs->rlo)
            return 4;
    }
    obs->dr = O_DR;
    obs->zhi = obs->zmax;
    obs->zlo = -obs->zmax;
    for (int ii = 0; ii < obs->numharmstages; ii++) {
        if (obs->numz == 1 && obs->numw == 0)
            obs->numindep[ii] = (long long) ((obs->rhi - obs->rlo) / (1 << ii));
        else if (obs->numz > 1 && obs->numw == 0)
            obs->numindep[ii] = (long long) ((obs->rhi - obs->rlo) * (obs->numz + 1) *
                                             (obs->dz / 6.95) / (1 << ii));
        else
            obs->numindep[ii] = (long long) ((obs->rhi - obs->rlo) * (obs->numz + 1) *
                                             (obs->dz / 6.95) * (obs->numw + 1) *
                                             (obs->dw / 44.2) / (1 << ii));
        obs->powcut[ii] = (float) o_power_for_sigma(obs->sigma, 1 << ii,
                                                    (double) obs->numindep[ii]);
    }
    obs->maxkernlen = 2 * O_NUMBETWEEN * o_w_resp_halfwidth(obs->zhi, obs->whi, 0);
    obs->fftlen = o_fftlen_from_kernwidth(obs->maxkernlen);
    if (obs->fftlen < 2048)
        obs->fftlen = 2048;
    obs->corr_uselen = obs->fftlen - obs->maxkernlen;
    if (obs->corr_uselen % O_RDR)
        obs->corr_uselen = obs->corr_uselen / O_RDR * O_RDR;
    {
        long long memuse;
        if (obs->wmaxP)
            memuse = 4LL * (obs->highestbin + obs->corr_uselen) * obs->numbetween *
                     obs->numz * obs->numw;
        else
            memuse = 4LL * (obs->highestbin + obs->corr_uselen) * obs->numbetween *
                     obs->numz;
        obs->inmem = (!obs->wmaxP && (memuse < O_MAXRAMUSE || obs->inmemP)) ? 1 : 0;
        if (obs->force_standard)
            obs->inmem = 0;
    }
    return 0;
}

/* ---- template banks: src/accel_utils.c:137-288, src/corr_prep.c:92-114 ---- */

static int calc_fftlen(int numharm, int harmnum, int max_zfull, int max_wfull,
                       const o_obs *obs)
{
    double hf = (double) harmnum / (double) numharm;
    int bins_needed = (int) ceil(obs->corr_uselen * hf) + 2;
    int end_effects = 2 * O_NUMBETWEEN *
        o_w_resp_halfwidth(calc_required_z(hf, max_zfull), calc_required_w(hf, max_wfull), 0);
    return o_next_good_fftlen(bins_needed + end_effects);
}

static void init_kernel(int z, int w, int fftlen, o_kernel *k, int fp64)
{
    k->z = z;
    k->w = w;
    k->fftlen = fftlen;
    k->numbetween = O_NUMBETWEEN;
    k->kern_half_width = o_w_resp_halfwidth((double) z, (double) w, 0);
    int numkern = 2 * k->numbetween * k->kern_half_width;
    int half = numkern / 2;
    k->numgoodbins = fftlen - numkern;
    k->data = NULL;
    k->ddata = NULL;
    if (fp64) {
        o_dcomplex *resp = (o_dcomplex *) malloc(sizeof(o_dcomplex) * numkern);
        o_gen_w_response(0.0, k->numbetween, z, w, numkern, NULL, resp);
        k->ddata = (o_dcomplex *) calloc(fftlen, sizeof(o_dcomplex));
        memcpy(k->ddata, resp + half, sizeof(o_dcomplex) * half);
        memcpy(k->ddata + fftlen - half, resp, sizeof(o_dcomplex) * half);
        free(resp);
        o_cfft_d(k->ddata, fftlen, -1);
    } else {
        o_fcomplex *resp = (o_fcomplex *) malloc(sizeof(o_fcomplex) * numkern);
        o_gen_w_response(0.0, k->numbetween, z, w, numkern, resp, NULL);
        k->data = (o_fcomplex *) calloc(fftlen, sizeof(o_fcomplex));
        memcpy(k->data, resp + half, sizeof(o_fcomplex) * half);
        memcpy(k->data + fftlen - half, resp, sizeof(o_fcomplex) * half);
        free(resp);
        o_cfft_f(k->data, fftlen, -1);
    }
}

static void init_subharminfo(int numharm, int harmnum, int zmax, int wmax, o_shi *shi,
                             const o_obs *obs, int fp64)
{
    double hf = (double) harmnum / (double) numharm;
    shi->numharm = numharm;
    shi->harmnum = harmnum;
    shi->zmax = calc_required_z(hf, zmax);
    shi->wmax = calc_required_w(hf, wmax);
    shi->rinds = shi->zinds = NULL;
    if (numharm > 1) {
        int n = obs->corr_uselen > obs->numz ? obs->corr_uselen : obs->numz;
        shi->rinds = (unsigned short *) malloc(n * sizeof(unsigned short));
        shi->zinds = (unsigned short *) malloc(n * sizeof(unsigned short));
    }
    int fftlen = (numharm == 1 && harmnum == 1) ? obs->fftlen
                                                : calc_fftlen(numharm, harmnum, zmax, wmax, obs);
    shi->numkern_zdim = (shi->zmax / O_DZ) * 2 + 1;
    shi->numkern_wdim = (shi->wmax / O_DW) * 2 + 1;
    shi->numkern = shi->numkern_zdim * shi->numkern_wdim;
    shi->kern = (o_kernel *) malloc(sizeof(o_kernel) * shi->numkern);
    /* kernels are independent (the FFT plan cache is locked): threads
     * only shorten the oracle's set-up */
#pragma omp parallel for schedule(dynamic)
    for (int k = 0; k < shi->numkern; k++) {
        const int ii = k / shi->numkern_zdim, jj = k % shi->numkern_zdim;
        init_kernel(-shi->zmax + jj * O_DZ, -shi->wmax + ii * O_DW, fftlen, &shi->kern[k],
                    fp64);
    }
}

o_shi **o_create_subharminfos(const o_obs *obs, int fp64)
{
    o_shi **shis = (o_shi **) calloc(obs->numharmstages, sizeof(o_shi *));
    shis[0] = (o_shi *) calloc(2, sizeof(o_shi));
    init_subharminfo(1, 1, (int) obs->zhi, (int) obs->whi, &shis[0][0], obs, fp64);
    if (!obs->inmem) {
        for (int ii = 1; ii < obs->numharmstages; ii++) {
            int h = 1 << ii;
            shis[ii] = (o_shi *) calloc(h, sizeof(o_shi));
            for (int jj = 1; jj < h; jj += 2)
                init_subharminfo(h, jj, (int) obs->zhi, (int) obs->whi, &shis[ii][jj - 1],
                                 obs, fp64);
        }
    }
    return shis;
}

static void free_shi(o_shi *shi)
{
    for (int k = 0; k < shi->numkern; k++) {
        free(shi->kern[k].data);
        free(shi->kern[k].ddata);
    }
    free(shi->kern);
    free(shi->rinds);
    free(shi->zinds);
}

void o_free_subharminfos(const o_obs *obs, o_shi **shis)
{
    if (!obs->inmem) {
        for (int ii = 1; ii < obs->numharmstages; ii++) {
            int h = 1 << ii;
            for (int jj = 1; jj < h; jj += 2)
                free_shi(&shis[ii][jj - 1]);
            free(shis[ii]);
        }
    }
    free_shi(&shis[0][0]);
    free(shis[0]);
    free(shis);
}

/* src/accel_utils.c:993-1022: zero-padded copy of [lobin, lobin+numbins) */
static void get_fourier_amplitudes(const o_obs *obs, const o_fcomplex *fft,
                                   long long lobin, int numbins, o_fcomplex *out)
{
    for (int ii = 0; ii < numbins; ii++) {
        long long b = lobin - obs->lobin + ii;
        if (b < 0 || b >= obs->numbins) {
            out[ii].r = 0.0f;
            out[ii].i = 0.0f;
        } else {
            out[ii] = fft[b];
        }
    }
}

/* src/accel_utils.c:1080-1099: 1/sqrt(nph) with -photon, else
 * 1/sqrt(median(|x|^2)/ln 2) with FP32 powers */
double o_block_norm(const o_obs *obs, const o_fcomplex *fft, long long lobin,
                    int numdata)
{
    if (obs->nph > 0.0)
        return 1.0 / sqrt(obs->nph);
    o_fcomplex *data = (o_fcomplex *) malloc(sizeof(o_fcomplex) * numdata);
    float *powers = (float *) malloc(sizeof(float) * numdata);
    get_fourier_amplitudes(obs, fft, lobin, numdata, data);
    for (int ii = 0; ii < numdata; ii++) {
        float pr = data[ii].r, pi = data[ii].i;
        powers[ii] = pr * pr + pi * pi;
    }
    double norm = 1.0 / sqrt(o_median(powers, numdata) / log(2.0));
    free(powers);
    free(data);
    return norm;
}

void o_corr_loc_pow(const float *powers, int n, float *loc)
{
    const int khw = 10 + 5;                       /* NUMLOCPOWAVG / 2 + DELTAAVGBINS */
    const double scale = (2.0 / n) * (float) ((n / 2) / 20);
    for (int i = khw; i < n - khw; i++) {
        double s = 0.0;
        for (int j = 5 + 1; j <= khw; j++)
            s += (double) powers[i + j] + (double) powers[i - j];
        loc[i] = (float) (scale * s);
    }
    for (int i = 0; i < khw; i++) {
        loc[i] = loc[khw];
        loc[n - 1 - i] = loc[n - 1 - khw];
    }
}

/* src/accel_utils.c:1024-1198 */
o_ffdot *o_subharm_fderivs_vol(int numharm, int harmnum, double fullrlo, double fullrhi,
                               o_shi *shi, const o_obs *obs, const o_fcomplex *fft,
                               int fp64)
{
    o_ffdot *ffdot = (o_ffdot *) calloc(1, sizeof(o_ffdot));
    double hf = (double) harmnum / (double) numharm;
    double drlo = calc_required_r(hf, fullrlo);
    double drhi = calc_required_r(hf, fullrhi);
    ffdot->rlo = (long long) floor(drlo);
    ffdot->zlo = calc_required_z(hf, obs->zlo);
    ffdot->wlo = calc_required_w(hf, obs->wlo);
    if (numharm > 1 && !obs->inmem) {
        for (int ii = 0; ii < obs->corr_uselen; ii++) {
            double rr = fullrlo + ii * O_DR;
            double subr = calc_required_r(hf, rr);
            shi->rinds[ii] = (unsigned short) index_from_r(subr, ffdot->rlo);
        }
        for (int ii = 0; ii < obs->numz; ii++) {
            double zz = obs->zlo + ii * O_DZ;
            double subz = calc_required_z(hf, zz);
            shi->zinds[ii] = (unsigned short) index_from_z(subz, ffdot->zlo);
        }
    }
    ffdot->rinds = shi->rinds;
    ffdot->zinds = shi->zinds;
    ffdot->numrs = (int) ((ceil(drhi) - floor(drlo)) * O_RDR + O_DBLCORRECT) + 1;
    if (numharm == 1 && harmnum == 1) {
        ffdot->numrs = obs->corr_uselen;
    } else if (ffdot->numrs % O_RDR) {
        ffdot->numrs = (ffdot->numrs / O_RDR + 1) * O_RDR;
    }
    ffdot->numzs = shi->numkern_zdim;
    ffdot->numws = shi->numkern_wdim;

    int binoffset = shi->kern[0].kern_half_width;
    int fftlen = shi->kern[0].fftlen;
    long long lobin = ffdot->rlo - binoffset;
    int numdata = fftlen / O_NUMBETWEEN;
    int offset = binoffset * O_NUMBETWEEN;
    size_t nplane = (size_t) ffdot->numws * ffdot->numzs * ffdot->numrs;

    o_fcomplex *data = (o_fcomplex *) malloc(sizeof(o_fcomplex) * numdata);
    get_fourier_amplitudes(obs, fft, lobin, numdata, data);
    float *powers = (float *) malloc(sizeof(float) * numdata);
    for (int ii = 0; ii < numdata; ii++) {
        float pr = data[ii].r, pi = data[ii].i;
        powers[ii] = pr * pr + pi * pi;
    }
    double norm = obs->nph > 0.0 ? 1.0 / sqrt(obs->nph)
                                 : 1.0 / sqrt(o_median(powers, numdata) / log(2.0));
    /* -locpow: per-bin factors invsqrtf(loc) applied in FP32 first, the
     * scalar norm is then 1 (src/accel_utils.c:1100-1115) */
    if (obs->nph <= 0.0 && obs->norm_type == 1) {
        for (int ii = 0; ii < numdata; ii++) {
            float pr = data[ii].r, pi = data[ii].i;
            powers[ii] = pr * pr + pi * pi;
        }
        float *loc = (float *) malloc(sizeof(float) * numdata);
        o_corr_loc_pow(powers, numdata, loc);
        for (int ii = 0; ii < numdata; ii++) {
            float nl = o_invsqrtf(loc[ii]);
            data[ii].r *= nl;
            data[ii].i *= nl;
        }
        free(loc);
        norm = 1.0;
    }
    free(powers);

    if (!fp64) {
        for (int ii = 0; ii < numdata; ii++) {
            data[ii].r *= norm;
            data[ii].i *= norm;
        }
        o_fcomplex *pdata = (o_fcomplex *) calloc(fftlen, sizeof(o_fcomplex));
        for (int ii = 0; ii < numdata; ii++)
            pdata[ii * O_NUMBETWEEN] = data[ii];
        o_cfft_f(pdata, fftlen, -1);
        ffdot->powers = (float *) malloc(sizeof(float) * nplane);
        const float pnorm = 1.0 / ((double) fftlen * fftlen);
        /* templates in parallel, as the reference's OpenMP loop
         * (src/accel_utils.c:1141-1193) */
#pragma omp parallel for collapse(2) schedule(static)
        for (int ii = 0; ii < ffdot->numws; ii++) {
            for (int jj = 0; jj < ffdot->numzs; jj++) {
                o_fcomplex *tmp = (o_fcomplex *) malloc(sizeof(o_fcomplex) * fftlen);
                const o_fcomplex *k = shi->kern[ii * shi->numkern_zdim + jj].data;
                float *outp = ffdot->powers + ((size_t) ii * ffdot->numzs + jj) * ffdot->numrs;
                for (int kk = 0; kk < fftlen; kk++) {
                    const float dr = pdata[kk].r, di = pdata[kk].i;
                    const float kr = k[kk].r, ki = k[kk].i;
                    tmp[kk].r = fmaf(dr, kr, fmaf(di, ki, 0.0f));
                    tmp[kk].i = fmaf(di, kr, -fmaf(dr, ki, 0.0f));
                }
                o_cfft_f(tmp, fftlen, +1);
                for (int kk = 0; kk < ffdot->numrs; kk++) {
                    const int ind = kk + offset;
                    outp[kk] = fmaf(tmp[ind].r, tmp[ind].r,
                                    fmaf(tmp[ind].i, tmp[ind].i, 0.0f)) * pnorm;
                }
                free(tmp);
            }
        }
        free(pdata);
    } else {
        o_dcomplex *pdata = (o_dcomplex *) calloc(fftlen, sizeof(o_dcomplex));
        for (int ii = 0; ii < numdata; ii++) {
            pdata[ii * O_NUMBETWEEN].r = data[ii].r * norm;
            pdata[ii * O_NUMBETWEEN].i = data[ii].i * norm;
        }
        o_cfft_d(pdata, fftlen, -1);
        ffdot->dpowers = (double *) malloc(sizeof(double) * nplane);
        const double pnorm = 1.0 / ((double) fftlen * fftlen);
#pragma omp parallel for collapse(2) schedule(static)
        for (int ii = 0; ii < ffdot->numws; ii++) {
            for (int jj = 0; jj < ffdot->numzs; jj++) {
                o_dcomplex *tmp = (o_dcomplex *) malloc(sizeof(o_dcomplex) * fftlen);
                const o_dcomplex *k = shi->kern[ii * shi->numkern_zdim + jj].ddata;
                double *outp = ffdot->dpowers + ((size_t) ii * ffdot->numzs + jj) * ffdot->numrs;
                for (int kk = 0; kk < fftlen; kk++) {
                    tmp[kk].r = pdata[kk].r * k[kk].r + pdata[kk].i * k[kk].i;
                    tmp[kk].i = pdata[kk].i * k[kk].r - pdata[kk].r * k[kk].i;
                }
                o_cfft_d(tmp, fftlen, +1);
                for (int kk = 0; kk < ffdot->numrs; kk++) {
                    const int ind = kk + offset;
                    outp[kk] = (tmp[ind].r * tmp[ind].r + tmp[ind].i * tmp[ind].i) * pnorm;
                }
                free(tmp);
            }
        }
        free(pdata);
    }
    free(data);
    return ffdot;
}

void o_free_ffdot(o_ffdot *f)
{
    if (!f)
        return;
    free(f->powers);
    free(f->dpowers);
    free(f);
}

/* src/accel_utils.c:1261-1283 */
void o_add_ffdotpows(o_ffdot *fund, const o_ffdot *sub, int numharm, int harmnum)
{
    const double hf = (double) harmnum / (double) numharm;
    for (int ii = 0; ii < fund->numws; ii++) {
        int ww = fund->wlo + ii * O_DW;
        int subw = calc_required_w(hf, ww);
        int wind = index_from_w(subw, sub->wlo);
        for (int jj = 0; jj < fund->numzs; jj++) {
            int zind = sub->zinds[jj];
            size_t fo = ((size_t) ii * fund->numzs + jj) * fund->numrs;
            size_t so = ((size_t) wind * sub->numzs + zind) * sub->numrs;
            for (int kk = 0; kk < fund->numrs; kk++) {
                int rind = sub->rinds[kk];
                if (fund->powers)
                    fund->powers[fo + kk] += sub->powers[so + rind];
                else
                    fund->dpowers[fo + kk] += sub->dpowers[so + rind];
            }
        }
    }
}

/* src/accel_utils.c:1219-1232 */
void o_fund_to_ffdotplane(const o_ffdot *ffd, const o_obs *obs, float *plane,
                          double *dplane)
{
    long long rlen = (obs->highestbin + obs->corr_uselen) * O_RDR;
    for (int ii = 0; ii < ffd->numzs; ii++) {
        long long off = ii * rlen + ffd->rlo * O_RDR;
        if (ffd->powers)
            memcpy(plane + off, ffd->powers + (size_t) ii * ffd->numrs,
                   ffd->numrs * sizeof(float));
        else
            memcpy(dplane + off, ffd->dpowers + (size_t) ii * ffd->numrs,
                   ffd->numrs * sizeof(double));
    }
}

/* src/accel_utils.c:1312-1363 */
void o_inmem_add_ffdotpows(o_ffdot *fund, const o_obs *obs, const float *plane,
                           const double *dplane, int numharm, int harmnum)
{
    const int rlo = (int) fund->rlo;
    const int numrs = fund->numrs, numzs = fund->numzs;
    const double hf = (double) harmnum / (double) numharm;
    int *rinds = (int *) malloc(sizeof(int) * numrs);
    for (int ii = 0, rrint = O_RDR * rlo; ii < numrs; ii++, rrint++)
        rinds[ii] = (int) (rrint * hf + 0.5);
    const int zlo = fund->zlo;
    const long long rlen = (obs->highestbin + obs->corr_uselen) * O_RDR;
    for (int ii = 0; ii < numzs; ii++) {
        int zz = zlo + ii * O_DZ;
        int subz = calc_required_z(hf, zz);
        int zind = index_from_z(subz, zlo);
        long long off = zind * rlen;
        if (fund->powers) {
            float *outp = fund->powers + (size_t) ii * numrs;
            for (int jj = 0; jj < numrs; jj++)
                outp[jj] += plane[off + rinds[jj]];
        } else {
            double *outp = fund->dpowers + (size_t) ii * numrs;
            for (int jj = 0; jj < numrs; jj++)
                outp[jj] += dplane[off + rinds[jj]];
        }
    }
    free(rinds);
}

/* ---- candidate list (src/accel_utils.c:330-549) ---- */

static o_cand *new_cand(float power, float sigma, int numharm, double r, double z,
                        double w)
{
    o_cand *c = (o_cand *) malloc(sizeof(o_cand));
    c->power = power;
    c->sigma = sigma;
    c->numharm = numharm;
    c->r = r;
    c->z = z;
    c->w = w;
    c->next = NULL;
    return c;
}

static void set_cand(o_cand *c, float power, float sigma, int numharm, double r,
                     double z, double w)
{
    c->power = power;
    c->sigma = sigma;
    c->numharm = numharm;
    c->r = r;
    c->z = z;
    c->w = w;
}

/* src/accel_utils.c:387-469 */
o_cand *o_insert_new_accelcand(o_cand *list, float power, float sigma, int numharm,
                               double rr, double zz, double ww, int *added)
{
    o_cand *tmp = list, *prev = NULL;
    double prev_diff_r = O_CLOSEST_R + 1.0, next_diff_r;
    *added = 0;
    if (!list) {
        *added = 1;
        return new_cand(power, sigma, numharm, rr, zz, ww);
    }
    while (tmp->next && tmp->r < rr) {
        prev = tmp;
        tmp = tmp->next;
    }
    next_diff_r = fabs(rr - tmp->r);
    if (prev)
        prev_diff_r = fabs(rr - prev->r);
    if (prev_diff_r < O_CLOSEST_R) {
        if (prev->sigma < sigma) {
            set_cand(prev, power, sigma, numharm, rr, zz, ww);
            *added = 1;
        }
        if (next_diff_r < O_CLOSEST_R) {
            if (tmp->sigma < sigma) {
                if (*added) {
                    prev->next = tmp->next;   /* prev precedes tmp */
                    free(tmp);
                } else {
                    set_cand(tmp, power, sigma, numharm, rr, zz, ww);
                    *added = 1;
                }
            }
        }
    } else if (next_diff_r < O_CLOSEST_R) {
        if (tmp->sigma < sigma) {
            set_cand(tmp, power, sigma, numharm, rr, zz, ww);
            *added = 1;
        }
    } else {
        o_cand *nw = new_cand(power, sigma, numharm, rr, zz, ww);
        *added = 1;
        if (!tmp->next && tmp->r < (rr - O_CLOSEST_R)) {
            tmp->next = nw;
            return list;
        }
        if (prev) {
            prev->next = nw;
            nw->next = tmp;
        } else {
            nw->next = list;
            return nw;
        }
    }
    return list;
}

/* src/accel_utils.c:364-377: sigma descending, then power descending */
static int cmp_cand(const o_cand *a, const o_cand *b)
{
    int result = (a->sigma < b->sigma) - (a->sigma > b->sigma);
    if (result)
        return result;
    return (a->power < b->power) - (a->power > b->power);
}

/* stable merge sort, as g_slist_sort */
o_cand *o_sort_accelcands(o_cand *list)
{
    if (!list || !list->next)
        return list;
    o_cand *slow = list, *fast = list->next;
    while (fast && fast->next) {
        slow = slow->next;
        fast = fast->next->next;
    }
    o_cand *b = slow->next;
    slow->next = NULL;
    o_cand *a = o_sort_accelcands(list);
    b = o_sort_accelcands(b);
    o_cand head, *t = &head;
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

/* src/accel_utils.c:472-549 */
o_cand *o_eliminate_harmonics(o_cand *cands, int *numcands)
{
    const int maxharm = 16;
    const double tooclose = 1.5;
    o_cand *cur = cands;
    if (!cur)
        return cands;
    while (cur->next) {
        o_cand *other = cur->next, *before = cur;
        do {
            int rem = 0;
            const double cr = cur->r, orr = other->r;
            for (int ii = 1; ii <= maxharm; ii++)
                if (fabs(cr / ii - orr) < tooclose) {
                    rem = 1;
                    break;
                }
            if (!rem)
                for (int ii = 1; ii <= maxharm; ii++)
                    if (fabs(cr * ii - orr) < tooclose) {
                        rem = 1;
                        break;
                    }
            if (!rem &&
                ((fabs(cr * 3.0 / 2.0 - orr) < tooclose) ||
                 (fabs(cr * 5.0 / 2.0 - orr) < tooclose) ||
                 (fabs(cr * 2.0 / 3.0 - orr) < tooclose) ||
                 (fabs(cr * 4.0 / 3.0 - orr) < tooclose) ||
                 (fabs(cr * 5.0 / 3.0 - orr) < tooclose) ||
                 (fabs(cr * 3.0 / 4.0 - orr) < tooclose) ||
                 (fabs(cr * 5.0 / 4.0 - orr) < tooclose) ||
                 (fabs(cr * 2.0 / 5.0 - orr) < tooclose) ||
                 (fabs(cr * 3.0 / 5.0 - orr) < tooclose) ||
                 (fabs(cr * 4.0 / 5.0 - orr) < tooclose) ||
                 (fabs(cr * 5.0 / 6.0 - orr) < tooclose) ||
                 (fabs(cr * 2.0 / 7.0 - orr) < tooclose) ||
                 (fabs(cr * 3.0 / 7.0 - orr) < tooclose) ||
                 (fabs(cr * 4.0 / 7.0 - orr) < tooclose) ||
                 (fabs(cr * 3.0 / 8.0 - orr) < tooclose) ||
                 (fabs(cr * 5.0 / 8.0 - orr) < tooclose) ||
                 (fabs(cr * 2.0 / 9.0 - orr) < tooclose) ||
                 (fabs(cr * 3.0 / 10.0 - orr) < tooclose) ||
                 (fabs(cr * 2.0 / 11.0 - orr) < tooclose) ||
                 (fabs(cr * 3.0 / 11.0 - orr) < tooclose) ||
                 (fabs(cr * 2.0 / 13.0 - orr) < tooclose) ||
                 (fabs(cr * 3.0 / 13.0 - orr) < tooclose) ||
                 (fabs(cr * 2.0 / 15.0 - orr) < tooclose)))
                rem = 1;
            if (rem) {
                o_cand *gone = other;
                other = other->next;
                before->next = other;
                free(gone);
                *numcands = *numcands - 1;
            } else {
                before = other;
                other = other->next;
            }
        } while (other);
        if (cur->next)
            cur = cur->next;
    }
    return cands;
}

void o_free_cands(o_cand *list)
{
    while (list) {
        o_cand *n = list->next;
        free(list);
        list = n;
    }
}

/* ---- search: src/accel_utils.c:1412-1452 ---- */
static o_cand *search_ffdotpows(const o_ffdot *ffdot, int numharm, const o_obs *obs,
                                o_cand *cands, int block, o_result *res)
{
    int stage = log2i(numharm);
    float powcut = obs->powcut[stage];
    long long numindep = obs->numindep[stage];
    for (int ii = 0; ii < ffdot->numws; ii++) {
        for (int jj = 0; jj < ffdot->numzs; jj++) {
            size_t base = ((size_t) ii * ffdot->numzs + jj) * ffdot->numrs;
            for (int kk = 0; kk < ffdot->numrs; kk++) {
                double pw = ffdot->powers ? ffdot->powers[base + kk] : ffdot->dpowers[base + kk];
                if (pw <= powcut)
                    continue;
                float pow = (float) pw;
                float sig = (float) o_candidate_sigma(pw, numharm, (double) numindep);
                double rr = (ffdot->rlo + kk * (double) O_DR) / (double) numharm;
                double zz = (ffdot->zlo + jj * (double) O_DZ) / (double) numharm;
                double ww = (ffdot->wlo + ii * (double) O_DW) / (double) numharm;
                int added = 0;
                cands = o_insert_new_accelcand(cands, pow, sig, numharm, rr, zz, ww, &added);
                if (res) {
                    if (res->nhits == res->caphits) {
                        res->caphits = res->caphits ? 2 * res->caphits : 4096;
                        res->hits = (o_hit *) realloc(res->hits, sizeof(o_hit) * res->caphits);
                    }
                    o_hit *h = &res->hits[res->nhits++];
                    h->block = block;
                    h->stage = stage;
                    h->wi = ii;
                    h->zi = jj;
                    h->ri = kk;
                    h->pow = pow;
                    h->sigma = sig;
                }
            }
        }
    }
    return cands;
}

/* ---- the block loop: src/accelsearch.c:161-228, then :230-246 ---- */
int o_accelsearch(const o_obs *obs, const o_fcomplex *fft, int fp64, int max_blocks,
                  o_result *res)
{
    return o_accelsearch_range(obs, fft, fp64, 0, max_blocks, res);
}

/* The block loop over blocks [first_block, first_block + max_blocks) only
 * (max_blocks < 0: to the end).  Block numbers stay those of the whole
 * search; on the in-memory path the earlier blocks' fundamentals are still
 * written into the plane, since later sub-harmonic sums read them. */
int o_accelsearch_range(const o_obs *obs, const o_fcomplex *fft, int fp64, int first_block,
                        int max_blocks, o_result *res)
{
    memset(res, 0, sizeof(*res));
    o_shi **shis = o_create_subharminfos(obs, fp64);
    int rstep = (int) (obs->corr_uselen * O_DR);
    float *plane = NULL;
    double *dplane = NULL;
    o_cand *cands = NULL;
    if (obs->inmem) {
        size_t n = (size_t) (obs->highestbin + obs->corr_uselen) * O_NUMBETWEEN * obs->numz;
        if (fp64)
            dplane = (double *) calloc(n, sizeof(double));
        else
            plane = (float *) calloc(n, sizeof(float));
        double startr = 8;
        while (startr < obs->rlo) {
            double nextr = startr + rstep, lastr = nextr - O_DR;
            o_ffdot *f = o_subharm_fderivs_vol(1, 1, startr, lastr, &shis[0][0], obs, fft, fp64);
            o_fund_to_ffdotplane(f, obs, plane, dplane);
            o_free_ffdot(f);
            startr = nextr;
        }
    }
    double startr = obs->rlo;
    int block = 0;
    while (startr + rstep < obs->highestbin) {
        if (max_blocks >= 0 && block >= first_block + max_blocks)
            break;
        double nextr = startr + rstep, lastr = nextr - O_DR;
        if (block < first_block) {
            if (obs->inmem && obs->numharmstages > 1) {
                o_ffdot *f = o_subharm_fderivs_vol(1, 1, startr, lastr, &shis[0][0], obs, fft,
                                                   fp64);
                o_fund_to_ffdotplane(f, obs, plane, dplane);
                o_free_ffdot(f);
            }
            startr = nextr;
            block++;
            continue;
        }
        o_ffdot *fund = o_subharm_fderivs_vol(1, 1, startr, lastr, &shis[0][0], obs, fft, fp64);
        cands = search_ffdotpows(fund, 1, obs, cands, block, res);
        if (obs->numharmstages > 1) {
            if (obs->inmem)
                o_fund_to_ffdotplane(fund, obs, plane, dplane);
            for (int stage = 1; stage < obs->numharmstages; stage++) {
                int h = 1 << stage;
                for (int harm = 1; harm < h; harm += 2) {
                    if (obs->inmem) {
                        o_inmem_add_ffdotpows(fund, obs, plane, dplane, h, harm);
                    } else {
                        o_ffdot *sub = o_subharm_fderivs_vol(h, harm, startr, lastr,
                                                             &shis[stage][harm - 1], obs,
                                                             fft, fp64);
                        o_add_ffdotpows(fund, sub, h, harm);
                        o_free_ffdot(sub);
                    }
                }
                cands = search_ffdotpows(fund, h, obs, cands, block, res);
            }
        }
        o_free_ffdot(fund);
        startr = nextr;
        block++;
    }
    res->nblocks = block;
    free(plane);
    free(dplane);
    o_free_subharminfos(obs, shis);
    if (cands) {
        int numcands = 0;
        for (o_cand *c = cands; c; c = c->next)
            numcands++;
        cands = o_sort_accelcands(cands);
        if (obs->numharm > 1)
            cands = o_eliminate_harmonics(cands, &numcands);
    }
    res->cands = cands;
    return 0;
}

void o_free_result(o_result *res)
{
    free(res->hits);
    o_free_cands(res->cands);
    memset(res, 0, sizeof(*res));
}

int o_result_ncands(const o_result *res)
{
    int n = 0;
    for (o_cand *c = res->cands; c; c = c->next)
        n++;
    return n;
}

void o_result_cands(const o_result *res, double *out7)
{
    int i = 0;
    for (o_cand *c = res->cands; c; c = c->next, i++) {
        out7[7 * i + 0] = c->power;
        out7[7 * i + 1] = c->sigma;
        out7[7 * i + 2] = c->numharm;
        out7[7 * i + 3] = c->r;
        out7[7 * i + 4] = c->z;
        out7[7 * i + 5] = c->w;
        out7[7 * i + 6] = 0.0;
    }
}

/* .dat input: src/accel_utils.c:1507-1564 (realfft, then deredden with the
 * window growing to 200 bins at Nyquist) and :1621-1624 (bin 0 := 1+1i).
 * Works in place on n floats; returns the number of complex bins. */
long long o_prepare_dat(float *series, long long n, double dt)
{
    o_realfft_f(series, (long) n, -1);
    o_fcomplex *fft = (o_fcomplex *) series;
    long long numbins = n / 2;
    o_deredden(fft, (long) numbins, 6, 200, 0.5 / dt, dt * n);
    fft[0].r = 1.0f;
    fft[0].i = 1.0f;
    return numbins;
}

/* threads of the template loop (the accelsearch -ncpus option) */
void o_set_threads(int n)
{
    omp_set_num_threads(n > 0 ? n : 1);
}
