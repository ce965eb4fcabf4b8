/*
 * o_refine.c -- CPU restatement of PRESTO's candidate refinement
 * (optimize_accelcand and what it calls).  TEST INFRASTRUCTURE ONLY: the
 * checker for the HIP refinement kernels; no product code uses it.
 *
 *   o_rz_interp / o_rzw_interp    src/rzinterp.c:144-237, src/rzwinterp.c:145-230
 *   rzw_vol_row                   src/rzwinterp.c:3-142 + src/corr_routines.c:5-272
 *                                 (the correlation those compute, as a direct sum)
 *   o_localpower3d, o_derivs3d    src/characteristics.c:79-192
 *   o_calc_props, o_calc_rzwerrs  src/characteristics.c:195-396
 *   o_nm_minimize                 GSL 2.7 multimin "nmsimplex2" (absent here;
 *                                 restated from its published algorithm)
 *   o_max_rz_arr[_harmonics]      src/maximize_rz.c:103-228
 *   o_max_rzw_arr_harmonics       src/maximize_rzw.c:84-191
 *   o_max_rzw_arr                 src/maximize_rzw.c:3-55 (-noharmpolish)
 *
 * Arithmetic follows the reference: FP32 spectrum and FP32 response taps,
 * products accumulated in FP64 and stored to an FP32 complex after every
 * term (ans->r += ... with ans an fcomplex, src/rzinterp.c:228-233).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define ZSCALE 4.0
#define PI 3.14159265358979323846
#define TWOPI 6.28318530717958647692

static double radian_phase(double r, double i)
{
    /* RADIAN_PHASE, include/presto.h:88-90 */
    double t = atan2(i, r);
    return t > 0.0 ? t : t + TWOPI;
}

/* sum data[k] * conj(resp[k - lodata + loresp]) the way rz_interp does */
static void direct_sum(const o_fcomplex *data, long numdata, long intfreq, int khw,
                       const o_fcomplex *resp, float *ar, float *ai)
{
    int numkern = 2 * khw;
    long lodata = intfreq - khw, loresp, hiresp;
    if (lodata < 0) {
        loresp = labs(lodata);
        lodata = 0;
    } else {
        loresp = 0;
    }
    long hidata = intfreq + khw - 1;
    if (hidata > numdata - 1)
        hiresp = numkern - hidata + numdata - 1;
    else
        hiresp = numkern;
    long nsum = hiresp - loresp;
    float rr = 0.0f, ii = 0.0f;
    for (long k = 0; k < nsum; k++) {
        const o_fcomplex d = data[lodata + k], q = resp[loresp + k];
        rr += (double) d.r * (double) q.r + (double) (d.i * q.i);
        ii += (double) d.i * (double) q.r - (double) q.i * (double) d.r;
    }
    *ar = rr;
    *ai = ii;
}

void o_rz_interp(const o_fcomplex *data, long numdata, double r, double z, int khw,
                 float *ar, float *ai)
{
    if (r > numdata - 1.0 || r < 0.0) {
        *ar = *ai = 0.0f;
        return;
    }
    double dint;
    double frac = modf(r, &dint);
    long intfreq = (long) dint;
    if (fabs(z) < 1E-4) {
        if (frac < 1E-5) {
            *ar = data[intfreq].r;
            *ai = data[intfreq].i;
            return;
        }
        if ((1.0 - frac) < 1E-5) {
            *ar = data[intfreq + 1].r;
            *ai = data[intfreq + 1].i;
            return;
        }
    }
    int numkern = 2 * khw;
    o_dcomplex *d = (o_dcomplex *) malloc(sizeof(o_dcomplex) * numkern);
    o_fcomplex *resp = (o_fcomplex *) malloc(sizeof(o_fcomplex) * numkern);
    o_gen_z_response(frac, 1, z, numkern, d);
    for (int k = 0; k < numkern; k++) {
        resp[k].r = (float) d[k].r;
        resp[k].i = (float) d[k].i;
    }
    direct_sum(data, numdata, intfreq, khw, resp, ar, ai);
    free(d);
    free(resp);
}

void o_rzw_interp(const o_fcomplex *data, long numdata, double r, double z, double w,
                  int khw, float *ar, float *ai)
{
    if (r > numdata - 1.0 || r < 0.0) {
        *ar = *ai = 0.0f;
        return;
    }
    if (fabs(w) < 1E-4) {
        o_rz_interp(data, numdata, r, z, khw, ar, ai);
        return;
    }
    double dint;
    double frac = modf(r, &dint);
    long intfreq = (long) dint;
    int numkern = 2 * khw;
    o_fcomplex *resp = (o_fcomplex *) malloc(sizeof(o_fcomplex) * numkern);
    o_gen_w_response(frac, 1, z, w, numkern, resp, NULL);
    direct_sum(data, numdata, intfreq, khw, resp, ar, ai);
    free(resp);
}

double o_localpower3d(const o_fcomplex *data, long numdata, double r, double z, double w)
{
    int binsperside = O_NUMLOCPOWAVG / 2;
    int khw = o_w_resp_halfwidth(z, w, 0);
    double lo1 = r - O_DELTAAVGBINS - binsperside, hi1 = lo1 + binsperside;
    double lo2 = r + O_DELTAAVGBINS, hi2 = lo2 + binsperside;
    double *b[4] = {&lo1, &lo2, &hi1, &hi2};
    for (int k = 0; k < 4; k++) {
        if (*b[k] < 0.0)
            *b[k] = 0.0;
        if (*b[k] > numdata)
            *b[k] = (double) numdata;
    }
    double sum = 0.0;
    float ar, ai;
    for (double f = lo1; f < hi1; f += 1.0) {
        o_rzw_interp(data, numdata, f, z, w, khw, &ar, &ai);
        sum += (double) ar * ar + (double) ai * ai;
    }
    for (double f = lo2; f < hi2; f += 1.0) {
        o_rzw_interp(data, numdata, f, z, w, khw, &ar, &ai);
        sum += (double) ar * ar + (double) ai * ai;
    }
    return sum / (double) O_NUMLOCPOWAVG;
}

void o_derivs3d(const o_fcomplex *data, long numdata, double r, double z, double w,
                double locpow, o_rderivs *res)
{
    const double h = 0.003, twoh = 0.006, twohsqrd = 0.000036;
    double pwr[5], phs[5];
    int khw = o_w_resp_halfwidth(z, w, 1);
    double f = r - twoh;
    for (int k = 0; k < 5; k++, f += h) {
        float ar, ai;
        o_rzw_interp(data, numdata, f, z, w, khw, &ar, &ai);
        pwr[k] = (double) ar * ar + (double) ai * ai;
        phs[k] = radian_phase(ar, ai);
    }
    for (int k = 0; k < 4; k++)
        if (fabs(phs[k + 1] - phs[k]) > PI)
            phs[k + 1] -= TWOPI;
    res->pow = pwr[2];
    res->phs = phs[2];
    res->dpow = (pwr[3] - pwr[1]) / twoh;
    res->dphs = (phs[3] - phs[1]) / twoh;
    res->d2pow = (pwr[4] - 2.0 * pwr[2] + pwr[0]) / twohsqrd;
    res->d2phs = (phs[4] - 2.0 * phs[2] + phs[0]) / twohsqrd;
    res->locpow = locpow;
}

int o_calc_props(const o_rderivs *d, double r, double z, double w, o_fourierprops *p)
{
    if (fabs(d->pow) < O_DBLCORRECT || d->locpow < O_DBLCORRECT)
        return -1;
    double tmppow = d->pow / d->locpow, tmpd2pow = d->d2pow / d->locpow;
    p->pur = sqrt(1.5 * fabs(tmpd2pow) / tmppow) / PI;
    if (fabs(tmppow) < O_DBLCORRECT)
        p->pur = 1.0;
    p->purerr = 1.0 / (p->pur * sqrt(10.0 * tmppow));
    p->r = r;
    p->rerr = 3.0 / (PI * p->pur * sqrt(6.0 * tmppow));
    p->z = z;
    p->zerr = 3.0 * sqrt(10.0) / (PI * p->pur * p->pur * sqrt(tmppow));
    p->w = w;
    p->werr = 6.0 * sqrt(105.0) / (PI * p->pur * p->pur * p->pur * sqrt(tmppow));
    p->pow = tmppow;
    p->powerr = sqrt(2.0 * tmppow);
    p->sig = sqrt(2.0 * tmppow - log(PI * tmppow));
    p->rawpow = d->pow;
    p->phs = d->phs;
    p->phserr = 1.0 / p->powerr;
    p->cen = -d->dphs / TWOPI;
    p->cenerr = 1.0 / sqrt(24.0 * tmppow);
    p->locpow = d->locpow;
    return 0;
}

int o_calc_rzwerrs(const o_fourierprops *p, double T, o_rzwerrs *e)
{
    if (T <= 0.0)
        return -1;
    double T2 = T * T, T3 = T2 * T;
    double r2 = p->r * p->r, sr2 = (double) (p->rerr * p->rerr), r4 = r2 * r2;
    double z2 = p->z * p->z, sz2 = (double) (p->zerr * p->zerr);
    double tmp = p->r * p->w - 3 * z2;
    e->f = p->r / T;
    e->ferr = p->rerr / T;
    e->fd = p->z / T2;
    e->fderr = p->zerr / T2;
    e->fdd = p->w / T3;
    e->fdderr = p->werr / T3;
    e->p = T / p->r;
    e->perr = T * p->rerr / r2;
    e->pd = -p->z / r2;
    e->pderr = sqrt(4 * z2 * sr2 / (r4 * r2) + sz2 / r4);
    if (p->w == 0.0)
        e->pdd = 0.0;
    else
        e->pdd = (2 * z2 - p->r * p->w) / (T * r2 * p->r);
    e->pdderr = sqrt(((double) (p->werr * p->werr) * r4 + 16 * sz2 * r2 + 4 * tmp * tmp * sr2) /
                     (r4 * r4 * T2));
    return 0;
}

/* ------------------------------------------------------------------ */
/* Nelder-Mead, restating GSL's nmsimplex2 for n = 2 (P = 3 corners):   */
/* reflection / expansion / 1-D contraction of the worst corner, else   */
/* shrink towards the best; centre and squared rms size S2 updated      */
/* incrementally, converged when sqrt(S2) < tol.                        */
/* ------------------------------------------------------------------ */
typedef double (*o_nmfun)(const double *x, void *par);

typedef struct {
    double x[3][2], y[3], c[2], S2;
} nm_state;

static void nm_center_size(nm_state *s)
{
    for (int j = 0; j < 2; j++)
        s->c[j] = (s->x[0][j] + s->x[1][j] + s->x[2][j]) / 3.0;
    double ss = 0.0;
    for (int i = 0; i < 3; i++) {
        double t = 0.0;
        for (int j = 0; j < 2; j++)
            t += (s->x[i][j] - s->c[j]) * (s->x[i][j] - s->c[j]);
        ss += t;
    }
    s->S2 = ss / 3.0;
}

static double nm_try(nm_state *s, double coeff, int corner, double *xc, o_nmfun f, void *par)
{
    const double P = 3.0;
    double alpha = (1 - coeff) * P / (P - 1.0), beta = (P * coeff - 1.0) / (P - 1.0);
    for (int j = 0; j < 2; j++)
        xc[j] = alpha * s->c[j] + beta * s->x[corner][j];
    return f(xc, par);
}

static void nm_update(nm_state *s, int i, const double *x, double val)
{
    const double P = 3.0;
    double d2 = 0.0, xmcd = 0.0;
    for (int j = 0; j < 2; j++) {
        double dl = x[j] - s->x[i][j];
        d2 += dl * dl;
        xmcd += (s->x[i][j] - s->c[j]) * dl;
    }
    s->S2 += (2.0 / P) * xmcd + ((P - 1.0) / P) * (d2 / P);
    for (int j = 0; j < 2; j++) {
        s->c[j] += -s->x[i][j] / P;
        s->c[j] += x[j] / P;
        s->x[i][j] = x[j];
    }
    s->y[i] = val;
}

/* returns the minimum; x in/out; iterations in *iters */
double o_nm_minimize(o_nmfun f, void *par, double *x, double step, double tol, int maxiter,
                     int *iters)
{
    nm_state s;
    s.x[0][0] = x[0];
    s.x[0][1] = x[1];
    s.y[0] = f(x, par);
    for (int i = 0; i < 2; i++) {
        double xt[2] = {x[0], x[1]};
        xt[i] += step;
        s.x[i + 1][0] = xt[0];
        s.x[i + 1][1] = xt[1];
        s.y[i + 1] = f(xt, par);
    }
    nm_center_size(&s);
    int it = 0;
    double size;
    int lo = 0;
    do {
        it++;
        double dhi = s.y[0], dlo = s.y[0], ds_hi = s.y[1];
        int hi = 0;
        lo = 0;
        for (int i = 1; i < 3; i++) {
            double v = s.y[i];
            if (v < dlo) {
                dlo = v;
                lo = i;
            } else if (v > dhi) {
                ds_hi = dhi;
                dhi = v;
                hi = i;
            } else if (v > ds_hi) {
                ds_hi = v;
            }
        }
        double xc[2], xc2[2];
        double val = nm_try(&s, -1.0, hi, xc, f, par);
        if (isfinite(val) && val < s.y[lo]) {
            double val2 = nm_try(&s, -2.0, hi, xc2, f, par);
            if (isfinite(val2) && val2 < s.y[lo])
                nm_update(&s, hi, xc2, val2);
            else
                nm_update(&s, hi, xc, val);
        } else if (!isfinite(val) || val > ds_hi) {
            if (isfinite(val) && val <= s.y[hi])
                nm_update(&s, hi, xc, val);
            double val2 = nm_try(&s, 0.5, hi, xc2, f, par);
            if (isfinite(val2) && val2 <= s.y[hi]) {
                nm_update(&s, hi, xc2, val2);
            } else {
                for (int i = 0; i < 3; i++) {
                    if (i == lo)
                        continue;
                    for (int j = 0; j < 2; j++)
                        s.x[i][j] = 0.5 * (s.x[i][j] + s.x[lo][j]);
                    s.y[i] = f(s.x[i], par);
                }
                nm_center_size(&s);
            }
        } else {
            nm_update(&s, hi, xc, val);
        }
        lo = 0;
        for (int i = 1; i < 3; i++)
            if (s.y[i] < s.y[lo])
                lo = i;
        if (s.S2 > 0)
            size = sqrt(s.S2);
        else {
            nm_center_size(&s);
            size = sqrt(s.S2);
        }
    } while (!(size < tol) && it < maxiter);
    x[0] = s.x[lo][0];
    x[1] = s.x[lo][1];
    if (iters)
        *iters = it;
    return s.y[lo];
}

/* ------------------------------------------------------------------ */
/* max_rz_arr / max_rz_arr_harmonics, src/maximize_rz.c:20-228          */
/* ------------------------------------------------------------------ */
typedef struct {
    const o_fcomplex *data;
    long numdata;
    const float *locpows;
    int numharm, kernhw;
    long nevals;
} rz_par;

static double rz_one(double r, double zs, rz_par *p)
{
    float ar, ai;
    o_rz_interp(p->data, p->numdata, r, zs * ZSCALE, p->kernhw, &ar, &ai);
    return -((double) ar * ar + (double) ai * ai);
}

static double rz_fun1(const double *x, void *vp)
{
    rz_par *p = (rz_par *) vp;
    p->nevals++;
    return rz_one(x[0], x[1], p);
}

static double rz_funh(const double *x, void *vp)
{
    rz_par *p = (rz_par *) vp;
    p->nevals++;
    double tot = 0.0;
    for (int k = 0; k < p->numharm; k++) {
        int n = k + 1;
        float ar, ai;
        o_rz_interp(p->data, p->numdata, x[0] * n, x[1] * ZSCALE * n, p->kernhw, &ar, &ai);
        tot += ((double) ar * ar + (double) ai * ai) / p->locpows[k];
    }
    if (!isfinite(tot))
        return 0.0;
    return -tot;
}

double o_max_rz_arr(const o_fcomplex *data, long numdata, double rin, double zin,
                    double *rout, double *zout, o_rderivs *derivs)
{
    rz_par par = {data, numdata, NULL, 1, 0, 0};
    double x[2] = {rin, zin / ZSCALE};
    par.kernhw = o_z_resp_halfwidth(fabs(zin) + 4.0, 0);
    o_nm_minimize(rz_fun1, &par, x, 0.4, 1.0e-4, 5000, NULL);
    par.kernhw = o_z_resp_halfwidth(fabs(x[1]) + 4.0, 1);
    double fval = o_nm_minimize(rz_fun1, &par, x, 0.01, 1.0e-7, 5000, NULL);
    *rout = x[0];
    *zout = x[1] * ZSCALE;
    double lp = o_localpower3d(data, numdata, *rout, *zout, 0.0);
    o_derivs3d(data, numdata, *rout, *zout, 0.0, lp, derivs);
    return -fval;
}

long o_max_rz_arr_harmonics(const o_fcomplex *data, long numdata, int numharm, double rin,
                            double zin, double *rout, double *zout, o_rderivs *derivs,
                            double *powers)
{
    float *locpow = (float *) malloc(sizeof(float) * numharm);
    for (int k = 0; k < numharm; k++) {
        int n = k + 1;
        locpow[k] = (float) o_localpower3d(data, numdata, rin * n, zin * n, 0.0);
    }
    rz_par par = {data, numdata, locpow, numharm, 0, 0};
    double x[2] = {rin, zin / ZSCALE};
    par.kernhw = o_z_resp_halfwidth(fabs(zin * numharm) + 4.0, 0);
    o_nm_minimize(rz_funh, &par, x, 0.4, 1.0e-4, 5000, NULL);
    par.kernhw = o_z_resp_halfwidth(fabs(x[1] * numharm) + 4.0, 1);
    o_nm_minimize(rz_funh, &par, x, 0.01, 1.0e-7, 5000, NULL);
    *rout = x[0];
    *zout = x[1] * ZSCALE;
    for (int k = 0; k < numharm; k++) {
        int n = k + 1;
        locpow[k] = (float) o_localpower3d(data, numdata, *rout * n, *zout * n, 0.0);
        powers[k] = -rz_one(*rout * n, *zout / ZSCALE * n, &par);
        o_derivs3d(data, numdata, *rout * n, *zout * n, 0.0, locpow[k], &derivs[k]);
    }
    free(locpow);
    return par.nevals;
}

/* ------------------------------------------------------------------ */
/* corr_rzw_vol as the direct correlation it computes:                  */
/* vol(z, w, j) = sum_k data[k] conj(R_{z,w}[numkern/2 + (k-lobin) nb - j]) */
/* with R the FP32 numbetween-sampled response of gen_w_response        */
/* (src/rzwinterp.c:88-100, src/corr_routines.c:146-258); FP64 sums.    */
/* ------------------------------------------------------------------ */
static void rzw_vol_row(const o_fcomplex *data, long numdata, int nb, long lobin, double z,
                        double w, int khw, int jlo, int nout, double *pw)
{
    int numkern = 2 * nb * khw;
    o_fcomplex *R = (o_fcomplex *) malloc(sizeof(o_fcomplex) * numkern);
    o_gen_w_response(0.0, nb, z, w, numkern, R, NULL);
    for (int jj = 0; jj < nout; jj++) {
        const int j = jlo + jj;
        double sr = 0.0, si = 0.0;
        /* indices into R: numkern/2 + (k - lobin) nb - j in [0, numkern) */
        long kmin = lobin + (long) floor((j - numkern / 2) / (double) nb);
        long kmax = lobin + (long) ceil((j + numkern / 2) / (double) nb);
        for (long k = kmin; k <= kmax; k++) {
            long idx = numkern / 2 + (k - lobin) * nb - j;
            if (idx < 0 || idx >= numkern || k < 0 || k >= numdata)
                continue;
            const o_fcomplex d = data[k], q = R[idx];
            sr += (double) d.r * q.r + (double) d.i * q.i;
            si += (double) d.i * q.r - (double) d.r * q.i;
        }
        float ar = (float) sr, ai = (float) si;
        pw[jj] = (double) ar * ar + (double) ai * ai;
    }
    free(R);
}

void o_max_rzw_arr_harmonics(const o_fcomplex *data, long numdata, int numharm, double rin,
                             double zin, double win, double *rout, double *zout, double *wout,
                             o_rderivs *derivs, double *powers)
{
    const int interpfac = 4, numwid = 2, nb = interpfac * O_RDR;
    const int numr = 3 * nb, numz = 2 * numwid * interpfac + 1, numw = numz;
    const double dr = 1.0 / (double) nb;
    float *powsum = (float *) calloc((size_t) numw * numz * numr, sizeof(float));
    long hhlobin = 0;
    for (int hind = 0; hind < numharm; hind++) {
        int n = numharm - hind;
        double rh = rin * n, zh = zin * n, wh = win * n;
        int khw = o_w_resp_halfwidth(fabs(zh) + 2, fabs(wh) + 20, 0);
        double hfrac = n / (double) numharm;
        float locpow = (float) o_localpower3d(data, numdata, rh, zh, wh);
        double rh_int;
        modf(rh, &rh_int);
        long lobin = (long) rh_int - 1;
        if (hind == 0)
            hhlobin = lobin;
        double zlo = zh - numwid * hfrac * O_DZ, zhi = zh + numwid * hfrac * O_DZ;
        double wlo = wh - numwid * hfrac * O_DW, whi = wh + numwid * hfrac * O_DW;
        if (zlo > zhi) {
            double t = zlo;
            zlo = zhi;
            zhi = t;
        }
        if (wlo > whi) {
            double t = wlo;
            wlo = whi;
            whi = t;
        }
        double dz = (zhi - zlo) / (numz - 1), dw = (whi - wlo) / (numw - 1);
        double maxz = (fabs(zlo) < fabs(zhi)) ? zhi : zlo;
        double maxw = (fabs(wlo) < fabs(whi)) ? whi : wlo;
        int vkhw = o_w_resp_halfwidth(maxz, maxw, 0);
        (void) khw;   /* sizes the reference's FFT only */
        /* the r positions of the harmonic grid this harmonic feeds; a lower
         * harmonic's grid can start below lobin, where the reference indexes
         * before the row it computed (src/maximize_rzw.c:148-150 with
         * rind < 0): the correlation at that r is used instead */
        int maxrind = -1000000, minrind = 1000000;
        int *rind = (int *) malloc(sizeof(int) * numr);
        for (int kk = 0; kk < numr; kk++) {
            rind[kk] = (int) (long) round(((hhlobin + kk * dr) * hfrac - lobin) * nb);
            if (rind[kk] > maxrind)
                maxrind = rind[kk];
            if (rind[kk] < minrind)
                minrind = rind[kk];
        }
        double *row = (double *) malloc(sizeof(double) * (maxrind - minrind + 1));
        for (int ii = 0; ii < numw; ii++) {
            for (int jj = 0; jj < numz; jj++) {
                rzw_vol_row(data, numdata, nb, lobin, zlo + jj * dz, wlo + ii * dw, vkhw,
                            minrind, maxrind - minrind + 1, row);
                for (int kk = 0; kk < numr; kk++) {
                    float *ps = &powsum[((size_t) ii * numz + jj) * numr + kk];
                    *ps += row[rind[kk] - minrind] / locpow;
                }
            }
        }
        free(row);
        free(rind);
    }
    float maxpow = 0.0f;
    int rind = 0, zind = 0, wind = 0;
    for (int ii = 0; ii < numw; ii++)
        for (int jj = 0; jj < numz; jj++)
            for (int kk = 0; kk < numr; kk++) {
                float p = powsum[((size_t) ii * numz + jj) * numr + kk];
                if (p > maxpow) {
                    maxpow = p;
                    wind = ii;
                    zind = jj;
                    rind = kk;
                }
            }
    double nh = numharm;
    *rout = (hhlobin + rind * dr) / nh;
    *zout = ((zin * nh) - numwid * O_DZ + zind * O_DZ / (double) interpfac) / nh;
    *wout = ((win * nh) - numwid * O_DW + wind * O_DW / (double) interpfac) / nh;
    for (int k = 0; k < numharm; k++) {
        int hh = k + 1;
        double rr = *rout * hh, zz = *zout * hh, ww = *wout * hh;
        float lp = (float) o_localpower3d(data, numdata, rr, zz, ww);
        o_derivs3d(data, numdata, rr, zz, ww, lp, &derivs[k]);
        powers[k] = derivs[k].pow;
    }
    free(powsum);
}

/* src/maximize_rzw.c:3-55: one harmonic, no polishing.  The 9 x 9 x 24
 * grid of r from (int) rin - 1 in steps of 1/8, z = zin +- DZ and
 * w = win +- DW in quarter steps; the first maximum of the raw power in
 * (w, z, r) order; derivatives at it with the local power of the start. */
double o_max_rzw_arr(const o_fcomplex *data, long numdata, double rin, double zin, double win,
                     double *rout, double *zout, double *wout, o_rderivs *derivs)
{
    const int interpfac = 4, nb = interpfac * O_RDR;
    const int numr = 3 * interpfac * O_RDR, numz = 2 * interpfac + 1, numw = numz;
    const float locpow = (float) o_localpower3d(data, numdata, rin, zin, win);
    const long startbin = (long) (int) rin - 1;
    const double zlo = zin - O_DZ, zhi = zin + O_DZ, wlo = win - O_DW, whi = win + O_DW;
    const double dz = (zhi - zlo) / (numz - 1), dw = (whi - wlo) / (numw - 1);
    const double maxz = (fabs(zlo) < fabs(zhi)) ? zhi : zlo;
    const double maxw = (fabs(wlo) < fabs(whi)) ? whi : wlo;
    const int vkhw = o_w_resp_halfwidth(maxz, maxw, 0);
    double *row = (double *) malloc(sizeof(double) * numr);
    float maxpow = 0.0f;
    int rind = 0, zind = 0, wind = 0;
    for (int ii = 0; ii < numw; ii++)
        for (int jj = 0; jj < numz; jj++) {
            rzw_vol_row(data, numdata, nb, startbin, zlo + jj * dz, wlo + ii * dw, vkhw, 0, numr,
                        row);
            for (int kk = 0; kk < numr; kk++) {
                const float p = (float) row[kk];
                if (p > maxpow) {
                    maxpow = p;
                    wind = ii;
                    zind = jj;
                    rind = kk;
                }
            }
        }
    free(row);
    *rout = startbin + rind * O_DR / (double) interpfac;
    *zout = zin - O_DZ + zind * O_DZ / (double) interpfac;
    *wout = win - O_DW + wind * O_DW / (double) interpfac;
    o_derivs3d(data, numdata, *rout, *zout, *wout, locpow, derivs);
    return maxpow;
}
