/*
 * o_math.c -- oracle: Fresnel integrals, response templates, kernel
 * widths, FFT-length policy, significance functions, median, invsqrtf.
 * Test infrastructure only (see oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define O_PI 3.1415926535897932384626433832795028841971693993751
#define O_TWOPI 6.2831853071795864769252867665590057683943387987502
#define O_SQRT2 1.4142135623730950488016887242096980785696718753769

/* ------------------------------------------------------------------ */
/* Fresnel integrals S(x), C(x) with the published Cephes rational     */
/* approximations (Moshier, Cephes Math Library 2.1).  The coefficient */
/* tables are the published ones; src/fresnl.c:76-229 uses the same.   */
/* ------------------------------------------------------------------ */
static const double FR_SN[6] = {
    -2.99181919401019853726E3, 7.08840045257738576863E5,
    -6.29741486205862506537E7, 2.54890880573376359104E9,
    -4.42979518059697779103E10, 3.18016297876567817986E11};
static const double FR_SD[6] = {
    2.81376268889994315696E2, 4.55847810806532581675E4,
    5.17343888770096400730E6, 4.19320245898111231129E8,
    2.24411795645340920940E10, 6.07366389490084639049E11};
static const double FR_CN[6] = {
    -4.98843114573573548651E-8, 9.50428062829859605134E-6,
    -6.45191435683965050962E-4, 1.88843319396703850064E-2,
    -2.05525900955013891793E-1, 9.99999999999999998822E-1};
static const double FR_CD[7] = {
    3.99982968972495980367E-12, 9.15439215774657478799E-10,
    1.25001862479598821474E-7, 1.22262789024179030997E-5,
    8.68029542941784300606E-4, 4.12142090722199792936E-2,
    1.00000000000000000118E0};
static const double FR_FN[10] = {
    4.21543555043677546506E-1, 1.43407919780758885261E-1,
    1.15220955073585758835E-2, 3.45017939782574027900E-4,
    4.63613749287867322088E-6, 3.05568983790257605827E-8,
    1.02304514164907233465E-10, 1.72010743268161828879E-13,
    1.34283276233062758925E-16, 3.76329711269987889006E-20};
static const double FR_FD[10] = {
    7.51586398353378947175E-1, 1.16888925859191382142E-1,
    6.44051526508858611005E-3, 1.55934409164153020873E-4,
    1.84627567348930545870E-6, 1.12699224763999035261E-8,
    3.60140029589371370404E-11, 5.88754533621578410010E-14,
    4.52001434074129701496E-17, 1.25443237090011264384E-20};
static const double FR_GN[11] = {
    5.04442073643383265887E-1, 1.97102833525523411709E-1,
    1.87648584092575249293E-2, 6.84079380915393090172E-4,
    1.15138826111884280931E-5, 9.82852443688422223854E-8,
    4.45344415861750144738E-10, 1.08268041139020870318E-12,
    1.37555460633261799868E-15, 8.36354435630677421531E-19,
    1.86958710162783235106E-22};
static const double FR_GD[11] = {
    1.47495759925128324529E0, 3.37748989120019970451E-1,
    2.53603741420338795122E-2, 8.14679107184306179049E-4,
    1.27545075667729118702E-5, 1.04314589657571990585E-7,
    4.60680728146520428211E-10, 1.10273215066240270757E-12,
    1.38796531259578871258E-15, 8.39158816283118707363E-19,
    1.86958710162783236342E-22};

/* Horner with leading coefficient c[0]; degree n */
static double horner(double x, const double *c, int n)
{
    double a = c[0];
    for (int k = 1; k <= n; k++)
        a = a * x + c[k];
    return a;
}

/* Horner for a monic polynomial: x^n + c[0] x^(n-1) + ... + c[n-1] */
static double horner_monic(double x, const double *c, int n)
{
    double a = x + c[0];
    for (int k = 1; k < n; k++)
        a = a * x + c[k];
    return a;
}

/* src/fresnl.c:365-406 */
int o_fresnl(double xa, double *s_out, double *c_out)
{
    double x = fabs(xa), x2 = x * x, cc, ss;
    if (x2 < 2.5625) {
        double t = x2 * x2;
        ss = x * x2 * horner(t, FR_SN, 5) / horner_monic(t, FR_SD, 6);
        cc = x * horner(t, FR_CN, 5) / horner(t, FR_CD, 6);
    } else if (x > 36974.0) {
        cc = 0.5;
        ss = 0.5;
    } else {
        double t = O_PI * x2;
        double u = 1.0 / (t * t);
        t = 1.0 / t;
        double f = 1.0 - u * horner(u, FR_FN, 9) / horner_monic(u, FR_FD, 10);
        double g = t * horner(u, FR_GN, 10) / horner_monic(u, FR_GD, 11);
        double ang = 0.5 * O_PI * x2;
        double c = cos(ang), s = sin(ang);
        double px = O_PI * x;
        cc = 0.5 + (f * s - g * c) / px;
        ss = 0.5 - (f * c + g * s) / px;
    }
    if (xa < 0.0) {
        cc = -cc;
        ss = -ss;
    }
    *c_out = cc;
    *s_out = ss;
    return 0;
}

/* ---- kernel half widths: src/responses.c:9-84 ---- */
int o_r_resp_halfwidth(int highacc)
{
    return highacc ? (O_NUMFINTBINS * 3) + (O_NUMLOCPOWAVG >> 1) + O_DELTAAVGBINS
                   : O_NUMFINTBINS;
}

int o_z_resp_halfwidth(double z, int highacc)
{
    int m = (int) (0.5 * 1.1 * fabs(z));
    return m + (highacc ? O_NUMFINTBINS * 3 : O_NUMFINTBINS);
}

int o_w_resp_halfwidth(double z, double w, int highacc)
{
    if (fabs(w) < 1.0e-7)
        return o_z_resp_halfwidth(z, highacc);
    double r0 = 0.5 * (w / 6.0 - z);
    double r1 = 0.5 * (w / 6.0 + z);
    double maxdev = fabs(r0) > fabs(r1) ? fabs(r0) : fabs(r1);
    double u_ext = 0.5 - z / w;
    if (u_ext > 0.0 && u_ext < 1.0) {
        double z0 = z - w / 2.0;
        double r_ext = 0.5 * w * u_ext * u_ext + z0 * u_ext + r0;
        if (fabs(r_ext) > maxdev)
            maxdev = fabs(r_ext);
    }
    return (int) (1.1 * maxdev) + (highacc ? O_NUMFINTBINS * 3 : O_NUMFINTBINS);
}

/* ---- responses ---- */

/* src/responses.c:159-225: sinc interpolation kernel by trig recursion */
void o_gen_r_response(double roffset, int numbetween, int numkern, o_dcomplex *out)
{
    double startr = O_PI * (numkern / (double) (2 * numbetween) + roffset);
    double delta = -O_PI / numbetween;
    double tmp = sin(0.5 * delta);
    double alpha = -2.0 * tmp * tmp;
    double beta = sin(delta);
    double c = cos(startr), s = sin(startr), r = startr;
    for (int ii = 0; ii < numkern; ii++, r += delta) {
        double sinc = (r == 0.0) ? 1.0 : s / r;
        out[ii].r = c * sinc;
        out[ii].i = s * sinc;
        tmp = c;
        c = alpha * tmp - beta * s + c;
        s = alpha * s + beta * tmp + s;
    }
    if (roffset < 1E-3) {
        double r2 = roffset * roffset;
        out[numkern / 2].r = 1 - 6.579736267392905746 * r2;
        out[numkern / 2].i = roffset * (O_PI - 10.335425560099940058 * r2);
    }
}

/* src/responses.c:228-316: constant-fdot response via Fresnel integrals */
void o_gen_z_response(double roffset, int numbetween, double z, int numkern,
                      o_dcomplex *out)
{
    double absz = fabs(z);
    if (absz < 1E-4) {
        o_gen_r_response(roffset, numbetween, numkern, out);
        return;
    }
    double ipart;
    double startr = roffset - (0.5 * z);
    double startroffset = (startr < 0) ? 1.0 + modf(startr, &ipart) : modf(startr, &ipart);
    int signz = (z < 0.0) ? -1 : 1;
    double zd = signz * O_SQRT2 / sqrt(absz);
    double cons = zd / 2.0;
    double pibyz = O_PI / z;
    startr += numkern / (double) (2 * numbetween);
    double delta = -1.0 / numbetween;
    double r = startr;
    for (int ii = 0; ii < numkern; ii++, r += delta) {
        double yy = r * zd;
        double zz = yy + z * zd;
        double xx = pibyz * r * r;
        double c = cos(xx), s = sin(xx);
        double fsy, fcy, fsz, fcz;
        o_fresnl(yy, &fsy, &fcy);
        o_fresnl(zz, &fsz, &fcz);
        double tre = signz * (fcz - fcy);
        double tim = fsy - fsz;
        out[ii].r = (tre * c - tim * s) * cons;
        out[ii].i = -(tre * s + tim * c) * cons;
    }
    if (startroffset < 1E-3 && absz < 1E-3) {
        double z2 = z * z, x2 = startroffset * startroffset;
        int h = numkern / 2;
        out[h].r = 1.0 - 0.16449340668482264365 * z2;
        out[h].i = -0.5235987755982988731 * z;
        out[h].r += startroffset * 1.6449340668482264365 * z;
        out[h].i += startroffset * (O_PI - 0.5167712780049970029 * z2);
        out[h].r += x2 * (-6.579736267392905746 + 0.9277056288952613070 * z2);
        out[h].i += x2 * (3.1006276680299820175 * z);
    }
}

/* src/responses.c:333-418: jerk response.  The reference samples a chirp
 * into FP32, zero-pads it numbetween-fold and real-FFTs it.  out32 follows
 * that construction with the oracle's FP32 FFT; out64 is the FP64 DFT of the
 * identical FP32 samples, evaluated only at the bins that are kept. */
void o_gen_w_response(double roffset, int numbetween, double z, double w,
                      int numkern, o_fcomplex *out32, o_dcomplex *out64)
{
    if (fabs(w) < 1E-4) {
        o_dcomplex *tmp = (o_dcomplex *) malloc(sizeof(o_dcomplex) * numkern);
        o_gen_z_response(roffset, numbetween, z, numkern, tmp);
        for (int ii = 0; ii < numkern; ii++) {
            if (out32) {
                out32[ii].r = (float) tmp[ii].r;
                out32[ii].i = (float) tmp[ii].i;
            }
            if (out64)
                out64[ii] = tmp[ii];
        }
        free(tmp);
        return;
    }
    int npts = (int) o_next2_to_n(6 * o_w_resp_halfwidth(z, w, 0) + 200 +
                                  numkern / numbetween);
    double dt = 1.0 / (double) npts;
    double amp = 2.0 * dt;
    int fbar = npts / 4;
    double dfbar = (double) fbar + roffset;
    double f = dfbar - 0.5 * z + w / 12.0;
    double fd = (z - 0.5 * w) / 2.0;
    double fdd = w / 6.0;
    long ntot = (long) npts * numbetween;
    float *data = (float *) calloc(ntot, sizeof(float));
    for (int ii = 0; ii < npts; ii++) {
        double t = ii * dt;
        double phase = O_TWOPI * (t * (t * (t * fdd + fd) + f));
        data[ii] = (float) (amp * cos(phase));
    }
    long first = (long) fbar * numbetween - numkern / 2;   /* complex bin index */
    if (out64) {
        for (int kk = 0; kk < numkern; kk++) {
            long m = first + kk;
            double sr = 0.0, si = 0.0;
            for (long t = 0; t < npts; t++) {
                double ang = -O_TWOPI * (double) ((m * t) % ntot) / (double) ntot;
                sr += data[t] * cos(ang);
                si += data[t] * sin(ang);
            }
            out64[kk].r = sr;
            out64[kk].i = si;
        }
    }
    if (out32) {
        o_realfft_f(data, ntot, -1);
        memcpy(out32, data + 2 * first, sizeof(o_fcomplex) * numkern);
    }
    free(data);
}

/* ---- FFT-length policy: src/corr_prep.c:3-35, src/misc_utils.c ---- */
long long o_next2_to_n(long long x)
{
    long long t = 1;
    while (t < x)
        t <<= 1;
    return t;
}

int o_next_good_fftlen(int n)
{
    static const int lens[17] = {128, 192, 256, 384, 512, 768, 1024, 1280, 2048,
                                 4096, 5120, 7680, 10240, 12288, 15360, 16384, 25600};
    if (n <= lens[0])
        return lens[0];
    if (n > lens[16])
        return (int) o_next2_to_n(n);
    int ii = 0;
    while (n > lens[ii])
        ii++;
    return lens[ii];
}

int o_fftlen_from_kernwidth(int kw)
{
    if (kw < 6) return 128;
    if (kw < 52) return 256;
    if (kw < 67) return 512;
    if (kw < 378) return 1024;
    if (kw < 664) return 2048;
    if (kw < 1672) return 4096;
    if (kw < 3015) return 10240;
    if (kw < 3554) return 15360;
    if (kw < 6000) return 25600;
    return (int) o_next2_to_n(kw * 5);
}

/* ---- significance (src/characteristics.c:399-559) ---- */

/* Upper-tail standard normal probability from libm's erfc. */
double o_gauss_Q(double x)
{
    return 0.5 * erfc(x / O_SQRT2);
}

/* Inverse of o_gauss_Q: Acklam's rational starting point (published,
 * 2003) refined by Halley steps on erfc to full double precision. */
double o_gauss_Qinv(double p)
{
    if (!(p >= 0.0) || p > 1.0)
        return NAN;
    if (p == 0.0)
        return INFINITY;
    if (p == 1.0)
        return -INFINITY;
    static const double a[6] = {-3.969683028665376e+01, 2.209460984245205e+02,
                                -2.759285104469687e+02, 1.383577518672690e+02,
                                -3.066479806614716e+01, 2.506628277459239e+00};
    static const double b[5] = {-5.447609879822406e+01, 1.615858368580409e+02,
                                -1.556989798598866e+02, 6.680131188771972e+01,
                                -1.328068155288572e+01};
    static const double c[6] = {-7.784894002430293e-03, -3.223964580411365e-01,
                                -2.400758277161838e+00, -2.549732539343734e+00,
                                4.374664141464968e+00, 2.938163982698783e+00};
    static const double d[4] = {7.784695709041462e-03, 3.224671290700398e-01,
                                2.445134137142996e+00, 3.754408661907416e+00};
    /* x such that P(X <= x) = 1 - p, i.e. Qinv(p) = -Pinv(p) */
    double q = p, x;
    if (q < 0.02425) {
        double t = sqrt(-2.0 * log(q));
        x = -(((((c[0] * t + c[1]) * t + c[2]) * t + c[3]) * t + c[4]) * t + c[5]) /
            ((((d[0] * t + d[1]) * t + d[2]) * t + d[3]) * t + 1.0);
    } else if (q > 1.0 - 0.02425) {
        double t = sqrt(-2.0 * log(1.0 - q));
        x = (((((c[0] * t + c[1]) * t + c[2]) * t + c[3]) * t + c[4]) * t + c[5]) /
            ((((d[0] * t + d[1]) * t + d[2]) * t + d[3]) * t + 1.0);
    } else {
        double u = q - 0.5, t = u * u;
        x = -(((((a[0] * t + a[1]) * t + a[2]) * t + a[3]) * t + a[4]) * t + a[5]) * u /
            (((((b[0] * t + b[1]) * t + b[2]) * t + b[3]) * t + b[4]) * t + 1.0);
    }
    for (int it = 0; it < 4; it++) {
        double e = o_gauss_Q(x) - p;                       /* Q decreasing */
        double pdf = exp(-0.5 * x * x) / sqrt(O_TWOPI);
        if (pdf == 0.0)
            break;
        double u = -e / pdf;                                 /* Newton step */
        x = x - u / (1.0 + 0.5 * x * u);                     /* Halley */
    }
    return x;
}

/* log of sum_{k<n} y^k/k! computed stably */
static double log_poisson_head(double y, int n)
{
    /* terms t_k = y^k/k!; use log-sum-exp over the terms */
    double logt = 0.0, maxlog = 0.0;
    double *lt = (double *) malloc(sizeof(double) * (n > 0 ? n : 1));
    for (int k = 0; k < n; k++) {
        if (k > 0)
            logt += log(y) - log((double) k);
        lt[k] = logt;
        if (k == 0 || logt > maxlog)
            maxlog = logt;
    }
    double s = 0.0;
    for (int k = 0; k < n; k++)
        s += exp(lt[k] - maxlog);
    free(lt);
    return maxlog + log(s);
}

/* Q(chi2 | dof) for even dof: e^{-y} sum_{k<dof/2} y^k/k!, y = chi2/2.
 * This closed form replaces gsl_cdf_chisq_Q (src/characteristics.c:498);
 * the search only ever uses dof = 2*numsum, which is even. */
double o_chisq_Q_even(double chi2, double dof)
{
    if (chi2 <= 0.0)
        return 1.0;
    int n = (int) (dof / 2.0 + 0.5);
    double y = 0.5 * chi2;
    return exp(-y + log_poisson_head(y, n));
}

static double log_chisq_Q_even(double chi2, int n)
{
    double y = 0.5 * chi2;
    return -y + log_poisson_head(y, n);
}

/* Inverse of o_chisq_Q_even: safeguarded Newton on log Q (replaces
 * gsl_cdf_chisq_Qinv, src/characteristics.c:557). */
double o_chisq_Qinv_even(double q, double dof)
{
    int n = (int) (dof / 2.0 + 0.5);
    if (q >= 1.0)
        return 0.0;
    if (q <= 0.0)
        return INFINITY;
    double target = log(q);
    double lo = 0.0, hi = 2.0;
    while (log_chisq_Q_even(hi, n) > target)
        hi *= 2.0;
    double x = 0.5 * (lo + hi);
    for (int it = 0; it < 200; it++) {
        double f = log_chisq_Q_even(x, n) - target;
        if (f > 0.0)
            lo = x;
        else
            hi = x;
        /* d/dx log Q = -pdf/Q; pdf/Q = (1/2) y^{n-1}/(n-1)! / sum_{k<n} y^k/k! */
        double y = 0.5 * x;
        double lterm = (n - 1) * log(y) - lgamma((double) n);
        double dlogq = -0.5 * exp(lterm - log_poisson_head(y, n));
        double xn = x - f / dlogq;
        if (!(xn > lo && xn < hi))
            xn = 0.5 * (lo + hi);
        if (fabs(xn - x) <= 1e-15 * x) {
            x = xn;
            break;
        }
        x = xn;
    }
    return x;
}

/* src/characteristics.c:389-406 (A&S 26.2.23 on log p) */
static double extended_equiv_gaussian_sigma(double logp)
{
    double t = sqrt(-2.0 * logp);
    double num = 2.515517 + t * (0.802853 + t * 0.010328);
    double denom = 1.0 + t * (1.432788 + t * (0.189269 + t * 0.001308));
    return t - num / denom;
}

/* src/characteristics.c:409-427 (A&S 6.5.32) */
static double log_asymtotic_incomplete_gamma(double a, double z)
{
    double x = 1.0, newxpart = 1.0, term = 1.0;
    int ii = 1;
    while (fabs(newxpart) > 1e-15) {
        term *= (a - ii);
        newxpart = term / pow(z, ii);
        x += newxpart;
        ii += 1;
    }
    return (a - 1.0) * log(z) - z + log(x);
}

/* src/characteristics.c:429-444 (A&S 6.1.41) */
static double log_asymtotic_gamma(double z)
{
    double x = (z - 0.5) * log(z) - z + 0.91893853320467267;
    double y = 1.0 / (z * z);
    x += (((-5.9523809523809529e-4 * y + 7.9365079365079365079365e-4) * y -
           2.7777777777777777777778e-3) * y + 8.3333333333333333333333e-2) / z;
    return x;
}

/* src/characteristics.c:458-470 */
double o_equivalent_gaussian_sigma(double logp)
{
    double x;
    if (logp < -600.0)
        x = extended_equiv_gaussian_sigma(logp);
    else
        x = o_gauss_Qinv(exp(logp));
    return (x < 0.0) ? 0.0 : x;
}

/* src/characteristics.c:473-506 */
double o_chi2_logp(double chi2, double dof)
{
    if (chi2 <= 0.0)
        return -INFINITY;
    if (chi2 > 1200.0 || (dof > 30 && chi2 > 100.0))
        return log_asymtotic_incomplete_gamma(0.5 * dof, 0.5 * chi2) -
               log_asymtotic_gamma(0.5 * dof);
    return log(o_chisq_Q_even(chi2, dof));
}

/* src/characteristics.c:526-546 */
double o_candidate_sigma(double power, int numsum, double numtrials)
{
    if (power <= 0.0)
        return 0.0;
    double logp = o_chi2_logp(2.0 * power, 2.0 * numsum);
    logp += log(numtrials);
    return o_equivalent_gaussian_sigma(logp);
}

/* src/characteristics.c:549-559 */
double o_power_for_sigma(double sigma, int numsum, double numtrials)
{
    double q = o_gauss_Q(sigma) / numtrials;
    return 0.5 * o_chisq_Qinv_even(q, 2.0 * numsum);
}

/* ---- median (src/median.c:505-511): exact order statistics ---- */
static float select_kth(float *a, int n, int k)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        float pivot = a[(lo + hi) / 2];
        int i = lo, j = hi;
        while (i <= j) {
            while (a[i] < pivot) i++;
            while (a[j] > pivot) j--;
            if (i <= j) {
                float t = a[i]; a[i] = a[j]; a[j] = t;
                i++; j--;
            }
        }
        if (k <= j)
            hi = j;
        else if (k >= i)
            lo = i;
        else
            return a[k];
    }
    return a[k];
}

float o_median(float *arr, int n)
{
    if (n <= 0)
        return 0.0f;
    int lhs = (n - 1) / 2, rhs = n / 2;
    if (lhs == rhs)
        return select_kth(arr, n, lhs);
    double a = select_kth(arr, n, lhs);
    double b = select_kth(arr, n, rhs);
    return (float) (0.5 * (a + b));
}

/* src/misc_utils.c:291-302 (compiled with -ffp-contract=off) */
float o_invsqrtf(float x)
{
    union { float f; int i; } tmp;
    tmp.f = x;
    tmp.i = 0x5f3759df - (tmp.i >> 1);
    float y = tmp.f;
    return y * (1.5f - 0.5f * x * y * y);
}
