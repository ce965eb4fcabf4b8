/*
 * accel_stats.c -- significance of summed powers
 * (src/characteristics.c:389-559).
 *
 * The search only sums powers over dof = 2*numsum, an even number, where
 * the chi-square tail has the closed form Q = e^{-y} sum_{k<n} y^k/k!
 * (y = chi2/2, n = dof/2); the quantiles are found by Newton iteration on
 * log Q.  The asymptotic branches of the reference are kept as they are,
 * since they decide the reported sigma of strong candidates.
 */
#include <math.h>
#include "presto_accel.h"

#define SQRT2 1.41421356237309504880
#define LN_SQRT_2PI 0.91893853320467274178

/* log(sum_{k<n} y^k/k!) */
static double log_head(double y, int n)
{
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < n; k++) {
        term *= y / k;
        sum += term;
    }
    return log(sum);
}

static double log_chisq_Q(double chi2, int n)
{
    double y = 0.5 * chi2;
    return -y + log_head(y, n);
}

/* log of the standard normal upper tail */
static double log_gauss_Q(double x)
{
    if (x < 5.0)
        return log(0.5 * erfc(x / SQRT2));
    /* continued fraction is unnecessary: erfc keeps full relative
     * precision until ~26 sigma; beyond, use the asymptotic series */
    if (x < 37.0)
        return log(0.5 * erfc(x / SQRT2));
    double x2 = x * x, s = 1.0 - 1.0 / x2 + 3.0 / (x2 * x2) - 15.0 / (x2 * x2 * x2);
    return -0.5 * x2 - log(x) - LN_SQRT_2PI + log(s);
}

/* x with Q(x) = p, Newton on log Q (concave), p in (0, 1) */
static double gauss_Qinv(double p)
{
    if (!(p > 0.0) || p > 1.0)
        return (p == 0.0) ? INFINITY : NAN;
    if (p == 1.0)
        return -INFINITY;
    if (p > 0.5)
        return -gauss_Qinv(1.0 - p);
    double lp = log(p);
    double x = sqrt(-2.0 * lp);       /* Q(x) <= exp(-x^2/2)/2: start above the root */
    for (int it = 0; it < 100; it++) {
        double lq = log_gauss_Q(x);
        double phi_over_q = exp(-0.5 * x * x - LN_SQRT_2PI - lq);
        double dx = (lq - lp) / phi_over_q;   /* g/g' with g' = -phi/Q */
        x += dx;
        if (fabs(dx) <= 1e-16 * fabs(x) + 1e-300)
            break;
    }
    return x;
}

/* chi2 with Q(chi2 | 2n) = q */
static double chisq_Qinv_even(double q, int n)
{
    if (q >= 1.0)
        return 0.0;
    double lq = log(q);
    double lo = 0.0, hi = 1.0;
    while (log_chisq_Q(hi, n) > lq)
        hi *= 2.0;
    double x = hi;
    for (int it = 0; it < 200; it++) {
        double g = log_chisq_Q(x, n) - lq;
        if (g > 0.0)
            lo = x;
        else
            hi = x;
        double y = 0.5 * x;
        double dlog = -0.5 * exp((n - 1) * log(y) - lgamma((double) n) - log_head(y, n));
        double xn = x - g / dlog;
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

/* src/characteristics.c:409-427 (A&S 6.5.32) */
static double log_asym_incomplete_gamma(double a, double z)
{
    double x = 1.0, part = 1.0, term = 1.0;
    int ii = 1;
    while (fabs(part) > 1e-15) {
        term *= (a - ii);
        part = term / pow(z, ii);
        x += part;
        ii += 1;
    }
    return (a - 1.0) * log(z) - z + log(x);
}

/* src/characteristics.c:429-444 (A&S 6.1.41) */
static double log_asym_gamma(double z)
{
    double x = (z - 0.5) * log(z) - z + 0.91893853320467267;
    double y = 1.0 / (z * z);
    x += (((-5.9523809523809529e-4 * y + 7.9365079365079365079365e-4) * y -
           2.7777777777777777777778e-3) * y + 8.3333333333333333333333e-2) / z;
    return x;
}

/* src/characteristics.c:473-506 */
static double chi2_logp(double chi2, double dof)
{
    if (chi2 <= 0.0)
        return -INFINITY;
    if (chi2 > 1200.0 || (dof > 30 && chi2 > 100.0))
        return log_asym_incomplete_gamma(0.5 * dof, 0.5 * chi2) - log_asym_gamma(0.5 * dof);
    return log_chisq_Q(chi2, (int) (0.5 * dof + 0.5));
}

/* src/characteristics.c:458-470 and :389-406 */
static double equivalent_gaussian_sigma(double logp)
{
    double x;
    if (logp < -600.0) {
        double t = sqrt(-2.0 * logp);
        double num = 2.515517 + t * (0.802853 + t * 0.010328);
        double den = 1.0 + t * (1.432788 + t * (0.189269 + t * 0.001308));
        x = t - num / den;
    } else {
        x = gauss_Qinv(exp(logp));
    }
    return (x < 0.0) ? 0.0 : x;
}

double accel_candidate_sigma(double power, int numsum, double numtrials)
{
    if (power <= 0.0)
        return 0.0;
    double logp = chi2_logp(2.0 * power, 2.0 * numsum) + log(numtrials);
    return equivalent_gaussian_sigma(logp);
}

double accel_power_for_sigma(double sigma, int numsum, double numtrials)
{
    double q = exp(log_gauss_Q(sigma)) / numtrials;
    return 0.5 * chisq_Qinv_even(q, numsum);
}
