/*
 * dev_fresnel.h -- FP64 Fresnel integrals S(x), C(x) for device code
 * (Cephes rational forms with the published Cephes tables, the algorithm
 * of src/fresnl.c:365-406; x > 36974 returns 1/2 as src/fresnl.c:385-389).
 * Included by each HIP translation unit that needs it.
 */
#ifndef PRESTO_DEV_FRESNEL_H
#define PRESTO_DEV_FRESNEL_H
#include <hip/hip_runtime.h>

/* ------------------------------------------------------------------ */
/* FP64 Fresnel integrals (Cephes rational forms, the published tables  */
/* used by src/fresnl.c:76-229)                                         */
/* ------------------------------------------------------------------ */
static __constant__ double c_sn[6] = {
    -2.99181919401019853726E3, 7.08840045257738576863E5, -6.29741486205862506537E7,
    2.54890880573376359104E9, -4.42979518059697779103E10, 3.18016297876567817986E11};
static __constant__ double c_sd[6] = {
    2.81376268889994315696E2, 4.55847810806532581675E4, 5.17343888770096400730E6,
    4.19320245898111231129E8, 2.24411795645340920940E10, 6.07366389490084639049E11};
static __constant__ double c_cn[6] = {
    -4.98843114573573548651E-8, 9.50428062829859605134E-6, -6.45191435683965050962E-4,
    1.88843319396703850064E-2, -2.05525900955013891793E-1, 9.99999999999999998822E-1};
static __constant__ double c_cd[7] = {
    3.99982968972495980367E-12, 9.15439215774657478799E-10, 1.25001862479598821474E-7,
    1.22262789024179030997E-5, 8.68029542941784300606E-4, 4.12142090722199792936E-2,
    1.00000000000000000118E0};
static __constant__ double c_fn[10] = {
    4.21543555043677546506E-1, 1.43407919780758885261E-1, 1.15220955073585758835E-2,
    3.45017939782574027900E-4, 4.63613749287867322088E-6, 3.05568983790257605827E-8,
    1.02304514164907233465E-10, 1.72010743268161828879E-13, 1.34283276233062758925E-16,
    3.76329711269987889006E-20};
static __constant__ double c_fd[10] = {
    7.51586398353378947175E-1, 1.16888925859191382142E-1, 6.44051526508858611005E-3,
    1.55934409164153020873E-4, 1.84627567348930545870E-6, 1.12699224763999035261E-8,
    3.60140029589371370404E-11, 5.88754533621578410010E-14, 4.52001434074129701496E-17,
    1.25443237090011264384E-20};
static __constant__ double c_gn[11] = {
    5.04442073643383265887E-1, 1.97102833525523411709E-1, 1.87648584092575249293E-2,
    6.84079380915393090172E-4, 1.15138826111884280931E-5, 9.82852443688422223854E-8,
    4.45344415861750144738E-10, 1.08268041139020870318E-12, 1.37555460633261799868E-15,
    8.36354435630677421531E-19, 1.86958710162783235106E-22};
static __constant__ double c_gd[11] = {
    1.47495759925128324529E0, 3.37748989120019970451E-1, 2.53603741420338795122E-2,
    8.14679107184306179049E-4, 1.27545075667729118702E-5, 1.04314589657571990585E-7,
    4.60680728146520428211E-10, 1.10273215066240270757E-12, 1.38796531259578871258E-15,
    8.39158816283118707363E-19, 1.86958710162783236342E-22};

__device__ static double d_poly(double x, const double *c, int n)
{
    double a = c[0];
    for (int k = 1; k <= n; k++)
        a = a * x + c[k];
    return a;
}

__device__ static double d_poly1(double x, const double *c, int n)
{
    double a = x + c[0];
    for (int k = 1; k < n; k++)
        a = a * x + c[k];
    return a;
}

__device__ static void d_fresnl(double xa, double *s_out, double *c_out)
{
    const double PI = 3.14159265358979323846;
    double x = fabs(xa), x2 = x * x, cc, ss;
    if (x2 < 2.5625) {
        double t = x2 * x2;
        ss = x * x2 * d_poly(t, c_sn, 5) / d_poly1(t, c_sd, 6);
        cc = x * d_poly(t, c_cn, 5) / d_poly(t, c_cd, 6);
    } else if (x > 36974.0) {
        cc = 0.5;
        ss = 0.5;
    } else {
        double t = PI * x2;
        double u = 1.0 / (t * t);
        t = 1.0 / t;
        double f = 1.0 - u * d_poly(u, c_fn, 9) / d_poly1(u, c_fd, 10);
        double g = t * d_poly(u, c_gn, 10) / d_poly1(u, c_gd, 11);
        double sv, cv;
        sincospi(0.5 * x2, &sv, &cv);
        double px = PI * x;
        cc = 0.5 + (f * sv - g * cv) / px;
        ss = 0.5 - (f * cv + g * sv) / px;
    }
    if (xa < 0.0) {
        cc = -cc;
        ss = -ss;
    }
    *c_out = cc;
    *s_out = ss;
}

#endif
