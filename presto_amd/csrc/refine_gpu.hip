/*
 * refine_gpu.hip -- FP64 candidate refinement on the GPU: the work of
 * optimize_accelcand (src/accel_utils.c:552-626) for the default
 * harmonic-polishing path, one workgroup per candidate.
 *
 *   k_refine_rz : max_rz_arr_harmonics (src/maximize_rz.c:171-228): local
 *       powers, a two-pass Nelder-Mead simplex over (r, z/4) maximising the
 *       locally normalised summed power of all harmonics, then per
 *       harmonic the power and the 5-point derivatives (get_localpower3d
 *       and get_derivs3d, src/characteristics.c:79-192).
 *
 * Every function value is a set of Fourier interpolations (rz_interp,
 * src/rzinterp.c:144-237): the response taps are generated on the fly in
 * FP64 (Fresnel integrals), rounded to FP32 as the reference stores them,
 * multiplied by the FP32 spectrum and summed in FP64 over the workgroup
 * with a fixed-order tree, so every thread holds the same value and the
 * simplex runs in lockstep, deterministically.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include "agpu.h"
#include "dev_fresnel.h"

#define RF_TPB 256
#define RF_MAXH AGPU_REFINE_MAXH

#define RF_CHECK(call)                                                            \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                   \
            char m_[512];                                                         \
            snprintf(m_, sizeof(m_), "%s:%d %s: %s", __FILE__, __LINE__, #call,   \
                     hipGetErrorString(e_));                                      \
            agpu_set_error(m_);                                                   \
            return -1;                                                            \
        }                                                                         \
    } while (0)

#include "dev_resp.h"

__device__ static float2 resp_tap(double roffset, double z, int numkern, int m)
{
    return resp_tap_nb(roffset, 1, z, numkern, m);
}

/* Workgroup sum in a fixed order; every thread gets the result. */
__device__ static double2 wg_sum(double2 v, double2 *red)
{
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int s = RF_TPB / 2; s > 0; s >>= 1) {
        if (t < s) {
            red[t].x += red[t + s].x;
            red[t].y += red[t + s].y;
        }
        __syncthreads();
    }
    double2 out = red[0];
    __syncthreads();
    return out;
}

/* rz_interp (src/rzinterp.c:144-237) by one workgroup */
__device__ static double2 rz_interp_wg(const float2 *spec, long long numdata, double r, double z,
                                       int khw, double2 *red)
{
    if (r > numdata - 1.0 || r < 0.0)
        return make_double2(0.0, 0.0);
    double dint;
    const double frac = modf(r, &dint);
    const long long intfreq = (long long) dint;
    if (fabs(z) < 1e-4) {
        if (frac < 1e-5) {
            float2 d = spec[intfreq];
            return make_double2(d.x, d.y);
        }
        if ((1.0 - frac) < 1e-5) {
            float2 d = spec[intfreq + 1];
            return make_double2(d.x, d.y);
        }
    }
    const int numkern = 2 * khw;
    long long lodata = intfreq - khw, loresp, hiresp;
    if (lodata < 0) {
        loresp = -lodata;
        lodata = 0;
    } else {
        loresp = 0;
    }
    const long long hidata = intfreq + khw - 1;
    hiresp = (hidata > numdata - 1) ? numkern - hidata + numdata - 1 : numkern;
    const long long nsum = hiresp - loresp;
    double2 acc = make_double2(0.0, 0.0);
    for (long long ii = threadIdx.x; ii < nsum; ii += RF_TPB) {
        const float2 q = resp_tap(frac, z, numkern, (int) (loresp + ii));
        const float2 d = spec[lodata + ii];
        acc.x += (double) d.x * q.x + (double) d.y * q.y;
        acc.y += (double) d.y * q.x - (double) q.y * d.x;
    }
    return wg_sum(acc, red);
}

/* get_localpower3d with w = 0 (src/characteristics.c:79-136) */
__device__ static double localpower_wg(const float2 *spec, long long numdata, double r, double z,
                                       double2 *red)
{
    const int binsperside = 10;   /* NUMLOCPOWAVG / 2 */
    const int khw = zr_halfwidth(z, 0);
    double lo1 = r - 5 - binsperside, hi1 = lo1 + binsperside;   /* DELTAAVGBINS = 5 */
    double lo2 = r + 5, hi2 = lo2 + binsperside;
    const double nd = (double) numdata;
    lo1 = lo1 < 0.0 ? 0.0 : (lo1 > nd ? nd : lo1);
    lo2 = lo2 < 0.0 ? 0.0 : (lo2 > nd ? nd : lo2);
    hi1 = hi1 < 0.0 ? 0.0 : (hi1 > nd ? nd : hi1);
    hi2 = hi2 < 0.0 ? 0.0 : (hi2 > nd ? nd : hi2);
    double sum = 0.0;
    for (double f = lo1; f < hi1; f += 1.0) {
        double2 a = rz_interp_wg(spec, numdata, f, z, khw, red);
        sum += a.x * a.x + a.y * a.y;
    }
    for (double f = lo2; f < hi2; f += 1.0) {
        double2 a = rz_interp_wg(spec, numdata, f, z, khw, red);
        sum += a.x * a.x + a.y * a.y;
    }
    return sum / 20.0;
}

/* get_derivs3d with w = 0 (src/characteristics.c:139-192) */
__device__ static void derivs_wg(const float2 *spec, long long numdata, double r, double z,
                                 double locpow, double *out7, double2 *red)
{
    const double PI = 3.14159265358979323846, TWOPI = 6.28318530717958647692;
    const double h = 0.003, twoh = 0.006, twohsqrd = 0.000036;
    const int khw = zr_halfwidth(z, 1);
    double pwr[5], phs[5];
    double f = r - twoh;
#pragma unroll
    for (int k = 0; k < 5; k++, f += h) {
        double2 a = rz_interp_wg(spec, numdata, f, z, khw, red);
        pwr[k] = a.x * a.x + a.y * a.y;
        double t = atan2(a.y, a.x);
        phs[k] = t > 0.0 ? t : t + TWOPI;
    }
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (fabs(phs[k + 1] - phs[k]) > PI)
            phs[k + 1] -= TWOPI;
    out7[0] = pwr[2];
    out7[1] = phs[2];
    out7[2] = (pwr[3] - pwr[1]) / twoh;
    out7[3] = (phs[3] - phs[1]) / twoh;
    out7[4] = (pwr[4] - 2.0 * pwr[2] + pwr[0]) / twohsqrd;
    out7[5] = (phs[4] - 2.0 * phs[2] + phs[0]) / twohsqrd;
    out7[6] = locpow;
}

struct rf_ctx {
    const float2 *spec;
    long long numdata;
    const float *locpow;   /* LDS, numharm entries */
    int numharm, kernhw;
    double2 *red;
    int nevals;
};

/* power_call_rz_harmonics_gsl (src/maximize_rz.c:43-63): minus the summed
 * locally normalised power at (r, z = zs * 4) over all harmonics */
__device__ static double rf_fun(rf_ctx &c, double r, double zs)
{
    c.nevals++;
    double tot = 0.0;
    for (int k = 0; k < c.numharm; k++) {
        const int n = k + 1;
        double2 a = rz_interp_wg(c.spec, c.numdata, r * n, zs * 4.0 * n, c.kernhw, c.red);
        tot += (a.x * a.x + a.y * a.y) / c.locpow[k];
    }
    if (!isfinite(tot))
        return 0.0;
    return -tot;
}

/* Nelder-Mead over 2 parameters, the nmsimplex2 scheme: reflect / expand /
 * contract the worst corner, else shrink towards the best; incremental
 * centre and squared rms size; stop when the size drops below tol or after
 * maxiter iterations (src/maximize_rz.c:65-100). */
__device__ static void rf_simplex(rf_ctx &c, double *x, double step, double tol, int maxiter)
{
    const double P = 3.0;
    double X[3][2], Y[3], C[2], S2;
    X[0][0] = x[0];
    X[0][1] = x[1];
    Y[0] = rf_fun(c, x[0], x[1]);
    for (int i = 0; i < 2; i++) {
        double xt0 = x[0] + (i == 0 ? step : 0.0), xt1 = x[1] + (i == 1 ? step : 0.0);
        X[i + 1][0] = xt0;
        X[i + 1][1] = xt1;
        Y[i + 1] = rf_fun(c, xt0, xt1);
    }
    auto center_size = [&]() {
        for (int j = 0; j < 2; j++)
            C[j] = (X[0][j] + X[1][j] + X[2][j]) / 3.0;
        double ss = 0.0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 2; j++)
                ss += (X[i][j] - C[j]) * (X[i][j] - C[j]);
        S2 = ss / 3.0;
    };
    auto update = [&](int i, double v0, double v1, double val) {
        double d0 = v0 - X[i][0], d1 = v1 - X[i][1];
        double xmcd = (X[i][0] - C[0]) * d0 + (X[i][1] - C[1]) * d1;
        S2 += (2.0 / P) * xmcd + ((P - 1.0) / P) * ((d0 * d0 + d1 * d1) / P);
        C[0] += -X[i][0] / P;
        C[0] += v0 / P;
        C[1] += -X[i][1] / P;
        C[1] += v1 / P;
        X[i][0] = v0;
        X[i][1] = v1;
        Y[i] = val;
    };
    center_size();
    int it = 0, lo = 0;
    double size;
    do {
        it++;
        double dhi = Y[0], dlo = Y[0], ds_hi = Y[1];
        int hi = 0;
        lo = 0;
        for (int i = 1; i < 3; i++) {
            double v = Y[i];
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
        /* corner moves: xc = a*centre + b*x_hi, a = (1-k) P/(P-1), b = (P k - 1)/(P-1) */
        const double ar = 3.0, br = -2.0;      /* k = -1: reflection */
        double r0 = ar * C[0] + br * X[hi][0], r1 = ar * C[1] + br * X[hi][1];
        double val = rf_fun(c, r0, r1);
        if (isfinite(val) && val < Y[lo]) {
            const double ae = 4.5, be = -3.5;  /* k = -2: expansion */
            double e0 = ae * C[0] + be * X[hi][0], e1 = ae * C[1] + be * X[hi][1];
            double val2 = rf_fun(c, e0, e1);
            if (isfinite(val2) && val2 < Y[lo])
                update(hi, e0, e1, val2);
            else
                update(hi, r0, r1, val);
        } else if (!isfinite(val) || val > ds_hi) {
            if (isfinite(val) && val <= Y[hi])
                update(hi, r0, r1, val);
            const double ac = 0.75, bc = 0.25; /* k = 1/2: contraction */
            double k0 = ac * C[0] + bc * X[hi][0], k1 = ac * C[1] + bc * X[hi][1];
            double val2 = rf_fun(c, k0, k1);
            if (isfinite(val2) && val2 <= Y[hi]) {
                update(hi, k0, k1, val2);
            } else {
                for (int i = 0; i < 3; i++) {
                    if (i == lo)
                        continue;
                    X[i][0] = 0.5 * (X[i][0] + X[lo][0]);
                    X[i][1] = 0.5 * (X[i][1] + X[lo][1]);
                    Y[i] = rf_fun(c, X[i][0], X[i][1]);
                }
                center_size();
            }
        } else {
            update(hi, r0, r1, val);
        }
        lo = 0;
        for (int i = 1; i < 3; i++)
            if (Y[i] < Y[lo])
                lo = i;
        if (!(S2 > 0))
            center_size();
        size = sqrt(S2);
    } while (!(size < tol) && it < maxiter);
    x[0] = X[lo][0];
    x[1] = X[lo][1];
}

__global__ __launch_bounds__(RF_TPB) void k_refine_rz(agpu_refine_args a)
{
    __shared__ double2 red[RF_TPB];
    __shared__ float s_lp[RF_MAXH];
    const int ci = blockIdx.x;
    const int H = a.numharm[ci];
    const float2 *spec = (const float2 *) a.spec;
    const double rin = a.rin[ci], zin = a.zin[ci];
    for (int k = 0; k < H; k++) {
        const int n = k + 1;
        double lp = localpower_wg(spec, a.numdata, rin * n, zin * n, red);
        if (threadIdx.x == 0)
            s_lp[k] = a.raw ? 1.0f : (float) lp;
    }
    __syncthreads();
    rf_ctx c{spec, a.numdata, s_lp, H, 0, red, 0};
    double x[2] = {rin, zin / 4.0};
    c.kernhw = zr_halfwidth(fabs(zin * H) + 4.0, 0);
    rf_simplex(c, x, 0.4, 1.0e-4, 5000);
    c.kernhw = zr_halfwidth(fabs(x[1] * H) + 4.0, 1);
    rf_simplex(c, x, 0.01, 1.0e-7, 5000);
    const double rout = x[0], zout = x[1] * 4.0;
    __syncthreads();
    for (int k = 0; k < H; k++) {
        const int n = k + 1;
        const float lp = (float) localpower_wg(spec, a.numdata, rout * n, zout * n, red);
        double2 amp = rz_interp_wg(spec, a.numdata, rout * n, zout / 4.0 * n * 4.0, c.kernhw,
                                   red);
        double d7[7];
        derivs_wg(spec, a.numdata, rout * n, zout * n, lp, d7, red);
        if (threadIdx.x == 0) {
            a.powers[(size_t) ci * RF_MAXH + k] = amp.x * amp.x + amp.y * amp.y;
            for (int j = 0; j < 7; j++)
                a.derivs[((size_t) ci * RF_MAXH + k) * 7 + j] = d7[j];
        }
    }
    if (threadIdx.x == 0) {
        a.rout[ci] = rout;
        a.zout[ci] = zout;
        a.nevals[ci] = c.nevals;
    }
}

int agpu_refine_rz(const agpu_refine_args *args, void *stream)
{
    if (args->ncand <= 0)
        return 0;
    if (args->ncand > 1 << 20) {
        agpu_set_error("agpu_refine_rz: too many candidates");
        return -2;
    }
    hipLaunchKernelGGL(k_refine_rz, dim3(args->ncand), dim3(RF_TPB), 0, (hipStream_t) stream,
                       *args);
    RF_CHECK(hipGetLastError());
    return 0;
}

/* ------------------------------------------------------------------ */
/* Jerk refinement: max_rzw_arr_harmonics (src/maximize_rzw.c:84-191)   */
/* ------------------------------------------------------------------ */

/* w_resp_halfwidth, src/responses.c:48-84 */
__device__ static int wr_halfwidth(double z, double w, int highacc)
{
    if (fabs(w) < 1.0e-7)
        return zr_halfwidth(z, highacc);
    double r0 = 0.5 * (w / 6.0 - z), r1 = 0.5 * (w / 6.0 + z);
    double maxdev = fabs(r0) > fabs(r1) ? fabs(r0) : fabs(r1);
    double u_ext = 0.5 - z / w;
    if (u_ext > 0.0 && u_ext < 1.0) {
        double z0 = z - w / 2.0;
        double r_ext = 0.5 * w * u_ext * u_ext + z0 * u_ext + r0;
        maxdev = fabs(r_ext) > maxdev ? fabs(r_ext) : maxdev;
    }
    return (int) (1.1 * maxdev) + (highacc ? 48 : 16);
}

__device__ static long long next2_dev(long long x)
{
    long long t = 1;
    while (t < x)
        t <<= 1;
    return t;
}

/* Sample n of the chirp gen_w_response builds (src/responses.c:384-395):
 * FP64 phase, stored as FP32. */
__device__ static float chirp_x(int n, int npts, double roffset, double z, double w)
{
    const double TWOPI = 6.28318530717958647692;
    const double dt = 1.0 / (double) npts;
    const double f = (double) (npts / 4) + roffset - 0.5 * z + w / 12.0;
    const double fd = (z - 0.5 * w) / 2.0, fdd = w / 6.0;
    const double t = n * dt;
    return (float) (2.0 * dt * cos(TWOPI * (t * (t * (t * fdd + fd) + f))));
}

/* i^n */
__device__ static double2 ipow(int n)
{
    switch (n & 3) {
    case 0: return make_double2(1.0, 0.0);
    case 1: return make_double2(0.0, 1.0);
    case 2: return make_double2(-1.0, 0.0);
    default: return make_double2(0.0, -1.0);
    }
}

/* S(n) = sum_{m=mlo}^{mhi-1} d[k0+m] e^{2 pi i m n / npts}, bins outside
 * [0, numdata) taken as zero (the zero padding of src/corr_routines.c:
 * 159-170).  The phasor is re-seeded exactly every 128 steps. */
__device__ static double2 window_dft(const float2 *spec, long long numdata, long long k0, int mlo,
                                     int mhi, int n, int npts)
{
    long long a = k0 + mlo < 0 ? -k0 : mlo;
    long long b = k0 + mhi > numdata ? numdata - k0 : mhi;
    double2 acc = make_double2(0.0, 0.0);
    double ss, cs;
    sincospi(2.0 * (double) n / (double) npts, &ss, &cs);
    double pr = 0.0, pi = 0.0;
    for (long long m = a; m < b; m++) {
        if (((m - a) & 127) == 0) {
            long long q = (m * (long long) n) % npts;
            if (q < 0)
                q += npts;
            sincospi(2.0 * (double) q / (double) npts, &pi, &pr);
        }
        const float2 d = spec[k0 + m];
        acc.x += d.x * pr - d.y * pi;
        acc.y += d.x * pi + d.y * pr;
        const double t = pr * cs - pi * ss;
        pi = pr * ss + pi * cs;
        pr = t;
    }
    return acc;
}

/* Forward DFT (e^{-2 pi i}) of a[0..npts), npts a power of two, by the
 * whole workgroup: radix-2 Stockham passes between a and b in global
 * memory (the candidate's scratch), FP64 twiddles.  Returns the buffer
 * holding the result. */
__device__ static double2 *wg_fft64(double2 *a, double2 *b, int npts)
{
    const int half = npts >> 1;
    double2 *src = a, *dst = b;
    for (int Ns = 1; Ns < npts; Ns <<= 1) {
        for (int j = threadIdx.x; j < half; j += RF_TPB) {
            const int k = j & (Ns - 1);
            double s, c;
            sincospi(-(double) k / (double) Ns, &s, &c);
            const double2 x0 = src[j], x1 = src[j + half];
            const double2 y = make_double2(x1.x * c - x1.y * s, x1.x * s + x1.y * c);
            const int d = ((j - k) << 1) + k;
            dst[d] = make_double2(x0.x + y.x, x0.y + y.y);
            dst[d + Ns] = make_double2(x0.x - y.x, x0.y - y.y);
        }
        __syncthreads();
        double2 *t = src;
        src = dst;
        dst = t;
    }
    return src;
}

/* S(n) = sum_{m=mlo}^{mhi-1} d[k0+m] e^{+2 pi i m n / npts} for all n <
 * npts (bins outside [0, numdata) as zero; mhi - mlo <= npts), through
 * wg_fft64 of the conjugated window: S = conj(FFT(conj(window))).  The
 * result is left conjugated in the returned buffer (read it with conj). */
__device__ static double2 *window_spectrum(const float2 *spec, long long numdata, long long k0,
                                           int mlo, int mhi, int npts, double2 *a, double2 *b)
{
    for (int j = threadIdx.x; j < npts; j += RF_TPB)
        a[j] = make_double2(0.0, 0.0);
    __syncthreads();
    for (int m = mlo + (int) threadIdx.x; m < mhi; m += RF_TPB) {
        const long long k = k0 + m;
        if (k < 0 || k >= numdata)
            continue;
        const float2 d = spec[k];
        const int j = m >= 0 ? m : m + npts;
        a[j] = make_double2(d.x, -d.y);
    }
    __syncthreads();
    return wg_fft64(a, b, npts);
}

/* rzw_interp (src/rzwinterp.c:145-230) by one workgroup.  With R the
 * response of gen_w_response(frac, 1, z, w, 2 khw) -- the DFT of the
 * chirp x at bins npts/4 - khw + kk -- the sum over data bins
 * sum_k d_k conj(R) regroups into sum_n x_n i^n S(n), S the DFT of the
 * data window around intfreq. */
__device__ static double2 rzw_interp_wg(const float2 *spec, long long numdata, double r,
                                        double z, double w, int khw, double2 *red,
                                        double2 *fa, double2 *fb, int npts_cap, int *bad)
{
    if (r > numdata - 1.0 || r < 0.0)
        return make_double2(0.0, 0.0);
    if (fabs(w) < 1e-4)
        return rz_interp_wg(spec, numdata, r, z, khw, red);
    double dint;
    const double frac = modf(r, &dint);
    const long long intfreq = (long long) dint;
    const int numkern = 2 * khw;
    const int npts = (int) next2_dev(6 * wr_halfwidth(z, w, 0) + 200 + numkern);
    if (npts > npts_cap) {
        *bad = 1;
        return make_double2(0.0, 0.0);
    }
    const double2 *S = window_spectrum(spec, numdata, intfreq, -khw, khw, npts, fa, fb);
    double2 acc = make_double2(0.0, 0.0);
    for (int n = threadIdx.x; n < npts; n += RF_TPB) {
        const double x = chirp_x(n, npts, frac, z, w);
        const double2 Sc = S[n];   /* conj of S(n) */
        const double2 ph = ipow(n);
        /* x i^n S(n), S(n) = (Sc.x, -Sc.y) */
        acc.x += x * (ph.x * Sc.x + ph.y * Sc.y);
        acc.y += x * (ph.y * Sc.x - ph.x * Sc.y);
    }
    __syncthreads();   /* fa/fb are reused by the next call */
    return wg_sum(acc, red);
}

__device__ static double localpower3d_wg(const float2 *spec, long long numdata, double r,
                                         double z, double w, double2 *red, double2 *fa,
                                         double2 *fb, int npts_cap, int *bad)
{
    const int khw = wr_halfwidth(z, w, 0);
    double lo1 = r - 5 - 10, hi1 = lo1 + 10;
    double lo2 = r + 5, hi2 = lo2 + 10;
    const double nd = (double) numdata;
    lo1 = lo1 < 0.0 ? 0.0 : (lo1 > nd ? nd : lo1);
    lo2 = lo2 < 0.0 ? 0.0 : (lo2 > nd ? nd : lo2);
    hi1 = hi1 < 0.0 ? 0.0 : (hi1 > nd ? nd : hi1);
    hi2 = hi2 < 0.0 ? 0.0 : (hi2 > nd ? nd : hi2);
    double sum = 0.0;
    for (double f = lo1; f < hi1; f += 1.0) {
        double2 a = rzw_interp_wg(spec, numdata, f, z, w, khw, red, fa, fb, npts_cap, bad);
        sum += a.x * a.x + a.y * a.y;
    }
    for (double f = lo2; f < hi2; f += 1.0) {
        double2 a = rzw_interp_wg(spec, numdata, f, z, w, khw, red, fa, fb, npts_cap, bad);
        sum += a.x * a.x + a.y * a.y;
    }
    return sum / 20.0;
}

__device__ static void derivs3d_wg(const float2 *spec, long long numdata, double r, double z,
                                   double w, double locpow, double *out7, double2 *red,
                                   double2 *fa, double2 *fb, int npts_cap, int *bad)
{
    const double PI = 3.14159265358979323846, TWOPI = 6.28318530717958647692;
    const double h = 0.003, twoh = 0.006, twohsqrd = 0.000036;
    const int khw = wr_halfwidth(z, w, 1);
    double pwr[5], phs[5];
    double f = r - twoh;
    for (int k = 0; k < 5; k++, f += h) {
        double2 a = rzw_interp_wg(spec, numdata, f, z, w, khw, red, fa, fb, npts_cap, bad);
        pwr[k] = a.x * a.x + a.y * a.y;
        double t = atan2(a.y, a.x);
        phs[k] = t > 0.0 ? t : t + TWOPI;
    }
    for (int k = 0; k < 4; k++)
        if (fabs(phs[k + 1] - phs[k]) > PI)
            phs[k + 1] -= TWOPI;
    out7[0] = pwr[2];
    out7[1] = phs[2];
    out7[2] = (pwr[3] - pwr[1]) / twoh;
    out7[3] = (phs[3] - phs[1]) / twoh;
    out7[4] = (pwr[4] - 2.0 * pwr[2] + pwr[0]) / twohsqrd;
    out7[5] = (phs[4] - 2.0 * phs[2] + phs[0]) / twohsqrd;
    out7[6] = locpow;
}

#define RZW_NR 24    /* 3 * interpfac * ACCEL_RDR */
#define RZW_NZ 17    /* 2 * numwid * interpfac + 1 */
#define RZW_NB 8     /* interpfac * ACCEL_RDR */
#define RZW_NC 7     /* data windows c = -1 .. 5 */
#define RZW_SCR (RZW_NR + 2)   /* scratch rows per candidate: Q tables + 2 FFT buffers */

/* Output j (in 1/8 bins from lobin) correlates the data bins
 * [c - khw, c + khw) around lobin, c = ceil(j / 8).  j in [-8, 47]: a
 * lower harmonic's grid can start up to one bin below its lobin; the
 * reference then indexes before the row it computed (src/maximize_rzw.c:
 * 148-150 with rind < 0) -- here the correlation at that r is used. */
__device__ static int rzw_window(int j)
{
    return j >= 0 ? (j + RZW_NB - 1) / RZW_NB : -((-j) / RZW_NB);
}

/* Workgroup sum of RZW_NR complex values per thread (wave shuffles, then
 * one LDS pass over the 4 waves); lane 0 of wave 0 ... every thread gets
 * the result in out. */
__device__ static void wg_sum_vec(double2 *v, double2 *out, double2 *red4)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int k = 0; k < RZW_NR; k++) {
        double x = v[k].x, y = v[k].y;
        for (int off = 32; off > 0; off >>= 1) {
            x += __shfl_xor(x, off);
            y += __shfl_xor(y, off);
        }
        if (lane == 0)
            red4[wv * RZW_NR + k] = make_double2(x, y);
    }
    __syncthreads();
    if (threadIdx.x < RZW_NR) {
        const int k = threadIdx.x;
        double2 s = red4[k];
        for (int q = 1; q < RF_TPB / 64; q++) {
            s.x += red4[q * RZW_NR + k].x;
            s.y += red4[q * RZW_NR + k].y;
        }
        out[k] = s;
    }
    __syncthreads();
}

__global__ __launch_bounds__(RF_TPB) void k_refine_rzw(agpu_refine_args a)
{
    __shared__ double2 red[RF_TPB];
    __shared__ double2 red4[(RF_TPB / 64) * RZW_NR];
    __shared__ double2 vol[RZW_NR];
    __shared__ float powsum[RZW_NZ * RZW_NZ * RZW_NR];
    __shared__ int rind[RZW_NR];
    const int ci = blockIdx.x;
    const int H = a.numharm[ci];
    const float2 *spec = (const float2 *) a.spec;
    const long long numdata = a.numdata;
    const double rin = a.rin[ci], zin = a.zin[ci], win = a.win[ci];
    /* scratch of this candidate: Q tables [RZW_NR][cap], two FFT buffers */
    double2 *Q = (double2 *) a.scratch + (size_t) ci * RZW_SCR * a.npts_cap;
    double2 *fa = Q + (size_t) RZW_NR * a.npts_cap, *fb = fa + a.npts_cap;
    const double dr = 1.0 / RZW_NB;
    for (int i = threadIdx.x; i < RZW_NZ * RZW_NZ * RZW_NR; i += RF_TPB)
        powsum[i] = 0.0f;
    __syncthreads();
    long long hhlobin = 0;
    int bad = 0;
    /* grid half-width in peak widths: 2 (harmonics), 1 (max_rzw_arr) */
    const int numwid = a.raw ? 1 : 2, ng = 2 * numwid * 4 + 1;
    float locpow0 = 0.0f;
    for (int hind = 0; hind < H && !bad; hind++) {
        const int n = H - hind;
        const double rh = rin * n, zh = zin * n, wh = win * n;
        const double hfrac = n / (double) H;
        const float locpow =
            (float) localpower3d_wg(spec, numdata, rh, zh, wh, red, fa, fb, a.npts_cap, &bad);
        if (hind == 0)
            locpow0 = locpow;
        double rh_int;
        modf(rh, &rh_int);
        const long long lobin = (long long) rh_int - 1;
        if (hind == 0)
            hhlobin = lobin;
        double zlo = zh - numwid * hfrac * 2.0, zhi = zh + numwid * hfrac * 2.0;
        double wlo = wh - numwid * hfrac * 20.0, whi = wh + numwid * hfrac * 20.0;
        const double dz = (zhi - zlo) / (ng - 1), dw = (whi - wlo) / (ng - 1);
        const double maxz = (fabs(zlo) < fabs(zhi)) ? zhi : zlo;
        const double maxw = (fabs(wlo) < fabs(whi)) ? whi : wlo;
        const int khw = wr_halfwidth(maxz, maxw, 0);
        const int numkern = 2 * RZW_NB * khw;
        if (threadIdx.x < RZW_NR)
            rind[threadIdx.x] =
                (int) llround(((hhlobin + threadIdx.x * dr) * hfrac - lobin) * RZW_NB);
        __syncthreads();
        for (int kk = 0; kk < RZW_NR; kk++)
            if (rind[kk] < -RZW_NB || rzw_window(rind[kk]) > RZW_NC - 2)
                bad = 1;
        if (bad)
            break;
        int npts_cur = -1;   /* npts of the Q tables in scratch */
        for (int wi = 0; wi < ng && !bad; wi++) {
            for (int zi = 0; zi < ng; zi++) {
                const double z = zlo + zi * dz, w = wlo + wi * dw;
                if (fabs(w) < 1e-4) {
                    /* z-response template: direct sums with Fresnel taps,
                     * window [ceil(j/8) - khw, ceil(j/8) + khw) around lobin */
                    double2 part[RZW_NR];
                    for (int kk = 0; kk < RZW_NR; kk++) {
                        const int j = rind[kk];
                        const int c = rzw_window(j);
                        double2 acc = make_double2(0.0, 0.0);
                        for (int m = c - khw + (int) threadIdx.x; m < c + khw; m += RF_TPB) {
                            const long long k = lobin + m;
                            if (k < 0 || k >= numdata)
                                continue;
                            const int idx = numkern / 2 + m * RZW_NB - j;
                            const float2 q = resp_tap_nb(0.0, RZW_NB, z, numkern, idx);
                            const float2 d = spec[k];
                            acc.x += (double) d.x * q.x + (double) d.y * q.y;
                            acc.y += (double) d.y * q.x - (double) q.y * d.x;
                        }
                        part[kk] = acc;
                    }
                    wg_sum_vec(part, vol, red4);
                } else {
                    const int npts =
                        (int) next2_dev(6 * wr_halfwidth(z, w, 0) + 200 + numkern / RZW_NB);
                    if (npts > a.npts_cap) {
                        bad = 1;
                        break;
                    }
                    if (npts != npts_cur) {
                        /* Q_kk(n) = i^n e^{-2 pi i j n / (8 npts)} S_c(n), S_c the
                         * DFT of the data window [c - khw, c + khw) around lobin,
                         * c = ceil(j / 8) in [-1, 5] */
                        const double2 *D = window_spectrum(spec, numdata, lobin, -1 - khw,
                                                           khw - 1, npts, fa, fb);
                        for (int nn = threadIdx.x; nn < npts; nn += RF_TPB) {
                            double2 S[RZW_NC];
                            S[0] = make_double2(D[nn].x, -D[nn].y);
                            for (int c = 0; c < RZW_NC - 1; c++) {
                                double2 e0 = window_dft(spec, numdata, lobin, c - 1 - khw,
                                                        c - khw, nn, npts);
                                double2 e1 = window_dft(spec, numdata, lobin, khw + c - 1,
                                                        khw + c, nn, npts);
                                S[c + 1] = make_double2(S[c].x - e0.x + e1.x,
                                                        S[c].y - e0.y + e1.y);
                            }
                            const double2 ph = ipow(nn);
                            for (int kk = 0; kk < RZW_NR; kk++) {
                                const int j = rind[kk];
                                const int c = rzw_window(j);
                                const long long q = ((long long) j * nn) % (RZW_NB * (long long) npts);
                                double sv, cv;
                                sincospi(-2.0 * (double) q / (RZW_NB * (double) npts), &sv, &cv);
                                const double er = ph.x * cv - ph.y * sv, ei = ph.x * sv + ph.y * cv;
                                const double2 Sc = S[c + 1];
                                Q[(size_t) kk * a.npts_cap + nn] =
                                    make_double2(er * Sc.x - ei * Sc.y, er * Sc.y + ei * Sc.x);
                            }
                        }
                        __syncthreads();
                        npts_cur = npts;
                    }
                    double2 part[RZW_NR];
                    for (int kk = 0; kk < RZW_NR; kk++)
                        part[kk] = make_double2(0.0, 0.0);
                    for (int nn = threadIdx.x; nn < npts; nn += RF_TPB) {
                        const double x = chirp_x(nn, npts, 0.0, z, w);
                        for (int kk = 0; kk < RZW_NR; kk++) {
                            const double2 q = Q[(size_t) kk * a.npts_cap + nn];
                            part[kk].x += x * q.x;
                            part[kk].y += x * q.y;
                        }
                    }
                    wg_sum_vec(part, vol, red4);
                }
                /* powsum += |vol|^2 / locpow (src/maximize_rzw.c:145-153),
                 * the raw |vol|^2 for max_rzw_arr (:24-37); the reference's
                 * volume is FP32 */
                if (threadIdx.x < RZW_NR) {
                    const int kk = threadIdx.x;
                    const float vr = (float) vol[kk].x, vi = (float) vol[kk].y;
                    float *ps = &powsum[(wi * RZW_NZ + zi) * RZW_NR + kk];
                    const double pw = (double) vr * vr + (double) vi * vi;
                    *ps += a.raw ? pw : pw / locpow;
                }
                __syncthreads();
            }
        }
        __syncthreads();
    }
    if (bad) {
        if (threadIdx.x == 0) {
            a.rout[ci] = NAN;
            a.nevals[ci] = -1;
        }
        return;
    }
    /* argmax in (w, z, r) order, first maximum wins (src/maximize_rzw.c:160-173);
     * cells outside an ng x ng grid stay 0 and never win */
    float maxpow = 0.0f;
    int best = 0;
    for (int i = 0; i < RZW_NZ * RZW_NZ * RZW_NR; i++)
        if (powsum[i] > maxpow) {
            maxpow = powsum[i];
            best = i;
        }
    const int wind = best / (RZW_NZ * RZW_NR), zind = (best / RZW_NR) % RZW_NZ,
              rix = best % RZW_NR;
    const double nh = H;
    const double rout = (hhlobin + rix * dr) / nh;
    const double zout = ((zin * nh) - numwid * 2.0 + zind * 2.0 / 4.0) / nh;
    const double wout = ((win * nh) - numwid * 20.0 + wind * 20.0 / 4.0) / nh;
    for (int k = 0; k < H; k++) {
        const int hh = k + 1;
        const double rr = rout * hh, zz = zout * hh, ww = wout * hh;
        const float lp = a.raw ? locpow0
                               : (float) localpower3d_wg(spec, numdata, rr, zz, ww, red, fa, fb,
                                                         a.npts_cap, &bad);
        double d7[7];
        derivs3d_wg(spec, numdata, rr, zz, ww, lp, d7, red, fa, fb, a.npts_cap, &bad);
        if (bad)
            break;
        if (threadIdx.x == 0) {
            a.powers[(size_t) ci * RF_MAXH + k] = a.raw ? (double) maxpow : d7[0];
            for (int j = 0; j < 7; j++)
                a.derivs[((size_t) ci * RF_MAXH + k) * 7 + j] = d7[j];
        }
    }
    if (threadIdx.x == 0) {
        a.rout[ci] = bad ? NAN : rout;
        a.zout[ci] = zout;
        a.wout[ci] = wout;
        a.nevals[ci] = bad ? -1 : 0;
    }
}

int agpu_refine_rzw(const agpu_refine_args *args, void *stream)
{
    if (args->ncand <= 0)
        return 0;
    if (args->ncand > 1 << 20 || args->npts_cap <= 0) {
        agpu_set_error("agpu_refine_rzw: bad candidate count or scratch size");
        return -2;
    }
    hipLaunchKernelGGL(k_refine_rzw, dim3(args->ncand), dim3(RF_TPB), 0, (hipStream_t) stream,
                       *args);
    RF_CHECK(hipGetLastError());
    return 0;
}
