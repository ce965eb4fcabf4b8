/*
 * refine_gpu.hip -- FP64 candidate refinement on the GPU: the work of
 * optimize_accelcand (src/accel_utils.c:552-626), spread over many
 * workgroups per candidate.
 *
 *   k_refine_rz_mw : max_rz_arr_harmonics (src/maximize_rz.c:171-228):
 *       local powers, a two-pass Nelder-Mead simplex over (r, z/4)
 *       maximising the locally normalised summed power of all harmonics,
 *       then per harmonic the power and the 5-point derivatives
 *       (get_localpower3d and get_derivs3d, src/characteristics.c:79-192);
 *       one workgroup per harmonic, in lockstep.
 *   k_rzw_* : max_rzw_arr_harmonics (src/maximize_rzw.c:84-191), a grid
 *       search, as a sequence of kernels per harmonic.
 *
 * Every function value is a set of Fourier interpolations (rz_interp,
 * src/rzinterp.c:144-237): the response taps are generated on the fly in
 * FP64 (Fresnel integrals), rounded to FP32 as the reference stores them,
 * multiplied by the FP32 spectrum and summed in FP64 over a workgroup with
 * a fixed-order tree, and the harmonics' terms are summed in harmonic
 * order, so the simplex runs deterministically.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <pthread.h>
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

/* ---- max_rz_arr_harmonics over one workgroup per harmonic ----
 *
 * The simplex is run by H workgroups per candidate in lockstep: at every
 * function evaluation workgroup k forms harmonic k's locally normalised
 * power, stores it in the candidate's group record and waits at the
 * group's barrier; then every workgroup sums the H terms in harmonic
 * order, so all of them hold the same value and take the same simplex
 * step.  The local powers before and the powers and derivatives after
 * the simplex are per harmonic too.  The launcher sizes each launch so
 * that all its workgroups can be resident at once; the barrier's wait is
 * bounded (RZ_SPIN_LIMIT polls), a group whose wait runs out is marked bad
 * and stops, and the launcher runs that candidate again on its own. */
#define RZ_SPIN_LIMIT (1LL << 24)

/* log2 of the first attempt's wait bound; lowered only by tests of the
 * retry path (agpu_refine_set_first_wait) */
static int g_rz_first_log2 = 24;

static pthread_mutex_t g_rz_launch_mu = PTHREAD_MUTEX_INITIALIZER;

struct rz_group {
    unsigned count;              /* barrier arrivals, monotonic */
    int bad;
    double term[2][RF_MAXH];     /* one evaluation's terms, by parity */
};

struct rz_mw {
    rz_group *g;
    int k, H;                    /* this workgroup's harmonic, the candidate's count */
    unsigned arrivals;           /* barriers passed by this workgroup */
    int bad;
    long long spin_limit;        /* polls before the wait gives up */
    float lp;                    /* this harmonic's local power */
};

/* every thread of the workgroup calls it */
__device__ static void rz_group_barrier(rz_mw &m, int *s_flag)
{
    __syncthreads();
    m.arrivals++;
    if (threadIdx.x == 0) {
        int bad = __hip_atomic_load(&m.g->bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!bad) {
            __hip_atomic_fetch_add(&m.g->count, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned target = m.arrivals * (unsigned) m.H;
            long long spins = 0;
            while (__hip_atomic_load(&m.g->count, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) <
                   target) {
                if (__hip_atomic_load(&m.g->bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                    ++spins > m.spin_limit) {
                    __hip_atomic_store(&m.g->bad, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    bad = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        *s_flag = bad;
    }
    __syncthreads();
    m.bad = *s_flag;
}

struct rf_ctx_mw {
    const float2 *spec;
    long long numdata;
    int kernhw;
    double2 *red;
    int nevals;
    rz_mw *m;
    int *s_flag;
    double *s_tot;
};

/* power_call_rz_harmonics_gsl (src/maximize_rz.c:43-63) with harmonic k
 * formed by workgroup k */
__device__ static double rf_fun_mw(rf_ctx_mw &c, double r, double zs)
{
    c.nevals++;
    rz_mw &m = *c.m;
    if (m.bad)
        return 0.0;
    const int n = m.k + 1;
    double2 a = rz_interp_wg(c.spec, c.numdata, r * n, zs * 4.0 * n, c.kernhw, c.red);
    const double term = (a.x * a.x + a.y * a.y) / m.lp;
    if (threadIdx.x == 0)
        __hip_atomic_store(&m.g->term[c.nevals & 1][m.k], term, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    rz_group_barrier(m, c.s_flag);
    if (m.bad)
        return 0.0;
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int k = 0; k < m.H; k++)
            tot += __hip_atomic_load(&m.g->term[c.nevals & 1][k], __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
        *c.s_tot = tot;
    }
    __syncthreads();
    const double tot = *c.s_tot;
    __syncthreads();
    if (!isfinite(tot))
        return 0.0;
    return -tot;
}

/* Nelder-Mead over 2 parameters, the nmsimplex2 scheme: reflect / expand /
 * contract the worst corner, else shrink towards the best; incremental
 * centre and squared rms size; stop when the size drops below tol or after
 * maxiter iterations (src/maximize_rz.c:65-100). */
__device__ static void rf_simplex_mw(rf_ctx_mw &c, double *x, double step, double tol,
                                     int maxiter)
{
    const double P = 3.0;
    double X[3][2], Y[3], C[2], S2;
    X[0][0] = x[0];
    X[0][1] = x[1];
    Y[0] = rf_fun_mw(c, x[0], x[1]);
    for (int i = 0; i < 2; i++) {
        double xt0 = x[0] + (i == 0 ? step : 0.0), xt1 = x[1] + (i == 1 ? step : 0.0);
        X[i + 1][0] = xt0;
        X[i + 1][1] = xt1;
        Y[i + 1] = rf_fun_mw(c, xt0, xt1);
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
        const double ar = 3.0, br = -2.0;
        double r0 = ar * C[0] + br * X[hi][0], r1 = ar * C[1] + br * X[hi][1];
        double val = rf_fun_mw(c, r0, r1);
        if (isfinite(val) && val < Y[lo]) {
            const double ae = 4.5, be = -3.5;
            double e0 = ae * C[0] + be * X[hi][0], e1 = ae * C[1] + be * X[hi][1];
            double val2 = rf_fun_mw(c, e0, e1);
            if (isfinite(val2) && val2 < Y[lo])
                update(hi, e0, e1, val2);
            else
                update(hi, r0, r1, val);
        } else if (!isfinite(val) || val > ds_hi) {
            if (isfinite(val) && val <= Y[hi])
                update(hi, r0, r1, val);
            const double ac = 0.75, bc = 0.25;
            double k0 = ac * C[0] + bc * X[hi][0], k1 = ac * C[1] + bc * X[hi][1];
            double val2 = rf_fun_mw(c, k0, k1);
            if (isfinite(val2) && val2 <= Y[hi]) {
                update(hi, k0, k1, val2);
            } else {
                for (int i = 0; i < 3; i++) {
                    if (i == lo)
                        continue;
                    X[i][0] = 0.5 * (X[i][0] + X[lo][0]);
                    X[i][1] = 0.5 * (X[i][1] + X[lo][1]);
                    Y[i] = rf_fun_mw(c, X[i][0], X[i][1]);
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
    } while (!(size < tol) && it < maxiter && !c.m->bad);
    x[0] = X[lo][0];
    x[1] = X[lo][1];
}

/* blockIdx.x = candidate * hmax + harmonic index; a.scratch: one rz_group
 * per candidate, zeroed by the launcher */
__global__ __launch_bounds__(RF_TPB) void k_refine_rz_mw(agpu_refine_args a, int hmax,
                                                         long long spin_limit)
{
    __shared__ double2 red[RF_TPB];
    __shared__ int s_flag;
    __shared__ double s_tot;
    const int ci = blockIdx.x / hmax, k = blockIdx.x % hmax;
    const int H = a.numharm[ci];
    if (k >= H)
        return;
    const float2 *spec = (const float2 *) a.spec;
    const double rin = a.rin[ci], zin = a.zin[ci];
    rz_mw m;
    m.g = (rz_group *) a.scratch + ci;
    m.k = k;
    m.H = H;
    m.arrivals = 0;
    m.bad = 0;
    m.spin_limit = spin_limit;
    const int n = k + 1;
    {
        double lp = localpower_wg(spec, a.numdata, rin * n, zin * n, red);
        m.lp = a.raw ? 1.0f : (float) lp;
    }
    rf_ctx_mw c{spec, a.numdata, 0, red, 0, &m, &s_flag, &s_tot};
    double x[2] = {rin, zin / 4.0};
    c.kernhw = zr_halfwidth(fabs(zin * H) + 4.0, 0);
    rf_simplex_mw(c, x, 0.4, 1.0e-4, 5000);
    c.kernhw = zr_halfwidth(fabs(x[1] * H) + 4.0, 1);
    rf_simplex_mw(c, x, 0.01, 1.0e-7, 5000);
    const double rout = x[0], zout = x[1] * 4.0;
    __syncthreads();
    const float lp = (float) localpower_wg(spec, a.numdata, rout * n, zout * n, red);
    double2 amp = rz_interp_wg(spec, a.numdata, rout * n, zout / 4.0 * n * 4.0, c.kernhw, red);
    double d7[7];
    derivs_wg(spec, a.numdata, rout * n, zout * n, lp, d7, red);
    if (threadIdx.x == 0) {
        a.powers[(size_t) ci * RF_MAXH + k] = amp.x * amp.x + amp.y * amp.y;
        for (int j = 0; j < 7; j++)
            a.derivs[((size_t) ci * RF_MAXH + k) * 7 + j] = d7[j];
        if (k == 0) {
            a.rout[ci] = m.bad ? NAN : rout;
            a.zout[ci] = zout;
            a.nevals[ci] = m.bad ? -1 : c.nevals;
        }
    }
}

size_t agpu_refine_rz_scratch(int ncand)
{
    return sizeof(rz_group) * (size_t) (ncand > 0 ? ncand : 1);
}

static hipError_t rz_launch(const agpu_refine_args *args, int c0, int m, int hmax,
                            long long spin_limit, hipStream_t st);

void agpu_refine_set_first_wait(int log2_polls)
{
    g_rz_first_log2 = log2_polls < 0 ? 0 : (log2_polls > 24 ? 24 : log2_polls);
}

int agpu_refine_rz(const agpu_refine_args *args, void *stream)
{
    if (args->ncand <= 0)
        return 0;
    if (args->ncand > 1 << 20) {
        agpu_set_error("agpu_refine_rz: too many candidates");
        return -2;
    }
    if (!args->scratch) {
        agpu_set_error("agpu_refine_rz: no group scratch (agpu_refine_rz_scratch bytes)");
        return -2;
    }
    hipStream_t st = (hipStream_t) stream;
    int hmax = args->maxharm;
    if (hmax < 1 || hmax > RF_MAXH) {
        agpu_set_error("agpu_refine_rz: harmonic bound out of range");
        return -2;
    }
    /* The workgroups of a candidate wait for each other, so every
     * workgroup of a launch must be resident at once: a launch holds at
     * most as many workgroups as the device keeps resident, and launches of
     * this kernel are serialised across the process (lock held until the
     * launch completes).  Other kernels never wait, so they only delay a
     * workgroup's start. */
    int dev = 0, ncu = 0, per_cu = 0;
    RF_CHECK(hipGetDevice(&dev));
    RF_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    RF_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_refine_rz_mw, RF_TPB, 0));
    const long long resident = (long long) ncu * per_cu;
    const int chunk = (int) (resident / hmax);
    if (chunk < 1) {
        agpu_set_error("agpu_refine_rz: a candidate's workgroups do not fit on the device");
        return -2;
    }
    pthread_mutex_lock(&g_rz_launch_mu);
    hipError_t e = hipMemsetAsync(args->scratch, 0, agpu_refine_rz_scratch(args->ncand), st);
    int *nev = NULL;
    for (int c0 = 0; c0 < args->ncand && e == hipSuccess; c0 += chunk) {
        const int m = args->ncand - c0 < chunk ? args->ncand - c0 : chunk;
        e = rz_launch(args, c0, m, hmax, 1LL << g_rz_first_log2, st);
    }
    /* a candidate whose group's wait ran out (its workgroups could not all
     * be resident, e.g. under another process's kernels) runs again on its
     * own, up to three times; one that still fails is reported by its
     * nevals of -1 */
    if (e == hipSuccess && !(nev = (int *) malloc(sizeof(int) * (size_t) args->ncand)))
        e = hipErrorOutOfMemory;
    if (e == hipSuccess)
        e = hipMemcpyAsync(nev, args->nevals, sizeof(int) * (size_t) args->ncand,
                           hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    for (int c = 0; c < args->ncand && e == hipSuccess; c++)
        for (int attempt = 0; attempt < 3 && nev[c] < 0 && e == hipSuccess; attempt++) {
            e = hipMemsetAsync((rz_group *) args->scratch + c, 0, sizeof(rz_group), st);
            if (e == hipSuccess)
                e = rz_launch(args, c, 1, hmax, RZ_SPIN_LIMIT, st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(nev + c, args->nevals + c, sizeof(int),
                                   hipMemcpyDeviceToHost, st);
            if (e == hipSuccess)
                e = hipStreamSynchronize(st);
        }
    free(nev);
    pthread_mutex_unlock(&g_rz_launch_mu);
    if (e != hipSuccess) {
        char m_[256];
        snprintf(m_, sizeof(m_), "agpu_refine_rz: %s", hipGetErrorString(e));
        agpu_set_error(m_);
        return -1;
    }
    return 0;
}

/* One launch of k_refine_rz_mw over candidates [c0, c0 + m), waited for. */
static hipError_t rz_launch(const agpu_refine_args *args, int c0, int m, int hmax,
                            long long spin_limit, hipStream_t st)
{
    agpu_refine_args a = *args;
    a.ncand = m;
    a.numharm += c0;
    a.rin += c0;
    a.zin += c0;
    if (a.win)
        a.win += c0;
    a.rout += c0;
    a.zout += c0;
    if (a.wout)
        a.wout += c0;
    a.derivs += (size_t) c0 * RF_MAXH * 7;
    a.powers += (size_t) c0 * RF_MAXH;
    a.nevals += c0;
    a.scratch = (void *) ((rz_group *) args->scratch + c0);
    hipLaunchKernelGGL(k_refine_rz_mw, dim3(m * hmax), dim3(RF_TPB), 0, st, a, hmax,
                       spin_limit);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    return e;
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
#define RZW_NLP 20   /* interpolations of a local power (NUMLOCPOWAVG) */
#define RZW_NPT 25   /* per candidate: the local power's 20 plus the 5 of the derivatives */
#define RZW_NV 21    /* distinct chirp lengths of one grid: powers of two up to 2^20 */
#define RZW_CELLS (RZW_NZ * RZW_NZ * RZW_NR)

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
 * one LDS pass over the 4 waves); every thread gets the result in out. */
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

/* max_rzw_arr_harmonics (src/maximize_rzw.c:84-191) is spread over many
 * workgroups per candidate, harmonic by harmonic, from highest to lowest
 * as the reference loops:
 *   k_rzw_points  the 20 interpolations of the harmonic's local power (one
 *                 workgroup each);
 *   k_rzw_setup   the local power, lobin, the (z, w) grid, the r indices
 *                 and the distinct chirp lengths of the grid;
 *   k_rzw_qtab    per chirp length, the tables Q_j(n) = i^n
 *                 e^{-2 pi i j n / (8 npts)} S_c(n), S_c the DFT of the data
 *                 window [c - khw, c + khw) around lobin (c = ceil(j / 8) in
 *                 [-1, 5]): with R the FFT of the zero-padded FP32 chirp x,
 *                 the reference's correlation sum_k d_k conj(R(k - r_j))
 *                 regroups exactly into sum_n x_n Q_j(n);
 *   k_rzw_grid    one workgroup per (w, z) grid point: the 24 correlations
 *                 and their powers over the local power.
 * Then k_rzw_pick sums the powers over the harmonics in the reference's
 * order (an FP32 sum) and takes the first maximum in (w, z, r) order, and
 * per harmonic of the result k_rzw_points and k_rzw_derivs give the local
 * power, the power and the 5-point derivatives.  Every value is formed
 * exactly as by one workgroup running the whole loop: same expressions,
 * same fixed-order workgroup sums. */
typedef struct {
    int bad, H;
    long long hhlobin;
    float locpow0, maxpow;
    double rout, zout, wout;
    /* the current harmonic */
    int khw, numkern, ng, nvar;
    long long lobin;
    double zlo, dz, wlo, dw;
    float locpow;
    int npts[RZW_NV];
    int rind[RZW_NR];
} rzw_state;

static size_t rzw_align(size_t b) { return (b + 255) & ~(size_t) 255; }

/* scratch of a launch for ncand candidates, by part */
struct rzw_layout {
    size_t st, pts, terms, q, fft, total;
};

static rzw_layout rzw_layout_of(int ncand, int npts_cap)
{
    rzw_layout l;
    size_t o = 0;
    l.st = o;
    o += rzw_align(sizeof(rzw_state) * (size_t) ncand);
    l.pts = o;
    o += rzw_align(sizeof(double2) * RZW_NPT * (size_t) ncand);
    l.terms = o;
    o += rzw_align(sizeof(double) * (size_t) RF_MAXH * RZW_CELLS * ncand);
    l.q = o;
    o += rzw_align(sizeof(double2) * RZW_NR * 2 * (size_t) npts_cap * ncand);
    l.fft = o;
    o += rzw_align(sizeof(double2) * 2 * (size_t) npts_cap * RZW_NPT * ncand);
    l.total = o;
    return l;
}

struct rzw_ptrs {
    rzw_state *st;
    double2 *pts;
    double *terms;
    double2 *q, *fft;
};

__device__ static rzw_ptrs rzw_ptrs_of(const agpu_refine_args &a, size_t o_st, size_t o_pts,
                                       size_t o_terms, size_t o_q, size_t o_fft)
{
    char *b = (char *) a.scratch;
    rzw_ptrs p;
    p.st = (rzw_state *) (b + o_st);
    p.pts = (double2 *) (b + o_pts);
    p.terms = (double *) (b + o_terms);
    p.q = (double2 *) (b + o_q);
    p.fft = (double2 *) (b + o_fft);
    return p;
}

#define RZW_LAYOUT_ARGS size_t o_st, size_t o_pts, size_t o_terms, size_t o_q, size_t o_fft
#define RZW_PTRS rzw_ptrs_of(a, o_st, o_pts, o_terms, o_q, o_fft)

/* the FFT buffers of slot k of candidate c */
__device__ static double2 *rzw_fa(const rzw_ptrs &P, const agpu_refine_args &a, int c, int k)
{
    return P.fft + ((size_t) c * RZW_NPT + k) * 2 * (size_t) a.npts_cap;
}

__global__ __launch_bounds__(RF_TPB) void k_rzw_init(agpu_refine_args a, RZW_LAYOUT_ARGS)
{
    const rzw_ptrs P = RZW_PTRS;
    const int c = blockIdx.x;
    rzw_state *st = P.st + c;
    double *terms = P.terms + (size_t) c * RF_MAXH * RZW_CELLS;
    for (int i = threadIdx.x; i < RF_MAXH * RZW_CELLS; i += RF_TPB)
        terms[i] = 0.0;
    if (threadIdx.x == 0) {
        st->bad = 0;
        st->H = a.numharm[c];
        st->locpow0 = 0.0f;
        st->maxpow = 0.0f;
    }
}

/* mode 0: the local-power interpolations of harmonic n = H - hind at the
 * start (rin, zin, win) n; mode 1: those of harmonic hind + 1 of the
 * result (slots 0-19, unless raw) and its 5 derivative interpolations
 * (slots 20-24, src/characteristics.c:140-192).  One workgroup per
 * (candidate, slot); each point is formed by the loop arithmetic of
 * get_localpower3d / get_derivs3d (src/characteristics.c:79-136). */
__global__ __launch_bounds__(RF_TPB) void k_rzw_points(agpu_refine_args a, RZW_LAYOUT_ARGS,
                                                       int hind, int mode)
{
    __shared__ double2 red[RF_TPB];
    const rzw_ptrs P = RZW_PTRS;
    const int c = blockIdx.x / RZW_NPT, slot = blockIdx.x % RZW_NPT;
    rzw_state *st = P.st + c;
    const int H = st->H;
    if (st->bad || hind >= H)
        return;
    if (mode == 0 && slot >= RZW_NLP)
        return;
    if (mode == 1 && slot < RZW_NLP && a.raw)
        return;
    double r, z, w;
    if (mode == 0) {
        const int n = H - hind;
        r = a.rin[c] * n;
        z = a.zin[c] * n;
        w = a.win[c] * n;
    } else {
        const int hh = hind + 1;
        r = st->rout * hh;
        z = st->zout * hh;
        w = st->wout * hh;
    }
    const float2 *spec = (const float2 *) a.spec;
    double2 *fa = rzw_fa(P, a, c, slot), *fb = fa + a.npts_cap;
    int bad = 0;
    double2 v = make_double2(0.0, 0.0);
    int have = 0;
    if (slot < RZW_NLP) {
        const int khw = wr_halfwidth(z, w, 0);
        double lo1 = r - 5 - 10, hi1 = lo1 + 10;
        double lo2 = r + 5, hi2 = lo2 + 10;
        const double nd = (double) a.numdata;
        lo1 = lo1 < 0.0 ? 0.0 : (lo1 > nd ? nd : lo1);
        lo2 = lo2 < 0.0 ? 0.0 : (lo2 > nd ? nd : lo2);
        hi1 = hi1 < 0.0 ? 0.0 : (hi1 > nd ? nd : hi1);
        hi2 = hi2 < 0.0 ? 0.0 : (hi2 > nd ? nd : hi2);
        /* point slot of the two loops, stepped as they step */
        int k = 0;
        for (double f = lo1; f < hi1; f += 1.0, k++)
            if (k == slot) {
                v = rzw_interp_wg(spec, a.numdata, f, z, w, khw, red, fa, fb, a.npts_cap, &bad);
                have = 1;
            }
        for (double f = lo2; f < hi2 && !have; f += 1.0, k++)
            if (k == slot) {
                v = rzw_interp_wg(spec, a.numdata, f, z, w, khw, red, fa, fb, a.npts_cap, &bad);
                have = 1;
            }
    } else {
        const double h = 0.003, twoh = 0.006;
        const int khw = wr_halfwidth(z, w, 1);
        double f = r - twoh;
        for (int k = 0; k < 5; k++, f += h)
            if (k == slot - RZW_NLP)
                v = rzw_interp_wg(spec, a.numdata, f, z, w, khw, red, fa, fb, a.npts_cap, &bad);
        have = 1;
    }
    if (threadIdx.x == 0) {
        P.pts[(size_t) c * RZW_NPT + slot] = v;
        if (bad)
            st->bad = 1;
    }
}

/* local power of the points of slots 0-19 (the loops' sum, over 20) */
__device__ static double rzw_locpow_sum(const double2 *pts, double r, long long numdata)
{
    double lo1 = r - 5 - 10, hi1 = lo1 + 10;
    double lo2 = r + 5, hi2 = lo2 + 10;
    const double nd = (double) numdata;
    lo1 = lo1 < 0.0 ? 0.0 : (lo1 > nd ? nd : lo1);
    lo2 = lo2 < 0.0 ? 0.0 : (lo2 > nd ? nd : lo2);
    hi1 = hi1 < 0.0 ? 0.0 : (hi1 > nd ? nd : hi1);
    hi2 = hi2 < 0.0 ? 0.0 : (hi2 > nd ? nd : hi2);
    double sum = 0.0;
    int k = 0;
    for (double f = lo1; f < hi1; f += 1.0, k++) {
        const double2 v = pts[k];
        sum += v.x * v.x + v.y * v.y;
    }
    for (double f = lo2; f < hi2; f += 1.0, k++) {
        const double2 v = pts[k];
        sum += v.x * v.x + v.y * v.y;
    }
    return sum / 20.0;
}

/* harmonic n = H - hind: local power, lobin, grid, r indices, chirp lengths */
__global__ __launch_bounds__(RF_TPB) void k_rzw_setup(agpu_refine_args a, RZW_LAYOUT_ARGS,
                                                      int hind)
{
    const rzw_ptrs P = RZW_PTRS;
    const int c = blockIdx.x;
    rzw_state *st = P.st + c;
    const int H = st->H;
    if (st->bad || hind >= H || threadIdx.x != 0)
        return;
    const int n = H - hind;
    const double rh = a.rin[c] * n, zh = a.zin[c] * n, wh = a.win[c] * n;
    const double hfrac = n / (double) H;
    const double dr = 1.0 / RZW_NB;
    const int numwid = a.raw ? 1 : 2, ng = 2 * numwid * 4 + 1;
    const float locpow = (float) rzw_locpow_sum(P.pts + (size_t) c * RZW_NPT, rh, a.numdata);
    if (hind == 0)
        st->locpow0 = locpow;
    double rh_int;
    modf(rh, &rh_int);
    const long long lobin = (long long) rh_int - 1;
    if (hind == 0)
        st->hhlobin = lobin;
    const long long hhlobin = st->hhlobin;
    const double zlo = zh - numwid * hfrac * 2.0, zhi = zh + numwid * hfrac * 2.0;
    const double wlo = wh - numwid * hfrac * 20.0, whi = wh + numwid * hfrac * 20.0;
    const double dz = (zhi - zlo) / (ng - 1), dw = (whi - wlo) / (ng - 1);
    const double maxz = (fabs(zlo) < fabs(zhi)) ? zhi : zlo;
    const double maxw = (fabs(wlo) < fabs(whi)) ? whi : wlo;
    const int khw = wr_halfwidth(maxz, maxw, 0);
    const int numkern = 2 * RZW_NB * khw;
    int bad = 0;
    for (int kk = 0; kk < RZW_NR; kk++) {
        const int rk = (int) llround(((hhlobin + kk * dr) * hfrac - lobin) * RZW_NB);
        st->rind[kk] = rk;
        if (rk < -RZW_NB || rzw_window(rk) > RZW_NC - 2)
            bad = 1;
    }
    /* the distinct chirp lengths of the grid, ascending */
    int nvar = 0;
    for (int wi = 0; wi < ng && !bad; wi++)
        for (int zi = 0; zi < ng; zi++) {
            const double z = zlo + zi * dz, w = wlo + wi * dw;
            if (fabs(w) < 1e-4)
                continue;
            const int npts = (int) next2_dev(6 * wr_halfwidth(z, w, 0) + 200 + numkern / RZW_NB);
            if (npts > a.npts_cap) {
                bad = 1;
                break;
            }
            int at = 0;
            while (at < nvar && st->npts[at] < npts)
                at++;
            if (at < nvar && st->npts[at] == npts)
                continue;
            for (int q = nvar; q > at; q--)
                st->npts[q] = st->npts[q - 1];
            st->npts[at] = npts;
            nvar++;
        }
    st->nvar = nvar;
    st->khw = khw;
    st->numkern = numkern;
    st->ng = ng;
    st->lobin = lobin;
    st->zlo = zlo;
    st->dz = dz;
    st->wlo = wlo;
    st->dw = dw;
    st->locpow = locpow;
    if (bad)
        st->bad = 1;
}

/* offset (complex values) of chirp-length variant v's Q tables */
__device__ static size_t rzw_qoff(const rzw_state *st, int v)
{
    size_t o = 0;
    for (int u = 0; u < v; u++)
        o += (size_t) RZW_NR * st->npts[u];
    return o;
}

__global__ __launch_bounds__(RF_TPB) void k_rzw_qtab(agpu_refine_args a, RZW_LAYOUT_ARGS,
                                                     int hind)
{
    const rzw_ptrs P = RZW_PTRS;
    const int c = blockIdx.x / RZW_NV, v = blockIdx.x % RZW_NV;
    const rzw_state *st = P.st + c;
    if (st->bad || hind >= st->H || v >= st->nvar)
        return;
    const float2 *spec = (const float2 *) a.spec;
    const long long numdata = a.numdata, lobin = st->lobin;
    const int khw = st->khw, npts = st->npts[v];
    double2 *fa = rzw_fa(P, a, c, v), *fb = fa + a.npts_cap;
    double2 *Q = P.q + (size_t) c * RZW_NR * 2 * a.npts_cap + rzw_qoff(st, v);
    const double2 *D = window_spectrum(spec, numdata, lobin, -1 - khw, khw - 1, npts, fa, fb);
    for (int nn = threadIdx.x; nn < npts; nn += RF_TPB) {
        double2 S[RZW_NC];
        S[0] = make_double2(D[nn].x, -D[nn].y);
        for (int cc = 0; cc < RZW_NC - 1; cc++) {
            double2 e0 = window_dft(spec, numdata, lobin, cc - 1 - khw, cc - khw, nn, npts);
            double2 e1 = window_dft(spec, numdata, lobin, khw + cc - 1, khw + cc, nn, npts);
            S[cc + 1] = make_double2(S[cc].x - e0.x + e1.x, S[cc].y - e0.y + e1.y);
        }
        const double2 ph = ipow(nn);
        for (int kk = 0; kk < RZW_NR; kk++) {
            const int j = st->rind[kk];
            const int cw = rzw_window(j);
            const long long q = ((long long) j * nn) % (RZW_NB * (long long) npts);
            double sv, cv;
            sincospi(-2.0 * (double) q / (RZW_NB * (double) npts), &sv, &cv);
            const double er = ph.x * cv - ph.y * sv, ei = ph.x * sv + ph.y * cv;
            const double2 Sc = S[cw + 1];
            Q[(size_t) kk * npts + nn] =
                make_double2(er * Sc.x - ei * Sc.y, er * Sc.y + ei * Sc.x);
        }
    }
}

/* one (w, z) grid point of harmonic n = H - hind: powsum's term
 * |vol|^2 / locpow (src/maximize_rzw.c:145-153), the raw |vol|^2 for
 * max_rzw_arr (:24-37); the reference's volume is FP32 */
__global__ __launch_bounds__(RF_TPB) void k_rzw_grid(agpu_refine_args a, RZW_LAYOUT_ARGS,
                                                     int hind)
{
    __shared__ double2 red4[(RF_TPB / 64) * RZW_NR];
    __shared__ double2 vol[RZW_NR];
    const rzw_ptrs P = RZW_PTRS;
    const int c = blockIdx.x / (RZW_NZ * RZW_NZ), g = blockIdx.x % (RZW_NZ * RZW_NZ);
    const int wi = g / RZW_NZ, zi = g % RZW_NZ;
    const rzw_state *st = P.st + c;
    if (st->bad || hind >= st->H || wi >= st->ng || zi >= st->ng)
        return;
    const float2 *spec = (const float2 *) a.spec;
    const long long numdata = a.numdata, lobin = st->lobin;
    const int khw = st->khw, numkern = st->numkern;
    const double z = st->zlo + zi * st->dz, w = st->wlo + wi * st->dw;
    double2 part[RZW_NR];
    if (fabs(w) < 1e-4) {
        /* z-response template: direct sums with Fresnel taps, window
         * [ceil(j/8) - khw, ceil(j/8) + khw) around lobin */
        for (int kk = 0; kk < RZW_NR; kk++) {
            const int j = st->rind[kk];
            const int cw = rzw_window(j);
            double2 acc = make_double2(0.0, 0.0);
            for (int m = cw - khw + (int) threadIdx.x; m < cw + khw; m += RF_TPB) {
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
    } else {
        const int npts = (int) next2_dev(6 * wr_halfwidth(z, w, 0) + 200 + numkern / RZW_NB);
        int v = 0;
        while (v < st->nvar && st->npts[v] != npts)
            v++;
        if (v == st->nvar)
            return;   /* not reached: setup listed every length of the grid */
        const double2 *Q = P.q + (size_t) c * RZW_NR * 2 * a.npts_cap + rzw_qoff(st, v);
        for (int kk = 0; kk < RZW_NR; kk++)
            part[kk] = make_double2(0.0, 0.0);
        for (int nn = threadIdx.x; nn < npts; nn += RF_TPB) {
            const double x = chirp_x(nn, npts, 0.0, z, w);
            for (int kk = 0; kk < RZW_NR; kk++) {
                const double2 q = Q[(size_t) kk * npts + nn];
                part[kk].x += x * q.x;
                part[kk].y += x * q.y;
            }
        }
    }
    wg_sum_vec(part, vol, red4);
    if (threadIdx.x < RZW_NR) {
        const int kk = threadIdx.x;
        const float vr = (float) vol[kk].x, vi = (float) vol[kk].y;
        const double pw = (double) vr * vr + (double) vi * vi;
        P.terms[((size_t) c * RF_MAXH + hind) * RZW_CELLS + (wi * RZW_NZ + zi) * RZW_NR + kk] =
            a.raw ? pw : pw / st->locpow;
    }
}

/* powsum in the reference's harmonic order, first maximum in (w, z, r)
 * order (src/maximize_rzw.c:160-173); cells outside an ng x ng grid stay
 * 0 and never win */
__global__ __launch_bounds__(RF_TPB) void k_rzw_pick(agpu_refine_args a, RZW_LAYOUT_ARGS)
{
    __shared__ float powsum[RZW_CELLS];
    const rzw_ptrs P = RZW_PTRS;
    const int c = blockIdx.x;
    rzw_state *st = P.st + c;
    if (st->bad)
        return;
    const int H = st->H;
    const double *terms = P.terms + (size_t) c * RF_MAXH * RZW_CELLS;
    for (int i = threadIdx.x; i < RZW_CELLS; i += RF_TPB) {
        float ps = 0.0f;
        for (int hind = 0; hind < H; hind++)
            ps += terms[(size_t) hind * RZW_CELLS + i];
        powsum[i] = ps;
    }
    __syncthreads();
    if (threadIdx.x != 0)
        return;
    float maxpow = 0.0f;
    int best = 0;
    for (int i = 0; i < RZW_CELLS; i++)
        if (powsum[i] > maxpow) {
            maxpow = powsum[i];
            best = i;
        }
    const int numwid = a.raw ? 1 : 2;
    const int wind = best / (RZW_NZ * RZW_NR), zind = (best / RZW_NR) % RZW_NZ,
              rix = best % RZW_NR;
    const double nh = H, dr = 1.0 / RZW_NB;
    const double zin = a.zin[c], win = a.win[c];
    st->rout = (st->hhlobin + rix * dr) / nh;
    st->zout = ((zin * nh) - numwid * 2.0 + zind * 2.0 / 4.0) / nh;
    st->wout = ((win * nh) - numwid * 20.0 + wind * 20.0 / 4.0) / nh;
    st->maxpow = maxpow;
}

/* harmonic k + 1 of the result: local power (or the start's, raw), power
 * and the 5-point derivatives (src/characteristics.c:140-192) */
__global__ void k_rzw_derivs(agpu_refine_args a, RZW_LAYOUT_ARGS, int k)
{
    const rzw_ptrs P = RZW_PTRS;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.ncand)
        return;
    const rzw_state *st = P.st + c;
    if (st->bad || k >= st->H)
        return;
    const double PI = 3.14159265358979323846, TWOPI = 6.28318530717958647692;
    const double twoh = 0.006, twohsqrd = 0.000036;
    const int hh = k + 1;
    const double2 *pts = P.pts + (size_t) c * RZW_NPT;
    const float lp = a.raw ? st->locpow0
                           : (float) rzw_locpow_sum(pts, st->rout * hh, a.numdata);
    const double locpow = lp;
    double pwr[5], phs[5];
    for (int q = 0; q < 5; q++) {
        const double2 v = pts[RZW_NLP + q];
        pwr[q] = v.x * v.x + v.y * v.y;
        double t = atan2(v.y, v.x);
        phs[q] = t > 0.0 ? t : t + TWOPI;
    }
    for (int q = 0; q < 4; q++)
        if (fabs(phs[q + 1] - phs[q]) > PI)
            phs[q + 1] -= TWOPI;
    double d7[7];
    d7[0] = pwr[2];
    d7[1] = phs[2];
    d7[2] = (pwr[3] - pwr[1]) / twoh;
    d7[3] = (phs[3] - phs[1]) / twoh;
    d7[4] = (pwr[4] - 2.0 * pwr[2] + pwr[0]) / twohsqrd;
    d7[5] = (phs[4] - 2.0 * phs[2] + phs[0]) / twohsqrd;
    d7[6] = locpow;
    a.powers[(size_t) c * RF_MAXH + k] = a.raw ? (double) st->maxpow : d7[0];
    for (int j = 0; j < 7; j++)
        a.derivs[((size_t) c * RF_MAXH + k) * 7 + j] = d7[j];
}

__global__ void k_rzw_done(agpu_refine_args a, RZW_LAYOUT_ARGS)
{
    const rzw_ptrs P = RZW_PTRS;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.ncand)
        return;
    const rzw_state *st = P.st + c;
    a.rout[c] = st->bad ? NAN : st->rout;
    a.zout[c] = st->zout;
    a.wout[c] = st->wout;
    a.nevals[c] = st->bad ? -1 : 0;
}

size_t agpu_refine_rzw_scratch(int ncand, int npts_cap)
{
    return rzw_layout_of(ncand, npts_cap).total;
}

int agpu_refine_rzw(const agpu_refine_args *args, void *stream)
{
    const int maxharm = args->maxharm;
    if (args->ncand <= 0)
        return 0;
    if (args->ncand > 1 << 16 || args->npts_cap <= 0 || maxharm < 1 || maxharm > RF_MAXH) {
        agpu_set_error("agpu_refine_rzw: bad candidate count, harmonic count or scratch size");
        return -2;
    }
    const agpu_refine_args a = *args;
    const rzw_layout l = rzw_layout_of(a.ncand, a.npts_cap);
    hipStream_t st = (hipStream_t) stream;
    const int m = a.ncand, nb1 = (m + 63) / 64;
#define RZW_L l.st, l.pts, l.terms, l.q, l.fft
    hipLaunchKernelGGL(k_rzw_init, dim3(m), dim3(RF_TPB), 0, st, a, RZW_L);
    for (int hind = 0; hind < maxharm; hind++) {
        hipLaunchKernelGGL(k_rzw_points, dim3(m * RZW_NPT), dim3(RF_TPB), 0, st, a, RZW_L, hind, 0);
        hipLaunchKernelGGL(k_rzw_setup, dim3(m), dim3(64), 0, st, a, RZW_L, hind);
        hipLaunchKernelGGL(k_rzw_qtab, dim3(m * RZW_NV), dim3(RF_TPB), 0, st, a, RZW_L, hind);
        hipLaunchKernelGGL(k_rzw_grid, dim3(m * RZW_NZ * RZW_NZ), dim3(RF_TPB), 0, st, a, RZW_L,
                           hind);
    }
    hipLaunchKernelGGL(k_rzw_pick, dim3(m), dim3(RF_TPB), 0, st, a, RZW_L);
    for (int k = 0; k < maxharm; k++) {
        hipLaunchKernelGGL(k_rzw_points, dim3(m * RZW_NPT), dim3(RF_TPB), 0, st, a, RZW_L, k, 1);
        hipLaunchKernelGGL(k_rzw_derivs, dim3(nb1), dim3(64), 0, st, a, RZW_L, k);
    }
    hipLaunchKernelGGL(k_rzw_done, dim3(nb1), dim3(64), 0, st, a, RZW_L);
#undef RZW_L
    RF_CHECK(hipGetLastError());
    return 0;
}
