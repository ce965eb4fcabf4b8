/*
 * fold_gpu.hip -- prepfold's fold of a time series for an ACCEL candidate
 * (prepfold -accelcand on a .dat), on the GPU.
 *
 * The reference (src/fold.c:46-107, 443-653; src/prepfold_pipeline.c:
 * 1375-1594) folds records of 1024 points one after another: each point
 * covers the phase interval between its start and end times, and its value
 * is "drizzled" over the profile bins that interval covers, in proportion
 * to the overlap.  Contributions first go to a buffer that is added to the
 * current sub-integration's profile each time one full turn of phase has
 * been added since the fold started (the buffer and the phase added carry
 * across records and sub-integrations); what the last, incomplete turn
 * holds is never added.
 *
 * Here every point is independent: its phase interval [phi(t), phi(t+dt))
 * comes from the FP64 phase polynomial, and each piece of it (split at bin
 * edges and at whole turns of the phase added since the start) goes to the
 * profile of the sub-integration during which its turn completes -- the
 * one whose end phase first reaches the turn's end.  One workgroup folds
 * one record (all its points lie in one sub-integration) into LDS
 * accumulators for that sub-integration and the next, with FP64 LDS atomics,
 * then adds them to the global profiles; turns that complete two or more
 * sub-integrations later (periods longer than a record) go straight to
 * global memory.  The sums are FP64, in an order that varies from run to
 * run in the last bits.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include "agpu.h"

#define FD_TPB 256
#define FD_REC 1024            /* SUBSBLOCKLEN */
#define FD_LDS_MAXPROF 2048    /* profiles longer than this go to global atomics */

#define FD_CHECK(call)                                                            \
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

/* phase at time t: t (t (t fdd/6 + fd/2) + f) (src/fold.c:530, 557) */
__device__ static double fd_phase(double t, double f, double fd2, double fdd6)
{
    return t * (t * (t * fdd6 + fd2) + f);
}

__global__ __launch_bounds__(FD_TPB) void k_fold(agpu_fold_args a)
{
    __shared__ double acc[2 * FD_LDS_MAXPROF];
    __shared__ double red[2][FD_TPB / 64];
    const int rec = blockIdx.x;                  /* record */
    const int part = rec / a.rpp, jj = rec % a.rpp;
    const int n = a.proflen;
    const bool use_lds = n <= FD_LDS_MAXPROF;
    if (use_lds)
        for (int k = threadIdx.x; k < 2 * n; k += FD_TPB)
            acc[k] = 0.0;
    __syncthreads();
    const double proftime = FD_REC * a.dt;
    /* fold_time0 of the record, as src/prepfold_pipeline.c:1515, 1577 */
    const double tlo = (double) ((long long) part * a.rpp) * proftime + jj * proftime;
    const double fd2 = a.fd / 2.0, fdd6 = a.fdd / 6.0;
    const double phi0 = a.phi_start;
    double s1 = 0.0, s2 = 0.0;
    for (int l = threadIdx.x; l < FD_REC; l += FD_TPB) {
        const long long i = (long long) rec * FD_REC + l;
        const double x = a.series[i];
        s1 += x;
        s2 += x * x;
        const double t = tlo + l * a.dt, tn = tlo + (l + 1) * a.dt;
        const double ph = fd_phase(t, a.f, fd2, fdd6), phn = fd_phase(tn, a.f, fd2, fdd6);
        double dph = phn - ph;
        if (!(dph > 1e-12))
            continue;
        const double valperphs = x / dph;
        double lophase = ph - floor(ph);
        double u = ph - phi0;                 /* phase added before this point */
        double curbin = lophase * n;
        while (dph > 1e-12) {
            int ib = (int) floor(curbin + 1e-12);
            double dphs = ((ib + 1.0) - curbin) / n;
            ib %= n;
            if (dphs > dph)
                dphs = dph;
            /* split the piece [u, u + dphs) at whole turns of phase added */
            double lo = u;
            const double hi = u + dphs;
            while (lo < hi) {
                const double turn = floor(lo);
                const double end = fmin(hi, turn + 1.0);
                const double v = (end - lo) * valperphs;
                /* the sub-integration during which this turn completes */
                int tp = part;
                while (tp < a.npart && a.part_end[tp] < turn + 1.0)
                    tp++;
                if (tp < a.npart) {
                    if (use_lds && tp - part < 2)
                        atomicAdd(&acc[(tp - part) * n + ib], v);
                    else
                        atomicAdd(&a.rawfolds[(size_t) tp * n + ib], v);
                }
                lo = end;
            }
            u = hi;
            dph -= dphs;
            curbin += dphs * n;
        }
    }
    /* per-record sums of the data for the statistics */
    for (int off = 32; off > 0; off >>= 1) {
        s1 += __shfl_xor(s1, off);
        s2 += __shfl_xor(s2, off);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = s1;
        red[1][threadIdx.x >> 6] = s2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t1 = 0.0, t2 = 0.0;
        for (int w = 0; w < FD_TPB / 64; w++) {
            t1 += red[0][w];
            t2 += red[1][w];
        }
        a.recsums[2 * rec] = t1;
        a.recsums[2 * rec + 1] = t2;
    }
    if (use_lds)
        for (int k = threadIdx.x; k < 2 * n; k += FD_TPB) {
            const int tp = part + k / n;
            if (tp < a.npart && acc[k] != 0.0)
                atomicAdd(&a.rawfolds[(size_t) tp * n + k % n], acc[k]);
        }
}

extern "C" int agpu_fold(const agpu_fold_args *args, void *stream)
{
    if (args->npart < 1 || args->rpp < 1 || args->proflen < 2) {
        agpu_set_error("agpu_fold: bad sub-integration or profile size");
        return -1;
    }
    const long long nrec = (long long) args->npart * args->rpp;
    if (nrec > 0x7fffffffLL) {
        agpu_set_error("agpu_fold: too many records");
        return -1;
    }
    hipLaunchKernelGGL(k_fold, dim3((unsigned) nrec), dim3(FD_TPB), 0, (hipStream_t) stream,
                       *args);
    FD_CHECK(hipGetLastError());
    return 0;
}

/* prepfold's period / period-derivative search over the sub-integrations
 * of one subband (src/prepfold_pipeline.c:1696-1868; combine_profs,
 * src/fold.c:696-754; chisqr, src/characteristics.c:562-575).  One
 * workgroup per trial (ip, ipd) of the launch's p-dotdot trial: part ii
 * is delayed by fdotdot2phasedelay(fdotdots[ipdd], parttimes[ii]) proflen
 * (0 without a p-dotdot search) + fdot2phasedelay(fdots[ipd],
 * parttimes[ii]) proflen + ii tot / npart bins
 * (tot = pstep (ip - centre)), the delay rounded to whole bins as
 * combine_profs does; thread k sums bin k of the shifted profiles in part
 * order, and the reduced chi-squared is summed over the bins in order by
 * one thread, so every trial's value is the reference's operation for
 * operation. */
__global__ __launch_bounds__(256) void k_fold_search(agpu_fold_search_args a)
{
#pragma clang fp contract(off)
    extern __shared__ double fs_lds[];
    double *prof = fs_lds;                     /* proflen */
    int *offs = (int *) (fs_lds + a.proflen);  /* npart */
    const int ip = blockIdx.x, ipd = blockIdx.y, nt = a.numtrials;
    const int L = a.proflen;
    const int tot = a.pstep * (ip - (nt - 1) / 2);
    for (int ii = threadIdx.x; ii < a.npart; ii += blockDim.x) {
        const double t = a.parttimes[ii];
        double pdd_delay = 0.0;
        if (a.fdotdots) {
            double pdd = a.fdotdots[a.ipdd] * t * t * t / 6.0;   /* fdotdot2phasedelay */
            if (pdd == -0)
                pdd = 0.0;
            pdd_delay = pdd * L;
        }
        double pd = a.fdots[ipd] * t * t / 2.0;   /* fdot2phasedelay */
        if (pd == -0)
            pd = 0.0;
        const double delay = (pdd_delay + pd * L) + (double) (ii * tot) / a.npart;
        double ld = fmod(-delay, (double) L);
        if (ld < 0.0)
            ld += L;
        offs[ii] = (int) (ld + 0.5);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < L; k += blockDim.x) {
        double s = 0.0;
        for (int ii = 0; ii < a.npart; ii++) {
            int j = k + offs[ii];
            if (j >= L)
                j -= L;
            s += a.profs[(size_t) ii * L + j];
        }
        prof[k] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double chi = 0.0;
        for (int k = 0; k < L; k++) {
            const double d = prof[k] - a.prof_avg;
            chi += d * d;
        }
        a.redchi[(size_t) ipd * nt + ip] = chi / a.prof_var / (L - 1.0);
    }
}

extern "C" int agpu_fold_search(const agpu_fold_search_args *args, void *stream)
{
    if (args->npart < 1 || args->proflen < 2 || args->numtrials < 1 ||
        args->numtrials > 65535) {
        agpu_set_error("agpu_fold_search: bad sizes");
        return -1;
    }
    const size_t lds = sizeof(double) * (size_t) args->proflen +
                       sizeof(int) * (size_t) args->npart;
    if (lds > 64 * 1024) {
        agpu_set_error("agpu_fold_search: profile and parts too large for LDS");
        return -1;
    }
    hipLaunchKernelGGL(k_fold_search, dim3((unsigned) args->numtrials, (unsigned) args->numtrials),
                       dim3(64), lds, (hipStream_t) stream, *args);
    FD_CHECK(hipGetLastError());
    return 0;
}
