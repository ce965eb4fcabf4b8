/*
 * accel_gpu.hip -- HIP kernels for the accelsearch hot path on MI355X
 * (gfx950) and the thin C ABI (agpu.h) that the C host calls.
 *
 *   k_resp_z / k_template_dft : FP64 z-response templates, FFT'd to the
 *       correlation length in FP64 and kept resident as FP32
 *       (src/accel_utils.c:153-170, src/responses.c:228-316).
 *   k_block_norms_wave / k_block_norms : exact block medians (one
 *       wavefront per window by ballot counts for windows of <= 1024
 *       bins, a workgroup radix select in LDS above) and the block-median
 *       normalisation of src/accel_utils.c:1087-1099.
 *   k_corr4096 : fused overlap-save correlation -- spread spectrum
 *       segment -> Stockham radix-16 FFT in LDS -> multiply by each
 *       template -> inverse FFT -> |.|^2 -> raw plane row
 *       (the template loop of src/accel_utils.c:1141-1190).
 *   k_harmsum : harmonic summing with the reference's index rules and
 *       per-block normalisation, and the threshold of search_ffdotpows
 *       (src/accel_utils.c:1261-1452).
 */
#include <hip/hip_runtime.h>
#include <pthread.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "agpu.h"

static thread_local char g_err[512];   /* per host thread */

#define AGPU_CHECK(call)                                                          \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                   \
            snprintf(g_err, sizeof(g_err), "%s:%d %s: %s", __FILE__, __LINE__,    \
                     #call, hipGetErrorString(e_));                               \
            return -1;                                                            \
        }                                                                         \
    } while (0)

void agpu_set_error(const char *msg)
{
    snprintf(g_err, sizeof(g_err), "%s", msg);
}

static int fail(const char *msg)
{
    snprintf(g_err, sizeof(g_err), "%s", msg);
    return -2;
}

#include "dev_fresnel.h"

/* One tap of the z-response kernel with roffset = 0 and numbetween = 2
 * (src/responses.c:159-225 for |z| < 1e-4, :228-316 otherwise). */
__global__ void k_resp_z(const int *zs, const int *nks, int nkmax, double2 *resp)
{
    const double PI = 3.14159265358979323846;
    const int t = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int nk = nks[t];
    if (k >= nkmax)
        return;
    double2 out = make_double2(0.0, 0.0);
    if (k < nk) {
        const double z = (double) zs[t];
        if (fabs(z) < 1e-4) {
            /* exp(-i pi f) sinc(f), f = k/2 - nk/4 */
            double f = 0.5 * k - 0.25 * nk;
            double r = -PI * f;
            double sv, cv;
            sincospi(-f, &sv, &cv);
            double sinc = (f == 0.0) ? 1.0 : sv / r;
            out = make_double2(cv * sinc, sv * sinc);
        } else {
            int signz = (z < 0.0) ? -1 : 1;
            double zd = signz * 1.41421356237309504880 / sqrt(fabs(z));
            double cons = zd / 2.0;
            double r = -0.5 * z + 0.25 * nk - 0.5 * k;   /* exact on this grid */
            double yy = r * zd;
            double zz = yy + z * zd;
            double sv, cv;
            sincospi(r * r / z, &sv, &cv);               /* cos/sin(pi r^2 / z) */
            double fsy, fcy, fsz, fcz;
            d_fresnl(yy, &fsy, &fcy);
            d_fresnl(zz, &fsz, &fcz);
            double tre = signz * (fcz - fcy);
            double tim = fsy - fsz;
            out = make_double2((tre * cv - tim * sv) * cons, -(tre * sv + tim * cv) * cons);
        }
    }
    resp[(size_t) t * nkmax + k] = out;
}

/* Jerk template (|w| >= 1e-4), src/responses.c:333-418: the reference
 * samples a chirp of npts points into FP32, zero-pads it to 2 npts and
 * real-FFTs it, keeping numkern bins around 2 fbar.  Here: the same FP32
 * samples (phase in FP64, rounded like the reference's data[] array), then
 * the FP64 DFT at exactly the kept bins.  One workgroup per template. */
__global__ __launch_bounds__(256) void k_resp_w(const int *zs, const int *ws, const int *nks,
                                                const int *npts_arr, const int *tlist,
                                                int nkmax, double2 *resp)
{
    extern __shared__ float xs[];
    const double TWOPI = 6.28318530717958647692;
    const int t = tlist[blockIdx.x];
    const double z = (double) zs[t], w = (double) ws[t];
    const int nk = nks[t], npts = npts_arr[t];
    const double dt = 1.0 / (double) npts;
    const double amp = 2.0 * dt;
    const int fbar = npts / 4;
    const double f = (double) fbar - 0.5 * z + w / 12.0;
    const double fd = (z - 0.5 * w) / 2.0;
    const double fdd = w / 6.0;
    for (int ii = threadIdx.x; ii < npts; ii += blockDim.x) {
        const double tt = ii * dt;
        const double phase = TWOPI * (tt * (tt * (tt * fdd + fd) + f));
        xs[ii] = (float) (amp * cos(phase));
    }
    __syncthreads();
    const long long ntot = 2LL * npts;
    const long long first = 2LL * fbar - nk / 2;
    for (int k = threadIdx.x; k < nk; k += blockDim.x) {
        const long long m = first + k;
        double ws_, wc_;
        sincospi(-2.0 * (double) (m % ntot) / (double) ntot, &ws_, &wc_);
        double ar = 0.0, ai = 0.0, cr = 1.0, ci = 0.0;
        for (int i = 0; i < npts; i++) {
            if ((i & 511) == 0 && i) {   /* re-seed the rotation exactly */
                sincospi(-2.0 * (double) ((m * i) % ntot) / (double) ntot, &ci, &cr);
            }
            const double x = (double) xs[i];
            ar += x * cr;
            ai += x * ci;
            const double nr = cr * wc_ - ci * ws_;
            ci = cr * ws_ + ci * wc_;
            cr = nr;
        }
        resp[(size_t) t * nkmax + k] = make_double2(ar, ai);
    }
}

/* T[t][f] = conj(sum_m resp[nk/2+m] e^{-2 pi i f m / L}) / L in FP64 */
__global__ void k_template_dft(const double2 *resp, const int *nks, int nkmax, int L,
                               float2 *T)
{
    const int t = blockIdx.y;
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= L)
        return;
    const int nk = nks[t], half = nk / 2;
    const double2 *rp = resp + (size_t) t * nkmax;
    double ws, wc, s0, c0;
    sincospi(-2.0 * (double) f / (double) L, &ws, &wc);
    long long e0 = ((long long) f * half) % L;              /* m = -half */
    sincospi(2.0 * (double) e0 / (double) L, &s0, &c0);
    double cr = c0, ci = s0, ar = 0.0, ai = 0.0;
    for (int m = 0; m < nk; m++) {
        double2 v = rp[m];
        ar += v.x * cr - v.y * ci;
        ai += v.x * ci + v.y * cr;
        double nr = cr * wc - ci * ws;
        ci = cr * ws + ci * wc;
        cr = nr;
    }
    const double inv = 1.0 / (double) L;
    T[(size_t) t * L + f] = make_float2((float) (ar * inv), (float) (-ai * inv));
}

/* ------------------------------------------------------------------ */
/* Exact block medians: radix select over the FP32 power bit patterns  */
/* ------------------------------------------------------------------ */
#define NORM_MAXN 15360   /* keys in LDS: 60 KiB, with the 1 KiB histogram under 64 KiB */

/* Inclusive prefix sum of hist[0..255] by a 256-thread block (one bucket
 * per thread); returns the sum up to and including this thread's bucket. */
__device__ static unsigned block_scan256(unsigned v, unsigned *wsum)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        unsigned o = __shfl_up(v, d, 64);
        if (lane >= d)
            v += o;
    }
    if (lane == 63)
        wsum[wv] = v;
    __syncthreads();
    unsigned add = 0;
    for (int w = 0; w < wv; w++)
        add += wsum[w];
    __syncthreads();
    return v + add;
}

/* k-th smallest (0-based) of keys[0..n) by 4 passes of 8-bit buckets */
__device__ static unsigned radix_select(const unsigned *keys, int n, int k, unsigned *hist,
                                        unsigned *sh)
{
    unsigned prefix = 0, mask = 0;
    for (int shift = 24; shift >= 0; shift -= 8) {
        hist[threadIdx.x] = 0;
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += 256) {
            unsigned key = keys[i];
            if ((key & mask) == prefix)
                atomicAdd(&hist[(key >> shift) & 255u], 1u);
        }
        __syncthreads();
        unsigned h = hist[threadIdx.x];
        unsigned incl = block_scan256(h, sh + 2);
        unsigned excl = incl - h;
        if ((unsigned) k >= excl && (unsigned) k < incl) {
            sh[0] = prefix | ((unsigned) threadIdx.x << shift);
            sh[1] = (unsigned) k - excl;
        }
        __syncthreads();
        prefix = sh[0];
        k = (int) sh[1];
        mask |= 255u << shift;
        __syncthreads();
    }
    return prefix;
}

/* Exact median of the FP32 powers |x|^2 of spectrum bins [lo, lo+n)
 * (lo relative to the array; bins outside [0, numbins) count as 0), as
 * PRESTO's median(): the middle order statistic, or for even n the mean of
 * the two middle ones formed in FP64 and rounded to FP32.  Called by the
 * whole workgroup; keys holds n entries of dynamic LDS. */
__device__ static float block_median(const float2 *spec, long long numbins, long long lo, int n,
                                     unsigned *keys, unsigned *hist, unsigned *sh)
{
    /* the reference forms r*r + i*i with two roundings (no FMA) */
#pragma clang fp contract(off)
    for (int i = threadIdx.x; i < n; i += 256) {
        long long b = lo + i;
        float p = 0.0f;
        if (b >= 0 && b < numbins) {
            float2 v = spec[b];
            p = v.x * v.x + v.y * v.y;
        }
        keys[i] = __float_as_uint(p);
    }
    __syncthreads();
    const int lhs = (n - 1) / 2, rhs = n / 2;
    const unsigned ka = radix_select(keys, n, lhs, hist, sh);
    float med = __uint_as_float(ka);
    if (lhs != rhs) {
        /* the (lhs+1)-th key: ka again if more than lhs+1 keys are <= ka,
         * else the smallest key above ka */
        unsigned cnt = 0, mn = 0xffffffffu;
        for (int i = threadIdx.x; i < n; i += 256) {
            unsigned key = keys[i];
            cnt += (key <= ka);
            if (key > ka && key < mn)
                mn = key;
        }
        if (threadIdx.x == 0) {
            sh[4] = 0;
            sh[5] = 0xffffffffu;
        }
        __syncthreads();
        atomicAdd(&sh[4], cnt);
        atomicMin(&sh[5], mn);
        __syncthreads();
        const unsigned kb = (sh[4] >= (unsigned) rhs + 1) ? ka : sh[5];
        med = (float) (0.5 * ((double) __uint_as_float(ka) + (double) __uint_as_float(kb)));
    }
    __syncthreads();
    return med;
}

__global__ __launch_bounds__(256) void k_block_norms(const float2 *spec, long long numbins,
                                                     long long lobin_file,
                                                     const agpu_window *win, double *norm,
                                                     float *nsq)
{
    extern __shared__ unsigned keys[];
    __shared__ unsigned hist[256];
    __shared__ unsigned sh[8];
    const agpu_window w = win[blockIdx.x];
    const float med = block_median(spec, numbins, w.lobin - lobin_file, w.numdata, keys, hist, sh);
    if (threadIdx.x == 0) {
        double nv = 1.0 / sqrt((double) med / 0.69314718055994530942);
        norm[blockIdx.x] = nv;
        nsq[blockIdx.x] = (float) (nv * nv);
    }
}

/* Windows of at most NW_KPL * 64 bins: one wavefront per window, four
 * windows per workgroup, and no workgroup barrier.  Each lane keeps
 * NW_KPL power bit patterns in registers (bins past the window hold
 * 0xffffffff, above every power); the k-th smallest pattern is built
 * bit by bit from the top: a bit is kept when fewer than k + 1 patterns
 * lie below the value with it set, the counts coming from wave ballots.
 * The selected patterns, hence the medians and norms, are those of
 * k_block_norms. */
#define NW_KPL 16

__device__ static unsigned wave_select(const unsigned *key, unsigned k)
{
    unsigned v = 0;
    for (int bit = 31; bit >= 0; bit--) {
        const unsigned c = v | (1u << bit);
        unsigned below = 0;
#pragma unroll
        for (int j = 0; j < NW_KPL; j++)
            below += (unsigned) __popcll(__ballot(key[j] < c));
        if (below <= k)
            v = c;
    }
    return v;
}

/* The exact median of window [lo, lo+n) (n <= NW_KPL * 64) by one
 * wavefront; every lane returns it. */
__device__ static float wave_block_median(const float2 *spec, long long numbins, long long lo,
                                          int n, int lane)
{
#pragma clang fp contract(off)
    unsigned key[NW_KPL];
#pragma unroll
    for (int j = 0; j < NW_KPL; j++) {
        const int i = lane + 64 * j;
        unsigned kv = 0xffffffffu;
        if (i < n) {
            const long long b = lo + i;
            float p = 0.0f;
            if (b >= 0 && b < numbins) {
                const float2 v = spec[b];
                p = v.x * v.x + v.y * v.y;
            }
            kv = __float_as_uint(p);
        }
        key[j] = kv;
    }
    const unsigned lhs = (unsigned) (n - 1) / 2, rhs = (unsigned) n / 2;
    const unsigned ka = wave_select(key, lhs);
    if (lhs == rhs)
        return __uint_as_float(ka);
    /* the (lhs+1)-th pattern: ka again if more than lhs + 1 patterns are
     * <= ka, else the smallest pattern above ka */
    unsigned le = 0, mn = 0xffffffffu;
#pragma unroll
    for (int j = 0; j < NW_KPL; j++) {
        le += (unsigned) __popcll(__ballot(key[j] <= ka));
        if (key[j] > ka && key[j] < mn)
            mn = key[j];
    }
    for (int d = 32; d >= 1; d >>= 1) {
        const unsigned o = __shfl_xor(mn, d, 64);
        mn = o < mn ? o : mn;
    }
    const unsigned kb = le >= rhs + 1 ? ka : mn;
    return (float) (0.5 * ((double) __uint_as_float(ka) + (double) __uint_as_float(kb)));
}

__global__ __launch_bounds__(256) void k_block_norms_wave(const float2 *spec, long long numbins,
                                                          long long lobin_file,
                                                          const agpu_window *win, int nwin,
                                                          double *norm, float *nsq)
{
    const int wid = (int) blockIdx.x * 4 + (int) (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (wid >= nwin)
        return;
    const agpu_window w = win[wid];
    const float med = wave_block_median(spec, numbins, w.lobin - lobin_file, w.numdata, lane);
    if (lane == 0) {
        const double nv = 1.0 / sqrt((double) med / 0.69314718055994530942);
        norm[wid] = nv;
        nsq[wid] = (float) (nv * nv);
    }
}

/* One window given by value (the accel.h surface normalises one block
 * per subharm_fderivs_vol call). */
__global__ __launch_bounds__(256) void k_block_norm1(const float2 *spec, long long numbins,
                                                     long long lobin_file, long long lobin,
                                                     int numdata, double *norm, float *nsq)
{
    extern __shared__ unsigned keys[];
    __shared__ unsigned hist[256];
    __shared__ unsigned sh[8];
    const float med = block_median(spec, numbins, lobin - lobin_file, numdata, keys, hist, sh);
    if (threadIdx.x == 0) {
        double nv = 1.0 / sqrt((double) med / 0.69314718055994530942);
        norm[0] = nv;
        nsq[0] = (float) (nv * nv);
    }
}

/* De-reddening (src/misc_utils.c:912-1067), block medians: mean_b =
 * median(|x|^2 of block b) / ln 2, rounded to FP32 as the reference. */
__global__ __launch_bounds__(256) void k_dered_means(const float2 *spec, long long numbins,
                                                     const agpu_window *blk, float *mean)
{
    extern __shared__ unsigned keys[];
    __shared__ unsigned hist[256];
    __shared__ unsigned sh[8];
    const agpu_window w = blk[blockIdx.x];
    const float med = block_median(spec, numbins, w.lobin, w.numdata, keys, hist, sh);
    if (threadIdx.x == 0)
        mean[blockIdx.x] = (float) ((double) med / 0.69314718055994530942);
}

/* The same block means, one wavefront per block (blocks of at most
 * NW_KPL * 64 bins, as the de-reddening's always are). */
__global__ __launch_bounds__(256) void k_dered_means_wave(const float2 *spec, long long numbins,
                                                          const agpu_window *blk, int nblk,
                                                          float *mean)
{
    const int wid = (int) blockIdx.x * 4 + (int) (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (wid >= nblk)
        return;
    const agpu_window w = blk[wid];
    const float med = wave_block_median(spec, numbins, w.lobin, w.numdata, lane);
    if (lane == 0)
        mean[wid] = (float) ((double) med / 0.69314718055994530942);
}

/* invsqrtf of src/misc_utils.c:291-302: one Newton step from the
 * 0x5f3759df estimate, in FP32 without fused operations. */
__device__ static float dered_invsqrt(float x)
{
#pragma clang fp contract(off)
    float y = __int_as_float(0x5f3759df - (__float_as_int(x) >> 1));
    return y * (1.5f - 0.5f * x * y * y);
}

/* De-reddening, scaling: bin 0 -> (1, 0); the first half of block 0 by
 * invsqrt(mean_0); then each span [middle of block b, middle of block
 * b+1) by invsqrt(mean_b + dslope_b ii), ii counted from the span start,
 * dslope_b = (mean_{b+1} - mean_b) / ((n_b + n_{b+1}) / 2); the last
 * half-block continues with the last slope (1 if there is one block). */
__global__ void k_dered_apply(float2 *spec, long long numbins, const agpu_window *blk,
                              const float *mean, int nblk)
{
#pragma clang fp contract(off)
    const long long k = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= numbins)
        return;
    if (k == 0) {
        spec[0] = make_float2(1.0f, 0.0f);
        return;
    }
    const long long mid0 = blk[0].lobin + blk[0].numdata / 2;
    float norm;
    if (k < mid0) {
        norm = dered_invsqrt(mean[0]);
    } else {
        /* the span holding k: largest b with lobin_b + n_b/2 <= k */
        int lo = 0, hi = nblk - 1;
        while (lo < hi) {
            int m = (lo + hi + 1) >> 1;
            if (blk[m].lobin + blk[m].numdata / 2 <= k)
                lo = m;
            else
                hi = m - 1;
        }
        const int b = lo;
        const int ii = (int) (k - (blk[b].lobin + blk[b].numdata / 2));
        float dslope;
        if (b + 1 < nblk)
            dslope = (float) ((mean[b + 1] - mean[b]) /
                              (0.5 * (blk[b].numdata + blk[b + 1].numdata)));
        else if (nblk > 1)
            dslope = (float) ((mean[b] - mean[b - 1]) /
                              (0.5 * (blk[b - 1].numdata + blk[b].numdata)));
        else
            dslope = 1.0f;
        norm = dered_invsqrt(mean[b] + dslope * (float) ii);
    }
    float2 v = spec[k];
    spec[k] = make_float2(v.x * norm, v.y * norm);
}

int agpu_deredden(void *d_spec, long long numbins, const agpu_window *d_blk, int nblk,
                  int maxn, float *d_mean, void *stream)
{
    hipStream_t st = (hipStream_t) stream;
    if (nblk <= 0 || maxn <= 0 || maxn > NORM_MAXN)
        return fail("agpu_deredden: bad block table");
    if (maxn <= NW_KPL * 64)
        hipLaunchKernelGGL(k_dered_means_wave, dim3((unsigned) ((nblk + 3) / 4)), dim3(256), 0,
                           st, (const float2 *) d_spec, numbins, d_blk, nblk, d_mean);
    else
        hipLaunchKernelGGL(k_dered_means, dim3(nblk), dim3(256), sizeof(unsigned) * maxn, st,
                           (const float2 *) d_spec, numbins, d_blk, d_mean);
    AGPU_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_dered_apply, dim3((unsigned) ((numbins + 255) / 256)), dim3(256), 0, st,
                       (float2 *) d_spec, numbins, d_blk, (const float *) d_mean, nblk);
    AGPU_CHECK(hipGetLastError());
    return 0;
}

/* In-memory path: norm^2 of the fundamental block that last wrote each
 * half-bin -- main blocks from rh0 on, pre-population blocks from half-bin
 * 16 (r = 8) on, nothing below (src/accelsearch.c:166-205). */
__global__ void k_inmem_nsq_table(const float *nsq, int npre, int nblocks, long long rh0,
                                  int cu, long long hb_end, float *out)
{
    long long hb = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (hb >= hb_end)
        return;
    float v = 0.0f;
    if (hb >= rh0) {
        long long b = (hb - rh0) / cu;
        if (b < nblocks)
            v = nsq[npre + b];
    } else if (hb >= 16 && npre > 0) {
        long long b = (hb - 16) / cu;
        if (b < npre)
            v = nsq[b];
    }
    out[hb] = v;
}

/* ------------------------------------------------------------------ */
/* Stockham radix-16 FFT of length 4096 over 256 threads               */
/* ------------------------------------------------------------------ */
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 w)
{
    return make_float2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
}

/* 4-point DFT with W4 = S*i */
template <int S>
__device__ __forceinline__ void dft4(float2 &a0, float2 &a1, float2 &a2, float2 &a3)
{
    float2 s02 = cadd(a0, a2), d02 = csub(a0, a2);
    float2 s13 = cadd(a1, a3), d13 = csub(a1, a3);
    float2 e = make_float2(-S * d13.y, S * d13.x);
    a0 = cadd(s02, s13);
    a2 = csub(s02, s13);
    a1 = cadd(d02, e);
    a3 = csub(d02, e);
}

/* 16-point DFT as 4x4: X[c+4d] ends in slot 4c+d */
template <int S>
__device__ __forceinline__ void dft16(float2 *v)
{
    const float C1 = 0.92387953251128675613f, S1 = 0.38268343236508977173f;
    const float R2 = 0.70710678118654752440f;
#pragma unroll
    for (int b = 0; b < 4; b++)
        dft4<S>(v[b], v[4 + b], v[8 + b], v[12 + b]);
    /* slot 4c+b *= W16^(b c) */
    v[5] = cmul(v[5], make_float2(C1, S * S1));                              /* W^1 */
    v[9] = make_float2(R2 * (v[9].x - S * v[9].y), R2 * (S * v[9].x + v[9].y));   /* W^2 (b1,c2) */
    v[13] = cmul(v[13], make_float2(S1, S * C1));                            /* W^3 */
    v[6] = make_float2(R2 * (v[6].x - S * v[6].y), R2 * (S * v[6].x + v[6].y));   /* W^2 (b2,c1) */
    v[10] = make_float2(-S * v[10].y, S * v[10].x);                          /* W^4 */
    v[14] = make_float2(R2 * (-v[14].x - S * v[14].y), R2 * (S * v[14].x - v[14].y)); /* W^6 */
    v[7] = cmul(v[7], make_float2(S1, S * C1));                              /* W^3 */
    v[11] = make_float2(R2 * (-v[11].x - S * v[11].y), R2 * (S * v[11].x - v[11].y)); /* W^6 */
    v[15] = cmul(v[15], make_float2(-C1, -S * S1));                          /* W^9 */
#pragma unroll
    for (int c = 0; c < 4; c++)
        dft4<S>(v[4 * c], v[4 * c + 1], v[4 * c + 2], v[4 * c + 3]);
}

#define OUTSLOT(k) (4 * ((k) & 3) + ((k) >> 2))
#define LP(i) ((i) + ((i) >> 4))
#define FFT_LDS (4096 + 256)

/* Input v[r] = x[j + 256 r]; output v[r] = X[j + 256 r].  tw1l is the
 * LDS table W_256^{r m} at [16 r + m] (m = j % 16, the first inter-pass
 * twiddles); tw2 holds this thread's second inter-pass twiddles
 * W_4096^{r j}.  Twiddles are the inverse-sign (S = +1) values. */
template <int S>
__device__ __forceinline__ void fft4096(float2 *v, float2 *lds, const float2 *tw1l,
                                        const float2 *tw2)
{
    const int j = threadIdx.x;
    dft16<S>(v);
#pragma unroll
    for (int r = 0; r < 16; r++)
        lds[LP(16 * j + r)] = v[OUTSLOT(r)];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; r++)
        v[r] = lds[LP(j + 256 * r)];
    __syncthreads();
#pragma unroll
    for (int r = 1; r < 16; r++) {
        float2 w = tw1l[16 * r + (j & 15)];
        v[r] = cmul(v[r], make_float2(w.x, S * w.y));
    }
    dft16<S>(v);
    const int base1 = (j >> 4) * 256 + (j & 15);
#pragma unroll
    for (int r = 0; r < 16; r++)
        lds[LP(base1 + 16 * r)] = v[OUTSLOT(r)];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; r++)
        v[r] = lds[LP(j + 256 * r)];
    __syncthreads();
#pragma unroll
    for (int r = 1; r < 16; r++) {
        float2 w = tw2[r];
        v[r] = cmul(v[r], make_float2(w.x, S * w.y));
    }
    dft16<S>(v);
    float2 o[16];
#pragma unroll
    for (int r = 0; r < 16; r++)
        o[r] = v[OUTSLOT(r)];
#pragma unroll
    for (int r = 0; r < 16; r++)
        v[r] = o[r];
}

/* Fused correlation for segments of L = 4096 half-bins.  Block id ->
 * (segment, template group); groups is a multiple of 8 so each XCD (block
 * ids are dealt round-robin over the 8 XCDs) touches one residue class of
 * template groups and keeps those templates in its own L2.  The forward
 * FFT of a segment is shared by the group's templates.  The first
 * inter-pass twiddles come from a 2 KiB LDS table (16 distinct values per
 * pass and thread), the second stay in registers: 196 VGPRs, two waves
 * per SIMD. */
__global__ __launch_bounds__(256) void k_corr4096(const float2 *spec, long long numbins,
                                                  long long lobin_file, const float2 *T,
                                                  int nt, int hw, int uselen, long long hb_end,
                                                  float *plane, int groups, int tpg,
                                                  long long seg0, long long nsr)
{
    __shared__ float2 lds[FFT_LDS];
    __shared__ float2 tw1l[256];
    const int j = threadIdx.x;
    /* XCD x (block ids x, x + 8, ...) takes the template groups x, x + 8,
     * ... one after another, each over all nsr segments, so the workgroups
     * resident on an XCD share one group's templates in its L2 (a jerk
     * bank is ~200 MB; with 8 groups this is one group per XCD) */
    const int g = (int) (blockIdx.x & 7) + 8 * (int) ((blockIdx.x >> 3) / nsr);
    const long long seg = seg0 + (long long) (blockIdx.x >> 3) % nsr;
    const long long seg_hb = seg * (long long) uselen;
    const long long lobin = seg_hb / 2 - hw;
    const int t0 = g * tpg;
    if (t0 >= nt)
        return;
    const int t1 = min(nt, t0 + tpg);
    {
        double s, c;
        sincospi(2.0 * (double) ((j >> 4) * (j & 15)) / 256.0, &s, &c);
        tw1l[j] = make_float2((float) c, (float) s);
    }
    float2 tw2[16];
#pragma unroll
    for (int r = 1; r < 16; r++) {
        double s, c;
        sincospi(2.0 * (double) (r * j) / 4096.0, &s, &c);
        tw2[r] = make_float2((float) c, (float) s);
    }

    float2 v[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        v[r] = make_float2(0.0f, 0.0f);
        if (!(j & 1)) {
            long long bin = lobin + ((j + 256 * r) >> 1) - lobin_file;
            if (bin >= 0 && bin < numbins)
                v[r] = spec[bin];
        }
    }
    __syncthreads();
    fft4096<-1>(v, lds, tw1l, tw2);
    float2 D[16];
#pragma unroll
    for (int r = 0; r < 16; r++)
        D[r] = v[r];

    for (int t = t0; t < t1; t++) {
        const float2 *Tt = T + (size_t) t * 4096;
#pragma unroll
        for (int r = 0; r < 16; r++)
            v[r] = cmul(D[r], Tt[j + 256 * r]);
        fft4096<1>(v, lds, tw1l, tw2);
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int q = j + 256 * r - 2 * hw;
            const long long hb = seg_hb + q;
            if (q >= 0 && q < uselen && hb < hb_end) {
                const float p = __builtin_fmaf(v[r].x, v[r].x, v[r].y * v[r].y);
                plane[AGPU_PLANE_IDX(t, hb, nt)] = p;
                /* the first AGPU_HALO half-bins of a tile are repeated in
                 * the previous tile's halo */
                if ((hb & (AGPU_TILE - 1)) < AGPU_HALO && hb >= AGPU_TILE)
                    plane[AGPU_PLANE_IDX(t, hb - AGPU_TILE, nt) + AGPU_TILE] = p;
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Harmonic summing and thresholding                                    */
/* ------------------------------------------------------------------ */
#define HS_TPB 256
#define HS_WAVES (HS_TPB / 64)
#define HS_RT AGPU_HS_RT
#define HS_RT_ROW AGPU_HS_RT_ROW
#define HS_RT_SUB AGPU_HS_RT_SUB
#define HS_CT AGPU_HS_CT
#define HS_CPT 4
#define HS_CHUNK (HS_TPB * HS_CPT)

__device__ __forceinline__ void emit_hit(const agpu_harmsum_args &a, int b, int s, int wi,
                                         int zi, int ii, float p)
{
    unsigned idx = atomicAdd(a.nhits, 1u);
    if (idx < a.cap) {
        agpu_hit h;
        h.block = b;
        h.stage = s;
        h.wi = wi;
        h.zi = zi;
        h.ii = ii;
        h.pow = p;
        a.hits[idx] = h;
    }
}

/* 64-float wave loads that cover one chunk's sub-harmonic span for fraction
 * k/H (the span is at most CHUNK k/H + 1 half-bins), rounded up to a
 * multiple of the workgroup's waves so that every wave issues the same
 * loads without a bound check; the extra floats stay inside the tile row
 * and its halo. */
__host__ __device__ constexpr int hs_frac_wl_span(int k, int H)
{
    return (HS_CHUNK * k / H + 1 + 63) / 64;
}

__host__ __device__ constexpr int hs_frac_wl(int k, int H)
{
    return ((HS_CHUNK * k / H + 1 + 63) / 64 + HS_WAVES - 1) / HS_WAVES * HS_WAVES;
}

__host__ __device__ constexpr int hs_stage_wl(int S)
{
    int n = 0;
    for (int k = 1; k < (1 << S); k += 2)
        n += hs_frac_wl(k, 1 << S);
    return n;
}

/* first LDS wave-load slot (64 floats) of stage S; stages 1-3 are resident
 * together, stage 4 and stage 5 (32 harmonics) each reuse the buffer from
 * slot 0 */
__host__ __device__ constexpr int hs_stage_base(int S)
{
    int n = 0;
    for (int s = 1; s < S && s <= 4; s++)
        n += hs_stage_wl(s);
    return n;
}

__host__ __device__ constexpr int hs_max(int a, int b) { return a > b ? a : b; }

/* LDS wave-load slots of a search with NSTAGES stages */
__host__ __device__ constexpr int hs_lds_wl(int nstages)
{
    return nstages <= 4 ? hs_stage_base(nstages)
                        : hs_max(hs_max(hs_stage_base(4), hs_stage_wl(4)),
                                 nstages > 5 ? hs_stage_wl(5) : 0);
}
static_assert(hs_frac_wl(15, 16) * 64 <= AGPU_HALO && hs_frac_wl(31, 32) * 64 <= AGPU_HALO &&
                  hs_frac_wl(1, 2) * 64 <= AGPU_HALO,
              "every padded sub-harmonic span fits in a tile-row halo");

/* Index-rule variants of the harmonic sum:
 *   HS_STD   standard path, sub-harmonic half-bin rint((hb + rh_off) k / H)
 *            (src/accel_utils.c:59, 1042-1047 with the block start r on the
 *            half-bin grid);
 *   HS_INMEM in-memory path, (int)(hb k / H + 0.5) (src/accel_utils.c:1326);
 *   HS_FRAC  standard path with a block start r off the half-bin grid: the
 *            reference's FP64 expression rint(2 (startr + ii/2) k / H),
 *            evaluated per cell. */
#define HS_STD 0
#define HS_INMEM 1
#define HS_FRAC 2

/* rint (half-even) or, for HS_INMEM, (int)(x+0.5) of rh * k / 2^s.
 * 32-bit: the launcher checks that every rh * k fits. */
template <int MODE>
__host__ __device__ __forceinline__ int hs_round(int rh, int k, int s)
{
    const unsigned q = (unsigned) rh * (unsigned) k;
    const unsigned H = 1u << s;
    if (MODE == HS_INMEM)
        return (int) ((q + (H >> 1)) >> s);
    return (int) ((q + (H >> 1) - 1 + ((q >> s) & 1u)) >> s);
}

/* per-thread staging registers of one fraction: its wave loads are dealt
 * round-robin to the HS_WAVES waves */
__host__ __device__ constexpr int hs_frac_regs(int k, int H)
{
    return hs_frac_wl(k, H) / HS_WAVES;
}

__host__ __device__ constexpr int hs_stage_regs(int S)
{
    int n = 0;
    for (int k = 1; k < (1 << S); k += 2)
        n += hs_frac_regs(k, 1 << S);
    return n;
}

__host__ __device__ constexpr int hs_regs_upto(int S)   /* stages 1..S-1 */
{
    int n = 0;
    for (int s = 1; s < S; s++)
        n += hs_stage_regs(s);
    return n;
}

/* Load every sub-harmonic span of stage S for this chunk into registers:
 * wave load i (64 consecutive half-bins) goes to wave i % HS_WAVES.  The
 * span's plane offset is the chunk table's tile offset of its first
 * half-bin plus the row table's offset of its (w, z) row, both loaded as
 * scalars: every load is base + lane + immediate. */
template <int S>
__device__ __forceinline__ void hs_load(const agpu_harmsum_args &a, float *v, const long long *cb,
                                        const unsigned *rt, int wv, int lane)
{
    constexpr int H = 1 << S;
    constexpr int F0 = (H >> 1) - 1;   /* index of fraction 1/H */
    int r = 0;
#pragma unroll
    for (int k = 1; k < H; k += 2) {
        const int f = F0 + (k - 1) / 2;
        const float *b1 = a.plane + cb[1 + f] + rt[HS_RT_SUB + f];
        /* wave loads past the span's own (the padding to a multiple of the
         * waves) repeat its last one: cache hits, no extra HBM traffic */
        const int real = hs_frac_wl_span(k, H);
#pragma unroll
        for (int u = 0; u < hs_frac_regs(k, H); u++) {
            const int wl = min(wv + HS_WAVES * u, real - 1);
            v[r + u] = b1[wl * 64 + lane];
        }
        r += hs_frac_regs(k, H);
    }
}

template <int S>
__device__ __forceinline__ void hs_store(float *lds, int base, const float *v, int wv, int lane)
{
    constexpr int H = 1 << S;
    int r = 0, wl = base;
#pragma unroll
    for (int k = 1; k < H; k += 2) {
#pragma unroll
        for (int u = 0; u < hs_frac_regs(k, H); u++)
            lds[(wl + wv + HS_WAVES * u) * 64 + lane] = v[r + u];
        r += hs_frac_regs(k, H);
        wl += hs_frac_wl(k, H);
    }
}

/* Accumulate stage S from LDS, in the reference's fraction order.  nr[j]
 * is the LDS row of norm^2 values of cell j's block; clo[1 + f] the first
 * half-bin of fraction f's span. */
template <int S, int MODE>
__device__ __forceinline__ void hs_consume4(const agpu_harmsum_args &a, const float *lds,
                                            int base, const float *nr0, const float *nr1,
                                            const float *nr2, const float *nr3, const int *clo,
                                            int rh_first, const double *rf, float *acc)
{
    /* each normalised sub-harmonic power is rounded to FP32 before it is
     * added, as the reference adds stored FP32 planes
     * (src/accel_utils.c:1261-1282, 1312-1362): no fused multiply-add */
#pragma clang fp contract(off)
    static_assert(HS_CPT == 4, "four cells per thread");
    constexpr int H = 1 << S;
    constexpr int F0 = (H >> 1) - 1;
    int wl = base;
#pragma unroll
    for (int k = 1; k < H; k += 2) {
        const int f = F0 + (k - 1) / 2;
        const int lo = clo[1 + f];
        float nn[4];
        if (MODE != HS_INMEM) {
            nn[0] = nr0[1 + f];
            nn[1] = nr1[1 + f];
            nn[2] = nr2[1 + f];
            nn[3] = nr3[1 + f];
        }
        if (MODE == HS_FRAC) {
            /* rint((2 rf) (k / H)): the reference's rounding steps, FP64 */
            const double hf = (double) k / (double) H;
            const int top = hs_frac_wl(k, H) * 64 - 1;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                int o = (int) __builtin_rint((2.0 * rf[j]) * hf) - lo;
                o = o < 0 ? 0 : (o > top ? top : o);
                acc[j] = acc[j] + lds[wl * 64 + o] * nn[j];
            }
        } else {
            const int rs0 = hs_round<MODE>(rh_first + (MODE == HS_STD ? a.rh_off : 0), k, S);
            const int step = (HS_TPB / H) * k;   /* HS_TPB k / H, even */
            const float *lp = lds + wl * 64 + (rs0 - lo);
            if (MODE == HS_INMEM) {
                const float *np = a.nsq_hb + rs0;
#pragma unroll
                for (int j = 0; j < 4; j++)
                    nn[j] = np[j * step];
            }
#pragma unroll
            for (int j = 0; j < 4; j++)
                acc[j] = acc[j] + lp[j * step] * nn[j];
        }
        wl += hs_frac_wl(k, H);
    }
}

/* One workgroup = one (w,z) row x one chunk of 1024 fundamental half-bins,
 * 4 cells per thread (cell = chunk base + t + 256 j).  The grid is
 * (8 rows' slots x groups of 8 chunks): blockIdx.x = 8 row + c, chunk =
 * c_lo + 8 blockIdx.y + c.  Workgroups are dealt round-robin to the 8
 * XCDs in linear order, so every row of a chunk runs on one XCD, one row
 * after another, and the sub-harmonic rows that the rows of a chunk share
 * are read once into that XCD's L2.  Row and chunk descriptors (plane
 * offsets of every fraction's span) come from tables built with the
 * search, so the per-workgroup scalar work is a few loads and adds.  The
 * fundamental values and every sub-harmonic span of stages 1-4 are loaded
 * with plain global loads, all in flight at once.  The spans of stages 1-3
 * go to LDS; stage 4's wait in registers and take the same LDS once stages
 * 1-3 are summed.  A 32-harmonic search loads stage 5 in a third round.
 * Adding 256 to a fundamental half-bin adds an even 256 k / 2^s to the
 * rounded quotient, so one rounding per thread and fraction serves all 4
 * cells. */
template <int NSTAGES, int MODE>
__global__ __launch_bounds__(HS_TPB) void k_harmsum(agpu_harmsum_args a, int c_lo, int c_hi)
{
    __shared__ float s_norm[3 * 32];   /* norm^2 of blocks b0..b0+2 x fraction */
    constexpr int NWL = hs_lds_wl(NSTAGES);
    __shared__ float lds[(NWL > 0 ? NWL : 1) * 64];
    const int row = blockIdx.x >> 3;
    const int chunk = c_lo + 8 * (int) blockIdx.y + (int) (blockIdx.x & 7);
    if (chunk >= c_hi)
        return;
    const unsigned *rt = a.row_tab + (size_t) row * HS_RT;
    const long long *cb = a.chunk_base + (size_t) chunk * HS_CT;
    const int *clo = a.chunk_lo + (size_t) chunk * HS_CT;
    const int wi = (int) rt[0], zi = (int) rt[1];
    const int ncell = a.nblocks * a.corr_uselen;
    const int rh_c0 = ((int) (a.rh0 / HS_CHUNK) + chunk) * HS_CHUNK;
    const int cell0 = rh_c0 - (int) a.rh0;   /* negative in the first chunk */
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
    const int cu = a.corr_uselen;
    const int b0 = cell0 > 0 ? cell0 / cu : 0;
    const int e1 = (b0 + 1) * cu, e2 = (b0 + 2) * cu;
    const int nfr = a.nfrac + 1;
    const float *prow = a.plane + cb[0] + rt[HS_RT_ROW];
    float fund[HS_CPT];
#pragma unroll
    for (int j = 0; j < HS_CPT; j++)
        fund[j] = prow[t + HS_TPB * j];
    constexpr int NREG = hs_regs_upto(NSTAGES < 5 ? NSTAGES : 5);
    float v[NREG > 0 ? NREG : 1];
    if (NSTAGES > 1) hs_load<1>(a, v + hs_regs_upto(1), cb, rt, wv, lane);
    if (NSTAGES > 2) hs_load<2>(a, v + hs_regs_upto(2), cb, rt, wv, lane);
    if (NSTAGES > 3) hs_load<3>(a, v + hs_regs_upto(3), cb, rt, wv, lane);
    if (NSTAGES > 4) hs_load<4>(a, v + hs_regs_upto(4), cb, rt, wv, lane);
    if (MODE != HS_INMEM && t < 3 * nfr) {
        const int bb = min(b0 + t / nfr, a.nblocks - 1);
        s_norm[(t / nfr) * 32 + t % nfr] = a.nsq[(size_t) bb * nfr + t % nfr];
    }
    if (NSTAGES > 1) hs_store<1>(lds, hs_stage_base(1), v + hs_regs_upto(1), wv, lane);
    if (NSTAGES > 2) hs_store<2>(lds, hs_stage_base(2), v + hs_regs_upto(2), wv, lane);
    if (NSTAGES > 3) hs_store<3>(lds, hs_stage_base(3), v + hs_regs_upto(3), wv, lane);
    /* cells outside [0, ncell) read valid plane memory (the chunk before the
     * first cell or the padding tile) and are never emitted */
    const int rh_first = rh_c0 + t;
    int nb[HS_CPT];      /* block offset (0..2) of the cell */
    float acc[HS_CPT];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < HS_CPT; j++) {
        const int c = cell0 + t + HS_TPB * j;
        nb[j] = (c >= e1) + (c >= e2);
    }
    /* HS_FRAC: each cell's start r + ii/2 in FP64 (src/accel_utils.c:1044) */
    double rf[HS_CPT];
#pragma unroll
    for (int j = 0; j < HS_CPT; j++) {
        rf[j] = 0.0;
        if (MODE == HS_FRAC) {
            const int c = cell0 + t + HS_TPB * j;
            const int b = min(b0 + nb[j], a.nblocks - 1);
            rf[j] = a.bstart[b] + (double) (c - b * cu) * 0.5;
        }
    }
#pragma unroll
    for (int j = 0; j < HS_CPT; j++) {
        const float nn = MODE == HS_INMEM ? a.nsq_hb[rh_first + HS_TPB * j]
                                          : s_norm[nb[j] * 32];
        acc[j] = fund[j] * nn;
    }
#define HS_EMIT(ST)                                                                    \
    {                                                                                  \
        const float cut_ = a.powcut[ST];                                               \
        const bool any_ = acc[0] > cut_ || acc[1] > cut_ || acc[2] > cut_ ||           \
                          acc[3] > cut_;                                               \
        if (__ballot(any_)) {                                                          \
            _Pragma("unroll") for (int j = 0; j < HS_CPT; j++) {                      \
                const int c = cell0 + t + HS_TPB * j;                                  \
                if (c >= 0 && c < ncell && acc[j] > cut_) {                            \
                    const int b = b0 + nb[j];                                          \
                    emit_hit(a, b, ST, wi, zi, c - b * cu, acc[j]);                    \
                }                                                                      \
            }                                                                          \
        }                                                                              \
    }
    HS_EMIT(0)
    /* per-cell norm row bases: cell j reads s_norm[nb[j]*32 + 1 + f] */
    const float *nrow0 = s_norm + nb[0] * 32;
    const float *nrow1 = s_norm + nb[1] * 32;
    const float *nrow2 = s_norm + nb[2] * 32;
    const float *nrow3 = s_norm + nb[3] * 32;
#define HS_STAGE(ST)                                                                   \
    hs_consume4<ST, MODE>(a, lds, hs_stage_base(ST), nrow0, nrow1, nrow2, nrow3, clo,  \
                          rh_first, rf, acc);                                         \
    HS_EMIT(ST)
    if (NSTAGES > 1) { HS_STAGE(1) }
    if (NSTAGES > 2) { HS_STAGE(2) }
    if (NSTAGES > 3) { HS_STAGE(3) }
    if (NSTAGES > 4) {
        /* stage 4's spans waited in registers; they take the LDS of
         * stages 1-3 once those are consumed */
        __syncthreads();
        hs_store<4>(lds, 0, v + hs_regs_upto(4), wv, lane);
        __syncthreads();
        hs_consume4<4, MODE>(a, lds, 0, nrow0, nrow1, nrow2, nrow3, clo, rh_first, rf, acc);
        HS_EMIT(4)
    }
    if (NSTAGES > 5) {
        float v5[hs_stage_regs(5)];
        hs_load<5>(a, v5, cb, rt, wv, lane);
        __syncthreads();   /* stages 1-4 are done with LDS */
        hs_store<5>(lds, 0, v5, wv, lane);
        __syncthreads();
        hs_consume4<5, MODE>(a, lds, 0, nrow0, nrow1, nrow2, nrow3, clo, rh_first, rf, acc);
        HS_EMIT(5)
    }
#undef HS_STAGE
#undef HS_EMIT
}

/* First half-bin of a chunk's span for fraction k / 2^s.  HS_FRAC starts
 * one half-bin below the grid rule's: its lookups lie within
 * (hb - 1/2, hb + 2) k / H, so the span [lo, lo + CHUNK k / H + 4) holds
 * them, inside the hs_frac_wl slots (whose 64-float rounding leaves at
 * least 31 spare half-bins). */
static int hs_span_lo_host(int mode, int rh_off, int rh_c0, int k, int s)
{
    if (mode == HS_STD)
        return hs_round<HS_STD>(rh_c0 + rh_off, k, s);
    if (mode == HS_INMEM)
        return hs_round<HS_INMEM>(rh_c0, k, s);
    const int lo = hs_round<HS_STD>(rh_c0, k, s) - 1;
    return lo > 0 ? lo : 0;
}

/* ------------------------------------------------------------------ */
/* C ABI                                                                */
/* ------------------------------------------------------------------ */
extern "C" {

const char *agpu_last_error(void) { return g_err; }

int agpu_init(int device)
{
    int n = 0;
    AGPU_CHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
        return fail("agpu_init: no such HIP device");
    AGPU_CHECK(hipSetDevice(device));
    return 0;
}

int agpu_device_name(char *buf, int len)
{
    hipDeviceProp_t p;
    int dev = 0;
    AGPU_CHECK(hipGetDevice(&dev));
    AGPU_CHECK(hipGetDeviceProperties(&p, dev));
    snprintf(buf, len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Device memory arena.  Freed device memory is not immediately cheap to  */
/* hand out again (the next large allocation can stall for seconds), and  */
/* every hipFree waits for the whole device.  A long-running caller (the  */
/* request loop) therefore reserves one block once and every allocation   */
/* is carved from it: best fit, free neighbours coalesced, 64 KiB units.  */
/* Without an arena, allocations go straight to hipMalloc / hipFree.     */
/* ------------------------------------------------------------------ */
#define ARENA_UNIT ((size_t) 64 << 10)

struct arena_seg {
    size_t off, len;
};

static pthread_mutex_t g_arena_mu = PTHREAD_MUTEX_INITIALIZER;
static char *g_arena = NULL;
static size_t g_arena_len = 0, g_arena_used = 0;
static arena_seg *g_free = NULL, *g_used = NULL;
static int g_nfree = 0, g_nused = 0, g_capfree = 0, g_capused = 0;

static void seg_push(arena_seg **v, int *n, int *cap, arena_seg s)
{
    if (*n == *cap) {
        *cap = *cap ? 2 * *cap : 64;
        *v = (arena_seg *) realloc(*v, sizeof(arena_seg) * *cap);
    }
    (*v)[(*n)++] = s;
}

int agpu_arena_init(size_t bytes)
{
    pthread_mutex_lock(&g_arena_mu);
    int rc = 0;
    if (g_arena) {
        rc = (g_arena_len >= bytes) ? 0 : -2;
        if (rc)
            snprintf(g_err, sizeof(g_err), "agpu_arena_init: an arena of %zu bytes exists",
                     g_arena_len);
    } else {
        bytes = (bytes + ARENA_UNIT - 1) / ARENA_UNIT * ARENA_UNIT;
        hipError_t e = hipMalloc((void **) &g_arena, bytes);
        if (e != hipSuccess) {
            snprintf(g_err, sizeof(g_err), "agpu_arena_init: hipMalloc(%zu): %s", bytes,
                     hipGetErrorString(e));
            g_arena = NULL;
            rc = -1;
        } else {
            g_arena_len = bytes;
            g_arena_used = 0;
            g_nfree = g_nused = 0;
            seg_push(&g_free, &g_nfree, &g_capfree, arena_seg{0, bytes});
        }
    }
    pthread_mutex_unlock(&g_arena_mu);
    return rc;
}

int agpu_arena_release(void)
{
    pthread_mutex_lock(&g_arena_mu);
    int rc = 0;
    if (g_arena && g_nused) {
        snprintf(g_err, sizeof(g_err), "agpu_arena_release: %d blocks still allocated", g_nused);
        rc = -2;
    } else if (g_arena) {
        (void) hipFree(g_arena);
        g_arena = NULL;
        g_arena_len = g_arena_used = 0;
        g_nfree = 0;
    }
    pthread_mutex_unlock(&g_arena_mu);
    return rc;
}

size_t agpu_arena_used(void)
{
    pthread_mutex_lock(&g_arena_mu);
    size_t u = g_arena_used;
    pthread_mutex_unlock(&g_arena_mu);
    return u;
}

int agpu_malloc(void **p, size_t bytes)
{
    pthread_mutex_lock(&g_arena_mu);
    if (!g_arena) {
        pthread_mutex_unlock(&g_arena_mu);
        AGPU_CHECK(hipMalloc(p, bytes));
        return 0;
    }
    const size_t len = (bytes + ARENA_UNIT - 1) / ARENA_UNIT * ARENA_UNIT;
    int best = -1;
    for (int i = 0; i < g_nfree; i++)
        if (g_free[i].len >= len && (best < 0 || g_free[i].len < g_free[best].len))
            best = i;
    if (best < 0) {
        pthread_mutex_unlock(&g_arena_mu);
        snprintf(g_err, sizeof(g_err), "device arena: no free block of %zu bytes", len);
        return -1;
    }
    arena_seg blk = {g_free[best].off, len};
    g_free[best].off += len;
    g_free[best].len -= len;
    if (g_free[best].len == 0)
        g_free[best] = g_free[--g_nfree];
    seg_push(&g_used, &g_nused, &g_capused, blk);
    g_arena_used += len;
    *p = g_arena + blk.off;
    pthread_mutex_unlock(&g_arena_mu);
    return 0;
}

int agpu_free(void *p)
{
    if (!p)
        return 0;
    pthread_mutex_lock(&g_arena_mu);
    if (g_arena && (char *) p >= g_arena && (char *) p < g_arena + g_arena_len) {
        const size_t off = (size_t) ((char *) p - g_arena);
        int u = -1;
        for (int i = 0; i < g_nused; i++)
            if (g_used[i].off == off) {
                u = i;
                break;
            }
        if (u < 0) {
            pthread_mutex_unlock(&g_arena_mu);
            return fail("device arena: free of an unknown block");
        }
        arena_seg blk = g_used[u];
        g_used[u] = g_used[--g_nused];
        g_arena_used -= blk.len;
        /* coalesce with the free blocks just before and after */
        for (int i = 0; i < g_nfree;) {
            if (g_free[i].off + g_free[i].len == blk.off) {
                blk.off = g_free[i].off;
                blk.len += g_free[i].len;
                g_free[i] = g_free[--g_nfree];
            } else if (blk.off + blk.len == g_free[i].off) {
                blk.len += g_free[i].len;
                g_free[i] = g_free[--g_nfree];
            } else {
                i++;
            }
        }
        seg_push(&g_free, &g_nfree, &g_capfree, blk);
        pthread_mutex_unlock(&g_arena_mu);
        return 0;
    }
    pthread_mutex_unlock(&g_arena_mu);
    AGPU_CHECK(hipFree(p));
    return 0;
}

int agpu_h2d(void *dst, const void *src, size_t bytes, void *stream)
{
    AGPU_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t) stream));
    AGPU_CHECK(hipStreamSynchronize((hipStream_t) stream));
    return 0;
}

int agpu_d2h(void *dst, const void *src, size_t bytes, void *stream)
{
    AGPU_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t) stream));
    AGPU_CHECK(hipStreamSynchronize((hipStream_t) stream));
    return 0;
}

int agpu_memset(void *p, int v, size_t bytes, void *stream)
{
    AGPU_CHECK(hipMemsetAsync(p, v, bytes, (hipStream_t) stream));
    return 0;
}

int agpu_stream_create(void **s)
{
    hipStream_t st;
    AGPU_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    *s = (void *) st;
    return 0;
}

int agpu_stream_wait_event(void *s, void *e)
{
    AGPU_CHECK(hipStreamWaitEvent((hipStream_t) s, (hipEvent_t) e, 0));
    return 0;
}

int agpu_stream_destroy(void *s)
{
    AGPU_CHECK(hipStreamDestroy((hipStream_t) s));
    return 0;
}

int agpu_stream_sync(void *s)
{
    AGPU_CHECK(hipStreamSynchronize((hipStream_t) s));
    return 0;
}

int agpu_event_create(void **e)
{
    hipEvent_t ev;
    AGPU_CHECK(hipEventCreate(&ev));
    *e = (void *) ev;
    return 0;
}

int agpu_event_destroy(void *e)
{
    AGPU_CHECK(hipEventDestroy((hipEvent_t) e));
    return 0;
}

int agpu_event_record(void *e, void *stream)
{
    AGPU_CHECK(hipEventRecord((hipEvent_t) e, (hipStream_t) stream));
    return 0;
}

int agpu_event_elapsed(float *ms, void *start, void *stop)
{
    AGPU_CHECK(hipEventSynchronize((hipEvent_t) stop));
    AGPU_CHECK(hipEventElapsedTime(ms, (hipEvent_t) start, (hipEvent_t) stop));
    return 0;
}

/* Records a HIP failure in g_err and jumps to the function's cleanup. */
#define AGPU_TRY(call)                                                            \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                   \
            snprintf(g_err, sizeof(g_err), "%s:%d %s: %s", __FILE__, __LINE__,    \
                     #call, hipGetErrorString(e_));                               \
            goto cleanup;                                                         \
        }                                                                         \
    } while (0)

/* The taps of every template rounded to FP32, as the reference stores
 * gen_z_response / gen_w_response (src/responses.c:159-418): row t at
 * R + t*stride, zero past its numkern taps. */
__global__ void k_resp_round(const double2 *resp, const int *nks, int nkmax, int stride,
                             float2 *R)
{
    const int t = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= stride)
        return;
    float2 v = make_float2(0.0f, 0.0f);
    if (k < nks[t] && k < nkmax) {
        const double2 d = resp[(size_t) t * nkmax + k];
        v = make_float2((float) d.x, (float) d.y);
    }
    R[(size_t) t * stride + k] = v;
}

/* The response stage shared by the template bank (d_T, length L) and the
 * FP32 tap table (d_R, row stride Rstride); either may be NULL. */
static int gen_responses(const int *zs, const int *ws, const int *numkern, int nt, int L,
                         void *d_T, int Rstride, void *d_R, void *stream)
{
    hipStream_t st = (hipStream_t) stream;
    int rc = -1;
    int nkmax = 0, nw = 0, nptsmax = 0;
    /* scratch comes from agpu_malloc: inside a request loop's arena no
     * hipMalloc / hipFree (and no device-wide wait) happens here */
    int *d_z = NULL, *d_w = NULL, *d_nk = NULL, *d_np = NULL, *d_tl = NULL;
    double2 *d_resp = NULL;
    int *npts = (int *) malloc(sizeof(int) * nt);
    int *tlist = (int *) malloc(sizeof(int) * nt);
    if (!npts || !tlist) {
        snprintf(g_err, sizeof(g_err), "agpu_gen_templates: out of host memory");
        goto cleanup;
    }
    for (int t = 0; t < nt; t++) {
        if (numkern[t] <= 0 || numkern[t] % 4 || (d_T && numkern[t] > L) ||
            (d_R && numkern[t] > Rstride)) {
            snprintf(g_err, sizeof(g_err),
                     "agpu_gen_templates: kernel width must be a positive multiple of 4 <= L");
            goto cleanup;
        }
        if (numkern[t] > nkmax)
            nkmax = numkern[t];
        npts[t] = 0;
        if (abs(ws[t]) > 0) {
            /* src/responses.c:381-382: next2(6 hw + 200 + numkern/2) */
            long long want = 6LL * (numkern[t] / 4) + 200 + numkern[t] / 2, p2 = 1;
            while (p2 < want)
                p2 <<= 1;
            npts[t] = (int) p2;
            if (p2 > nptsmax)
                nptsmax = (int) p2;
            tlist[nw++] = t;
        }
    }
    if (nptsmax * (long long) sizeof(float) > 60 * 1024) {
        snprintf(g_err, sizeof(g_err), "agpu_gen_templates: jerk chirp longer than 15360 samples");
        goto cleanup;
    }
    if (agpu_malloc((void **) &d_z, sizeof(int) * nt) ||
        agpu_malloc((void **) &d_w, sizeof(int) * nt) ||
        agpu_malloc((void **) &d_nk, sizeof(int) * nt) ||
        agpu_malloc((void **) &d_np, sizeof(int) * nt) ||
        agpu_malloc((void **) &d_tl, sizeof(int) * (nw > 0 ? nw : 1)) ||
        agpu_malloc((void **) &d_resp, sizeof(double2) * (size_t) nt * nkmax))
        goto cleanup;
    AGPU_TRY(hipMemcpyAsync(d_z, zs, sizeof(int) * nt, hipMemcpyHostToDevice, st));
    AGPU_TRY(hipMemcpyAsync(d_w, ws, sizeof(int) * nt, hipMemcpyHostToDevice, st));
    AGPU_TRY(hipMemcpyAsync(d_nk, numkern, sizeof(int) * nt, hipMemcpyHostToDevice, st));
    AGPU_TRY(hipMemcpyAsync(d_np, npts, sizeof(int) * nt, hipMemcpyHostToDevice, st));
    if (nw)
        AGPU_TRY(hipMemcpyAsync(d_tl, tlist, sizeof(int) * nw, hipMemcpyHostToDevice, st));
    {
        /* z responses for every template; jerk templates then overwrite theirs */
        dim3 g1((nkmax + 255) / 256, nt);
        hipLaunchKernelGGL(k_resp_z, g1, dim3(256), 0, st, d_z, d_nk, nkmax, d_resp);
        AGPU_TRY(hipGetLastError());
        if (nw) {
            hipLaunchKernelGGL(k_resp_w, dim3(nw), dim3(256), sizeof(float) * nptsmax, st, d_z,
                               d_w, d_nk, d_np, d_tl, nkmax, d_resp);
            AGPU_TRY(hipGetLastError());
        }
        if (d_T) {
            dim3 g2((L + 255) / 256, nt);
            hipLaunchKernelGGL(k_template_dft, g2, dim3(256), 0, st, d_resp, d_nk, nkmax, L,
                               (float2 *) d_T);
            AGPU_TRY(hipGetLastError());
        }
        if (d_R) {
            dim3 g3((Rstride + 255) / 256, nt);
            hipLaunchKernelGGL(k_resp_round, g3, dim3(256), 0, st, d_resp, d_nk, nkmax, Rstride,
                               (float2 *) d_R);
            AGPU_TRY(hipGetLastError());
        }
    }
    rc = 0;
cleanup:
    /* the scratch may be returned only once the kernels that read it ended */
    if (hipStreamSynchronize(st) != hipSuccess && rc == 0) {
        snprintf(g_err, sizeof(g_err), "agpu_gen_templates: stream synchronisation failed");
        rc = -1;
    }
    agpu_free(d_z);
    agpu_free(d_w);
    agpu_free(d_nk);
    agpu_free(d_np);
    agpu_free(d_tl);
    agpu_free(d_resp);
    free(npts);
    free(tlist);
    return rc;
}

int agpu_gen_templates(const int *zs, const int *ws, const int *numkern, int nt, int L,
                       void *d_T, void *stream)
{
    if (nt <= 0 || L <= 0 || nt > 65535 || !d_T)
        return fail("agpu_gen_templates: bad template count or length");
    return gen_responses(zs, ws, numkern, nt, L, d_T, 0, NULL, stream);
}

int agpu_gen_response_taps(const int *zs, const int *ws, const int *numkern, int nt,
                           int stride, void *d_R, void *stream)
{
    if (nt <= 0 || stride <= 0 || nt > 65535 || !d_R)
        return fail("agpu_gen_response_taps: bad template count or stride");
    return gen_responses(zs, ws, numkern, nt, 0, NULL, stride, d_R, stream);
}

int agpu_block_norms(const void *d_spec, long long numbins, long long lobin_file,
                     const agpu_window *d_win, int nwin, int maxn, double *d_norm,
                     float *d_nsq, void *stream)
{
    if (nwin <= 0)
        return 0;
    if (maxn <= 0 || maxn > NORM_MAXN)
        return fail("agpu_block_norms: window longer than the LDS selection buffer");
    if (maxn <= NW_KPL * 64) {
        hipLaunchKernelGGL(k_block_norms_wave, dim3((unsigned) ((nwin + 3) / 4)), dim3(256), 0,
                           (hipStream_t) stream, (const float2 *) d_spec, numbins, lobin_file,
                           d_win, nwin, d_norm, d_nsq);
        AGPU_CHECK(hipGetLastError());
        return 0;
    }
    hipLaunchKernelGGL(k_block_norms, dim3(nwin), dim3(256), sizeof(unsigned) * maxn,
                       (hipStream_t) stream, (const float2 *) d_spec, numbins, lobin_file,
                       d_win, d_norm, d_nsq);
    AGPU_CHECK(hipGetLastError());
    return 0;
}

int agpu_block_norm1(const void *d_spec, long long numbins, long long lobin_file,
                     long long lobin, int numdata, double *d_norm, float *d_nsq, void *stream)
{
    if (numdata <= 0 || numdata > NORM_MAXN)
        return fail("agpu_block_norm1: window longer than the LDS selection buffer");
    hipLaunchKernelGGL(k_block_norm1, dim3(1), dim3(256), sizeof(unsigned) * numdata,
                       (hipStream_t) stream, (const float2 *) d_spec, numbins, lobin_file,
                       lobin, numdata, d_norm, d_nsq);
    AGPU_CHECK(hipGetLastError());
    return 0;
}

int agpu_correlate(const void *d_spec, long long numbins, long long lobin_file,
                   const void *d_T, int nt, int L, int hw, int uselen, long long hb_end,
                   float *d_plane, long long ntiles, int groups, long long seg_lo,
                   long long seg_hi, void *stream)
{
    if (L != 4096)
        return fail("agpu_correlate: only L = 4096 is built");
    if (uselen <= 0 || uselen % 2 || uselen > L - 4 * hw + 1)
        return fail("agpu_correlate: uselen must be even and <= L - 4*hw + 1");
    if (groups <= 0 || groups % 8)
        return fail("agpu_correlate: groups must be a positive multiple of 8");
    if (hb_end > ntiles * AGPU_TILE)
        return fail("agpu_correlate: plane tiles do not cover the half-bin range");
    long long nseg = (hb_end + uselen - 1) / uselen;
    if (seg_hi < 0 || seg_hi > nseg)
        seg_hi = nseg;
    if (seg_lo < 0 || seg_lo > seg_hi)
        return fail("agpu_correlate: bad segment range");
    if (seg_lo == seg_hi)
        return 0;
    int tpg = (nt + groups - 1) / groups;
    long long nblk = (seg_hi - seg_lo) * groups;
    if (nblk >= (1LL << 24))   /* a dispatch counts work-items in 32 bits */
        return fail("agpu_correlate: grid too large");
    hipLaunchKernelGGL(k_corr4096, dim3((unsigned) nblk), dim3(256), 0, (hipStream_t) stream,
                       (const float2 *) d_spec, numbins, lobin_file, (const float2 *) d_T, nt,
                       hw, uselen, hb_end, d_plane, groups, tpg, seg_lo, seg_hi - seg_lo);
    AGPU_CHECK(hipGetLastError());
    return 0;
}

long long agpu_harmsum_nchunks(const agpu_harmsum_args *a)
{
    const long long ncell = (long long) a->nblocks * a->corr_uselen;
    if (ncell <= 0)
        return 0;
    return (a->rh0 + ncell - 1) / HS_CHUNK - a->rh0 / HS_CHUNK + 1;
}

long long agpu_harmsum_chunk_hb(const agpu_harmsum_args *a, long long c)
{
    return (a->rh0 / HS_CHUNK + c) * HS_CHUNK;
}

int agpu_harmsum(const agpu_harmsum_args *a, long long c_lo, long long c_hi, void *stream)
{
    long long ncell = (long long) a->nblocks * a->corr_uselen;
    if (ncell <= 0)
        return 0;
    if (a->nstages < 1 || a->nstages > AGPU_MAX_STAGES)
        return fail("agpu_harmsum: 1..6 stages");
    if (a->nfrac + 1 > 32)
        return fail("agpu_harmsum: at most 31 sub-harmonic fractions");
    if (a->corr_uselen < HS_CHUNK / 2)
        return fail("agpu_harmsum: corr_uselen below 1024 half-bins");
    if (a->inmem && !a->nsq_hb)
        return fail("agpu_harmsum: in-memory path needs the per-half-bin norm table");
    if (a->rh_off < 0 || a->rh_off > 1 || (a->inmem && (a->rh_off || a->bstart)))
        return fail("agpu_harmsum: sub-harmonic offset is 0 or 1 half-bin, standard path only");
    const long long rh_end = a->rh0 + ncell;
    if (a->ntiles < ((rh_end - 1) >> AGPU_TILE_LOG2) + 2 || a->ntiles * AGPU_TILE > 0x7fffffffLL)
        return fail("agpu_harmsum: plane tiles must pad the half-bin range by one tile "
                    "and stay below 2^31 half-bins");
    long long nchunks = (rh_end - 1) / HS_CHUNK - a->rh0 / HS_CHUNK + 1;
    if ((rh_end + HS_CHUNK) * (long long) ((1 << (a->nstages - 1)) - 1) > 0x7fffffffLL)
        return fail("agpu_harmsum: spectrum too long for 32-bit sub-harmonic indices "
                    "(half-bins x highest harmonic must stay below 2^31)");
    if (c_hi < 0 || c_hi > nchunks)
        c_hi = nchunks;
    if (c_lo < 0 || c_lo > c_hi)
        return fail("agpu_harmsum: bad chunk range");
    if (c_lo == c_hi)
        return 0;
    const long long rows = (long long) a->numz * a->numw;
    if (!a->row_tab || !a->chunk_base || !a->chunk_lo)
        return fail("agpu_harmsum: address tables missing");
    /* grid (8 rows' slots, groups of 8 chunks); each dimension stays far
     * below the dispatch packet's 32-bit work-item counts */
    if (8 * rows * HS_TPB > 0xffffffffLL || (c_hi - c_lo + 7) / 8 > 0xffffLL)
        return fail("agpu_harmsum: grid too large");
    dim3 grid((unsigned) (8 * rows), (unsigned) ((c_hi - c_lo + 7) / 8));
    hipStream_t st = (hipStream_t) stream;
    const int cl = (int) c_lo, ch = (int) c_hi;
#define HS_CASE(NS)                                                                         \
    case NS:                                                                                \
        if (a->inmem)                                                                       \
            hipLaunchKernelGGL((k_harmsum<NS, HS_INMEM>), grid, dim3(HS_TPB), 0, st, *a, cl,  \
                               ch);                                                         \
        else if (a->bstart)                                                                 \
            hipLaunchKernelGGL((k_harmsum<NS, HS_FRAC>), grid, dim3(HS_TPB), 0, st, *a, cl,   \
                               ch);                                                         \
        else                                                                                \
            hipLaunchKernelGGL((k_harmsum<NS, HS_STD>), grid, dim3(HS_TPB), 0, st, *a, cl,    \
                               ch);                                                         \
        break;
    switch (a->nstages) {
        HS_CASE(1)
        HS_CASE(2)
        HS_CASE(3)
        HS_CASE(4)
        HS_CASE(5)
        HS_CASE(6)
    }
#undef HS_CASE
    AGPU_CHECK(hipGetLastError());
    return 0;
}

int agpu_harmsum_row_table(int numz, int numw, int nfrac, const int *zmap, const int *wmap,
                           unsigned *out)
{
    const long long rows = (long long) numz * numw;
    if (nfrac > AGPU_MAX_FRACS - 1 || AGPU_HS_RT_SUB + nfrac > AGPU_HS_RT)
        return fail("agpu_harmsum_row_table: too many fractions");
    if (rows * AGPU_ROWLEN > 0xffffffffLL)
        return fail("agpu_harmsum_row_table: row offsets exceed 32 bits");
    for (long long r = 0; r < rows; r++) {
        const int wi = (int) (r / numz), zi = (int) (r % numz);
        unsigned *o = out + r * AGPU_HS_RT;
        memset(o, 0, sizeof(unsigned) * AGPU_HS_RT);
        o[0] = (unsigned) wi;
        o[1] = (unsigned) zi;
        o[AGPU_HS_RT_ROW] = (unsigned) (r * AGPU_ROWLEN);
        for (int f = 0; f < nfrac; f++) {
            const long long srow = (long long) wmap[f * numw + wi] * numz + zmap[f * numz + zi];
            o[AGPU_HS_RT_SUB + f] = (unsigned) (srow * AGPU_ROWLEN);
        }
    }
    return 0;
}

int agpu_harmsum_chunk_tables(int mode, int rh_off, long long rh0, long long nchunks,
                              int nstages, int nrows, long long *base, int *lo)
{
    if (mode < 0 || mode > 2 || nstages < 1 || nstages > AGPU_MAX_STAGES)
        return fail("agpu_harmsum_chunk_tables: bad mode or stage count");
    const long long tile_stride = (long long) nrows * AGPU_ROWLEN;
    for (long long c = 0; c < nchunks; c++) {
        long long *b = base + c * AGPU_HS_CT;
        int *l = lo + c * AGPU_HS_CT;
        memset(b, 0, sizeof(long long) * AGPU_HS_CT);
        memset(l, 0, sizeof(int) * AGPU_HS_CT);
        const long long rh_c0 = (rh0 / HS_CHUNK + c) * HS_CHUNK;
        if (rh_c0 > 0x7fffffffLL)
            return fail("agpu_harmsum_chunk_tables: half-bins beyond 2^31");
        b[0] = (rh_c0 >> AGPU_TILE_LOG2) * tile_stride + (rh_c0 & (AGPU_TILE - 1));
        l[0] = (int) rh_c0;
        int f = 0;
        for (int st = 1; st < nstages; st++)
            for (int k = 1; k < (1 << st); k += 2, f++) {
                const int x = hs_span_lo_host(mode, rh_off, (int) rh_c0, k, st);
                l[1 + f] = x;
                b[1 + f] = ((long long) x >> AGPU_TILE_LOG2) * tile_stride +
                           (x & (AGPU_TILE - 1));
            }
    }
    return 0;
}

int agpu_inmem_nsq_table(const float *d_nsq, int npre, int nblocks, long long rh0, int cu,
                         long long hb_end, float *d_out, void *stream)
{
    if (hb_end <= 0)
        return 0;
    long long nb = (hb_end + 255) / 256;
    hipLaunchKernelGGL(k_inmem_nsq_table, dim3((unsigned) nb), dim3(256), 0,
                       (hipStream_t) stream, d_nsq, npre, nblocks, rh0, cu, hb_end, d_out);
    AGPU_CHECK(hipGetLastError());
    return 0;
}

} /* extern "C" */
