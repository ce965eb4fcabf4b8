/*
 * fft_gpu.hip -- forward real FFT of a time series on the GPU, in
 * PRESTO's packing (src/fastffts.c:163-243 realfft with isign = -1):
 * out[k] = sum_n x[n] e^{-2 pi i k n / N} for k = 1 .. N/2-1, and
 * out[0] = (DC, Nyquist), both real.  N must be a power of two.
 *
 * The N/2-point complex FFT of z[n] = x[2n] + i x[2n+1] runs as radix-2
 * Stockham passes through HBM (each pass streams the array once, output in
 * natural order), with twiddles computed in FP64 and stored FP32; a final
 * pass splits Z into the spectrum of the real series.
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include "agpu.h"

#define FFT_TPB 256

#define FFT_CHECK(call)                                                           \
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

/* tw[t] = e^{-2 pi i t / M}, t < M/2 */
__global__ void k_fft_twiddles(float2 *tw, long long half, long long M)
{
    long long t = (long long) blockIdx.x * FFT_TPB + threadIdx.x;
    if (t >= half)
        return;
    double s, c;
    sincospi(-2.0 * (double) t / (double) M, &s, &c);
    tw[t] = make_float2((float) c, (float) s);
}

/* One radix-2 Stockham pass: sub-transforms of length Ns merge into 2 Ns. */
__global__ void k_fft_pass(const float2 *in, float2 *out, const float2 *tw, long long M,
                           long long Ns)
{
    long long j = (long long) blockIdx.x * FFT_TPB + threadIdx.x;
    const long long half = M >> 1;
    if (j >= half)
        return;
    const long long k = j & (Ns - 1);
    const float2 a = in[j], b0 = in[j + half];
    const float2 w = tw[k * (half / Ns)];
    const float2 b = make_float2(b0.x * w.x - b0.y * w.y, b0.x * w.y + b0.y * w.x);
    const long long d = ((j - k) << 1) + k;
    out[d] = make_float2(a.x + b.x, a.y + b.y);
    out[d + Ns] = make_float2(a.x - b.x, a.y - b.y);
}

/* X[k] = (Z[k] + conj Z[M-k]) / 2 - i e^{-2 pi i k / N} (Z[k] - conj Z[M-k]) / 2,
 * N = 2M; X[0] = (Re Z0 + Im Z0, Re Z0 - Im Z0).  Arithmetic in FP64. */
__global__ void k_realfft_post(const float2 *Z, float2 *X, long long M)
{
    long long k = (long long) blockIdx.x * FFT_TPB + threadIdx.x;
    if (k >= M)
        return;
    if (k == 0) {
        X[0] = make_float2(Z[0].x + Z[0].y, Z[0].x - Z[0].y);
        return;
    }
    const float2 zk = Z[k], zm = Z[M - k];
    const double er = 0.5 * ((double) zk.x + zm.x), ei = 0.5 * ((double) zk.y - zm.y);
    const double or_ = 0.5 * ((double) zk.y + zm.y), oi = -0.5 * ((double) zk.x - zm.x);
    double s, c;
    sincospi(-(double) k / (double) M, &s, &c);   /* e^{-2 pi i k / N} */
    X[k] = make_float2((float) (er + c * or_ - s * oi), (float) (ei + c * oi + s * or_));
}

/* Inverse split (realfft with isign = +1, src/fastffts.c:186-242): from
 * X (PRESTO packing) form Z[k] = E[k] + i O[k], E = (X[k] + conj X[M-k]) / 2,
 * O = (X[k] - conj X[M-k]) / (2 e^{-2 pi i k / N}), Z[0] = ((DC + Nyq) / 2,
 * (DC - Nyq) / 2); written conjugated, so that the forward passes give
 * conj of the inverse transform.  Arithmetic in FP64. */
__global__ void k_realfft_pre_inv(const float2 *X, float2 *Zc, long long M)
{
    long long k = (long long) blockIdx.x * FFT_TPB + threadIdx.x;
    if (k >= M)
        return;
    if (k == 0) {
        const double dc = X[0].x, ny = X[0].y;
        Zc[0] = make_float2((float) (0.5 * (dc + ny)), (float) (-0.5 * (dc - ny)));
        return;
    }
    const float2 xk = X[k], xm = X[M - k];
    const double er = 0.5 * ((double) xk.x + xm.x), ei = 0.5 * ((double) xk.y - xm.y);
    const double dr = 0.5 * ((double) xk.x - xm.x), di = 0.5 * ((double) xk.y + xm.y);
    double s, c;
    sincospi((double) k / (double) M, &s, &c);   /* 1 / e^{-2 pi i k / N} */
    const double or_ = dr * c - di * s, oi = dr * s + di * c;
    /* Z = E + i O */
    Zc[k] = make_float2((float) (er - oi), (float) (-(ei + or_)));
}

/* x[2n] + i x[2n+1] = conj(F[n]) / M, F the forward transform of conj Z */
__global__ void k_realfft_post_inv(const float2 *F, float2 *x, long long M)
{
    long long n = (long long) blockIdx.x * FFT_TPB + threadIdx.x;
    if (n >= M)
        return;
    const double sc = 1.0 / (double) M;
    const float2 f = F[n];
    x[n] = make_float2((float) (f.x * sc), (float) (-f.y * sc));
}

int agpu_realfft_inverse(const void *d_X, long long n, float *d_out, void *d_work, void *stream)
{
    hipStream_t st = (hipStream_t) stream;
    if (n < 4 || (n & (n - 1))) {
        agpu_set_error("agpu_realfft_inverse: the length must be a power of two >= 4");
        return -2;
    }
    const long long M = n / 2, half = M / 2;
    float2 *a = (float2 *) d_out, *b = (float2 *) d_work;
    float2 *tw = b + M;
    const dim3 blk(FFT_TPB);
    hipLaunchKernelGGL(k_fft_twiddles, dim3((half + FFT_TPB - 1) / FFT_TPB), blk, 0, st, tw,
                       half, M);
    FFT_CHECK(hipGetLastError());
    /* conj Z into b (the input may be d_out itself) */
    hipLaunchKernelGGL(k_realfft_pre_inv, dim3((M + FFT_TPB - 1) / FFT_TPB), blk, 0, st,
                       (const float2 *) d_X, b, M);
    FFT_CHECK(hipGetLastError());
    float2 *src = b, *dst = a;
    for (long long Ns = 1; Ns < M; Ns <<= 1) {
        hipLaunchKernelGGL(k_fft_pass, dim3((half + FFT_TPB - 1) / FFT_TPB), blk, 0, st, src,
                           dst, tw, M, Ns);
        FFT_CHECK(hipGetLastError());
        float2 *t = src;
        src = dst;
        dst = t;
    }
    /* src holds F; the post pass writes x into the other buffer */
    float2 *xo = (src == a) ? b : a;
    hipLaunchKernelGGL(k_realfft_post_inv, dim3((M + FFT_TPB - 1) / FFT_TPB), blk, 0, st, src,
                       xo, M);
    FFT_CHECK(hipGetLastError());
    if (xo != a)
        FFT_CHECK(hipMemcpyAsync(a, xo, sizeof(float2) * M, hipMemcpyDeviceToDevice, st));
    return 0;
}

int agpu_realfft(const float *d_x, long long n, void *d_out, void *d_work, void *stream)
{
    hipStream_t st = (hipStream_t) stream;
    if (n < 4 || (n & (n - 1))) {
        agpu_set_error("agpu_realfft: the length must be a power of two >= 4");
        return -2;
    }
    const long long M = n / 2, half = M / 2;
    float2 *a = (float2 *) d_out, *b = (float2 *) d_work;
    float2 *tw = b + M;   /* work holds M + M/2 complex values */
    const dim3 blk(FFT_TPB);
    hipLaunchKernelGGL(k_fft_twiddles, dim3((half + FFT_TPB - 1) / FFT_TPB), blk, 0, st, tw,
                       half, M);
    FFT_CHECK(hipGetLastError());
    if ((const void *) d_x != (const void *) a)
        FFT_CHECK(hipMemcpyAsync(a, d_x, sizeof(float) * n, hipMemcpyDeviceToDevice, st));
    float2 *src = a, *dst = b;
    for (long long Ns = 1; Ns < M; Ns <<= 1) {
        hipLaunchKernelGGL(k_fft_pass, dim3((half + FFT_TPB - 1) / FFT_TPB), blk, 0, st, src,
                           dst, tw, M, Ns);
        FFT_CHECK(hipGetLastError());
        float2 *t = src;
        src = dst;
        dst = t;
    }
    /* src holds Z; the post pass writes X into the other buffer */
    float2 *X = (src == a) ? b : a;
    hipLaunchKernelGGL(k_realfft_post, dim3((M + FFT_TPB - 1) / FFT_TPB), blk, 0, st, src, X,
                       M);
    FFT_CHECK(hipGetLastError());
    if (X != a)
        FFT_CHECK(hipMemcpyAsync(a, X, sizeof(float2) * M, hipMemcpyDeviceToDevice, st));
    return 0;
}
