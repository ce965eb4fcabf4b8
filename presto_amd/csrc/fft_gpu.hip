/*
 * fft_gpu.hip -- real FFTs of a time series on the GPU, in PRESTO's packing
 * (src/fastffts.c:163-243 realfft): forward, out[k] = sum_n x[n]
 * e^{-2 pi i k n / N} for k = 1 .. N/2-1 and out[0] = (DC, Nyquist), both
 * real; inverse, the series back from that packing, scaled by 2/N.  N is
 * any even length (FFTW transforms any length for the reference,
 * src/fftcalls.c:60-159).
 *
 * The N/2 = M-point complex FFT of z[n] = x[2n] + i x[2n+1] runs as
 * Stockham passes through HBM (each pass streams the array once and leaves
 * it in natural order): radix 8, 4 and 2 for the power of two in M and
 * radix 3, 5, 7, 11 and 13 for the other small primes.  Twiddles come from
 * one table e^{-2 pi i m / M}, generated in FP64 and stored FP32 (half a
 * table and its antisymmetry when M is even).  An M with a prime factor
 * above 13 goes through Bluestein's chirp-z identity: nk = (n^2 + k^2 -
 * (k-n)^2) / 2 turns the DFT into a circular convolution of length L, a
 * power of two >= 2M - 1, computed with three power-of-two FFTs; the chirp
 * phases are reduced exactly (n^2 mod 2M in integers) before the FP64
 * sine and cosine.  A last pass splits Z into the spectrum of the real
 * series, pairs (k, M-k) per thread, in place.
 */
#include <hip/hip_runtime.h>
#include <pthread.h>
#include <stdio.h>
#include "agpu.h"
#include "accel_internal.h"

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

/* ---- the plan: radices of M, or Bluestein ---- */

#define FFT_MAXRAD 64
typedef struct {
    long long M;         /* complex length */
    int nrad, rad[FFT_MAXRAD];
    long long L;         /* Bluestein convolution length, 0 for a direct transform */
} fft_plan;

static const int SMALL_PRIMES[] = {3, 5, 7, 11, 13};

/* Radices of a power of two 2^e: 8s, then one 4 or one 2. */
static void add_pow2(fft_plan *p, int e)
{
    while (e >= 3) {
        p->rad[p->nrad++] = 8;
        e -= 3;
    }
    if (e == 2)
        p->rad[p->nrad++] = 4;
    else if (e == 1)
        p->rad[p->nrad++] = 2;
}

static void plan_fft(long long M, fft_plan *p)
{
    p->M = M;
    p->nrad = 0;
    p->L = 0;
    long long r = M;
    int e = 0;
    while (r % 2 == 0) {
        r /= 2;
        e++;
    }
    int nr = 0, tmp[FFT_MAXRAD];
    for (int i = 0; i < 5; i++)
        while (r % SMALL_PRIMES[i] == 0) {
            r /= SMALL_PRIMES[i];
            tmp[nr++] = SMALL_PRIMES[i];
        }
    if (r == 1) {
        add_pow2(p, e);
        for (int i = 0; i < nr; i++)
            p->rad[p->nrad++] = tmp[i];
        return;
    }
    /* a larger prime factor: Bluestein over a power of two */
    long long L = 1;
    int le = 0;
    while (L < 2 * M - 1) {
        L <<= 1;
        le++;
    }
    p->L = L;
    add_pow2(p, le);
}

/* Twiddle-table entries a direct transform of length M keeps. */
static long long tw_len(long long M)
{
    return (M % 2 == 0) ? M / 2 : M;
}

/* Complex work values of a transform of length M: the ping-pong buffer
 * and the twiddle table, or Bluestein's three length-L buffers and its
 * table. */
static long long cfft_work(long long M)
{
    fft_plan p;
    plan_fft(M, &p);
    if (p.L)
        return 3 * p.L + tw_len(p.L);
    return M + tw_len(M);
}

long long agpu_realfft_work(long long n)
{
    return cfft_work(n / 2);
}

/* ---- kernels ---- */

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
/* (-i) a */
__device__ __forceinline__ float2 cmi(float2 a) { return make_float2(a.y, -a.x); }

/* tw[t] = e^{-2 pi i t / M}, t < count */
__global__ void k_fft_twiddles(float2 *tw, long long count, long long M)
{
    long long t = (long long) blockIdx.x * FFT_TPB + threadIdx.x;
    if (t >= count)
        return;
    double s, c;
    sincospi(-2.0 * (double) t / (double) M, &s, &c);
    tw[t] = make_float2((float) c, (float) s);
}

/* e^{-2 pi i m / M} for 0 <= m < M from the table (half a table when M is
 * even: e^{-2 pi i (m + M/2) / M} = -e^{-2 pi i m / M}) */
__device__ __forceinline__ float2 twid(const float2 *__restrict__ tw, unsigned m, unsigned M)
{
    if (!(M & 1u) && m >= (M >> 1)) {
        const float2 t = tw[m - (M >> 1)];
        return make_float2(-t.x, -t.y);
    }
    return tw[m];
}

/* In-register DFT of P points: y_q = sum_t v_t e^{-2 pi i t q / P}. */
template <int P>
__device__ __forceinline__ void dft(float2 *v, const float2 *__restrict__ tw, unsigned M)
{
    /* odd primes: the P-th roots of unity from the table (P divides M) */
    float2 w[P];
    const unsigned step = M / P;
#pragma unroll
    for (int m = 0; m < P; m++)
        w[m] = twid(tw, (unsigned) m * step, M);
    float2 y[P];
#pragma unroll
    for (int q = 0; q < P; q++) {
        float2 acc = v[0];
#pragma unroll
        for (int t = 1; t < P; t++)
            acc = cadd(acc, cmul(v[t], w[(t * q) % P]));
        y[q] = acc;
    }
#pragma unroll
    for (int q = 0; q < P; q++)
        v[q] = y[q];
}

template <>
__device__ __forceinline__ void dft<2>(float2 *v, const float2 *__restrict__, unsigned)
{
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

__device__ __forceinline__ void dft4(float2 &v0, float2 &v1, float2 &v2, float2 &v3)
{
    const float2 a = cadd(v0, v2), b = csub(v0, v2), c = cadd(v1, v3), d = cmi(csub(v1, v3));
    v0 = cadd(a, c);
    v2 = csub(a, c);
    v1 = cadd(b, d);
    v3 = csub(b, d);
}

template <>
__device__ __forceinline__ void dft<4>(float2 *v, const float2 *__restrict__, unsigned)
{
    dft4(v[0], v[1], v[2], v[3]);
}

template <>
__device__ __forceinline__ void dft<8>(float2 *v, const float2 *__restrict__, unsigned)
{
    /* even and odd halves, then y_q = E_q + W8^q O_q, y_{q+4} = E_q - W8^q O_q */
    float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4(e0, e1, e2, e3);
    dft4(o0, o1, o2, o3);
    const float h = 0.70710678118654752f;
    const float2 t1 = make_float2(h * (o1.x + o1.y), h * (o1.y - o1.x));    /* (1 - i)/sqrt2 */
    const float2 t2 = cmi(o2);                                              /* -i */
    const float2 t3 = make_float2(h * (o3.y - o3.x), -h * (o3.x + o3.y));   /* (-1 - i)/sqrt2 */
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, t1);
    v[5] = csub(e1, t1);
    v[2] = cadd(e2, t2);
    v[6] = csub(e2, t2);
    v[3] = cadd(e3, t3);
    v[7] = csub(e3, t3);
}

/* One Stockham pass of radix P: sub-transforms of length Ns merge into
 * P Ns.  Thread j (< M/P) reads in[j + t M/P], t < P, twiddles them by
 * e^{-2 pi i t k / (P Ns)} (k = j mod Ns), transforms them and writes
 * out[(j - k) P + k + q Ns], q < P. */
template <int P>
__global__ void __launch_bounds__(FFT_TPB)
k_stockham(const float2 *__restrict__ in, float2 *__restrict__ out,
           const float2 *__restrict__ tw, unsigned M, unsigned Ns)
{
    const unsigned R = M / P;
    const unsigned j = blockIdx.x * FFT_TPB + threadIdx.x;
    if (j >= R)
        return;
    const unsigned k = j % Ns;
    const unsigned step = M / (P * Ns);
    float2 v[P];
#pragma unroll
    for (int t = 0; t < P; t++)
        v[t] = in[j + (unsigned) t * R];
    if (k) {
#pragma unroll
        for (int t = 1; t < P; t++)
            v[t] = cmul(v[t], twid(tw, (unsigned) t * k * step, M));
    }
    dft<P>(v, tw, M);
    const unsigned base = (j - k) * P + k;
#pragma unroll
    for (int q = 0; q < P; q++)
        out[base + (unsigned) q * Ns] = v[q];
}

/* X[k] = (Z[k] + conj Z[M-k]) / 2 - i e^{-2 pi i k / N} (Z[k] - conj Z[M-k]) / 2,
 * N = 2M; X[0] = (Re Z0 + Im Z0, Re Z0 - Im Z0).  Thread k <= M/2 forms
 * X[k] and X[M-k] from Z[k] and Z[M-k] (so Z may be X).  FP64 arithmetic. */
__device__ __forceinline__ float2 split_fwd(float2 zk, float2 zm, long long k, long long M)
{
    const double er = 0.5 * ((double) zk.x + zm.x), ei = 0.5 * ((double) zk.y - zm.y);
    const double or_ = 0.5 * ((double) zk.y + zm.y), oi = -0.5 * ((double) zk.x - zm.x);
    double s, c;
    sincospi(-(double) k / (double) M, &s, &c);   /* e^{-2 pi i k / N} */
    return make_float2((float) (er + c * or_ - s * oi), (float) (ei + c * oi + s * or_));
}

__global__ void k_realfft_post(const float2 *Z, float2 *X, long long M)
{
    const long long k = (long long) blockIdx.x * FFT_TPB + threadIdx.x;
    if (k > M / 2)
        return;
    const float2 zk = Z[k];
    if (k == 0) {
        X[0] = make_float2(zk.x + zk.y, zk.x - zk.y);
        return;
    }
    const float2 zm = Z[M - k];
    const float2 xk = split_fwd(zk, zm, k, M);
    const float2 xm = split_fwd(zm, zk, M - k, M);
    X[k] = xk;
    X[M - k] = xm;
}

/* Inverse split (realfft with isign = +1, src/fastffts.c:186-242): from
 * X (PRESTO packing) form Z[k] = E[k] + i O[k], E = (X[k] + conj X[M-k]) / 2,
 * O = (X[k] - conj X[M-k]) / (2 e^{-2 pi i k / N}), Z[0] = ((DC + Nyq) / 2,
 * (DC - Nyq) / 2); written conjugated, so that a forward transform gives
 * conj of the inverse transform.  Pairs (k, M-k) per thread, FP64. */
__device__ __forceinline__ float2 split_inv(float2 xk, float2 xm, long long k, long long M)
{
    const double er = 0.5 * ((double) xk.x + xm.x), ei = 0.5 * ((double) xk.y - xm.y);
    const double dr = 0.5 * ((double) xk.x - xm.x), di = 0.5 * ((double) xk.y + xm.y);
    double s, c;
    sincospi((double) k / (double) M, &s, &c);   /* 1 / e^{-2 pi i k / N} */
    const double or_ = dr * c - di * s, oi = dr * s + di * c;
    return make_float2((float) (er - oi), (float) (-(ei + or_)));
}

__global__ void k_realfft_pre_inv(const float2 *X, float2 *Zc, long long M)
{
    const long long k = (long long) blockIdx.x * FFT_TPB + threadIdx.x;
    if (k > M / 2)
        return;
    const float2 xk = X[k];
    if (k == 0) {
        const double dc = xk.x, ny = xk.y;
        Zc[0] = make_float2((float) (0.5 * (dc + ny)), (float) (-0.5 * (dc - ny)));
        return;
    }
    const float2 xm = X[M - k];
    const float2 zk = split_inv(xk, xm, k, M);
    const float2 zm = split_inv(xm, xk, M - k, M);
    Zc[k] = zk;
    Zc[M - k] = zm;
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

/* Bluestein: the chirp w[n] = e^{-i pi n^2 / M}, its phase reduced exactly */
__device__ __forceinline__ double2 chirp(long long n, long long M)
{
    const unsigned long long q = ((unsigned long long) n * (unsigned long long) n) %
                                 (unsigned long long) (2 * M);
    double s, c;
    sincospi(-(double) q / (double) M, &s, &c);
    return make_double2(c, s);
}

/* a[n] = z[n] w[n] for n < M, 0 up to L */
__global__ void k_bs_pre(const float2 *z, float2 *a, long long M, long long L)
{
    const long long n = (long long) blockIdx.x * FFT_TPB + threadIdx.x;
    if (n >= L)
        return;
    if (n >= M) {
        a[n] = make_float2(0.0f, 0.0f);
        return;
    }
    const double2 w = chirp(n, M);
    const float2 v = z[n];
    a[n] = make_float2((float) (v.x * w.x - v.y * w.y), (float) (v.x * w.y + v.y * w.x));
}

/* b[m] = conj w[m] for |m| < M, index m mod L; 0 elsewhere */
__global__ void k_bs_kernel(float2 *b, long long M, long long L)
{
    const long long i = (long long) blockIdx.x * FFT_TPB + threadIdx.x;
    if (i >= L)
        return;
    long long m = -1;
    if (i < M)
        m = i;
    else if (i > L - M)
        m = L - i;
    if (m < 0) {
        b[i] = make_float2(0.0f, 0.0f);
        return;
    }
    const double2 w = chirp(m, M);
    b[i] = make_float2((float) w.x, (float) -w.y);
}

/* a = conj(a b) / L: the inverse transform of a b runs as a forward one */
__global__ void k_bs_mul(float2 *a, const float2 *b, long long L)
{
    const long long i = (long long) blockIdx.x * FFT_TPB + threadIdx.x;
    if (i >= L)
        return;
    const float sc = 1.0f / (float) L;   /* a power of two: exact */
    const float2 p = cmul(a[i], b[i]);
    a[i] = make_float2(p.x * sc, -p.y * sc);
}

/* Z[k] = w[k] conj(c[k]), k < M */
__global__ void k_bs_post(const float2 *c, float2 *Z, long long M)
{
    const long long k = (long long) blockIdx.x * FFT_TPB + threadIdx.x;
    if (k >= M)
        return;
    const double2 w = chirp(k, M);
    const float2 v = c[k];
    /* w conj(v) */
    Z[k] = make_float2((float) (w.x * v.x + w.y * v.y), (float) (w.y * v.x - w.x * v.y));
}

/* ---- host side ---- */

static inline dim3 blocks_for(long long n)
{
    return dim3((unsigned) ((n + FFT_TPB - 1) / FFT_TPB));
}

static int launch_pass(int P, const float2 *in, float2 *out, const float2 *tw, long long M,
                       long long Ns, hipStream_t st)
{
    const dim3 g = blocks_for(M / P), b(FFT_TPB);
    const unsigned m = (unsigned) M, ns = (unsigned) Ns;
    switch (P) {
    case 2: hipLaunchKernelGGL(k_stockham<2>, g, b, 0, st, in, out, tw, m, ns); break;
    case 3: hipLaunchKernelGGL(k_stockham<3>, g, b, 0, st, in, out, tw, m, ns); break;
    case 4: hipLaunchKernelGGL(k_stockham<4>, g, b, 0, st, in, out, tw, m, ns); break;
    case 5: hipLaunchKernelGGL(k_stockham<5>, g, b, 0, st, in, out, tw, m, ns); break;
    case 7: hipLaunchKernelGGL(k_stockham<7>, g, b, 0, st, in, out, tw, m, ns); break;
    case 8: hipLaunchKernelGGL(k_stockham<8>, g, b, 0, st, in, out, tw, m, ns); break;
    case 11: hipLaunchKernelGGL(k_stockham<11>, g, b, 0, st, in, out, tw, m, ns); break;
    case 13: hipLaunchKernelGGL(k_stockham<13>, g, b, 0, st, in, out, tw, m, ns); break;
    default:
        agpu_set_error("fft: no kernel for this radix");
        return -1;
    }
    FFT_CHECK(hipGetLastError());
    return 0;
}

/* Direct transform of length M (its plan has no Bluestein step) of a in
 * place, with pong (M values) and the twiddle table tw (filled here).
 * Returns the buffer holding the result (a or pong), NULL on failure. */
static float2 *direct_fft(const fft_plan *p, float2 *a, float2 *pong, float2 *tw, hipStream_t st)
{
    const long long M = p->M;
    if (M > 0x7fffffffLL) {
        agpu_set_error("fft: length beyond 2^31");
        return NULL;
    }
    hipLaunchKernelGGL(k_fft_twiddles, blocks_for(tw_len(M)), dim3(FFT_TPB), 0, st, tw,
                       tw_len(M), M);
    if (hipGetLastError() != hipSuccess) {
        agpu_set_error("fft: twiddle kernel launch failed");
        return NULL;
    }
    float2 *src = a, *dst = pong;
    long long Ns = 1;
    for (int i = 0; i < p->nrad; i++) {
        if (launch_pass(p->rad[i], src, dst, tw, M, Ns, st))
            return NULL;
        Ns *= p->rad[i];
        float2 *t = src;
        src = dst;
        dst = t;
    }
    return src;
}

/* Complex forward FFT of z (M values, left intact only for direct plans:
 * its buffer is used as working space) into the buffer returned. */
static float2 *cfft(long long M, float2 *z, float2 *work, hipStream_t st)
{
    fft_plan p;
    plan_fft(M, &p);
    if (!p.L)
        return direct_fft(&p, z, work, work + M, st);
    /* Bluestein: three transforms of length L */
    const long long L = p.L;
    float2 *W0 = work, *W1 = work + L, *W2 = work + 2 * L, *tw = work + 3 * L;
    fft_plan q;
    plan_fft(L, &q);
    hipLaunchKernelGGL(k_bs_kernel, blocks_for(L), dim3(FFT_TPB), 0, st, W1, M, L);
    float2 *B = direct_fft(&q, W1, W2, tw, st);
    if (!B)
        return NULL;
    float2 *free1 = (B == W1) ? W2 : W1;
    hipLaunchKernelGGL(k_bs_pre, blocks_for(L), dim3(FFT_TPB), 0, st, z, W0, M, L);
    float2 *A = direct_fft(&q, W0, free1, tw, st);
    if (!A)
        return NULL;
    float2 *free2 = (A == W0) ? free1 : W0;
    hipLaunchKernelGGL(k_bs_mul, blocks_for(L), dim3(FFT_TPB), 0, st, A, B, L);
    float2 *Cc = direct_fft(&q, A, free2, tw, st);
    if (!Cc)
        return NULL;
    hipLaunchKernelGGL(k_bs_post, blocks_for(M), dim3(FFT_TPB), 0, st, Cc, z, M);
    if (hipGetLastError() != hipSuccess) {
        agpu_set_error("fft: Bluestein kernel launch failed");
        return NULL;
    }
    return z;
}

static int check_len(const char *who, long long n)
{
    if (n < 4 || (n & 1)) {
        char m[160];
        snprintf(m, sizeof(m), "%s: the length must be even and >= 4", who);
        agpu_set_error(m);
        return -2;
    }
    return 0;
}

int agpu_realfft_inverse(const void *d_X, long long n, float *d_out, void *d_work, void *stream)
{
    hipStream_t st = (hipStream_t) stream;
    if (check_len("agpu_realfft_inverse", n))
        return -2;
    const long long M = n / 2;
    float2 *a = (float2 *) d_out, *work = (float2 *) d_work;
    /* conj Z into a (pairs (k, M-k) per thread, so d_X may be d_out) */
    hipLaunchKernelGGL(k_realfft_pre_inv, blocks_for(M / 2 + 1), dim3(FFT_TPB), 0, st,
                       (const float2 *) d_X, a, M);
    FFT_CHECK(hipGetLastError());
    float2 *F = cfft(M, a, work, st);
    if (!F)
        return -1;
    hipLaunchKernelGGL(k_realfft_post_inv, blocks_for(M), dim3(FFT_TPB), 0, st, F, a, M);
    FFT_CHECK(hipGetLastError());
    return 0;
}

int agpu_realfft(const float *d_x, long long n, void *d_out, void *d_work, void *stream)
{
    hipStream_t st = (hipStream_t) stream;
    if (check_len("agpu_realfft", n))
        return -2;
    const long long M = n / 2;
    float2 *a = (float2 *) d_out, *work = (float2 *) d_work;
    if ((const void *) d_x != (const void *) a)
        FFT_CHECK(hipMemcpyAsync(a, d_x, sizeof(float) * n, hipMemcpyDeviceToDevice, st));
    float2 *Z = cfft(M, a, work, st);
    if (!Z)
        return -1;
    /* X into d_out: in place (pairs per thread) or from the ping-pong buffer */
    hipLaunchKernelGGL(k_realfft_post, blocks_for(M / 2 + 1), dim3(FFT_TPB), 0, st, Z, a, M);
    FFT_CHECK(hipGetLastError());
    return 0;
}

/* ------------------------------------------------------------------ */
/* The forward real FFT without allocation (accel_realfft_forward_noalloc). */
/* Its work buffer is this file-scope device array; each device holds its  */
/* own copy, found once per device with hipGetSymbolAddress.               */

#define FFT_SCRATCH_COMPLEX (3LL << 20)   /* 24 MiB: the work of a 2^22-sample series */
static __device__ __align__(256) float2 scratch[FFT_SCRATCH_COMPLEX];

#define FFT_MAX_DEVICES 64
static float2 *g_scratch_dev[FFT_MAX_DEVICES];
static pthread_mutex_t g_scratch_mu = PTHREAD_MUTEX_INITIALIZER;

/* The device address of `scratch` on `device` (which must be current). */
static float2 *scratch_address(int device)
{
    if (device < 0 || device >= FFT_MAX_DEVICES)
        return NULL;
    pthread_mutex_lock(&g_scratch_mu);
    float2 *ptr = g_scratch_dev[device];
    if (!ptr) {
        void *p = NULL;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e == hipSuccess && p) {
            ptr = (float2 *) p;
            g_scratch_dev[device] = ptr;
        } else {
            char m[256];
            snprintf(m, sizeof(m), "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            agpu_set_error(m);
        }
    }
    pthread_mutex_unlock(&g_scratch_mu);
    return ptr;
}

extern "C" int accel_realfft_forward_noalloc(const float *d_series, long long n,
                                             accel_fcomplex *d_out, void *stream, int device)
{
    if (n < 4 || (n & 1)) {
        accel_set_error("accel_realfft_forward_noalloc: length %lld is not even and >= 4", n);
        return -1;
    }
    if (agpu_realfft_work(n) > FFT_SCRATCH_COMPLEX) {
        accel_set_error("accel_realfft_forward_noalloc: a series of %lld samples needs %lld "
                        "bytes of work, more than the %lld-byte scratch", n,
                        8 * agpu_realfft_work(n), 8 * FFT_SCRATCH_COMPLEX);
        return -1;
    }
    if (!d_series || !d_out) {
        accel_set_error("accel_realfft_forward_noalloc: null device buffer");
        return -1;
    }
    if (agpu_init(device)) {
        accel_set_error("accel_realfft_forward_noalloc: %s", agpu_last_error());
        return -1;
    }
    float2 *work = scratch_address(device);
    if (!work || agpu_realfft(d_series, n, d_out, work, stream)) {
        accel_set_error("accel_realfft_forward_noalloc: %s", agpu_last_error());
        return -1;
    }
    return 0;
}
