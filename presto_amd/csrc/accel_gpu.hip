/*
 * accel_gpu.hip -- HIP kernels for the accelsearch hot path on MI355X
 * (gfx950) and the thin C ABI (agpu.h) that the C host calls.
 *
 *   k_resp_z / k_template_dft : FP64 z-response templates, FFT'd to the
 *       correlation length in FP64 and kept resident as FP32
 *       (src/accel_utils.c:153-170, src/responses.c:228-316).
 *   k_block_norms : exact block medians (radix select in LDS) and the
 *       block-median normalisation of src/accel_utils.c:1087-1099.
 *   k_corr4096 : fused overlap-save correlation -- spread spectrum
 *       segment -> Stockham radix-16 FFT in LDS -> multiply by each
 *       template -> inverse FFT -> |.|^2 -> raw plane row
 *       (the template loop of src/accel_utils.c:1141-1190).
 *   k_harmsum : harmonic summing with the reference's index rules and
 *       per-block normalisation, and the threshold of search_ffdotpows
 *       (src/accel_utils.c:1261-1452).
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include "agpu.h"

static char g_err[512];

#define AGPU_CHECK(call)                                                          \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                   \
            snprintf(g_err, sizeof(g_err), "%s:%d %s: %s", __FILE__, __LINE__,    \
                     #call, hipGetErrorString(e_));                               \
            return -1;                                                            \
        }                                                                         \
    } while (0)

static int fail(const char *msg)
{
    snprintf(g_err, sizeof(g_err), "%s", msg);
    return -2;
}

/* ------------------------------------------------------------------ */
/* FP64 Fresnel integrals (Cephes rational forms, the published tables  */
/* used by src/fresnl.c:76-229)                                         */
/* ------------------------------------------------------------------ */
__constant__ double c_sn[6] = {
    -2.99181919401019853726E3, 7.08840045257738576863E5, -6.29741486205862506537E7,
    2.54890880573376359104E9, -4.42979518059697779103E10, 3.18016297876567817986E11};
__constant__ double c_sd[6] = {
    2.81376268889994315696E2, 4.55847810806532581675E4, 5.17343888770096400730E6,
    4.19320245898111231129E8, 2.24411795645340920940E10, 6.07366389490084639049E11};
__constant__ double c_cn[6] = {
    -4.98843114573573548651E-8, 9.50428062829859605134E-6, -6.45191435683965050962E-4,
    1.88843319396703850064E-2, -2.05525900955013891793E-1, 9.99999999999999998822E-1};
__constant__ double c_cd[7] = {
    3.99982968972495980367E-12, 9.15439215774657478799E-10, 1.25001862479598821474E-7,
    1.22262789024179030997E-5, 8.68029542941784300606E-4, 4.12142090722199792936E-2,
    1.00000000000000000118E0};
__constant__ double c_fn[10] = {
    4.21543555043677546506E-1, 1.43407919780758885261E-1, 1.15220955073585758835E-2,
    3.45017939782574027900E-4, 4.63613749287867322088E-6, 3.05568983790257605827E-8,
    1.02304514164907233465E-10, 1.72010743268161828879E-13, 1.34283276233062758925E-16,
    3.76329711269987889006E-20};
__constant__ double c_fd[10] = {
    7.51586398353378947175E-1, 1.16888925859191382142E-1, 6.44051526508858611005E-3,
    1.55934409164153020873E-4, 1.84627567348930545870E-6, 1.12699224763999035261E-8,
    3.60140029589371370404E-11, 5.88754533621578410010E-14, 4.52001434074129701496E-17,
    1.25443237090011264384E-20};
__constant__ double c_gn[11] = {
    5.04442073643383265887E-1, 1.97102833525523411709E-1, 1.87648584092575249293E-2,
    6.84079380915393090172E-4, 1.15138826111884280931E-5, 9.82852443688422223854E-8,
    4.45344415861750144738E-10, 1.08268041139020870318E-12, 1.37555460633261799868E-15,
    8.36354435630677421531E-19, 1.86958710162783235106E-22};
__constant__ double c_gd[11] = {
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

__device__ static unsigned radix_select(const unsigned *keys, int n, int k, unsigned *hist,
                                        unsigned *shared2)
{
    unsigned prefix = 0, mask = 0;
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int i = threadIdx.x; i < 256; i += blockDim.x)
            hist[i] = 0;
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            unsigned key = keys[i];
            if ((key & mask) == prefix)
                atomicAdd(&hist[(key >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned cum = 0, bucket = 0;
            for (int b = 0; b < 256; b++) {
                if ((unsigned) k < cum + hist[b]) {
                    bucket = b;
                    break;
                }
                cum += hist[b];
            }
            shared2[0] = prefix | (bucket << shift);
            shared2[1] = (unsigned) k - cum;
        }
        __syncthreads();
        prefix = shared2[0];
        k = (int) shared2[1];
        mask |= 255u << shift;
        __syncthreads();
    }
    return prefix;
}

__global__ __launch_bounds__(256) void k_block_norms(const float2 *spec, long long numbins,
                                                     long long lobin_file,
                                                     const agpu_window *win, double *norm,
                                                     float *nsq)
{
    /* the reference forms r*r + i*i with two roundings (no FMA) */
#pragma clang fp contract(off)
    extern __shared__ unsigned keys[];
    __shared__ unsigned hist[256];
    __shared__ unsigned sh2[2];
    const agpu_window w = win[blockIdx.x];
    const int n = w.numdata;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        long long b = w.lobin + i - lobin_file;
        float p = 0.0f;
        if (b >= 0 && b < numbins) {
            float2 v = spec[b];
            p = v.x * v.x + v.y * v.y;
        }
        keys[i] = __float_as_uint(p);
    }
    __syncthreads();
    const int lhs = (n - 1) / 2, rhs = n / 2;
    float a = __uint_as_float(radix_select(keys, n, lhs, hist, sh2));
    float med;
    if (lhs == rhs) {
        med = a;
    } else {
        float b = __uint_as_float(radix_select(keys, n, rhs, hist, sh2));
        med = (float) (0.5 * ((double) a + (double) b));
    }
    if (threadIdx.x == 0) {
        double nv = 1.0 / sqrt((double) med / 0.69314718055994530942);
        norm[blockIdx.x] = nv;
        nsq[blockIdx.x] = (float) (nv * nv);
    }
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

/* Input v[r] = x[j + 256 r]; output v[r] = X[j + 256 r]. tw1/tw2 hold the
 * inverse-sign pass twiddles W_256^{r (j%16)} and W_4096^{r j}. */
template <int S>
__device__ __forceinline__ void fft4096(float2 *v, float2 *lds, const float2 *tw1,
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
        float2 w = tw1[r];
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
 * template groups and keeps those templates in its own L2. */
__global__ __launch_bounds__(256) void k_corr4096(const float2 *spec, long long numbins,
                                                  long long lobin_file, const float2 *T,
                                                  int nt, int hw, int uselen, long long hb_end,
                                                  float *plane, long long stride, int groups,
                                                  int tpg)
{
    __shared__ float2 lds[FFT_LDS];
    const int j = threadIdx.x;
    const int g = blockIdx.x % groups;
    const long long seg = blockIdx.x / groups;
    const long long seg_hb = seg * (long long) uselen;
    const long long lobin = seg_hb / 2 - hw;
    const int t0 = g * tpg;
    if (t0 >= nt)
        return;
    const int t1 = min(nt, t0 + tpg);

    float2 tw1[16], tw2[16];
#pragma unroll
    for (int r = 1; r < 16; r++) {
        double s, c;
        sincospi(2.0 * (double) (r * (j & 15)) / 256.0, &s, &c);
        tw1[r] = make_float2((float) c, (float) s);
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
    fft4096<-1>(v, lds, tw1, tw2);
    float2 D[16];
#pragma unroll
    for (int r = 0; r < 16; r++)
        D[r] = v[r];

    for (int t = t0; t < t1; t++) {
        const float2 *Tt = T + (size_t) t * 4096;
#pragma unroll
        for (int r = 0; r < 16; r++)
            v[r] = cmul(D[r], Tt[j + 256 * r]);
        fft4096<1>(v, lds, tw1, tw2);
        float *row = plane + (size_t) t * stride;
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int q = j + 256 * r - 2 * hw;
            const long long hb = seg_hb + q;
            if (q >= 0 && q < uselen && hb < hb_end)
                row[hb] = __builtin_fmaf(v[r].x, v[r].x, v[r].y * v[r].y);
        }
    }
}

/* ------------------------------------------------------------------ */
/* Harmonic summing and thresholding                                    */
/* ------------------------------------------------------------------ */
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

__device__ __forceinline__ float inmem_nsq(const agpu_harmsum_args &a, long long rh)
{
    if (rh >= a.rh0) {
        long long bb = (rh - a.rh0) / a.corr_uselen;
        return a.nsq[a.npre + bb];
    }
    if (rh >= 16 && a.npre > 0) {
        long long pb = (rh - 16) / a.corr_uselen;
        if (pb < a.npre)
            return a.nsq[pb];
    }
    return 0.0f;
}

template <int NSTAGES, int INMEM>
__global__ __launch_bounds__(256) void k_harmsum(agpu_harmsum_args a)
{
    /* each sub-harmonic power is rounded before it is added, as in
     * add_ffdotpows (src/accel_utils.c:1276-1277) */
#pragma clang fp contract(off)
    const int row = blockIdx.y;
    const int wi = row / a.numz, zi = row - wi * a.numz;
    const long long cell = (long long) blockIdx.x * 256 + threadIdx.x;
    const long long ncell = (long long) a.nblocks * a.corr_uselen;
    if (cell >= ncell)
        return;
    const long long rh = a.rh0 + cell;
    const int b = (int) (cell / a.corr_uselen);
    const int ii = (int) (cell - (long long) b * a.corr_uselen);
    const float *nsq_b = a.nsq + (size_t) b * (a.nfrac + 1);
    const float f0 = INMEM ? inmem_nsq(a, rh) : nsq_b[0];
    float acc = a.plane[(size_t) row * a.stride + rh] * f0;
    if (acc > a.powcut[0])
        emit_hit(a, b, 0, wi, zi, ii, acc);
    int f = 0;
#pragma unroll
    for (int s = 1; s < NSTAGES; s++) {
        const int H = 1 << s;
#pragma unroll
        for (int k = 1; k < H; k += 2, f++) {
            const long long q = rh * k;
            long long rs;
            if (INMEM) {
                rs = (q + (H >> 1)) >> s;
            } else {
                rs = q >> s;
                const long long rem = q & (H - 1);
                if (rem > (H >> 1) || (rem == (H >> 1) && (rs & 1)))
                    rs++;
            }
            const int zs = a.zmap[f * a.numz + zi];
            const int ws = a.wmap[f * a.numw + wi];
            const float v = a.plane[((size_t) ws * a.numz + zs) * a.stride + rs];
            const float nn = INMEM ? inmem_nsq(a, rs) : nsq_b[1 + f];
            acc = acc + v * nn;
        }
        if (acc > a.powcut[s])
            emit_hit(a, b, s, wi, zi, ii, acc);
    }
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

int agpu_malloc(void **p, size_t bytes)
{
    AGPU_CHECK(hipMalloc(p, bytes));
    return 0;
}

int agpu_free(void *p)
{
    if (p)
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

int agpu_gen_templates_z(const int *zs, const int *numkern, int nt, int L, void *d_T,
                         void *stream)
{
    hipStream_t st = (hipStream_t) stream;
    if (nt <= 0 || L <= 0 || nt > 65535)
        return fail("agpu_gen_templates_z: bad template count or length");
    int nkmax = 0;
    for (int t = 0; t < nt; t++) {
        if (numkern[t] <= 0 || numkern[t] % 4 || numkern[t] > L)
            return fail("agpu_gen_templates_z: kernel width must be a positive multiple of 4 <= L");
        if (numkern[t] > nkmax)
            nkmax = numkern[t];
    }
    int *d_z = NULL, *d_nk = NULL;
    double2 *d_resp = NULL;
    AGPU_CHECK(hipMalloc(&d_z, sizeof(int) * nt));
    AGPU_CHECK(hipMalloc(&d_nk, sizeof(int) * nt));
    AGPU_CHECK(hipMalloc(&d_resp, sizeof(double2) * (size_t) nt * nkmax));
    AGPU_CHECK(hipMemcpyAsync(d_z, zs, sizeof(int) * nt, hipMemcpyHostToDevice, st));
    AGPU_CHECK(hipMemcpyAsync(d_nk, numkern, sizeof(int) * nt, hipMemcpyHostToDevice, st));
    dim3 g1((nkmax + 255) / 256, nt);
    hipLaunchKernelGGL(k_resp_z, g1, dim3(256), 0, st, d_z, d_nk, nkmax, d_resp);
    AGPU_CHECK(hipGetLastError());
    dim3 g2((L + 255) / 256, nt);
    hipLaunchKernelGGL(k_template_dft, g2, dim3(256), 0, st, d_resp, d_nk, nkmax, L,
                       (float2 *) d_T);
    AGPU_CHECK(hipGetLastError());
    AGPU_CHECK(hipStreamSynchronize(st));
    AGPU_CHECK(hipFree(d_z));
    AGPU_CHECK(hipFree(d_nk));
    AGPU_CHECK(hipFree(d_resp));
    return 0;
}

int agpu_block_norms(const void *d_spec, long long numbins, long long lobin_file,
                     const agpu_window *d_win, int nwin, int maxn, double *d_norm,
                     float *d_nsq, void *stream)
{
    if (nwin <= 0)
        return 0;
    if (maxn <= 0 || maxn > NORM_MAXN)
        return fail("agpu_block_norms: window longer than the LDS selection buffer");
    hipLaunchKernelGGL(k_block_norms, dim3(nwin), dim3(256), sizeof(unsigned) * maxn,
                       (hipStream_t) stream, (const float2 *) d_spec, numbins, lobin_file,
                       d_win, d_norm, d_nsq);
    AGPU_CHECK(hipGetLastError());
    return 0;
}

int agpu_correlate(const void *d_spec, long long numbins, long long lobin_file,
                   const void *d_T, int nt, int L, int hw, int uselen, long long hb_end,
                   float *d_plane, long long stride, int groups, void *stream)
{
    if (L != 4096)
        return fail("agpu_correlate: only L = 4096 is built");
    if (uselen <= 0 || uselen % 2 || uselen > L - 4 * hw + 1)
        return fail("agpu_correlate: uselen must be even and <= L - 4*hw + 1");
    if (groups <= 0)
        return fail("agpu_correlate: groups must be positive");
    if (hb_end > stride)
        return fail("agpu_correlate: plane stride shorter than the half-bin range");
    long long nseg = (hb_end + uselen - 1) / uselen;
    int tpg = (nt + groups - 1) / groups;
    long long nblk = nseg * groups;
    if (nblk > 0x7fffffffLL)
        return fail("agpu_correlate: grid too large");
    hipLaunchKernelGGL(k_corr4096, dim3((unsigned) nblk), dim3(256), 0, (hipStream_t) stream,
                       (const float2 *) d_spec, numbins, lobin_file, (const float2 *) d_T, nt,
                       hw, uselen, hb_end, d_plane, stride, groups, tpg);
    AGPU_CHECK(hipGetLastError());
    return 0;
}

int agpu_harmsum(const agpu_harmsum_args *a, void *stream)
{
    long long ncell = (long long) a->nblocks * a->corr_uselen;
    if (ncell <= 0)
        return 0;
    if (a->nstages < 1 || a->nstages > AGPU_MAX_STAGES)
        return fail("agpu_harmsum: 1..6 stages");
    long long gx = (ncell + 255) / 256;
    int rows = a->numz * a->numw;
    if (gx > 0x7fffffffLL || rows > 65535)
        return fail("agpu_harmsum: grid too large");
    dim3 grid((unsigned) gx, rows);
    hipStream_t st = (hipStream_t) stream;
#define HS_CASE(NS)                                                                     \
    case NS:                                                                            \
        if (a->inmem)                                                                   \
            hipLaunchKernelGGL((k_harmsum<NS, 1>), grid, dim3(256), 0, st, *a);         \
        else                                                                            \
            hipLaunchKernelGGL((k_harmsum<NS, 0>), grid, dim3(256), 0, st, *a);         \
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

} /* extern "C" */
