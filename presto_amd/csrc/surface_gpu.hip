/*
 * surface_gpu.hip -- device work of the accel.h surface (accel_surface.c):
 * block planes cut from the resident raw plane, the sub-harmonic
 * additions of the standard and in-memory searches, and the threshold
 * scan of search_ffdotpows.  Planes are [numws][numzs][numrs] FP32, as the
 * reference's ffdotpows powers (src/accel_utils.c:1024-1452).
 *
 * The harmonic index rules are evaluated here in FP64 with the reference's
 * operation order and no contraction into fused multiply-adds, so that
 * every index equals the one the reference's host code computes.
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include "agpu.h"
#include "dev_resp.h"

#define SF_TPB 256

#define SF_CHECK(call)                                                            \
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

/* ---- index rules: src/accel_utils.c:54-102 ---- */
__device__ static int sf_nearest_int(double x)
{
    return (int) (x < 0 ? x - 0.5 : x + 0.5);
}

__device__ static double sf_required_r(double hf, double rfull)
{
#pragma clang fp contract(off)
    return rint(2.0 * rfull * hf) * 0.5;
}

__device__ static int sf_required_z(double hf, double zfull)
{
#pragma clang fp contract(off)
    return sf_nearest_int(0.5 * zfull * hf) * 2;
}

__device__ static int sf_required_w(double hf, double wfull)
{
#pragma clang fp contract(off)
    return sf_nearest_int(0.05 * wfull * hf) * 20;
}

__device__ static int sf_index_r(double r, double lor)
{
#pragma clang fp contract(off)
    return (int) ((r - lor) * 2 + 1e-14);
}

__device__ static int sf_index_z(double z, double loz)
{
#pragma clang fp contract(off)
    return (int) ((z - loz) * 0.5 + 1e-14);
}

__device__ static int sf_index_w(double w, double low)
{
#pragma clang fp contract(off)
    return (int) ((w - low) * 0.05 + 1e-14);
}

/* out[(w numzs + z) numrs + r] = nsq * raw(row (w+woff) numz + z+zoff,
 * half-bin hb0 + r): grid (ceil(numrs/256), numzs * numws). */
__global__ void k_sf_extract(const float *plane, int nrows, int numz_plane, int woff, int zoff,
                             int numzs, int numrs, long long hb0, const float *nsq, float *out)
{
    const int r = blockIdx.x * SF_TPB + threadIdx.x;
    if (r >= numrs)
        return;
    const int wz = blockIdx.y;
    const int w = wz / numzs, z = wz - w * numzs;
    const int row = (w + woff) * numz_plane + (z + zoff);
    const long long hb = hb0 + r;
    out[(size_t) wz * numrs + r] = nsq[0] * plane[AGPU_PLANE_IDX(row, hb, nrows)];
}

/* add_ffdotpows (src/accel_utils.c:1261-1282): grid (ceil(numrs/256),
 * numzs * numws) over the fundamental. */
__global__ void k_sf_add(float *fund, const float *sub, agpu_sf_add_args a)
{
#pragma clang fp contract(off)
    const int kk = blockIdx.x * SF_TPB + threadIdx.x;
    if (kk >= a.fnumrs)
        return;
    const int wz = blockIdx.y;
    const int ii = wz / a.fnumzs, jj = wz - ii * a.fnumzs;
    /* the sub-harmonic's tables (src/accel_utils.c:1042-1053) */
    const double rr = a.sub_fullrlo + kk * 0.5;
    const int rind = (unsigned short) sf_index_r(sf_required_r(a.hf, rr), (double) a.sub_rlo);
    const double zz = a.obs_zlo + jj * 2;
    const int zind = (unsigned short) sf_index_z((double) sf_required_z(a.hf, zz),
                                                 (double) a.sub_zlo);
    const int ww = a.fund_wlo + ii * 20;
    const int wind = sf_index_w((double) sf_required_w(a.hf, (double) ww), (double) a.sub_wlo);
    fund[(size_t) wz * a.fnumrs + kk] +=
        sub[((size_t) wind * a.snumzs + zind) * a.snumrs + rind];
}

/* fund_to_ffdotplane (src/accel_utils.c:1219-1234): full[z][2 rlo + r] */
__global__ void k_sf_inmem_store(const float *fund, int numzs, int numrs, float *full,
                                 long long rlen, long long col0)
{
    const int r = blockIdx.x * SF_TPB + threadIdx.x;
    if (r >= numrs)
        return;
    const int z = blockIdx.y;
    full[(size_t) z * rlen + col0 + r] = fund[(size_t) z * numrs + r];
}

/* inmem_add_ffdotpows (src/accel_utils.c:1312-1362) */
__global__ void k_sf_inmem_add(float *fund, int numzs, int numrs, const float *full,
                               long long rlen, int rlo, double hf, int zlo)
{
#pragma clang fp contract(off)
    const int jj = blockIdx.x * SF_TPB + threadIdx.x;
    if (jj >= numrs)
        return;
    const int ii = blockIdx.y;
    const int rrint = 2 * rlo + jj;
    const int rind = (int) (rrint * hf + 0.5);
    const int zz = zlo + ii * 2;
    const int zind = sf_index_z((double) sf_required_z(hf, (double) zz), (double) zlo);
    fund[(size_t) ii * numrs + jj] += full[(size_t) zind * rlen + rind];
}

/* search_ffdotpows threshold (src/accel_utils.c:1424-1429): every element
 * above powcut, as (flat index, power) */
__global__ void k_sf_threshold(const float *pow, long long n, float powcut, agpu_sf_hit *hits,
                               unsigned *count, unsigned cap)
{
    for (long long i = (long long) blockIdx.x * SF_TPB + threadIdx.x; i < n;
         i += (long long) gridDim.x * SF_TPB) {
        const float p = pow[i];
        if (p > powcut) {
            const unsigned k = atomicAdd(count, 1u);
            if (k < cap) {
                hits[k].idx = i;
                hits[k].pow = p;
            }
        }
    }
}

extern "C" {

int agpu_sf_extract(const float *d_plane, int nrows, long long ntiles, int numz_plane, int woff,
                    int zoff, int numws, int numzs, int numrs, long long hb0, const float *d_nsq,
                    float *d_out, void *stream)
{
    if (numws <= 0 || numzs <= 0 || numrs <= 0)
        return 0;
    if (woff < 0 || zoff < 0 || (woff + numws - 1) * numz_plane + zoff + numzs - 1 >= nrows ||
        zoff + numzs > numz_plane)
        return agpu_set_error("agpu_sf_extract: (w, z) rows outside the plane"), -1;
    if (hb0 < 0 || hb0 + numrs > ntiles * AGPU_TILE)
        return agpu_set_error("agpu_sf_extract: half-bins outside the plane"), -1;
    if ((long long) numws * numzs > 65535)
        return agpu_set_error("agpu_sf_extract: too many rows for one launch"), -1;
    dim3 grid((numrs + SF_TPB - 1) / SF_TPB, numws * numzs);
    hipLaunchKernelGGL(k_sf_extract, grid, dim3(SF_TPB), 0, (hipStream_t) stream, d_plane,
                       nrows, numz_plane, woff, zoff, numzs, numrs, hb0, d_nsq, d_out);
    SF_CHECK(hipGetLastError());
    return 0;
}

int agpu_sf_add(float *d_fund, const float *d_sub, const agpu_sf_add_args *a, void *stream)
{
    if (a->fnumws <= 0 || a->fnumzs <= 0 || a->fnumrs <= 0)
        return 0;
    if ((long long) a->fnumws * a->fnumzs > 65535)
        return agpu_set_error("agpu_sf_add: too many rows for one launch"), -1;
    dim3 grid((a->fnumrs + SF_TPB - 1) / SF_TPB, a->fnumws * a->fnumzs);
    hipLaunchKernelGGL(k_sf_add, grid, dim3(SF_TPB), 0, (hipStream_t) stream, d_fund, d_sub,
                       *a);
    SF_CHECK(hipGetLastError());
    return 0;
}

int agpu_sf_inmem_store(const float *d_fund, int numzs, int numrs, float *d_full,
                        long long rlen, long long col0, void *stream)
{
    if (numzs <= 0 || numrs <= 0)
        return 0;
    if (numzs > 65535)
        return agpu_set_error("agpu_sf_inmem_store: too many rows"), -1;
    dim3 grid((numrs + SF_TPB - 1) / SF_TPB, numzs);
    hipLaunchKernelGGL(k_sf_inmem_store, grid, dim3(SF_TPB), 0, (hipStream_t) stream, d_fund,
                       numzs, numrs, d_full, rlen, col0);
    SF_CHECK(hipGetLastError());
    return 0;
}

int agpu_sf_inmem_add(float *d_fund, int numzs, int numrs, const float *d_full, long long rlen,
                      int rlo, double hf, int zlo, void *stream)
{
    if (numzs <= 0 || numrs <= 0)
        return 0;
    if (numzs > 65535)
        return agpu_set_error("agpu_sf_inmem_add: too many rows"), -1;
    dim3 grid((numrs + SF_TPB - 1) / SF_TPB, numzs);
    hipLaunchKernelGGL(k_sf_inmem_add, grid, dim3(SF_TPB), 0, (hipStream_t) stream, d_fund,
                       numzs, numrs, d_full, rlen, rlo, hf, zlo);
    SF_CHECK(hipGetLastError());
    return 0;
}

int agpu_sf_threshold(const float *d_pow, long long n, float powcut, agpu_sf_hit *d_hits,
                      unsigned *d_count, unsigned cap, void *stream)
{
    if (n <= 0)
        return 0;
    long long nb = (n + SF_TPB - 1) / SF_TPB;
    if (nb > 4096)
        nb = 4096;
    hipLaunchKernelGGL(k_sf_threshold, dim3((unsigned) nb), dim3(SF_TPB), 0,
                       (hipStream_t) stream, d_pow, n, powcut, d_hits, d_count, cap);
    SF_CHECK(hipGetLastError());
    return 0;
}

/* -locpow window normalisation (src/accel_utils.c:1100-1115 with
 * corr_loc_pow, src/corr_routines.c:309-370): the window's bins
 * spec[lobin - lobin_file + i] (zero outside the spectrum), their FP32
 * powers p, the local power loc(i) = (2/n) ((n/2)/20) sum_{j=6..15}
 * (p[c+j] + p[c-j]) with c clamped to [15, n-16] (the reference's copies at
 * both ends), accumulated in FP64, and the bin scaled by the
 * 0x5f3759df inverse square root of loc with one Newton step
 * (src/misc_utils.c:291-302). */
__device__ static float2 sf_bin(const float2 *spec, long long numbins, long long b)
{
    return (b >= 0 && b < numbins) ? spec[b] : make_float2(0.0f, 0.0f);
}

__global__ __launch_bounds__(SF_TPB) void k_sf_locpow_norm(const float2 *spec, long long numbins,
                                                           long long first, int n, float2 *wdata)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * SF_TPB + threadIdx.x;
    if (i >= n)
        return;
    const int c = i < 15 ? 15 : (i > n - 16 ? n - 16 : i);
    double s = 0.0;
    for (int j = 6; j <= 15; j++) {
        const float2 a = sf_bin(spec, numbins, first + c + j);
        const float2 b = sf_bin(spec, numbins, first + c - j);
        const float pa = a.x * a.x + a.y * a.y, pb = b.x * b.x + b.y * b.y;
        s += (double) pa + (double) pb;
    }
    const double scale = (2.0 / n) * (float) ((n / 2) / 20);
    const float loc = (float) (scale * s);
    const float y = __int_as_float(0x5f3759df - (__float_as_int(loc) >> 1));
    const float nl = y * (1.5f - 0.5f * loc * y * y);
    const float2 d = sf_bin(spec, numbins, first + i);
    wdata[i] = make_float2(d.x * nl, d.y * nl);
}

/* The block plane of one -locpow window (src/accel_utils.c:1100-1188):
 * the window's numdata bins, each already scaled by invsqrtf of its local
 * power, spread to half-bins and correlated with the z row's response
 * (gen_z_response, numbetween 2, its own 2 numbetween z_resp_halfwidth(z)
 * taps); cell kk is |sum|^2 at spread index kk + 2 binoffset.  The
 * reference gets the same sums from FFTs; here they are direct, with FP64
 * accumulation of the FP32 products.  One workgroup per (z row, 256 cells);
 * the row's taps are made in LDS. */
__global__ __launch_bounds__(SF_TPB) void k_sf_locpow_plane(const float2 *wdata, int numdata,
                                                            int binoffset, int zlo, int numrs,
                                                            float *out)
{
    extern __shared__ float2 taps[];
    const int zi = blockIdx.y;
    const double z = (double) zlo + 2.0 * zi;
    const int nk = 2 * 2 * zr_halfwidth(z, 0);
    for (int m = threadIdx.x; m < nk; m += SF_TPB)
        taps[m] = resp_tap_nb(0.0, 2, z, nk, m);
    __syncthreads();
    const int kk = blockIdx.x * SF_TPB + threadIdx.x;
    if (kk >= numrs)
        return;
    const int j = kk + 2 * binoffset;
    double ar = 0.0, ai = 0.0;
    /* taps m in [-nk/2, nk/2) with j + m even and inside the window */
    int m0 = -nk / 2;
    if ((j + m0) & 1)
        m0++;
    for (int m = m0; m < nk / 2; m += 2) {
        const int i = (j + m) >> 1;
        if (i < 0 || i >= numdata)
            continue;
        const float2 d = wdata[i], q = taps[nk / 2 + m];
        ar += (double) d.x * q.x + (double) d.y * q.y;
        ai += (double) d.y * q.x - (double) d.x * q.y;
    }
    const float fr = (float) ar, fi = (float) ai;
    out[(long long) zi * numrs + kk] = fr * fr + fi * fi;
}

/* The same plane for a window with w rows (jerk searches): row
 * r = w index * numzs + z index correlates with its own nks[r] FP32 taps,
 * R + r*stride (agpu_gen_response_taps), staged in LDS. */
__global__ __launch_bounds__(SF_TPB) void k_sf_locpow_plane_tab(const float2 *wdata, int numdata,
                                                                int binoffset, int numrs,
                                                                const float2 *R, const int *nks,
                                                                int stride, float *out)
{
    extern __shared__ float2 taps[];
    const int row = blockIdx.y;
    const int nk = nks[row];
    for (int m = threadIdx.x; m < nk; m += SF_TPB)
        taps[m] = R[(size_t) row * stride + m];
    __syncthreads();
    const int kk = blockIdx.x * SF_TPB + threadIdx.x;
    if (kk >= numrs)
        return;
    const int j = kk + 2 * binoffset;
    double ar = 0.0, ai = 0.0;
    int m0 = -nk / 2;
    if ((j + m0) & 1)
        m0++;
    for (int m = m0; m < nk / 2; m += 2) {
        const int i = (j + m) >> 1;
        if (i < 0 || i >= numdata)
            continue;
        const float2 d = wdata[i], q = taps[nk / 2 + m];
        ar += (double) d.x * q.x + (double) d.y * q.y;
        ai += (double) d.y * q.x - (double) d.x * q.y;
    }
    const float fr = (float) ar, fi = (float) ai;
    out[(long long) row * numrs + kk] = fr * fr + fi * fi;
}

int agpu_sf_locpow_norm(const void *d_spec, long long numbins, long long lobin_file,
                        long long lobin, int numdata, void *d_wdata, void *stream)
{
    if (numdata < 31)
        return agpu_set_error("agpu_sf_locpow_norm: window shorter than 31 bins"), -1;
    hipLaunchKernelGGL(k_sf_locpow_norm, dim3((unsigned) ((numdata + SF_TPB - 1) / SF_TPB)),
                       dim3(SF_TPB), 0, (hipStream_t) stream, (const float2 *) d_spec, numbins,
                       lobin - lobin_file, numdata, (float2 *) d_wdata);
    SF_CHECK(hipGetLastError());
    return 0;
}

int agpu_sf_locpow_plane(const void *d_wdata, int numdata, int binoffset, int zlo, int numzs,
                         int numrs, float *d_out, void *stream)
{
    if (numdata <= 0 || numzs <= 0 || numrs <= 0 || binoffset < 0)
        return agpu_set_error("agpu_sf_locpow_plane: bad geometry"), -1;
    /* the widest row's taps */
    const double zmax = fabs((double) zlo) > fabs((double) zlo + 2.0 * (numzs - 1))
                            ? fabs((double) zlo) : fabs((double) zlo + 2.0 * (numzs - 1));
    const int nkmax = 2 * 2 * ((int) (0.5 * 1.1 * zmax) + 16);
    const size_t lds = sizeof(float2) * (size_t) nkmax;
    if (lds > 64 * 1024)
        return agpu_set_error("agpu_sf_locpow_plane: response too long for LDS"), -1;
    dim3 grid((unsigned) ((numrs + SF_TPB - 1) / SF_TPB), (unsigned) numzs);
    hipLaunchKernelGGL(k_sf_locpow_plane, grid, dim3(SF_TPB), lds, (hipStream_t) stream,
                       (const float2 *) d_wdata, numdata, binoffset, zlo, numrs, d_out);
    SF_CHECK(hipGetLastError());
    return 0;
}

int agpu_sf_locpow_plane_tab(const void *d_wdata, int numdata, int binoffset, int nrows,
                             int numrs, const void *d_R, const int *d_nk, int stride,
                             float *d_out, void *stream)
{
    if (numdata <= 0 || nrows <= 0 || nrows > 65535 || numrs <= 0 || binoffset < 0 ||
        stride <= 0)
        return agpu_set_error("agpu_sf_locpow_plane_tab: bad geometry"), -1;
    const size_t lds = sizeof(float2) * (size_t) stride;
    if (lds > 64 * 1024)
        return agpu_set_error("agpu_sf_locpow_plane_tab: response too long for LDS"), -1;
    dim3 grid((unsigned) ((numrs + SF_TPB - 1) / SF_TPB), (unsigned) nrows);
    hipLaunchKernelGGL(k_sf_locpow_plane_tab, grid, dim3(SF_TPB), lds, (hipStream_t) stream,
                       (const float2 *) d_wdata, numdata, binoffset, numrs, (const float2 *) d_R,
                       d_nk, stride, d_out);
    SF_CHECK(hipGetLastError());
    return 0;
}

int agpu_copy_d2d(void *dst, const void *src, size_t bytes, void *stream)
{
    SF_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t) stream));
    return 0;
}

}  /* extern "C" */
