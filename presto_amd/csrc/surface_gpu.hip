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

int agpu_copy_d2d(void *dst, const void *src, size_t bytes, void *stream)
{
    SF_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t) stream));
    return 0;
}

}  /* extern "C" */
