/*
 * dedisp_gpu.hip -- subband de-dispersion of a filterbank held in HBM
 * (prepsubband's data path, src/prepsubband.c:918-1040 with
 * src/dispersion.c:164-224).  Two kernels:
 *
 *   k_dd_subbands  sub(t, s) = sum over the channels c of subband s, in
 *                  ascending order, of raw(t + idispdt[c], c); then the
 *                  downsampled ds(v, s) = 0 + (sum_k<D sub(vD + k, s)) / D
 *                  (the plain sub(v, s) when D = 1).  Output subband-major
 *                  [nsub][nds] so that the DM sums read along time.
 *   k_dd_dms       out(d, u) = -0 + sum over s, in ascending order, of
 *                  ds(u + offsets[d][s], s).
 *
 * Every sum is a chain of FP32 additions in the reference's order, so the
 * results are the reference's bits.  The caller sizes nds and nout so that
 * every index stays inside raw and ds (the launchers check it).
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include "agpu.h"

#define DD_TPB 256

#define DD_CHECK(call)                                                            \
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

/* one thread per (v, s); lanes run along v, so the channel loop walks one
 * row segment of raw at a time and its cache lines serve the next channels */
__global__ __launch_bounds__(DD_TPB) void k_dd_subbands(const float *raw, int numchan,
                                                        const int *idispdt, int cps, int nsub,
                                                        int downsamp, long long nds, float *ds)
{
#pragma clang fp contract(off)
    const long long v = (long long) blockIdx.x * DD_TPB + threadIdx.x;
    const int s = blockIdx.y;
    if (v >= nds)
        return;
    const int c0 = s * cps;
    float ftmp = 0.0f;
    for (int k = 0; k < downsamp; k++) {
        const long long t = v * downsamp + k;
        float sub = 0.0f;
        for (int j = 0; j < cps; j++) {
            const int c = c0 + j;
            sub = sub + raw[(t + idispdt[c]) * numchan + c];
        }
        if (downsamp == 1) {
            ds[(long long) s * nds + v] = sub;
            return;
        }
        ftmp = ftmp + sub;
    }
    ds[(long long) s * nds + v] = 0.0f + ftmp / (float) downsamp;
}

/* one workgroup = 256 output samples of one DM; the DMs of a time tile
 * are consecutive workgroups of one XCD (ids are dealt round-robin to the
 * 8 XCDs), so the tile's ds rows are read from HBM once per XCD */
__global__ __launch_bounds__(DD_TPB) void k_dd_dms(const float *ds, long long nds, int nsub,
                                                   const int *offsets, int numdms,
                                                   long long nwrite, long long numout,
                                                   long long ntiles, float *out)
{
#pragma clang fp contract(off)
    const long long bid = blockIdx.x;
    const int xcd = (int) (bid & 7);
    const long long local = bid >> 3;
    const int d = (int) (local % numdms);
    const long long tile = (local / numdms) * 8 + xcd;
    if (tile >= ntiles)
        return;
    const long long u = tile * DD_TPB + threadIdx.x;
    if (u >= nwrite)
        return;
    const int *off = offsets + (long long) d * nsub;
    float acc = -0.0f;
    for (int s = 0; s < nsub; s++)
        acc = acc + ds[(long long) s * nds + u + off[s]];
    out[(long long) d * numout + u] = acc;
}

__global__ __launch_bounds__(DD_TPB) void k_dd_fill(float *out, int numdms, long long numout,
                                                    long long from, float value)
{
    const long long n = numout - from;
    const long long i = (long long) blockIdx.x * DD_TPB + threadIdx.x;
    if (i >= n * numdms)
        return;
    const long long d = i / n, u = from + i % n;
    out[d * numout + u] = value;
}

extern "C" {

int agpu_dedisp_subbands(const float *d_raw, long long nspec, int numchan, const int *d_idispdt,
                         int maxdelay, int nsub, int downsamp, long long nds, float *d_ds,
                         void *stream)
{
    if (numchan <= 0 || nsub <= 0 || numchan % nsub || downsamp < 1 || nds < 0 || maxdelay < 0)
        return agpu_set_error("agpu_dedisp_subbands: bad geometry"), -1;
    /* the last sample read: (nds D - 1) + maxdelay */
    if (nds && nds * downsamp - 1 + maxdelay >= nspec)
        return agpu_set_error("agpu_dedisp_subbands: reads past the filterbank"), -1;
    if (nds == 0)
        return 0;
    const long long nb = (nds + DD_TPB - 1) / DD_TPB;
    if (nb > 0x7fffffffLL || nsub > 65535)
        return agpu_set_error("agpu_dedisp_subbands: grid too large"), -1;
    hipLaunchKernelGGL(k_dd_subbands, dim3((unsigned) nb, (unsigned) nsub), dim3(DD_TPB), 0,
                       (hipStream_t) stream, d_raw, numchan, d_idispdt, numchan / nsub, nsub,
                       downsamp, nds, d_ds);
    DD_CHECK(hipGetLastError());
    return 0;
}

int agpu_dedisp_dms(const float *d_ds, long long nds, int nsub, const int *d_offsets,
                    int maxoff, int numdms, long long nwrite, long long numout, float *d_out,
                    void *stream)
{
    if (nsub <= 0 || numdms <= 0 || nwrite < 0 || nwrite > numout || maxoff < 0)
        return agpu_set_error("agpu_dedisp_dms: bad geometry"), -1;
    if (nwrite && nwrite - 1 + maxoff >= nds)
        return agpu_set_error("agpu_dedisp_dms: reads past the subbands"), -1;
    if (nwrite == 0)
        return 0;
    const long long ntiles = (nwrite + DD_TPB - 1) / DD_TPB;
    const long long nblk = ((ntiles + 7) / 8) * 8 * numdms;
    if (nblk > 0x7fffffffLL)
        return agpu_set_error("agpu_dedisp_dms: grid too large"), -1;
    hipLaunchKernelGGL(k_dd_dms, dim3((unsigned) nblk), dim3(DD_TPB), 0, (hipStream_t) stream,
                       d_ds, nds, nsub, d_offsets, numdms, nwrite, numout, ntiles, d_out);
    DD_CHECK(hipGetLastError());
    return 0;
}

int agpu_dedisp_fill(float *d_out, int numdms, long long numout, long long from, float value,
                     void *stream)
{
    if (from >= numout)
        return 0;
    const long long n = (numout - from) * numdms;
    const long long nb = (n + DD_TPB - 1) / DD_TPB;
    if (nb > 0x7fffffffLL)
        return agpu_set_error("agpu_dedisp_fill: grid too large"), -1;
    hipLaunchKernelGGL(k_dd_fill, dim3((unsigned) nb), dim3(DD_TPB), 0, (hipStream_t) stream,
                       d_out, numdms, numout, from, value);
    DD_CHECK(hipGetLastError());
    return 0;
}

}  // extern "C"
