/*
 * dedisp_gpu.hip -- subband de-dispersion of a filterbank held in HBM
 * (prepsubband's data path, src/prepsubband.c:918-1040 with
 * src/dispersion.c:164-224).  Two stages (k_dd_subbands_tile does the
 * first straight from the filterbank through LDS when a tile fits, else
 * k_dd_transpose + k_dd_subbands):
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
#define DD_UNROLL 8

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

/* raw [nspec][numchan] -> rawT [numchan][ncols] for the first ncols
 * spectra, through 64 x 64 LDS tiles (both sides coalesced) */
__global__ __launch_bounds__(DD_TPB) void k_dd_transpose(const float *raw, int numchan,
                                                         long long ncols, float *rawT)
{
    __shared__ float tile[64][65];
    const long long t0 = (long long) blockIdx.x * 64;
    const int c0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;   /* 64 x 4 */
    for (int r = ty; r < 64; r += 4) {
        const long long t = t0 + r;
        const int c = c0 + tx;
        if (t < ncols && c < numchan)
            tile[r][tx] = raw[t * numchan + c];
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
        const int c = c0 + r;
        const long long t = t0 + tx;
        if (t < ncols && c < numchan)
            rawT[(long long) c * ncols + t] = tile[tx][r];
    }
}

/* one thread per (v, s) on the channel-major copy; lanes run along v, so
 * every channel's read is one coalesced run */
__global__ __launch_bounds__(DD_TPB) void k_dd_subbands(const float *rawT, long long ncols,
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
            sub = sub + rawT[(long long) c * ncols + t + idispdt[c]];
        }
        if (downsamp == 1) {
            ds[(long long) s * nds + v] = sub;
            return;
        }
        ftmp = ftmp + sub;
    }
    ds[(long long) s * nds + v] = 0.0f + ftmp / (float) downsamp;
}

/* The subband sums straight from the spectrum-major filterbank: a
 * workgroup takes DD_TV output samples of spw consecutive subbands
 * (gc = spw cps channels), loads the rows [v0 D, v0 D + DD_TV D + maxdelay)
 * of those channels into LDS once (each row's gc channels are one
 * contiguous run), and sums from LDS in the same order as k_dd_subbands.
 * The row stride gc + 1 keeps the lanes of one channel on distinct banks. */
#define DD_TV 256
__global__ __launch_bounds__(DD_TPB) void k_dd_subbands_tile(const float *raw, long long nspec,
                                                             int numchan, const int *idispdt,
                                                             int cps, int spw, int nsub,
                                                             int downsamp, int maxdelay,
                                                             long long nds, float *ds)
{
#pragma clang fp contract(off)
    extern __shared__ float tile[];   /* [rows][gc + 1] */
    __shared__ int idd[64];
    const int gc = spw * cps, stride = gc + 1;
    const int s0 = blockIdx.y * spw, c0 = s0 * cps;
    const long long v0 = (long long) blockIdx.x * DD_TV;
    const long long t0 = v0 * downsamp;
    const int rows = DD_TV * downsamp + maxdelay;
    for (int j = threadIdx.x; j < gc; j += DD_TPB)
        idd[j] = idispdt[c0 + j];
    /* DD_UNROLL independent loads in flight per thread before their LDS
     * stores */
    const int n = rows * gc;
    if (DD_TPB % gc == 0) {
        /* whole rows per pass: each thread keeps one channel, no division */
        const int rpp = DD_TPB / gc, j = threadIdx.x % gc;
        const float *src = raw + c0 + j;
        for (int r0 = threadIdx.x / gc; r0 < rows; r0 += rpp * DD_UNROLL) {
            float val[DD_UNROLL];
#pragma unroll
            for (int u = 0; u < DD_UNROLL; u++) {
                const int r = r0 + u * rpp;
                const long long t = t0 + r;
                val[u] = (r < rows && t < nspec) ? src[t * numchan] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < DD_UNROLL; u++) {
                const int r = r0 + u * rpp;
                if (r < rows)
                    tile[r * stride + j] = val[u];
            }
        }
    } else
    for (int i0 = threadIdx.x; i0 < n; i0 += DD_TPB * DD_UNROLL) {
        float val[DD_UNROLL];
#pragma unroll
        for (int u = 0; u < DD_UNROLL; u++) {
            const int i = i0 + u * DD_TPB;
            const int r = i / gc, j = i - r * gc;
            const long long t = t0 + r;
            val[u] = (i < n && t < nspec) ? raw[t * numchan + c0 + j] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < DD_UNROLL; u++) {
            const int i = i0 + u * DD_TPB;
            const int r = i / gc, j = i - r * gc;
            if (i < n)
                tile[r * stride + j] = val[u];
        }
    }
    __syncthreads();
    const long long v = v0 + threadIdx.x;
    if (v >= nds)
        return;
    const int vl = threadIdx.x;
    for (int si = 0; si < spw; si++) {
        const int jb = si * cps;
        float ftmp = 0.0f;
        for (int k = 0; k < downsamp; k++) {
            const int r0 = vl * downsamp + k;
            float sub = 0.0f;
            for (int j = 0; j < cps; j++)
                sub = sub + tile[(r0 + idd[jb + j]) * stride + jb + j];
            if (downsamp == 1) {
                ftmp = sub;
                break;
            }
            ftmp = ftmp + sub;
        }
        ds[(long long) (s0 + si) * nds + v] = downsamp == 1 ? ftmp : 0.0f + ftmp / (float) downsamp;
    }
}

/* subbands per workgroup for k_dd_subbands_tile, 0 when the tile does not
 * fit (the transpose path then runs) */
static int dd_tile_spw(int numchan, int nsub, int downsamp, int maxdelay)
{
    const int cps = numchan / nsub;
    if (cps > 64 || downsamp > 16)
        return 0;
    for (int spw = nsub; spw >= 1; spw--) {
        if (nsub % spw || spw * cps > 64)
            continue;
        const long long rows = (long long) DD_TV * downsamp + maxdelay;
        if (rows * (spw * cps + 1) * (long long) sizeof(float) <= 64 * 1024 &&
            (spw * cps >= 16 || spw == nsub))
            return spw;
    }
    return 0;
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

/* The same sums through LDS: a workgroup takes DD_DMT output samples (two
 * per thread) of a group of AGPU_DD_GROUP consecutive DMs.  For each
 * subband it loads the ds run that the group's offsets reach,
 * [u0 + lo_s, u0 + DD_DMT + hi_s), once (coalesced), and every DM of the
 * group reads its shifted window from LDS: one global read per sample and
 * subband instead of one per DM.  Subbands are staged DD_SB at a time
 * (the group's sums wait in registers between passes), so a workgroup
 * holds DD_SB runs of LDS and six workgroups share a CU; the sums still
 * add the subbands in ascending order. */
#define DD_DMT (2 * DD_TPB)
#define DD_SB 8
#define DD_COLS 4   /* loads per thread and run: width <= DD_COLS * DD_TPB */
__global__ __launch_bounds__(DD_TPB) void k_dd_dms_lds(const float *ds, long long nds, int nsub,
                                                       const int *offsets, int numdms,
                                                       long long nwrite, long long numout,
                                                       long long ntiles, int ngroups, int width,
                                                       float *out)
{
#pragma clang fp contract(off)
    extern __shared__ float lds[];   /* [DD_SB][width] */
    __shared__ int lo[256];
    const long long bid = blockIdx.x;
    const int xcd = (int) (bid & 7);
    const long long local = bid >> 3;
    const int g = (int) (local % ngroups);
    const long long tile = (local / ngroups) * 8 + xcd;
    if (tile >= ntiles)
        return;
    const int d0 = g * AGPU_DD_GROUP, nd = min(AGPU_DD_GROUP, numdms - d0);
    const long long u0 = tile * DD_DMT;
    for (int si = threadIdx.x; si < nsub; si += DD_TPB) {
        int m = offsets[(long long) d0 * nsub + si];
        for (int d = d0 + 1; d < d0 + nd; d++)
            m = min(m, offsets[(long long) d * nsub + si]);
        lo[si] = m;
    }
    const int x0 = threadIdx.x, x1 = threadIdx.x + DD_TPB;
    float a0[AGPU_DD_GROUP], a1[AGPU_DD_GROUP];
#pragma unroll
    for (int d = 0; d < AGPU_DD_GROUP; d++) {
        a0[d] = -0.0f;
        a1[d] = -0.0f;
    }
    __syncthreads();
    for (int s0 = 0; s0 < nsub; s0 += DD_SB) {
        const int nr = min(DD_SB, nsub - s0);
        /* half a pass of loads in flight before their LDS stores */
#pragma unroll
        for (int h = 0; h < DD_SB; h += DD_SB / 2) {
            float val[DD_SB / 2][DD_COLS];
#pragma unroll
            for (int r = 0; r < DD_SB / 2; r++) {
#pragma unroll
                for (int c = 0; c < DD_COLS; c++) {
                    const int col = threadIdx.x + c * DD_TPB;
                    const long long k =
                        h + r < nr && col < width ? u0 + lo[s0 + h + r] + col : nds;
                    val[r][c] = k < nds ? ds[(long long) (s0 + h + r) * nds + k] : 0.0f;
                }
            }
#pragma unroll
            for (int r = 0; r < DD_SB / 2; r++) {
#pragma unroll
                for (int c = 0; c < DD_COLS; c++) {
                    const int col = threadIdx.x + c * DD_TPB;
                    if (h + r < nr && col < width)
                        lds[(h + r) * width + col] = val[r][c];
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int d = 0; d < AGPU_DD_GROUP; d++) {
            if (d < nd) {
                const int *off = offsets + (long long) (d0 + d) * nsub + s0;
                for (int r = 0; r < nr; r++) {
                    const int b = r * width + (off[r] - lo[s0 + r]);
                    a0[d] = a0[d] + lds[b + x0];
                    a1[d] = a1[d] + lds[b + x1];
                }
            }
        }
        __syncthreads();   /* the next pass overwrites the runs */
    }
    const bool w0 = u0 + x0 < nwrite, w1 = u0 + x1 < nwrite;
#pragma unroll
    for (int d = 0; d < AGPU_DD_GROUP; d++) {
        if (d < nd) {
            if (w0)
                out[(long long) (d0 + d) * numout + u0 + x0] = a0[d];
            if (w1)
                out[(long long) (d0 + d) * numout + u0 + x1] = a1[d];
        }
    }
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

long long agpu_dedisp_ncols(long long nds, int downsamp, int maxdelay)
{
    return nds > 0 ? nds * downsamp + maxdelay : 0;
}

int agpu_dedisp_needs_transpose(int numchan, int nsub, int downsamp, int maxdelay)
{
    return dd_tile_spw(numchan, nsub, downsamp, maxdelay) == 0;
}

int agpu_dedisp_subbands(const float *d_raw, long long nspec, int numchan, const int *d_idispdt,
                         int maxdelay, int nsub, int downsamp, long long nds, float *d_rawT,
                         float *d_ds, void *stream)
{
    if (numchan <= 0 || nsub <= 0 || numchan % nsub || downsamp < 1 || nds < 0 || maxdelay < 0)
        return agpu_set_error("agpu_dedisp_subbands: bad geometry"), -1;
    /* the spectra read: [0, nds D + maxdelay) */
    const long long ncols = agpu_dedisp_ncols(nds, downsamp, maxdelay);
    if (ncols > nspec)
        return agpu_set_error("agpu_dedisp_subbands: reads past the filterbank"), -1;
    if (nds == 0)
        return 0;
    const int spw = dd_tile_spw(numchan, nsub, downsamp, maxdelay);
    if (spw) {
        const long long nbt = (nds + DD_TV - 1) / DD_TV;
        if (nbt > 0x7fffffffLL || nsub / spw > 65535)
            return agpu_set_error("agpu_dedisp_subbands: grid too large"), -1;
        const int cps = numchan / nsub;
        const size_t lds = sizeof(float) * (size_t) (DD_TV * downsamp + maxdelay) *
                           (size_t) (spw * cps + 1);
        hipLaunchKernelGGL(k_dd_subbands_tile, dim3((unsigned) nbt, (unsigned) (nsub / spw)),
                           dim3(DD_TPB), lds, (hipStream_t) stream, d_raw, nspec, numchan,
                           d_idispdt, cps, spw, nsub, downsamp, maxdelay, nds, d_ds);
        DD_CHECK(hipGetLastError());
        return 0;
    }
    if (!d_rawT)
        return agpu_set_error("agpu_dedisp_subbands: this geometry needs the transposed copy"), -1;
    const long long nb = (nds + DD_TPB - 1) / DD_TPB, nt = (ncols + 63) / 64;
    if (nb > 0x7fffffffLL || nt > 0x7fffffffLL || nsub > 65535 || numchan > 65535 * 64)
        return agpu_set_error("agpu_dedisp_subbands: grid too large"), -1;
    hipLaunchKernelGGL(k_dd_transpose, dim3((unsigned) nt, (unsigned) ((numchan + 63) / 64)),
                       dim3(DD_TPB), 0, (hipStream_t) stream, d_raw, numchan, ncols, d_rawT);
    DD_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_dd_subbands, dim3((unsigned) nb, (unsigned) nsub), dim3(DD_TPB), 0,
                       (hipStream_t) stream, d_rawT, ncols, d_idispdt, numchan / nsub, nsub,
                       downsamp, nds, d_ds);
    DD_CHECK(hipGetLastError());
    return 0;
}

int agpu_dedisp_dms(const float *d_ds, long long nds, int nsub, const int *d_offsets,
                    int maxoff, int span, int numdms, long long nwrite, long long numout,
                    float *d_out, void *stream)
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
    /* LDS path when every group's runs fit: width = DD_DMT + the largest
     * spread of one subband's offsets over AGPU_DD_GROUP consecutive DMs */
    const int width = DD_DMT + span;
    const long long ldsb = (long long) (nsub < DD_SB ? nsub : DD_SB) * width *
                           (long long) sizeof(float);
    if (span >= 0 && span <= maxoff && nsub <= 256 && width <= DD_COLS * DD_TPB) {
        const int ngroups = (numdms + AGPU_DD_GROUP - 1) / AGPU_DD_GROUP;
        const long long ntl = (nwrite + DD_DMT - 1) / DD_DMT;
        const long long nblk2 = ((ntl + 7) / 8) * 8 * ngroups;
        if (nblk2 > 0x7fffffffLL)
            return agpu_set_error("agpu_dedisp_dms: grid too large"), -1;
        hipLaunchKernelGGL(k_dd_dms_lds, dim3((unsigned) nblk2), dim3(DD_TPB),
                           (size_t) ldsb, (hipStream_t) stream, d_ds,
                           nds, nsub, d_offsets, numdms, nwrite, numout, ntl, ngroups, width,
                           d_out);
    } else {
        hipLaunchKernelGGL(k_dd_dms, dim3((unsigned) nblk), dim3(DD_TPB), 0,
                           (hipStream_t) stream, d_ds, nds, nsub, d_offsets, numdms, nwrite,
                           numout, ntiles, d_out);
    }
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
