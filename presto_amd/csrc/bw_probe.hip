/*
 * A streaming-read kernel that measures the HBM read bandwidth this GPU
 * reaches, the ceiling a read-bound kernel (the harmonic sums) is judged
 * against beside the 8 TB/s peak, and a known read for calibrating the
 * FETCH_SIZE counter.  Built into its own library (libpresto_probe.so);
 * nothing in the search calls it.
 *
 * k_read_probe: every lane reads 16-byte vectors, grid-stride, four loads
 * in flight per lane, and folds them into a sum that is stored only if it
 * equals a value the zero-filled buffer cannot give, so the loads stay and
 * the kernel writes nothing.
 */
#include <hip/hip_runtime.h>

#define RP_TPB 256

__global__ __launch_bounds__(RP_TPB) void k_read_probe(const float4 *__restrict__ x,
                                                        long long n4, float *out)
{
    const long long stride = (long long) gridDim.x * RP_TPB;
    long long i = (long long) blockIdx.x * RP_TPB + threadIdx.x;
    float s = 0.0f;
    for (; i + 3 * stride < n4; i += 4 * stride) {
        const float4 a = x[i], b = x[i + stride], c = x[i + 2 * stride], d = x[i + 3 * stride];
        s += (a.x + a.y + a.z + a.w) + (b.x + b.y + b.z + b.w) + (c.x + c.y + c.z + c.w) +
             (d.x + d.y + d.z + d.w);
    }
    for (; i < n4; i += stride) {
        const float4 a = x[i];
        s += a.x + a.y + a.z + a.w;
    }
    if (s == 1234.5f)
        out[0] = s;
}

/* Reads a zero-filled buffer of `bytes` (rounded down to 16) `reps` times
 * at each grid of {1024 ... 1048576} workgroups of 256 lanes and returns the
 * best rate in GB/s (1e9 bytes/s) with its grid and mean time per read.
 * Allocates and frees its own buffers.  0, or -1 on a HIP error. */
extern "C" int probe_read_bandwidth(long long bytes, int reps, double *best_gbps,
                                    int *best_grid, double *best_ms)
{
    if (bytes < (1 << 20) || reps < 1)
        return -1;
    const long long n4 = bytes / 16;
    float4 *x = nullptr;
    float *out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = -1;
    *best_gbps = 0.0;
    *best_grid = 0;
    *best_ms = 0.0;
    if (hipMalloc((void **) &x, (size_t) n4 * 16) != hipSuccess)
        return -1;
    if (hipMalloc((void **) &out, sizeof(float)) != hipSuccess ||
        hipMemset(x, 0, (size_t) n4 * 16) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        goto done;
    for (int grid = 1024; grid <= 1048576; grid *= 2) {
        hipLaunchKernelGGL(k_read_probe, dim3(grid), dim3(RP_TPB), 0, 0, x, n4, out);
        if (hipEventRecord(e0, 0) != hipSuccess)
            goto done;
        for (int r = 0; r < reps; r++)
            hipLaunchKernelGGL(k_read_probe, dim3(grid), dim3(RP_TPB), 0, 0, x, n4, out);
        float ms = 0.0f;
        if (hipGetLastError() != hipSuccess || hipEventRecord(e1, 0) != hipSuccess ||
            hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            goto done;
        const double per = (double) ms / reps;
        const double gbps = (double) n4 * 16 / (per * 1e-3) / 1e9;
        if (gbps > *best_gbps) {
            *best_gbps = gbps;
            *best_grid = grid;
            *best_ms = per;
        }
    }
    rc = 0;
done:
    if (e0)
        hipEventDestroy(e0);
    if (e1)
        hipEventDestroy(e1);
    hipFree(out);
    hipFree(x);
    return rc;
}
