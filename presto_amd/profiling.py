"""Profiling helpers.

`python -m presto_amd.profiling copy NBYTES` runs a plain device-to-device
copy of NBYTES (hipMemcpyAsync, HIP's own copy kernel) a few times.  Under
`rocprofv3 --pmc FETCH_SIZE` / `WRITE_SIZE` it gives the counter reading
for a known number of bytes read and written, the scale factor used to
turn those counters into HBM bytes for the search kernels."""
import ctypes as C
import sys

from presto_amd import accel as A


def copy(nbytes, reps=5):
    lib = A.lib
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) or \
            hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, C.c_size_t(nbytes))
    for _ in range(reps):
        hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3)   # device to device
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    del lib
    print(f"copied {nbytes} bytes x {reps}", flush=True)


if __name__ == "__main__":
    if len(sys.argv) >= 3 and sys.argv[1] == "copy":
        copy(int(sys.argv[2]))
    else:
        print(__doc__)
