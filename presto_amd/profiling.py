"""Profiling helpers.

python -m presto_amd.profiling copy NBYTES
    A plain device-to-device copy of NBYTES (HIP's own copy kernel), five
    times.  Under `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` it gives
    the counter reading for a known number of bytes read and written: the
    scale factor that turns those counters into HBM bytes.

python -m presto_amd.profiling read NBYTES
    The streaming-read probe (k_read_probe, libpresto_probe.so) over a
    zero-filled buffer of NBYTES at grids of 1024 ... 1048576 workgroups:
    prints the best read rate.  Under `rocprofv3 --pmc FETCH_SIZE` it is a
    known read without writes.

python -m presto_amd.profiling summarize DB [DB ...]
    Per-kernel averages from rocprofv3 result databases (*_results.db):
    dispatch count, mean duration (us) and the mean of every collected
    counter per dispatch.
"""
import collections
import ctypes as C
import os
import sqlite3
import sys


def copy(nbytes, reps=5):
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
    print(f"copied {nbytes} bytes x {reps}", flush=True)


def read_probe(nbytes, reps=5):
    """(best read GB/s, its grid, ms per read) of the streaming-read probe."""
    lib = C.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)), "libpresto_probe.so"))
    lib.probe_read_bandwidth.restype = C.c_int
    lib.probe_read_bandwidth.argtypes = [C.c_longlong, C.c_int, C.POINTER(C.c_double),
                                         C.POINTER(C.c_int), C.POINTER(C.c_double)]
    gbs, grid, ms = C.c_double(), C.c_int(), C.c_double()
    if lib.probe_read_bandwidth(int(nbytes), int(reps), C.byref(gbs), C.byref(grid),
                                C.byref(ms)):
        raise RuntimeError("read probe failed")
    return gbs.value, grid.value, ms.value


def _short(name):
    return name.split("(")[0].replace("void ", "")


def summarize(db):
    """{kernel: {"calls": n, "mean_us": t, counter: mean per dispatch}}"""
    con = sqlite3.connect(db)
    cur = con.cursor()
    out = collections.defaultdict(dict)
    durs = collections.defaultdict(list)
    for name, dur in cur.execute("select name, duration from kernels"):
        durs[_short(name)].append(dur)
    for k, d in durs.items():
        out[k]["calls"] = len(d)
        out[k]["mean_us"] = sum(d) / len(d) / 1e3
    tabs = {r[0] for r in cur.execute("select name from sqlite_master")}
    if "counters_collection" in tabs:
        cols = [r[1] for r in cur.execute("pragma table_info(counters_collection)")]
        ki, ci = cols.index("kernel_name"), cols.index("counter_name")
        vi, di = cols.index("value"), cols.index("dispatch_id")
        per = collections.defaultdict(float)
        for r in cur.execute("select * from counters_collection"):
            per[(_short(r[ki]), r[ci], r[di])] += r[vi]
        acc = collections.defaultdict(list)
        for (k, c, _), v in per.items():
            acc[(k, c)].append(v)
        for (k, c), v in acc.items():
            out[k][c] = sum(v) / len(v)
    return out


def main(argv):
    if len(argv) >= 2 and argv[0] == "copy":
        copy(int(argv[1]))
    elif len(argv) >= 2 and argv[0] == "read":
        gbs, grid, ms = read_probe(int(argv[1]))
        print(f"read {int(argv[1])} bytes: {gbs:.1f} GB/s at {grid} workgroups "
              f"({ms:.3f} ms per read)", flush=True)
    elif len(argv) >= 2 and argv[0] == "summarize":
        for db in argv[1:]:
            print(f"# {db}")
            for k, d in sorted(summarize(db).items(), key=lambda kv: -kv[1].get("mean_us", 0)
                               * kv[1].get("calls", 0)):
                fields = "  ".join(f"{c}={v:.6g}" for c, v in d.items())
                print(f"{k:40s} {fields}")
    else:
        print(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
