"""Benchmark of the accelsearch hot path on one MI355X.

Workload (the plain single-GPU configuration): one spectrum from 2^23 float32
samples (dt 64 us, seeded noise plus 3 injected pulsars), zmax 200,
numharm 16, sigma 2, FP32 planes, FP64 templates.  One step = one full
search of that spectrum with the spectrum already resident in HBM: block
normalisation, fused correlation, harmonic sums and thresholds on the GPU,
then the hit replay (candidate merge, sort, harmonic removal) on the host.

Prints one JSON line.  value = ffdot cells/s, where cells follow the
reference geometry (blocks x corr_uselen x numz x max(numw,1)), so CPU and
GPU figures are comparable; DM-trials/s and p50/p99 step latency are
reported beside it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DM-trials/s and ffdot cells/s at zmax=200,numharm=16; p50/p99 latency"
HBM_PEAK_GBS = 8000.0


def cpu_baseline(fft, N, dt, max_blocks):
    """The oracle (C restatement of the reference's search, 1 thread) on the
    first max_blocks blocks of the same spectrum."""
    from oracle import oracle as O
    o = O.make_obs(len(fft), N, dt, zmax=200, numharm=16, sigma=2.0)
    t0 = time.perf_counter()
    _, _, nb = O.accelsearch(o, fft, max_blocks=max_blocks)
    t = time.perf_counter() - t0
    cells = nb * o.corr_uselen * o.numz
    return dict(value=cells / t, unit="ffdot cells/s", cores=1, kind="port",
                sample=f"first {nb} of the workload's blocks (standard path, all 16 "
                       f"harmonic stages), {t:.1f} s on 1 thread")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-blocks", type=int, default=80,
                    help="blocks of the oracle CPU baseline sample (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); the "
              "single-GPU figure is the whole-job figure.", file=sys.stderr)
        sys.exit(2)

    from presto_amd import accel as A, synth

    dt = 64e-6
    x, pulsars = synth.config2_series()
    N = len(x)
    fft = A.spectrum_from_series(x, dt)
    p = A.make_params(len(fft), N, dt, zmax=200, numharm=16, sigma=2.0)
    s = A.Search(p, device=0)
    g = s.geom
    s.set_spectrum(fft)
    for _ in range(args.warmup):
        s.run()
    lat, corr_ms, hs_ms, norm_ms = [], [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        s.run()
        lat.append(time.perf_counter() - ts)
        t = s.timings()
        corr_ms.append(t["corr"])
        hs_ms.append(t["harmsum"])
        norm_ms.append(t["norms"])
    wall = time.perf_counter() - t0
    ncand = len(s.candidates())

    cells = g.nblocks * g.corr_uselen * g.numz * max(g.numw, 1)
    ms_step = 1e3 * wall / args.steps
    shape = s.corr_shape()
    plane_cells = s.plane_cells()
    nt = g.numz * max(g.numw, 1)
    corr_bytes = 4.0 * plane_cells + 8.0 * g.numbins + 8.0 * nt * shape["L"]
    corr_avg = float(np.mean(corr_ms))
    achieved = corr_bytes / (corr_avg * 1e-3) / 1e9
    s.close()

    line = {
        "metric": METRIC,
        "value": cells / (wall / args.steps),
        "unit": "ffdot cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain accelsearch of one spectrum from 2^23 float32 samples, "
                               "dt 64 us, 3 injected pulsars, zmax 200, numharm 16, sigma 2",
                   "numbins": int(g.numbins), "blocks": int(g.nblocks),
                   "corr_uselen": int(g.corr_uselen), "numz": int(g.numz),
                   "path": "in-memory" if g.inmem else "standard",
                   "segment_len": shape["L"], "segment_uselen": shape["uselen"]},
        "dm_trials_per_s": 1e3 / ms_step,
        "p50_ms": 1e3 * float(np.percentile(lat, 50)),
        "p99_ms": 1e3 * float(np.percentile(lat, 99)),
        "kernel_ms": {"norms": float(np.mean(norm_ms)), "correlation": corr_avg,
                      "harmonic_sum": float(np.mean(hs_ms))},
        "candidates": ncand,
        "roofline": {"bound": "hbm", "kernel": "k_corr4096", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None},
    }
    if args.cpu_blocks > 0:
        line["cpu_baseline"] = cpu_baseline(fft, N, dt, args.cpu_blocks)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
