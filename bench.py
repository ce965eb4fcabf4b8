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

Roofline: per-kernel durations come from HIP events recorded on the
search's own stream around each launch.  Algorithmic bytes per fundamental
cell (the reference geometry's cells): harmonic sum 4 (1 + sum_f (k/H)^2)
bytes (the fundamental plus each sub-harmonic fraction's share of the
plane), correlation 4 bytes per raw-plane cell written plus the spectrum
and template reads.  `traffic` is the measured HBM bytes per launch from
the committed rocprofv3 counter summary profiles/plain_2e23_counters.md
(FETCH_SIZE x 2 + WRITE_SIZE, the scale of a plain copy kernel).
`hbm_read_ceiling` is what a streaming-read kernel (k_read_probe, 16-byte
loads over 4 GiB, the best of several grids) reaches on the same GPU in the
same run; the harmonic sums' rate is also given as a fraction of it.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DM-trials/s and ffdot cells/s at zmax=200,numharm=16; p50/p99 latency"
HBM_PEAK_GBS = 8000.0
VALU_PEAK_TFLOPS = 78.6   # FP32 vector FMA without packed math (256 CU x 64 lanes x 2 x 2.4 GHz)
COUNTERS = os.path.join(ROOT, "profiles", "plain_2e23_counters.md")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary."""
    try:
        txt = open(COUNTERS).read()
    except OSError:
        return None
    m = re.search(r"^traffic\s+%s\s+read_bytes=([0-9.e+]+)\s+write_bytes=([0-9.e+]+)"
                  % re.escape(kernel), txt, re.M)
    return float(m.group(1)) + float(m.group(2)) if m else None


def _stage(msg):
    """Progress line on stderr with the process's peak resident memory."""
    import resource
    rss = resource.getrusage(resource.RUSAGE_SELF).ru_maxrss / 1024
    print(f"bench: {msg} (peak rss {rss:.0f} MiB)", file=sys.stderr, flush=True)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(fft, N, dt, max_blocks, threads):
    """The oracle (C restatement of the reference's search, templates of a
    block in parallel over OpenMP threads as the reference's loop at
    src/accel_utils.c:1141-1193) on the first max_blocks blocks of the plain
    workload's spectrum, and on the whole CPU-reference workload (2^20
    samples, zmax 50, numharm 8, in-memory) with `threads` threads and with
    one; with the host CPU model and the oracle's 2048-point FFT time."""
    from oracle import oracle as O
    from presto_amd import accel as A, synth
    O.lib.o_set_threads(threads)
    o = O.make_obs(len(fft), N, dt, zmax=200, numharm=16, sigma=2.0)
    t0 = time.perf_counter()
    _, _, nb = O.accelsearch(o, fft, max_blocks=max_blocks)
    t = time.perf_counter() - t0
    cells = nb * o.corr_uselen * o.numz
    x1 = synth.config1_series()
    f1 = A.spectrum_from_series(x1, 64e-6)
    o1 = O.make_obs(len(f1), len(x1), 64e-6, zmax=50, numharm=8, sigma=2.0)
    c1_cells = None
    c1 = {}
    for nt in (threads, 1):
        O.lib.o_set_threads(nt)
        t1 = time.perf_counter()
        _, cands, nb1 = O.accelsearch(o1, f1)
        el = time.perf_counter() - t1
        c1_cells = nb1 * o1.corr_uselen * o1.numz
        c1[f"threads_{nt}"] = {"seconds": el, "cells_per_s": c1_cells / el,
                               "dm_trials_per_s": 1.0 / el, "candidates": len(cands)}
    O.lib.o_set_threads(threads)
    return dict(value=cells / t, unit="ffdot cells/s", cores=threads, kind="port",
                sample=f"first {nb} of the plain workload's {o.numz}-template blocks "
                       f"(standard path, all 16 harmonic stages), {t:.1f} s on "
                       f"{threads} OpenMP threads",
                cpu_model=_cpu_model(), host_cpus=os.cpu_count(),
                fft2048_ns=O.lib.o_time_cfft_ns(2048, 101),
                cpu_reference_workload={
                    "workload": "2^20 samples, dt 64 us, one pulsar at 123.456 Hz (z 20), "
                                "zmax 50, numharm 8, sigma 2, in-memory, whole search",
                    "cells": c1_cells, **c1})


def batched_section(ntrials, threads):
    """The batched DM-trial workload (ntrials spectra of 2^23 / 2^22 / 2^21
    samples in a 2:1:1 mix, zmax 200, numharm 16): every trial taken through
    search, refinement and ACCEL / .cand files, first one spectrum at a time
    (upload, search, refine, write; handles reused per geometry), then as a
    pipelined batch (accel_batch_run).  Spectra are prepared before timing
    (GPU real FFT, de-reddening) and sit in host memory."""
    import tempfile
    from concurrent.futures import ThreadPoolExecutor
    from presto_amd import accel as A, synth

    trials = list(synth.config3_trials(ntrials))

    def gen(t):
        i, N, dt, seed, ps = t
        return synth.make_series_f32(N, dt, seed, ps)

    jobs = []
    with ThreadPoolExecutor(max(1, threads)) as ex:
        for t, x in zip(trials, ex.map(gen, trials)):
            i, N, dt, seed, ps = t
            fft = A.spectrum_from_series_gpu(x, dt)
            p = A.make_params(len(fft), N, dt, zmax=200, numharm=16, sigma=2.0)
            jobs.append((p, fft))
    _stage(f"batched workload: {len(jobs)} spectra prepared")
    cells = 0
    for p, fft in jobs:
        g = A.geometry(p)
        cells += g.nblocks * g.corr_uselen * g.numz
    with tempfile.TemporaryDirectory() as td:
        # one spectrum at a time
        handles = {}
        plain = []
        t0 = time.perf_counter()
        for k, (p, fft) in enumerate(jobs):
            key = bytes(p)
            if key not in handles:
                handles[key] = A.Search(p)
            s = handles[key]
            s.set_spectrum(fft)
            s.run()
            c = s.candidates()
            if len(c):
                s.optimize()
                s.write_outputs(os.path.join(td, f"p{k}_ACCEL_200"),
                                os.path.join(td, f"p{k}_ACCEL_200.cand"))
            plain.append(c)
        t_plain = time.perf_counter() - t0
        for s in handles.values():
            s.close()
        _stage(f"batched workload: one-at-a-time pass {t_plain:.2f} s")
        bjobs = [(p, fft, os.path.join(td, f"b{k}_ACCEL_200"),
                  os.path.join(td, f"b{k}_ACCEL_200.cand")) for k, (p, fft) in enumerate(jobs)]
        with A.Batch() as b:
            b.run(bjobs[:4])                     # handles and templates resident
            t0 = time.perf_counter()
            bc, gms = b.run(bjobs)
            t_batch = time.perf_counter() - t0
            nh = b.handles()
        same = all(np.array_equal(x, y) for x, y in zip(plain, bc))
        files_same = all(
            open(os.path.join(td, f"p{k}_ACCEL_200")).read() ==
            open(os.path.join(td, f"b{k}_ACCEL_200")).read()
            for k, c in enumerate(plain) if len(c))
    _stage(f"batched workload: pipelined pass {t_batch:.2f} s")
    n = len(jobs)
    return {"workload": f"{n} DM-trial spectra, 2^23/2^22/2^21 samples 2:1:1 (dt 64/128/256 us), "
                        "10% with a pulsar peaking at one trial, zmax 200, numharm 16, through "
                        "refinement and ACCEL/.cand files",
            "trials": n, "cells": cells,
            "one_at_a_time_dm_trials_per_s": n / t_plain,
            "batched_dm_trials_per_s": n / t_batch,
            "batched_cells_per_s": cells / t_batch,
            "speedup": t_plain / t_batch,
            "search_gpu_ms_sum": float(np.sum(gms)),
            "handles": nh,
            "identical_candidates": bool(same), "identical_accel_files": bool(files_same),
            "candidates": int(sum(len(c) for c in bc))}


def serving_section(nreq, rate, threads, budget_gib, workers):
    """The serving workload: requests (time series + search parameters)
    arrive as a Poisson stream and go to the request loop (accel_server),
    which admits them by device memory and runs them concurrently; each is
    taken through spectrum preparation on the GPU, search and refinement.
    Latency = arrival to result.  Series are generated by a thread pool a
    bounded number of requests ahead of their arrival."""
    import collections
    import threading
    from concurrent.futures import ThreadPoolExecutor
    from presto_amd import accel as A, synth

    reqs = list(synth.config5_requests(nreq, rate=rate))

    def gen(q):
        i, t, N, dt, zmax, wmax, nh, seed, ps = q
        return synth.make_series_f32(N, dt, seed, ps)

    params = [A.make_params(q[2] // 2, q[2], q[3], zmax=q[4], numharm=q[6], sigma=2.0,
                            wmax=(q[5] if q[5] else None)) for q in reqs]
    cells = []
    for p in params:
        g = A.geometry(p)
        cells.append(g.nblocks * g.corr_uselen * g.numz * max(g.numw, 1))
    _stage(f"serving workload: {nreq} requests drawn")
    lat, svc, status = [None] * nreq, [None] * nreq, [None] * nreq
    done_at, timing = [None] * nreq, [None] * nreq
    ahead = 96
    with ThreadPoolExecutor(max(1, threads)) as ex, \
            A.Server(budget_bytes=int(budget_gib) << 30, workers=workers) as sv:
        pending = collections.deque(ex.submit(gen, q) for q in reqs[:ahead])
        nxt = min(ahead, nreq)
        tickets = [None] * nreq
        ready = threading.Semaphore(0)

        def waiter():
            for i in range(nreq):
                ready.acquire()
                status[i], lat[i], svc[i] = sv.wait(tickets[i])
                done_at[i] = time.perf_counter()
                timing[i] = sv.timing(tickets[i])
                if (i + 1) % 500 == 0:
                    _stage(f"serving workload: {i + 1}/{nreq} requests done")

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        late = 0
        for i, q in enumerate(reqs):
            x = pending.popleft().result()
            if nxt < nreq:
                pending.append(ex.submit(gen, reqs[nxt]))
                nxt += 1
            dtw = t0 + q[1] - time.perf_counter()
            if dtw > 0:
                time.sleep(dtw)
            elif dtw < -0.05:
                late += 1
            tickets[i] = sv.submit(params[i], x)
            ready.release()
        th.join()
        wall = time.perf_counter() - t0
    ok = [i for i in range(nreq) if status[i] == 0]
    lat_ok = [lat[i] for i in ok]
    # the sustained rate: completions over the second half of the run
    half = sorted(done_at[i] for i in ok)
    mid = len(half) // 2
    sustained = (len(half) - mid - 1) / (half[-1] - half[mid]) if len(half) > 4 else None
    phases = {k: float(np.mean([timing[i][k] for i in ok]))
              for k in ("queued", "create", "prep", "search", "refine", "service")}
    lat2 = [lat[i] for i in ok[len(ok) // 2:]]
    _stage(f"serving workload: {len(ok)}/{nreq} served in {wall:.2f} s")
    return {"workload": f"{nreq} requests, Poisson arrivals at {rate:g}/s, lengths 2^20..2^24 "
                        "(log-uniform, nearest even length with factors 2,3,5,7), zmax "
                        "{0,50,200}, wmax {0,100}, numharm {8,16}, 20% with a pulsar; GPU "
                        "spectrum preparation, search and refinement per request",
            "requests": nreq, "served": len(ok), "workers": workers,
            "hbm_budget_gib": budget_gib,
            "dm_trials_per_s": len(ok) / wall, "sustained_dm_trials_per_s": sustained,
            "cells_per_s": sum(cells[i] for i in ok) / wall,
            "offered_rate_per_s": rate, "submitted_late": late,
            "p50_ms": float(np.percentile(lat_ok, 50)) if ok else None,
            "p99_ms": float(np.percentile(lat_ok, 99)) if ok else None,
            "p50_ms_second_half": float(np.percentile(lat2, 50)) if lat2 else None,
            "p99_ms_second_half": float(np.percentile(lat2, 99)) if lat2 else None,
            "service_p50_ms": float(np.percentile([svc[i] for i in ok], 50)) if ok else None,
            "mean_ms": phases}


def dedisp_section(numdms=64):
    """DM trials from a filterbank held in HBM: 128 channels (1200-1455
    MHz), 2^21 + spectra at dt 64 us with a dispersed pulsar (DM 45),
    prepsubband -nsub 16 -lodm 0 -dmstep 1.5 -numdms 64 -numout 2^21; both
    kernels timed with HIP events; then every trial searched from device
    memory (zmax 50, numharm 8), one by one and as a pipelined batch, and
    the DM at which the pulsar is strongest."""
    from presto_amd import accel as A, synth
    numchan, dt, lofreq, cw, B = 128, 64e-6, 1200.0, 2.0, 8192
    numout = 1 << 21
    nspec = numout + 4 * B
    nspec -= nspec % B
    raw = synth.filterbank(nspec, numchan, dt, lofreq, cw, 61,
                           [dict(f0=91.7, dm=45.0, duty=0.04, amp=0.05)])
    p = A.dedisp_params(numchan, dt, lofreq, cw, B, 16, 1, numdms, 0.0, 1.5, numout)
    idd, off, bpr, nw = A.dedisp_plan(p, nspec)
    ms = []
    with A.DeviceBuffer(raw.nbytes) as d_raw, \
            A.DeviceBuffer(4 * numdms * numout) as d_out:
        d_raw.upload(raw)
        for k in range(4):
            _, m = A.dedisperse_dev(p, d_raw.ptr, nspec, d_out.ptr)
            if k:
                ms.append(m)
        sub_ms = float(np.mean([m[0] for m in ms]))
        dm_ms = float(np.mean([m[1] for m in ms]))
        nds = nw + int(off.max())
        # algorithmic bytes: the spectra the subbands read once and the
        # subbands written; the subbands read once and the trials written
        sub_bytes = 4.0 * (nds * numchan + nds * 16)
        dm_bytes = 4.0 * (nds * 16 + numdms * nw)
        sp = A.make_params(numout // 2, numout, dt, zmax=50, numharm=8, sigma=2.0)
        best = []
        t0 = time.perf_counter()
        with A.Search(sp) as s:
            for d in range(numdms):
                s.set_device_series(d_out.ptr + d * numout * 4, numout)
                s.run()
                c = s.candidates()
                near = [x["sigma"] for x in c
                        if abs(x["r"] / (numout * dt) - 91.7) < 0.02 * 91.7]
                best.append(max(near) if near else 0.0)
        search_s = time.perf_counter() - t0
        # the same trials through the pipelined batch: one trial's host
        # replay overlaps the next one's search
        with A.Batch() as b:
            ptrs = [d_out.ptr + d * numout * 4 for d in range(numdms)]
            b.run_device(sp, ptrs[:2], numout, maxcands=1 << 16)
            t0 = time.perf_counter()
            bc, _ = b.run_device(sp, ptrs, numout, maxcands=1 << 16)
            batch_s = time.perf_counter() - t0
        best_b = []
        for c in bc:
            near = [float(x["sigma"]) for x in c
                    if abs(x["r"] / (numout * dt) - 91.7) < 0.02 * 91.7]
            best_b.append(max(near) if near else 0.0)
    del raw
    return {"workload": "128-channel filterbank (1200-1455 MHz), %d spectra at dt 64 us with "
                        "a pulsar at DM 45, prepsubband -nsub 16 -lodm 0 -dmstep 1.5 -numdms "
                        "%d -numout 2^21; every trial searched from device memory (zmax 50, "
                        "numharm 8)" % (nspec, numdms),
            "written": nw, "blocksperread": bpr,
            "kernel_ms": {"subbands": sub_ms, "dm_sums": dm_ms},
            "roofline": {"bound": "hbm", "unit": "GB/s", "peak": 8000.0,
                         "subbands_achieved": sub_bytes / sub_ms / 1e6,
                         "dm_sums_achieved": dm_bytes / dm_ms / 1e6,
                         "subbands_frac": sub_bytes / sub_ms / 1e6 / 8000.0,
                         "dm_sums_frac": dm_bytes / dm_ms / 1e6 / 8000.0},
            "dm_trials_per_s_dedispersion": numdms / ((sub_ms + dm_ms) / 1e3),
            "dm_trials_per_s_dedisp_and_search": numdms / (search_s + (sub_ms + dm_ms) / 1e3),
            "dm_trials_per_s_dedisp_and_batched_search":
                numdms / (batch_s + (sub_ms + dm_ms) / 1e3),
            "batched_equals_one_by_one": bool(best_b == [float(x) for x in best]),
            "best_dm": 1.5 * int(np.argmax(best)), "best_sigma": float(max(best))}


def jerk_section(n):
    """The jerk workload: n spectra of 2^22 samples (dt 64 us), zmax 200,
    wmax 300, numharm 8, each with one binary-like signal (average z in
    [-100, 100], w in [-300, 300]); every spectrum is prepared on the GPU,
    searched and refined on one handle (the raw plane is 104.5 GB, so one
    handle at a time).  Reports per-spectrum search and refinement times,
    the DM-trial rate through refinement and how many signals were found."""
    from presto_amd import accel as A, synth
    trials = list(synth.config4_trials(n))
    N, dt = trials[0][1], trials[0][2]
    p = A.make_params(N // 2, N, dt, zmax=200, numharm=8, sigma=2.0, wmax=300)
    g = A.geometry(p)
    cells = g.nblocks * g.corr_uselen * g.numz * g.numw
    t_search, t_opt, tims, hosts, found, ncands, nhits = [], [], [], [], 0, [], []
    best = None
    with A.Search(p) as s:
        for k, (i, N, dt, seed, pulsars, expect) in enumerate(trials):
            x = synth.make_series_f32(N, dt, seed, pulsars)
            s.set_series(x)
            if k == 0:
                tuned = s.tune()          # measured template-group size
                s.run()                   # warm
            t0 = time.perf_counter()
            s.run()
            t_search.append(time.perf_counter() - t0)
            tims.append(s.timings())
            hosts.append(s.host_timings())
            nhits.append(len(s.hits()))
            c = s.candidates()
            ncands.append(len(c))
            t0 = time.perf_counter()
            if len(c):
                s.optimize()
            t_opt.append(time.perf_counter() - t0)
            r_avg, z_avg, w = expect
            if any(abs(c["r"][j] - r_avg) < 2.0 for j in range(len(c))):
                found += 1
            if k == 0 and len(c):
                best = {"r": float(c["r"][0]), "z": float(c["z"][0]), "w": float(c["w"][0]),
                        "sigma": float(c["sigma"][0]), "numharm": int(c["numharm"][0]),
                        "expected_r_z_w": [r_avg, z_avg, w]}
        plane_gb = s.plane_cells() * 4 / 1e9
    _stage(f"jerk workload: {n} spectra, search {1e3 * np.mean(t_search):.1f} ms, refinement "
           f"{1e3 * np.mean(t_opt):.1f} ms per spectrum, {found}/{n} signals found")

    def avg(lst, k):
        return float(np.mean([t[k] for t in lst]))

    return {"workload": f"{n} spectra of 2^22 float32 samples, dt 64 us, zmax 200, wmax 300 "
                        "(201 x 31 templates), numharm 8, one signal each (average z in "
                        "[-100, 100], w in [-300, 300])",
            "spectra": n, "cells_per_spectrum": cells,
            "search_ms": 1e3 * float(np.mean(t_search)),
            "search_ms_p50": 1e3 * float(np.percentile(t_search, 50)),
            "search_ms_max": 1e3 * float(np.max(t_search)),
            "cells_per_s": cells / float(np.mean(t_search)),
            "kernel_ms": {"correlation": avg(tims, "corr"), "harmonic_sum": avg(tims, "harmsum"),
                          "block_norms": avg(tims, "norms"), "corr_to_harmsum_gap": avg(tims, "halo"),
                          "device_total": avg(tims, "total")},
            "host_ms": {k: avg(hosts, k) for k in hosts[0]},
            "hits_mean": float(np.mean(nhits)), "candidates_mean": float(np.mean(ncands)),
            "corr_tuning": {"groups_kept": tuned[0],
                            "timed_groups_ms": tuned[1]},
            "raw_plane_gb": plane_gb, "optimize_ms": 1e3 * float(np.mean(t_opt)),
            "dm_trials_per_s_through_refinement": n / float(np.sum(t_search) + np.sum(t_opt)),
            "signals_found": found, "first": best}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-blocks", type=int, default=2000,
                    help="blocks of the oracle CPU baseline sample (0: skip)")
    ap.add_argument("--batch-trials", type=int, default=256,
                    help="spectra of the batched DM-trial section (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=2000,
                    help="requests of the serving section (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=40.0)
    ap.add_argument("--serve-workers", type=int, default=4)
    ap.add_argument("--serve-budget-gib", type=int, default=190,
                    help="device memory for requests; allocations slow down sharply "
                         "above ~200 GB in use on a 288 GB MI355X")
    ap.add_argument("--dedisp-dms", type=int, default=64,
                    help="DM trials of the filterbank section (0 skips it)")
    ap.add_argument("--jerk", type=int, default=64,
                    help="spectra of the full-size jerk section (0: skip)")
    ap.add_argument("--read-probe-gib", type=int, default=4,
                    help="buffer of the streaming-read probe that measures the read "
                         "ceiling (0: skip)")
    ap.add_argument("--cpu-threads", type=int,
                    default=int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                    help="OpenMP threads of the CPU baseline")
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
    _stage("spectrum ready")
    p = A.make_params(len(fft), N, dt, zmax=200, numharm=16, sigma=2.0)
    s = A.Search(p, device=0)
    _stage("search handle created")
    g = s.geom
    s.set_spectrum(fft)
    tuned = s.tune()     # measured template-group size of the correlation
    for _ in range(args.warmup):
        s.run()
    lat, tim, host = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        s.run()          # returns after the stream has drained and hits are replayed
        lat.append(time.perf_counter() - ts)
        tim.append(s.timings())
        host.append(s.host_timings())
    wall = time.perf_counter() - t0
    _stage(f"{args.steps} timed steps done")
    ncand = len(s.candidates())
    nhits = len(s.hits())
    shape = s.corr_shape()
    plane_cells = s.plane_cells()
    # the rest of a DM trial: FP64 refinement of every candidate on the GPU
    # and the ACCEL / .cand files (src/accelsearch.c:243-284)
    import tempfile
    t_opt, t_out = [], []
    with tempfile.TemporaryDirectory() as td:
        for _ in range(3):
            ts = time.perf_counter()
            s.optimize()
            t_opt.append(time.perf_counter() - ts)
            ts = time.perf_counter()
            s.write_outputs(os.path.join(td, "bench_ACCEL_200"),
                            os.path.join(td, "bench_ACCEL_200.cand"))
            t_out.append(time.perf_counter() - ts)
    s.close()
    _stage("refinement and outputs timed")
    ceiling = None
    if args.read_probe_gib > 0:
        from presto_amd import profiling
        gbs, grid, pms = profiling.read_probe(args.read_probe_gib << 30)
        ceiling = {"kernel": "k_read_probe", "gbs": gbs, "grid": grid,
                   "bytes": args.read_probe_gib << 30, "ms_per_read": pms}
        _stage(f"read probe: {gbs:.0f} GB/s")
    opt_ms = 1e3 * float(np.median(t_opt))
    out_ms = 1e3 * float(np.median(t_out))

    def avg(k):
        return float(np.mean([t[k] for t in tim]))

    numw = max(g.numw, 1)
    cells = g.nblocks * g.corr_uselen * g.numz * numw
    ms_step = 1e3 * wall / args.steps
    nfrac = [(k, 1 << st) for st in range(1, g.numharmstages) for k in range(1, 1 << st, 2)]
    hs_bytes_per_cell = 4.0 * (1.0 + sum((k / h) ** 2 for k, h in nfrac))
    hs_bytes = hs_bytes_per_cell * cells
    hs_ms = avg("harmsum")
    corr_ms = avg("corr")
    nt = g.numz * numw
    corr_bytes = 4.0 * plane_cells + 8.0 * g.numbins + 8.0 * nt * shape["L"]
    L, U = shape["L"], shape["uselen"]
    fft_flops = 5.0 * L * np.log2(L)
    # per (segment, template): inverse FFT, spectrum x template, |.|^2 of the
    # kept outputs; per (segment, template group): the forward FFT
    corr_flops = shape["nseg"] * (nt * (fft_flops + 6.0 * L + 3.0 * U)
                                  + shape["groups"] * fft_flops)

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
                   "segment_len": L, "segment_uselen": U},
        "dm_trials_per_s": 1e3 / ms_step,
        "optimize_ms": opt_ms,
        "accel_output_ms": out_ms,
        "dm_trials_per_s_through_accel": 1e3 / (ms_step + opt_ms + out_ms),
        "p50_ms": 1e3 * float(np.percentile(lat, 50)),
        "p99_ms": 1e3 * float(np.percentile(lat, 99)),
        "kernel_ms": {"block_norms": avg("norms"), "correlation": corr_ms,
                      "corr_to_harmsum_gap": avg("halo"), "harmonic_sum": hs_ms,
                      "device_total": avg("total")},
        "host_ms": {k: float(np.mean([h[k] for h in host])) for k in host[0]},
        "candidates": ncand,
        "hits": nhits,
        "roofline": {"bound": "hbm", "kernel": "k_harmsum<5,0>",
                     "achieved": hs_bytes / (hs_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s",
                     "frac": hs_bytes / (hs_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": measured_traffic("k_harmsum<5, 0>"),
                     "bytes_per_cell": hs_bytes_per_cell,
                     "frac_of_read_ceiling": (hs_bytes / (hs_ms * 1e-3) / 1e9 / ceiling["gbs"]
                                              if ceiling else None)},
        "hbm_read_ceiling": ceiling,
        "corr_tuning": {"groups_kept": tuned[0],
                        "timed_groups_ms": tuned[1]},
        "roofline_correlation": {"bound": "valu", "kernel": "k_corr4096",
                                 "achieved_gbs": corr_bytes / (corr_ms * 1e-3) / 1e9,
                                 "hbm_frac": corr_bytes / (corr_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                 "achieved_tflops": corr_flops / (corr_ms * 1e-3) / 1e12,
                                 "valu_peak_tflops": VALU_PEAK_TFLOPS,
                                 "traffic": measured_traffic("k_corr4096")},
    }
    if args.batch_trials > 0:
        line["batched"] = batched_section(args.batch_trials, args.cpu_threads)
    if args.dedisp_dms > 0:
        _stage("filterbank de-dispersion workload")
        line["dedispersion"] = dedisp_section(args.dedisp_dms)
    if args.jerk > 0:
        line["jerk"] = jerk_section(args.jerk)
    if args.serve_requests > 0:
        line["serving"] = serving_section(args.serve_requests, args.serve_rate,
                                          args.cpu_threads, args.serve_budget_gib,
                                          args.serve_workers)
    if args.cpu_blocks > 0:
        _stage(f"CPU baseline on {args.cpu_threads} threads")
        line["cpu_baseline"] = cpu_baseline(fft, N, dt, args.cpu_blocks, args.cpu_threads)
        _stage("CPU baseline done")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
