"""The batched, jerk and serving workloads at their full shapes on the GPU,
against the oracle (bars as in test_gpu_parity.py: plane powers within
1e-4 of FP64, hits equal except within 1e-5 of a threshold, identical
candidate lists).

  * jerk: one spectrum of 2^22 samples, zmax 200, wmax 300 (201 x 31
    templates), numharm 8: the whole candidate list against the oracle's
    (tests/golden/config4_candidates.json), plane cells of three blocks and
    four harmonic fractions against FP64, the hits of the first blocks and
    of the blocks around the injected signal, which must be found;
  * batched: twelve trials of the 256-trial DM workload (2^23, 2^22 and the
    in-memory 2^21), through accel_batch_run: equal to separate runs, hits
    of each length against the oracle, the pulsar strongest at its peak
    trial;
  * serving: the first 16 requests of the Poisson request workload (lengths
    with factors 2, 3, 5, 7; zmax, wmax and numharm mixed): equal alone and
    in the mix, and the small ones' candidates equal to the oracle's.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from parity_util import compare_cands, compare_hits, plane_error
from presto_amd import synth

pytestmark = pytest.mark.gpu
A = pytest.importorskip("presto_amd.accel")

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _say(msg):
    print(msg, flush=True)


def _near(cands, r, tol=2.0):
    return [c for c in cands if abs(c["r"] - r) < tol]


# ------------------------------------------------------------------ jerk, full shape
def test_config4_full_shape_jerk():
    i, N, dt, seed, ps, (r_avg, z_avg, w_avg) = next(synth.config4_trials(1))
    _say(f"case: jerk full shape, N=2^22 zmax 200 wmax 300 numharm 8, signal r={r_avg:.1f} "
         f"z={z_avg:.1f} w={w_avg:.1f}")
    x = synth.make_series(N, dt, seed, ps)
    fft = A.spectrum_from_series(x, dt)
    kw = dict(zmax=200, wmax=300, numharm=8, sigma=2.0)
    p = A.make_params(len(fft), N, dt, **kw)
    o = O.make_obs(len(fft), N, dt, **kw)
    assert (o.numz, o.numw, o.inmem) == (201, 31, 0)
    fix = json.load(open(os.path.join(GOLDEN, "config4_candidates.json")))
    same_input = hashlib.sha256(fft.tobytes()).hexdigest() == fix["spectrum_sha256"]
    with A.Search(p) as s:
        s.set_spectrum(fft)
        s.run()
        gc, gh = s.candidates(), s.hits()
        want = np.array([[c["power"], c["sigma"], c["numharm"], c["r"], c["z"], c["w"]]
                         for c in fix["candidates"]])
        _say(f"full candidate list: {len(gc)} here, {len(want)} from the oracle over "
             f"{fix['blocks']} blocks (input spectrum {'identical' if same_input else 'DIFFERENT'})")
        compare_cands(gc, want)
        near = _near(gc, r_avg)
        assert near, "injected signal not found"
        best = max(near, key=lambda c: c["sigma"])
        _say(f"signal: r={best['r']:.2f} z={best['z']:.1f} w={best['w']:.1f} "
             f"sigma={best['sigma']:.2f}")
        assert abs(best["z"] - z_avg) < 6 and abs(best["w"] - w_avg) < 30
        b0 = int((r_avg - o.rlo) // (o.corr_uselen // 2))
        nb = s.geom.nblocks
        err = plane_error(s, o, fft, [0, b0, nb - 1], [(1, 1), (2, 1), (4, 3), (8, 5)])
        _say(f"jerk full-shape plane error vs FP64 oracle: {err:.3e}")
        assert err < 1e-4
        for first, count in ((0, 4), (b0 - 1, 4)):
            oh, _, _ = O.accelsearch(o, fft, first_block=first, max_blocks=count)
            sub = gh[(gh["block"] >= first) & (gh["block"] < first + count)]
            worst = compare_hits(sub, oh, list(o.powcut))
            _say(f"blocks {first}..{first + count - 1}: {len(sub)} hits, {len(oh)} in the "
                 f"oracle, worst power rel. diff {worst:.2e}")
            assert worst < 1e-4
        assert len(gh[(gh["block"] >= b0 - 1) & (gh["block"] < b0 + 3)]) > 0


# ------------------------------------------------------------------ batched DM trials
def test_config3_subset_through_the_batch(tmp_path):
    _say("case: 12 full-size trials of the batched DM workload (2^23, 2^22, 2^21)")
    trials = {t[0]: t for t in synth.config3_trials(256)}
    peak = 256 // 6
    pick = [peak - 4, peak - 2, peak, peak + 2, peak + 4, 0, 128, 129, 130, 192, 193, 194]
    jobs, obs, specs = [], [], []
    for i in pick:
        _, N, dt, seed, ps = trials[i]
        x = synth.make_series_f32(N, dt, seed, ps)
        fft = A.spectrum_from_series(x, dt)
        p = A.make_params(len(fft), N, dt, zmax=200, numharm=16, sigma=2.0)
        jobs.append((p, fft, str(tmp_path / f"b{i}_ACCEL_200"), str(tmp_path / f"b{i}_ACCEL_200.cand")))
        obs.append(O.make_obs(len(fft), N, dt, zmax=200, numharm=16, sigma=2.0))
        specs.append(fft)
    assert sorted({o.inmem for o in obs}) == [0, 1]
    sep = []
    handles = {}
    for (p, fft, _, _) in jobs:
        key = bytes(p)
        if key not in handles:
            handles[key] = A.Search(p)
        s = handles[key]
        s.set_spectrum(fft)
        s.run()
        sep.append((s.candidates(), s.hits()))
    for s in handles.values():
        s.close()
    with A.Batch() as b:
        bc, ms = b.run(jobs)
    for k, i in enumerate(pick):
        assert np.array_equal(bc[k], sep[k][0]), i
    # each length against the oracle: the first blocks, and the blocks
    # around the pulsar at its peak trial
    f0, z = 150.7, 30.0
    checked = set()
    for k, i in enumerate(pick):
        o, fft, (gc, gh) = obs[k], specs[k], sep[k]
        ranges = []
        if int(o.numbins) not in checked:
            ranges.append((0, 3))
            checked.add(int(o.numbins))
        if i == peak:
            r = f0 * o.N * o.dt + z / 2
            b0 = int((r - o.rlo) // (o.corr_uselen // 2))
            ranges.append((b0 - 1, 3))
        if int(o.numbins) == (1 << 20) and i == 192:
            ranges.append((900, 3))     # in-memory path, far from the start
        for first, count in ranges:
            oh, _, _ = O.accelsearch(o, fft, first_block=first, max_blocks=count)
            sub = gh[(gh["block"] >= first) & (gh["block"] < first + count)]
            worst = compare_hits(sub, oh, list(o.powcut))
            _say(f"trial {i} (2^{int(np.log2(o.N))}, {'in-memory' if o.inmem else 'standard'}) "
                 f"blocks {first}..{first + count - 1}: {len(sub)} hits, worst {worst:.2e}")
    # the pulsar is strongest at its peak trial
    strength = {}
    for k, i in enumerate(pick[:5]):
        o = obs[k]
        r = f0 * o.N * o.dt + z / 2
        _say(f"trial {i}: top candidates " + ", ".join(
            f"r={c['r']:.2f} z={c['z']:.1f} h={c['numharm']} sig={c['sigma']:.1f}"
            for c in bc[k][:4]) + f" (pulsar r={r:.2f})")
        # found at its frequency or at a harmonic / sub-harmonic of it (the
        # harmonic sums can peak at r/3 and remove r as related to it)
        near = [c["sigma"] for c in bc[k] for n in range(1, 17)
                if abs(c["r"] * n - r) < 2 or abs(c["r"] / n - r) < 2]
        strength[i] = max(near) if near else 0.0
    _say("pulsar sigma by trial: " + ", ".join(f"{i}: {v:.1f}" for i, v in strength.items()))
    assert max(strength, key=strength.get) == peak


# ------------------------------------------------------------------ serving
def _serve(reqs, workers, budget, together):
    out = []
    with A.Server(budget_bytes=budget, workers=workers) as sv:
        if together:
            tickets = [sv.submit(p, x) for p, x in reqs]
        else:
            tickets = []
        for k, (p, x) in enumerate(reqs):
            t = tickets[k] if together else sv.submit(p, x)
            st, _, _ = sv.wait(t)
            assert st == 0, k
            out.append((sv.candidates(t), [bytes(v) for v in sv.props(t)]))
    return out


def test_config5_requests_alone_and_mixed():
    _say("case: the first 16 requests of the serving workload, alone vs 4 workers")
    reqs, meta = [], []
    for q in synth.config5_requests(16):
        i, t, N, dt, zmax, wmax, nh, seed, ps = q
        assert N % 2 == 0
        x = synth.make_series_f32(N, dt, seed, ps)
        p = A.make_params(N // 2, N, dt, zmax=zmax, numharm=nh, sigma=2.0,
                          wmax=(wmax if wmax else None))
        reqs.append((p, x))
        meta.append(q)
    assert any(N & (N - 1) for _, _, N, *_ in meta)      # lengths that are not 2^k
    alone = _serve(reqs, 1, 200 << 30, together=False)
    mixed = _serve(reqs, 4, 200 << 30, together=True)
    for k, ((ca, pa), (cm, pm)) in enumerate(zip(alone, mixed)):
        assert np.array_equal(ca, cm), k
        assert pa == pm, k
    # the small requests against the oracle, on the spectrum the GPU prepares
    small = [k for k, q in enumerate(meta) if q[2] <= 1_600_000]
    assert len(small) >= 3
    for k in small:
        i, t, N, dt, zmax, wmax, nh, seed, ps = meta[k]
        fft = A.spectrum_from_series_gpu(reqs[k][1], dt)
        o = O.make_obs(len(fft), N, dt, zmax=zmax, numharm=nh, sigma=2.0,
                       wmax=(wmax if wmax else None))
        _, oc, nb = O.accelsearch(o, fft)
        _say(f"request {i}: N={N} zmax={zmax} wmax={wmax} numharm={nh}: {len(alone[k][0])} "
             f"candidates, {len(oc)} from the oracle over {nb} blocks")
        compare_cands(alone[k][0], oc)
