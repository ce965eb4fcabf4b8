"""Batches of DM-trial spectra (accel_batch_run): every spectrum's
candidates, refinement and files are exactly those of a separate run,
whatever the mix of lengths and the pipelining order."""
import os

import numpy as np
import pytest

from presto_amd import synth

pytestmark = pytest.mark.gpu
A = pytest.importorskip("presto_amd.accel")


def _trials():
    out = []
    for i, (N, dt) in enumerate([(1 << 16, 64e-6), (1 << 15, 128e-6), (1 << 16, 64e-6),
                                 (1 << 17, 32e-6), (1 << 15, 128e-6), (1 << 16, 64e-6)]):
        ps = [dict(f0=123.4 + 7 * i, z=4.0 * i - 10, duty=0.05, amp=0.5)] if i % 2 == 0 else []
        x = synth.make_series(N, dt, 600 + i, ps)
        fft = A.spectrum_from_series(x, dt)
        p = A.make_params(len(fft), N, dt, zmax=50, numharm=8, sigma=2.0)
        out.append((p, fft))
    return out


def test_batch_equals_separate_runs(tmp_path):
    print("case: batch of 6 mixed-length spectra vs separate runs", flush=True)
    trials = _trials()
    ref_cands, ref_files = [], []
    for k, (p, fft) in enumerate(trials):
        with A.Search(p) as s:
            s.set_spectrum(fft)
            s.run()
            c = s.candidates()
            ref_cands.append(c)
            acc = str(tmp_path / f"sep{k}_ACCEL_50")
            if len(c):
                s.optimize()
                s.write_outputs(acc, acc + ".cand")
                ref_files.append(open(acc).read())
            else:
                ref_files.append(None)
    jobs = [(p, fft, str(tmp_path / f"bat{k}_ACCEL_50"), str(tmp_path / f"bat{k}_ACCEL_50.cand"))
            for k, (p, fft) in enumerate(trials)]
    with A.Batch() as b:
        cands, ms = b.run(jobs)
        assert b.handles() >= 3       # three geometries, each double-buffered as needed
        cands2, _ = b.run(jobs[::-1]) # reused handles, other order
    assert sum(len(c) for c in ref_cands) >= 3
    for k in range(len(trials)):
        assert np.array_equal(cands[k], ref_cands[k]), k
        assert np.array_equal(cands2[len(trials) - 1 - k], ref_cands[k]), k
        if ref_files[k] is not None:
            assert open(tmp_path / f"bat{k}_ACCEL_50").read() == ref_files[k]
            assert open(tmp_path / f"bat{k}_ACCEL_50.cand", "rb").read() == \
                open(tmp_path / f"sep{k}_ACCEL_50.cand", "rb").read()
        else:
            assert not os.path.exists(tmp_path / f"bat{k}_ACCEL_50")
    assert all(t > 0 for t in ms)
