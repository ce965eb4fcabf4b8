"""Subband de-dispersion (prepsubband -nobary -noclip) into DM trials.

CPU: the library's plan (channel delays, DM offsets, blocksperread,
written count) against the oracle's restatement of prepsubband; the
oracle's block-by-block data path against the continuous shifted sums it
must equal, summed in the same FP32 order with numpy (bit for bit); and a
dispersed pulse lines up at its own DM.

GPU: the HIP kernels against the oracle bit for bit, written samples and
padding, over several subband / downsampling shapes; and DM trials of a
filterbank carrying a dispersed pulsar, searched on the GPU, find it
strongest at its DM."""
import numpy as np
import pytest

from oracle import oracle as O
from presto_amd import synth

A = pytest.importorskip("presto_amd.accel")

CASES = [
    # numchan, dt, lofreq, chanwidth, B, nsub, downsamp, numdms, lodm, dmstep
    (64, 256e-6, 1200.0, 4.0, 1024, 16, 1, 12, 0.0, 10.0),
    (64, 256e-6, 1200.0, 4.0, 1024, 8, 2, 9, 20.0, 5.0),
    (128, 128e-6, 1300.0, 2.0, 2048, 32, 4, 7, 50.0, 2.5),
    (32, 512e-6, 400.0, 1.0, 512, 32, 1, 5, 10.0, 1.0),
    # 3 channels per subband: LDS tiles 48 channels wide, staged element by
    # element (48 does not divide the 256 threads)
    (48, 256e-6, 1200.0, 4.0, 1024, 16, 1, 6, 0.0, 10.0),
    # one channel per subband, offsets spread by thousands of samples per
    # group of 16 DMs: the DM sums take the direct (non-LDS) kernel
    (64, 256e-6, 300.0, 1.0, 1024, 64, 1, 10, 0.0, 20.0),
]


def _fb(numchan, dt, lofreq, cw, nspec, seed, pulsars=()):
    return synth.filterbank(nspec, numchan, dt, lofreq, cw, seed, pulsars)


def _continuous(raw, idd, off, nsub, D, nwrite):
    """out[d][u] = -0 + sum_s ds[u + off[d][s]][s], ds the downsampled
    subbands of sub[t][s] = 0 + sum_c raw[t + idd[c]][c], in the
    reference's FP32 order."""
    nspec, numchan = raw.shape
    cps = numchan // nsub
    nds = nwrite + int(off.max())
    T = nds * D
    sub = np.zeros((T, nsub), dtype=np.float32)
    for c in range(numchan):
        sub[:, c // cps] += raw[idd[c]:idd[c] + T, c]
    if D > 1:
        ftmp = np.zeros((nds, nsub), dtype=np.float32)
        for k in range(D):
            ftmp += sub[k::D][:nds]
        ds = np.float32(0.0) + ftmp / np.float32(D)
    else:
        ds = sub
    out = np.full((off.shape[0], nwrite), -0.0, dtype=np.float32)
    for d in range(off.shape[0]):
        for s in range(nsub):
            out[d] += ds[off[d, s]:off[d, s] + nwrite, s]
    return out


@pytest.mark.parametrize("case", CASES)
def test_plan_matches_oracle(case):
    numchan, dt, lofreq, cw, B, nsub, D, nd, lodm, dmstep = case
    p = A.dedisp_params(numchan, dt, lofreq, cw, B, nsub, D, nd, lodm, dmstep, 1 << 15)
    idd, off, bpr, nw = A.dedisp_plan(p, 40 * B)
    _, oidd, ooff, obpr = O.dedisp_plan(numchan, dt, lofreq, cw, B, nsub, D, nd, lodm, dmstep)
    assert np.array_equal(idd, oidd) and np.array_equal(off, ooff) and bpr == obpr
    assert idd.min() >= 0 and off.min() >= 0 and off[:, -1].max() == 0
    # the reference writes floor((K - 1) / bpr) - 1 blocks of B bpr / D
    assert nw == min(((40 - 1) // bpr - 1) * B * bpr // D, 1 << 15)


def test_plan_refusals():
    p = A.dedisp_params(64, 256e-6, 1200.0, 4.0, 1024, 10, 1, 4, 0.0, 10.0, 1024)
    with pytest.raises(A.AccelError, match="subbands"):
        A.dedisp_plan(p, 1 << 16)
    # one subband: the delay across it exceeds a 64-spectrum block
    p = A.dedisp_params(64, 64e-6, 300.0, 1.0, 64, 1, 1, 2, 100.0, 10.0, 1024)
    with pytest.raises(A.AccelError, match="exceeds"):
        A.dedisp_plan(p, 1 << 16)


@pytest.mark.parametrize("case", CASES[:3])
def test_oracle_block_path_equals_continuous_sums(case):
    numchan, dt, lofreq, cw, B, nsub, D, nd, lodm, dmstep = case
    nspec = 24 * B
    raw = _fb(numchan, dt, lofreq, cw, nspec, 5)
    numout = 1 << 20   # longer than the data: the rest is padding
    out, n, pad = O.prepsubband(raw, dt, lofreq, cw, B, nsub, D, nd, lodm, dmstep, numout)
    _, idd, off, bpr = O.dedisp_plan(numchan, dt, lofreq, cw, B, nsub, D, nd, lodm, dmstep)
    assert n == ((24 - 1) // bpr - 1) * B * bpr // D
    ref = _continuous(raw, idd, off, nsub, D, n)
    assert np.array_equal(out[:, :n].view(np.uint32), ref.view(np.uint32))
    # padding: update_stats' running mean of the first trial
    avg = 0.0
    for i, x in enumerate(out[0, :n].astype(np.float64)):
        avg += (x - avg) / (i + 1.0)
    assert pad == avg and np.all(out[:, n:] == np.float32(avg))
    assert abs(avg - out[0, :n].mean()) < 1e-6 * max(1.0, abs(avg))


def test_dispersed_pulse_lines_up_at_its_dm():
    """One pulse, noise-free: at its own DM every channel adds at the same
    output sample (the offsets' rounding keeps the subbands within one
    sample), away from it the pulse smears."""
    numchan, dt, lofreq, cw, B = 64, 256e-6, 1200.0, 4.0, 1024
    nspec = 16 * B
    raw = np.zeros((nspec, numchan), dtype=np.float32)
    dm0, t0 = 60.0, 3000
    ftop = lofreq + (numchan - 1) * cw
    for c in range(numchan):
        f = lofreq + c * cw
        delay = dm0 / (0.000241 * f * f) - dm0 / (0.000241 * ftop * ftop)
        raw[t0 + int(round(delay / dt)), c] = 1.0
    out, n, _ = O.prepsubband(raw, dt, lofreq, cw, B, 64, 1, 13, 0.0, 10.0, 1 << 14)
    peaks = out[:, :n].max(axis=1)
    assert int(np.argmax(peaks)) == 6 and peaks[6] == 64.0
    assert int(np.argmax(out[6, :n])) == t0
    assert peaks[0] < 8 and peaks[12] < 16


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_gpu_matches_oracle_bit_for_bit(case):
    numchan, dt, lofreq, cw, B, nsub, D, nd, lodm, dmstep = case
    print(f"case: dedispersion {case}", flush=True)
    nspec = 40 * B
    raw = _fb(numchan, dt, lofreq, cw, nspec, 9, [dict(f0=13.7, dm=lodm + 2 * dmstep,
                                                        duty=0.05, amp=0.5)])
    numout = 1 << 16
    p = A.dedisp_params(numchan, dt, lofreq, cw, B, nsub, D, nd, lodm, dmstep, numout)
    got, n, pad = A.dedisperse(p, raw)
    ref, on, opad = O.prepsubband(raw, dt, lofreq, cw, B, nsub, D, nd, lodm, dmstep, numout)
    assert n == on and pad == opad
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.gpu
def test_dm_trials_feed_the_search():
    """A dispersed pulsar (17.3 Hz, DM 60) in a 64-channel filterbank: the
    DM trials (0..120 in steps of 10), each searched on the GPU, see it
    strongest at DM 60."""
    print("case: DM trials of a dispersed pulsar through the search", flush=True)
    numchan, dt, lofreq, cw, B = 64, 256e-6, 1200.0, 4.0, 2048
    numout = 1 << 16
    nspec = 36 * B
    raw = _fb(numchan, dt, lofreq, cw, nspec, 11, [dict(f0=17.3, dm=60.0, duty=0.04,
                                                         amp=0.08)])
    p = A.dedisp_params(numchan, dt, lofreq, cw, B, 16, 1, 13, 0.0, 10.0, numout)
    trials, n, _ = A.dedisperse(p, raw)
    assert n == numout
    best = []
    sp = A.make_params(numout // 2, numout, dt, zmax=0, numharm=8, sigma=2.0)
    with A.Search(sp) as s:
        for d in range(13):
            s.set_series(trials[d])
            s.run()
            c = s.candidates()
            near = [x["sigma"] for x in c if abs(x["r"] / (numout * dt) - 17.3) < 0.05]
            best.append(max(near) if near else 0.0)
    print("best sigma per DM:", ["%.1f" % b for b in best], flush=True)
    assert int(np.argmax(best)) == 6 and best[6] > 5.0
    assert best[0] < best[6]


@pytest.mark.gpu
def test_device_resident_trials_equal_host_path():
    """The filterbank uploaded once, its DM trials written to device memory
    and each searched from there (accel_search_set_device_series, one by one
    and as a pipelined batch): the trials and the candidates equal the host
    path's."""
    print("case: device-resident DM trials into the search", flush=True)
    numchan, dt, lofreq, cw, B = 64, 256e-6, 1200.0, 4.0, 2048
    numout, nspec = 1 << 16, 36 * 2048
    raw = _fb(numchan, dt, lofreq, cw, nspec, 12, [dict(f0=23.9, dm=40.0, duty=0.04,
                                                         amp=0.08)])
    p = A.dedisp_params(numchan, dt, lofreq, cw, B, 16, 1, 9, 0.0, 10.0, numout)
    host, n, pad = A.dedisperse(p, raw)
    sp = A.make_params(numout // 2, numout, dt, zmax=20, numharm=4, sigma=2.0)
    with A.DeviceBuffer(raw.nbytes) as d_raw, A.DeviceBuffer(host.nbytes) as d_out:
        d_raw.upload(raw)
        dpad, ms = A.dedisperse_dev(p, d_raw.ptr, nspec, d_out.ptr)
        assert dpad == pad and ms[0] > 0 and ms[1] > 0
        dev = d_out.download(np.zeros_like(host))
        assert np.array_equal(dev.view(np.uint32), host.view(np.uint32))
        with A.Search(sp) as s:
            for d in (0, 4, 8):
                s.set_device_series(d_out.ptr + d * numout * 4, numout)
                s.run()
                c_dev = s.candidates().copy()
                s.set_series(host[d])
                s.run()
                assert np.array_equal(c_dev, s.candidates())
        # every trial through the pipelined batch, straight from device memory
        ref = []
        with A.Search(sp) as s:
            for d in range(9):
                s.set_series(host[d])
                s.run()
                ref.append(s.candidates().copy())
        with A.Batch() as b:
            cands, gms = b.run_device(sp, [d_out.ptr + d * numout * 4 for d in range(9)], numout)
        assert sum(len(c) for c in ref) > 0 and all(m > 0 for m in gms)
        for d in range(9):
            assert np.array_equal(cands[d], ref[d]), d
