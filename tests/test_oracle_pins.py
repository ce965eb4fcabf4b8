"""Pins of the oracle's de-reddening (src/misc_utils.c:291-302, 912-1067)
and of calc_rzwerrs's period-derivative errors (src/characteristics.c:
353-396) by facts that do not come from restating their code:

* invsqrtf is the published fast inverse square root (one Newton step);
  it is computed here on the float32 bit pattern and bounded against
  1/sqrt(x);
* a spectrum of constant power c: every block median is c, so every bin
  is scaled by invsqrtf(c / ln 2) and bin 0 becomes (1, 0);
* a two-level step spectrum: the normaliser is constant outside the span
  between the midpoints of the last block whose median is the low level
  and the first whose median is the high level;
* block widths: startwidth * ln(binnum) bins, at most endwidth, endwidth
  once binnum / T passes endfreq -- computed here from that rule and
  checked through spectra whose power is constant on each block;
* pderr / pdderr: first-order error propagation through a finite-
  difference Jacobian of p, pdot, pddot as functions of (r, z, w).

Phases are multiples of 90 degrees and amplitudes multiples of 1/2, so
every bin's float32 power is exact and block medians are known."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from presto_amd import accel as A

LN2 = math.log(2.0)


def invsqrtf(x):
    """Fast inverse square root on float32 bit patterns (0x5f3759df and one
    Newton step, evaluated left to right in float32)."""
    x = np.asarray(x, dtype=np.float32)
    i = x.view(np.int32)
    y = (np.int32(0x5F3759DF) - (i >> 1)).astype(np.int32).view(np.float32)
    h = ((np.float32(0.5) * x) * y) * y
    return (y * (np.float32(1.5) - h)).astype(np.float32)


def level(c):
    """float32(median / ln 2), the block mean the de-reddening divides by."""
    return np.float32(np.float64(np.float32(c)) / LN2)


def blocks(numbins, sw, ew, endfreq, T):
    """(first bin, length) of every median block over bins [1, numbins):
    startwidth bins first, then floor(startwidth * ln(binnum)) capped at
    endwidth while binnum / T (float32) is below endfreq, endwidth after."""
    out, b = [(1, sw)], 1 + sw
    width = min(int(sw * math.log(b)), ew)
    while b < numbins:
        n = min(width, numbins - b)
        out.append((b, n))
        b += n
        if np.float32(b) / np.float32(T) < endfreq:
            width = min(int(sw * math.log(b)), ew)
        else:
            width = ew
    return out


def axis_spectrum(amps, seed):
    """Complex64 bins of the given amplitudes with random 90-degree phases."""
    rng = np.random.default_rng(seed)
    ph = np.array([1, 1j, -1, -1j])[rng.integers(0, 4, len(amps))]
    return (np.asarray(amps, dtype=np.float64) * ph).astype(np.complex64)


def test_invsqrtf_is_the_fast_inverse_square_root():
    x = np.float32(np.geomspace(1e-6, 1e6, 2001))
    got = np.array([O.lib.o_invsqrtf(float(v)) for v in x[::50]], dtype=np.float32)
    assert np.array_equal(got, invsqrtf(x[::50]))
    rel = np.abs(invsqrtf(x) * np.sqrt(x.astype(np.float64)) - 1.0)
    assert rel.max() < 1.8e-3 and rel.max() > 1e-4   # one Newton step's known error


@pytest.mark.parametrize("c,numbins", [(2.25, 1 << 16), (0.25, 50000), (9.0, 777)])
def test_constant_power_scales_every_bin_alike(c, numbins):
    x = axis_spectrum(np.full(numbins, math.sqrt(c)), 3)
    assert np.all(np.abs(x[1:]) ** 2 == np.float32(c))
    y, n = O.deredden(x, 6, 200, 1e9, 100.0)
    assert n == numbins
    norm = invsqrtf(level(c))
    assert y[0] == np.complex64(1.0)
    want = (x[1:].real * norm).astype(np.float32) + 1j * (x[1:].imag * norm).astype(np.float32)
    assert np.array_equal(y[1:], want.astype(np.complex64))


@pytest.mark.parametrize("step,lo,hi", [(10000, 1.0, 2.0), (333, 1.5, 0.5), (40001, 0.5, 4.0)])
def test_step_spectrum_interpolates_only_between_block_midpoints(step, lo, hi):
    numbins, T = 1 << 16, 100.0
    amps = np.where(np.arange(numbins) < step, lo, hi)
    x = axis_spectrum(amps, 5)
    y, _ = O.deredden(x, 6, 200, 1e9, T)
    scale = np.abs(y[1:].astype(np.complex128)) / np.abs(x[1:].astype(np.complex128))
    bl = blocks(numbins, 6, 200, 1e9, T)
    med = [np.median(amps[s:s + n] ** 2) for s, n in bl]
    last_lo = max(i for i, m in enumerate(med) if m == lo * lo)
    first_hi = min(i for i, m in enumerate(med) if m == hi * hi)
    mid_lo = bl[last_lo][0] + bl[last_lo][1] // 2
    mid_hi = bl[first_hi][0] + bl[first_hi][1] // 2
    nlo, nhi = float(invsqrtf(level(lo * lo))), float(invsqrtf(level(hi * hi)))
    k = np.arange(1, numbins)
    assert np.allclose(scale[k <= mid_lo], nlo, rtol=0, atol=1e-6 * nlo)
    assert np.allclose(scale[k >= mid_hi], nhi, rtol=0, atol=1e-6 * nhi)
    inside = scale[(k > mid_lo) & (k < mid_hi)]
    assert len(inside) > 0
    assert np.all(inside >= min(nlo, nhi) * (1 - 1e-6)) and np.all(inside <= max(nlo, nhi) * (1 + 1e-6))
    assert np.all(np.abs(inside - nlo) > 1e-7)   # the interpolation starts right after mid_lo


@pytest.mark.parametrize("numbins,sw,ew,endfreq,T", [
    (1 << 16, 6, 200, 1e9, 100.0),       # width grows to the cap
    (1 << 15, 6, 100, 2.0, 1000.0),      # the endfreq switch at bin 2000
    (20000, 10, 400, 30.0, 50.0),        # switch before the cap is reached
    (5000, 6, 30, 1e9, 10.0)])
def test_block_widths_follow_the_log_rule(numbins, sw, ew, endfreq, T):
    bl = blocks(numbins, sw, ew, endfreq, T)
    # the library's block list (the device de-reddening's plan) is this list
    st = np.zeros(len(bl) + 8, dtype=np.int64)
    ln = np.zeros(len(bl) + 8, dtype=np.int32)
    nb = A.lib.accel_dered_blocks(numbins, sw, ew, endfreq, T, st.ctypes.data, ln.ctypes.data,
                                  len(st))
    assert nb == len(bl)
    assert [(int(a), int(b)) for a, b in zip(st[:nb], ln[:nb])] == bl
    # the oracle: a spectrum of constant power on each block has block
    # medians equal to those powers only if its blocks are these
    rng = np.random.default_rng(numbins)
    lv = rng.integers(1, 9, len(bl)) * 0.5
    lv[-1] = lv[-2]   # the last slope, continued to the end, stays 0 (no negative mean)
    amps = np.ones(numbins)
    for (s, n), a in zip(bl, lv):
        amps[s:s + n] = a
    x = axis_spectrum(amps, 9)
    y, _ = O.deredden(x, sw, ew, endfreq, T)
    # expected normaliser: block means at the block midpoints, linear in
    # between (slope over half the two lengths), constant before the first
    # midpoint, the last slope continued after the last midpoint
    m = [level(a * a) for a in lv]
    want = np.zeros(numbins, dtype=np.float32)
    s0, n0 = bl[0]
    want[s0:s0 + n0 // 2] = m[0]
    slope = np.float32(1.0)
    for b in range(len(bl) - 1):
        (sa, na), (sb, nb_) = bl[b], bl[b + 1]
        slope = np.float32((m[b + 1] - m[b]) / (0.5 * (na + nb_)))
        ks = np.arange(sa + na // 2, sb + nb_ // 2)
        ii = (ks - (sa + na // 2)).astype(np.float32)
        want[ks] = m[b] + slope * ii
    sl, nl = bl[-1]
    ks = np.arange(sl + nl // 2, numbins)
    want[ks] = m[-1] + slope * (ks - (sl + nl // 2)).astype(np.float32)
    norm = invsqrtf(want[1:])
    exp = (x[1:].real * norm).astype(np.float32) + 1j * (x[1:].imag * norm).astype(np.float32)
    assert np.array_equal(y[1:], exp.astype(np.complex64))


def _period_derivs(r, z, w, T):
    """p, pdot, pddot from f = r/T, fdot = z/T^2, fddot = w/T^3."""
    f, fd, fdd = r / T, z / T ** 2, w / T ** 3
    return np.array([1.0 / f, -fd / f ** 2, 2.0 * fd ** 2 / f ** 3 - fdd / f ** 2])


@pytest.mark.parametrize("r,z,w,T,errs", [
    (1234.5, 12.0, 30.0, 100.0, (0.03, 0.4, 5.0)),
    (88000.25, -150.0, 280.0, 536.87, (0.002, 0.08, 3.0)),
    (205.75, 3.5, -60.0, 67.1, (0.1, 1.2, 11.0))])
def test_pderr_pdderr_are_first_order_error_propagation(r, z, w, T, errs):
    rd = O.ORDerivs(pow=400.0, phs=1.0, dpow=0.0, dphs=-np.pi, d2pow=-2000.0, d2phs=0.0,
                    locpow=2.0)
    p = O.calc_props(rd, r, z, w)
    p.rerr, p.zerr, p.werr = errs
    e = O.calc_rzwerrs(p, T)
    sig = np.array([p.rerr, p.zerr, p.werr], dtype=np.float64)
    x0 = np.array([r, z, w])
    J = np.zeros((3, 3))
    for j in range(3):
        h = 1e-4 * max(1.0, abs(x0[j]))
        xp, xm = x0.copy(), x0.copy()
        xp[j] += h
        xm[j] -= h
        J[:, j] = (_period_derivs(*xp, T) - _period_derivs(*xm, T)) / (2 * h)
    prop = np.sqrt((J ** 2) @ (sig ** 2))
    vals = _period_derivs(r, z, w, T)
    assert np.isclose(e.p, vals[0], rtol=1e-12) and np.isclose(e.pd, vals[1], rtol=1e-12)
    assert np.isclose(e.pdd, vals[2], rtol=1e-9)
    assert np.isclose(e.perr, prop[0], rtol=1e-5)
    assert np.isclose(e.pderr, prop[1], rtol=1e-5)
    assert np.isclose(e.pdderr, prop[2], rtol=1e-5)
    # the library's calc_rzwerrs gives the same numbers
    pl = A.FourierProps()
    for f, _ in A.FourierProps._fields_:
        setattr(pl, f, getattr(p, f))
    el = A.calc_rzwerrs(pl, T)
    for f, _ in A.RzwErrs._fields_:
        assert getattr(el, f) == getattr(e, f), f
