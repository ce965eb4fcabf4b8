"""CPU checks of the C ABI: the library loads, exports every symbol that
include/presto_accel.h declares, and its host-side logic (geometry,
significance, de-reddening) matches the oracle.  No GPU calls."""
import ctypes
import os
import re

import numpy as np
import pytest
import scipy.stats as st

from oracle import oracle as O
from presto_amd import accel as A, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "presto_accel.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(accel_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(A.LIB_PATH)
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(A.EXPORTED)


@pytest.mark.parametrize("n,dt,zmax,numharm,wmax", [
    (2**20, 64e-6, 50, 8, None), (2**23, 64e-6, 200, 16, None), (2**21, 256e-6, 200, 16, None),
    (2**22, 128e-6, 200, 16, None), (2**22, 64e-6, 200, 8, 300), (2**24, 64e-6, 200, 16, 100),
    (2**20, 64e-6, 0, 8, None), (2**16, 1e-3, 31, 32, None), (2**18, 1e-4, 7, 2, None)])
def test_geometry_matches_oracle(n, dt, zmax, numharm, wmax):
    g = A.geometry(A.make_params(n // 2, n, dt, zmax=zmax, numharm=numharm, wmax=wmax))
    o = O.make_obs(n // 2, n, dt, zmax=zmax, numharm=numharm, wmax=wmax)
    for f in ("numbins", "highestbin", "maxkernlen", "corr_uselen", "fftlen", "numharmstages",
              "numz", "numw", "inmem", "rlo", "rhi", "zlo", "zhi", "T"):
        assert getattr(g, f) == getattr(o, f), f
    assert list(g.powcut)[:g.numharmstages] == list(o.powcut)[:o.numharmstages]
    assert list(g.numindep) == list(o.numindep)


@pytest.mark.parametrize("kw", [
    dict(rlo=3000.0, rhi=9000.0, zmax=25), dict(rlo=1234.5, rhi=5000.5, zmax=7, inmem=True),
    dict(lobin=1500, dat_input=False, zmax=30), dict(lobin=1500, rlo=200.0, zmax=30),
    dict(flo=20.0, fhi=300.0, zmax=13, wmax=35), dict(rhi=1e9, zmax=3),
    dict(lobin=2000, fhi=120.0, zmax=1, numharm=32)])
def test_geometry_options_match_oracle(kw):
    """-rlo/-rhi/-flo/-fhi/-lobin and odd zmax/wmax (rounded up to the
    next multiple of the step), src/accel_utils.c:1594-1730."""
    kw = dict(dict(numharm=8), **kw)
    g = A.geometry(A.make_params(1 << 15, 1 << 16, 5e-4, **kw))
    o = O.make_obs(1 << 15, 1 << 16, 5e-4, **kw)
    for f in ("numbins", "lobin", "highestbin", "maxkernlen", "corr_uselen", "fftlen",
              "numharmstages", "numz", "numw", "inmem", "rlo", "rhi", "zlo", "zhi", "wlo",
              "whi", "T"):
        assert getattr(g, f) == getattr(o, f), f
    assert g.zhi % 2 == 0 and g.zhi >= kw["zmax"]
    assert list(g.powcut)[:g.numharmstages] == list(o.powcut)[:o.numharmstages]
    assert list(g.numindep) == list(o.numindep)


def test_geometry_errors():
    with pytest.raises(A.AccelError, match="power-of-two"):
        A.geometry(A.make_params(1 << 15, 1 << 16, 1e-3, numharm=3))
    with pytest.raises(A.AccelError):
        A.geometry(A.make_params(1 << 15, 1 << 16, 1e-3, flo=1e6))
    with pytest.raises(A.AccelError, match="lobin"):
        A.geometry(A.make_params(1 << 15, 1 << 16, 1e-3, lobin=1 << 15))
    with pytest.raises(A.AccelError, match="below the low"):
        A.geometry(A.make_params(1 << 15, 1 << 16, 1e-3, rlo=5000.0, rhi=4000.0))
    with pytest.raises(A.AccelError, match="zmax"):
        A.geometry(A.make_params(1 << 15, 1 << 16, 1e-3, zmax=-2))


@pytest.mark.parametrize("pw,ns,nt", [(30.0, 1, 2e8), (60.0, 8, 1e7), (200.0, 16, 1e6),
                                      (25.0, 4, 1e3), (45.0, 16, 1e6), (700.0, 1, 1e8),
                                      (3000.0, 32, 1e9), (23.1, 1, 2.4e8)])
def test_candidate_sigma_matches_oracle(pw, ns, nt):
    a = A.lib.accel_candidate_sigma(pw, ns, nt)
    b = O.lib.o_candidate_sigma(pw, ns, nt)
    assert abs(a - b) < 1e-9 * max(1.0, b)


@pytest.mark.parametrize("sig,ns,nt", [(2.0, 1, 2.4e8), (2.0, 16, 1.5e7), (5.0, 8, 1e10),
                                       (2.0, 32, 3e6), (1.0, 2, 10.0), (8.0, 4, 1e12)])
def test_power_for_sigma_matches_scipy(sig, ns, nt):
    p = A.lib.accel_power_for_sigma(sig, ns, nt)
    ref = 0.5 * st.chi2.isf(st.norm.sf(sig) / nt, 2 * ns)
    assert abs(p - ref) < 1e-11 * ref


@pytest.mark.parametrize("n,dt", [(1 << 16, 1e-3), (1 << 20, 64e-6), (4000, 0.01)])
def test_deredden_bit_identical_to_oracle(n, dt):
    x = synth.make_series(n, dt, 5)
    X = np.fft.rfft(x.astype(np.float64))
    n2 = n // 2
    sp = np.empty(n2, np.complex64)
    sp[1:] = X[1:n2]
    sp[0] = complex(X[0].real, X[n2].real)
    a = sp.copy()
    A.lib.accel_prepare_dat_spectrum(a.ctypes.data, n2, float(n), dt)
    b, _ = O.deredden(sp, 6, 200, 0.5 / dt, dt * n)
    b[0] = 1 + 1j
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_spectrum_from_series_close_to_oracle_realfft():
    x = synth.make_series(1 << 14, 1e-3, 9)
    a = A.spectrum_from_series(x, 1e-3)
    b = O.prepare_dat(x, 1e-3)
    assert np.max(np.abs(a[1:] - b[1:])) < 1e-3


@pytest.mark.gpu
def test_search_geometry_outlives_the_handle():
    """Search.geom is an owned copy, still valid after close() and after
    other allocations reuse the handle's memory."""
    import presto_amd.accel as A
    p = A.make_params(1 << 15, 1 << 16, 1e-3, zmax=20, numharm=8)
    s = A.Search(p)
    before = bytes(s.geom)
    s.close()
    junk = [bytearray(4096) for _ in range(256)]
    assert bytes(s.geom) == before and s.geom.numharmstages == 4
    del junk
