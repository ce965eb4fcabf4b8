"""numpy restatement of accelsearch's birdie zapping (test infrastructure
only; the checker for the library's accel_zap_birdies).

get_birdies  src/birdzap.c:6-90
zapbirds     src/zapping.c:58-160, as src/accelsearch.c:87-110 applies it

The noise comes from MT19937 seeded with init_genrand(1) -- numpy's legacy
RandomState seeding of the same published generator -- turned into
exponential powers by GSL 2.7's gsl_ran_exponential, -mu log1p(-u) with
u = draw / 2^32, and uniform phases 2 pi draw / 2^32 (GSL's source is not
on this machine: that transform is restated from its documentation and is
parity unpinned)."""
import numpy as np

MEDIANBINS, GUARDBINS = 200, 100


def read_birdies(path, T, baryv=0.0):
    birds = []
    for line in open(path):
        if line.startswith("#"):
            continue
        isb = line.startswith("B")
        f = line[1:].split() if isb else line.split()
        if len(f) < 2:
            continue
        freq, width = float(f[0]), float(f[1])
        scale = T if isb else T * (1.0 + baryv)
        lo, hi = (freq - 0.5 * width) * scale, (freq + 0.5 * width) * scale
        if hi - lo < 1.0:
            mid = 0.5 * (hi + lo)
            lo, hi = mid - 0.5, mid + 0.5
        birds.append((lo, hi))
    birds.sort(key=lambda b: b[0])
    return np.array([b[0] for b in birds]), np.array([b[1] for b in birds])


class _Stream:
    def __init__(self):
        self.rs = np.random.RandomState(1)

    def next(self):
        return int(self.rs.randint(0, 2 ** 32, dtype=np.uint64))


def _median_power(d):
    p = (d.real * d.real + d.imag * d.imag).astype(np.float32)
    return np.float32(np.median(p))


def zap_birdies(fft, N, lobins, hibins):
    out = np.array(fft, dtype=np.complex64, copy=True)
    numbins = len(out)
    g = _Stream()
    hibin = N / 2
    for lo, hi in zip(lobins, hibins):
        if lo >= hibin:
            break
        if hi >= hibin:
            hi = hibin - 1
        ilo, ihi = int(np.floor(lo)), int(np.ceil(hi))
        nz = ihi - ilo
        if nz < 1:
            continue
        lostart, histart = ilo - GUARDBINS - MEDIANBINS, ihi + GUARDBINS
        uselo, usehi = lostart >= 1, histart + MEDIANBINS <= numbins
        if not (uselo or usehi):
            continue
        med, sides = np.float32(0.0), 0
        if uselo:
            med = np.float32(med + _median_power(out[lostart:lostart + MEDIANBINS]))
            sides += 1
        if usehi:
            med = np.float32(med + _median_power(out[histart:histart + MEDIANBINS]))
            sides += 1
        med = np.float32(med / np.float32(sides))
        meanpow = np.float32(float(med) / -np.log(0.5))
        for i in range(nz):
            u = g.next() / 4294967296.0
            amp = np.sqrt(-float(meanpow) * np.log1p(-u))
            ph = 2 * np.pi * (g.next() / 4294967296.0)
            out[ilo + i] = np.complex64(complex(np.float32(amp * np.cos(ph)),
                                                np.float32(amp * np.sin(ph))))
    return out


def zapbirds_tool(fft, N, lobins, hibins, constamp=False):
    """The automatic mode of the zapbirds command (src/zapbirds.c:285-303):
    birdies zapped while lobin < N/2 - 1, over numbins = N/2 bins, and with
    -constamp every zapped bin set to sqrt(local mean power) with zero phase
    (src/zapping.c:124-128) instead of noise."""
    out = np.array(fft, dtype=np.complex64, copy=True)
    hibin = N / 2
    numbins = min(int(hibin), len(out))
    g = _Stream()
    for lo, hi in zip(lobins, hibins):
        if lo >= hibin - 1:
            break
        if hi >= hibin:
            hi = hibin - 1
        ilo, ihi = int(np.floor(lo)), int(np.ceil(hi))
        nz = ihi - ilo
        if nz < 1:
            continue
        lostart, histart = ilo - GUARDBINS - MEDIANBINS, ihi + GUARDBINS
        uselo, usehi = lostart >= 1, histart + MEDIANBINS <= numbins
        if not (uselo or usehi):
            continue
        med, sides = np.float32(0.0), 0
        if uselo:
            med = np.float32(med + _median_power(out[lostart:lostart + MEDIANBINS]))
            sides += 1
        if usehi:
            med = np.float32(med + _median_power(out[histart:histart + MEDIANBINS]))
            sides += 1
        med = np.float32(med / np.float32(sides))
        meanpow = np.float32(float(med) / -np.log(0.5))
        if constamp:
            out[ilo:ilo + nz] = np.complex64(complex(np.float32(np.sqrt(float(meanpow))), 0.0))
            continue
        for i in range(nz):
            u = g.next() / 4294967296.0
            amp = np.sqrt(-float(meanpow) * np.log1p(-u))
            ph = 2 * np.pi * (g.next() / 4294967296.0)
            out[ilo + i] = np.complex64(complex(np.float32(amp * np.cos(ph)),
                                                np.float32(amp * np.sin(ph))))
    return out
