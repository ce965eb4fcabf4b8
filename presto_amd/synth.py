"""Seeded synthetic time series for the accelsearch workloads: Gaussian
noise plus pulsars with a Gaussian pulse profile and constant f-dot /
f-dotdot.  PRESTO's Fourier derivatives are z = fdot*T^2 and
w = fdotdot*T^3 (src/accel_utils.c:1431-1433 works on the same grid)."""
import numpy as np


def pulsar_series(N, dt, f0, z=0.0, w=0.0, duty=0.05, amp=0.05):
    """Pulse train with average Fourier frequency f0*T + z/2 (+ w/6) bins.

    phase(t) = f0 t + fdot t^2/2 + fdotdot t^3/6, Gaussian pulse of FWHM
    duty/f0 in phase units, peak amplitude amp per sample."""
    T = N * dt
    t = np.arange(N, dtype=np.float64) * dt
    fdot = z / T**2
    fdd = w / T**3
    ph = f0 * t + 0.5 * fdot * t * t + fdd * t * t * t / 6.0
    d = ph - np.floor(ph + 0.5)          # distance to the nearest pulse centre
    sig = duty / (2.0 * np.sqrt(2.0 * np.log(2.0)))
    return amp * np.exp(-0.5 * (d / sig) ** 2)


def make_series(N, dt, seed, pulsars=()):
    """float32 series: N(0,1) noise from PCG64(seed) plus the given pulsars,
    each a dict of pulsar_series keyword arguments."""
    rng = np.random.Generator(np.random.PCG64(seed))
    x = rng.standard_normal(N)
    for p in pulsars:
        x = x + pulsar_series(N, dt, **p)
    return x.astype(np.float32)


def config1_series():
    """One 2^20-sample series, dt 64 us, one pulsar at 123.456 Hz, z=20."""
    return make_series(1 << 20, 64e-6, 1,
                       [dict(f0=123.456, z=20.0, duty=0.05, amp=0.05)])


def config2_series(N=1 << 23, dt=64e-6, seed=2):
    """2^23 samples, dt 64 us, 3 pulsars with f in [50,700] Hz, z in
    [-150,150], duty 2-10 %, amplitude drawn in [0.01,0.05] from the seed."""
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    ps = []
    for _ in range(3):
        ps.append(dict(f0=float(rng.uniform(50, 700)), z=float(rng.uniform(-150, 150)),
                       duty=float(rng.uniform(0.02, 0.10)),
                       amp=float(rng.uniform(0.01, 0.05))))
    return make_series(N, dt, seed, ps), ps


def make_series_f32(N, dt, seed, pulsars=()):
    """As make_series, with the noise drawn directly in float32 (the batched
    workload's generator: PCG64(seed) float32 normals)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    x = rng.standard_normal(N, dtype=np.float32)
    for p in pulsars:
        x += pulsar_series(N, dt, **p).astype(np.float32)
    return x


def config3_trials(n=256, seed0=3000):
    """The batched DM-trial workload: n trials, lengths 2^23 / 2^22 / 2^21
    samples (dt 64 / 128 / 256 us, so every trial spans 536.87 s) in a
    2:1:1 ratio as the DM (trial index) grows and the data are downsampled;
    seeded noise (seed0 + i); about 10 % of the trials carry one pulsar
    whose amplitude peaks at trial i0 and falls off as a Gaussian of width
    3 trials.  Yields (i, N, dt, seed, pulsars)."""
    i0 = n // 6
    for i in range(n):
        if i < n // 2:
            N, dt = 1 << 23, 64e-6
        elif i < 3 * n // 4:
            N, dt = 1 << 22, 128e-6
        else:
            N, dt = 1 << 21, 256e-6
        pulsars = []
        if abs(i - i0) <= max(1, int(0.05 * n)):
            amp = 0.04 * float(np.exp(-0.5 * ((i - i0) / 3.0) ** 2))
            pulsars = [dict(f0=150.7, z=30.0, duty=0.05, amp=amp)]
        yield i, N, dt, seed0 + i, pulsars


def config4_trials(n=64, seed0=4000):
    """The jerk workload: n spectra of 2^22 samples (dt 64 us, T = 268.4 s),
    seeded noise (seed0 + i) plus one binary-like signal each, drawn from
    PCG64(seed0 + 100000 + i): f0 in [50, 700] Hz, average z in
    [-100, 100], w in [-300, 300], duty 5 %, amplitude 0.1.  The series
    generator takes the starting f-dot, z0 = z_avg - w / 2.  Yields
    (i, N, dt, seed, pulsars, (r_avg, z_avg, w))."""
    N, dt = 1 << 22, 64e-6
    T = N * dt
    for i in range(n):
        rng = np.random.Generator(np.random.PCG64(seed0 + 100000 + i))
        f0 = float(rng.uniform(50.0, 700.0))
        z_avg = float(rng.uniform(-100.0, 100.0))
        w = float(rng.uniform(-300.0, 300.0))
        z0 = z_avg - 0.5 * w
        r_avg = f0 * T + 0.5 * z0 + w / 6.0
        yield i, N, dt, seed0 + i, [dict(f0=f0, z=z0, w=w, duty=0.05, amp=0.1)], \
            (r_avg, z_avg, w)


def smooth_even(n, primes=(2, 3, 5, 7)):
    """The even length nearest to n (ties to the smaller) whose prime
    factors are all in primes: the lengths real FFTs handle with small
    radices."""
    def smooth(m):
        for p in primes:
            while m % p == 0:
                m //= p
        return m == 1
    n = max(2, int(round(n)))
    for d in range(n):
        for m in (n - d, n + d):
            if m >= 2 and m % 2 == 0 and smooth(m):
                return m
    return 2


def config5_requests(n, seed_arrivals=5, seed0=5_000_000, rate=40.0):
    """The serving workload: n requests with Poisson arrival times (rate per
    second, PCG64(seed_arrivals)); per request (PCG64(seed0 + i)): a length
    log-uniform in 2^20..2^24 samples, rounded to the nearest even length
    whose prime factors are 2, 3, 5 and 7 (smooth_even), dt 64 us, zmax from
    {0, 50, 200}, wmax from {0, 100} (0: no jerk search), numharm from
    {8, 16}, and a pulsar in 20 % of them.  Yields (i, t_arrival, N, dt,
    zmax, wmax, numharm, seed, pulsars)."""
    ra = np.random.Generator(np.random.PCG64(seed_arrivals))
    t = np.cumsum(ra.exponential(1.0 / rate, n))
    for i in range(n):
        r = np.random.Generator(np.random.PCG64(seed0 + i))
        N = smooth_even(2.0 ** r.uniform(20.0, 24.0))
        zmax = int(r.choice([0, 50, 200]))
        wmax = int(r.choice([0, 100]))
        nh = int(r.choice([8, 16]))
        ps = []
        if r.uniform() < 0.2:
            ps = [dict(f0=float(r.uniform(20, 500)), z=float(r.uniform(-0.5, 0.5) * zmax),
                       w=float(r.uniform(-0.5, 0.5) * wmax), duty=0.05, amp=0.05)]
        yield i, float(t[i]), N, 64e-6, zmax, wmax, nh, seed0 + i, ps


def filterbank(nspec, numchan, dt, lofreq, chanwidth, seed, pulsars=()):
    """float32 spectra [nspec][numchan] in ascending frequency: N(0,1)
    noise from PCG64(seed) plus dispersed pulsars (dicts with f0, dm, duty,
    amp): in channel c the pulse train arrives dm / (0.000241 f_c^2) s
    later than at the top channel (src/dispersion.c:34-43)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    x = rng.standard_normal((nspec, numchan), dtype=np.float32)
    t = np.arange(nspec, dtype=np.float64) * dt
    ftop = lofreq + (numchan - 1) * chanwidth
    for p in pulsars:
        sig = p.get("duty", 0.05) / (2.0 * np.sqrt(2.0 * np.log(2.0)))
        for c in range(numchan):
            f = lofreq + c * chanwidth
            delay = p["dm"] / (0.000241 * f * f) - p["dm"] / (0.000241 * ftop * ftop)
            ph = p["f0"] * (t - delay)
            d = ph - np.floor(ph + 0.5)
            x[:, c] += (p["amp"] * np.exp(-0.5 * (d / sig) ** 2)).astype(np.float32)
    return x
