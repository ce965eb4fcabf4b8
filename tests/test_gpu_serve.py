"""The request loop (accel_server_*): requests of mixed sizes and search
parameters give the same candidates and refined properties whether each is
served alone or together with the others, and admission by memory keeps
the device within the budget."""
import numpy as np
import pytest

from presto_amd import synth

pytestmark = pytest.mark.gpu
A = pytest.importorskip("presto_amd.accel")


def _requests():
    spec = [(1 << 15, 1e-3, 0, None, 4), (1 << 16, 5e-4, 20, None, 8), (1 << 17, 2.5e-4, 50, None, 8),
            (1 << 16, 5e-4, 20, 40, 4), (1 << 15, 1e-3, 50, None, 2), (1 << 17, 2.5e-4, 0, None, 8),
            (1 << 16, 5e-4, 20, None, 8), (1 << 16, 5e-4, 10, 20, 2)]
    out = []
    for i, (N, dt, zmax, wmax, nh) in enumerate(spec):
        ps = [dict(f0=41.3 + 3 * i, z=2.0 * i - 6, duty=0.06, amp=0.4)] if i % 3 != 2 else []
        x = synth.make_series(N, dt, 900 + i, ps)
        p = A.make_params(N // 2, N, dt, zmax=zmax, numharm=nh, sigma=2.0, wmax=wmax)
        out.append((p, x))
    return out


def _serve(reqs, workers, budget, together):
    res = []
    with A.Server(budget_bytes=budget, workers=workers) as sv:
        if together:
            tickets = [sv.submit(p, x) for p, x in reqs]
            for t in tickets:
                st, lat, svc = sv.wait(t)
                assert st == 0
                res.append((sv.candidates(t), sv.props(t)))
        else:
            for p, x in reqs:
                t = sv.submit(p, x)
                st, lat, svc = sv.wait(t)
                assert st == 0
                res.append((sv.candidates(t), sv.props(t)))
    return res


def test_results_equal_alone_and_in_the_mix():
    print("case: 8 mixed requests, alone vs 4 workers together", flush=True)
    reqs = _requests()
    big = max(A.footprint(p) for p, _ in reqs)
    alone = _serve(reqs, 1, 64 << 30, together=False)
    mixed = _serve(reqs, 4, 64 << 30, together=True)
    tight = _serve(reqs, 4, big + (64 << 20), together=True)   # one big request at a time
    assert sum(len(c) for c, _ in alone) >= 4
    for (ca, pa), (cm, pm), (ct, pt) in zip(alone, mixed, tight):
        assert np.array_equal(ca, cm) and np.array_equal(ca, ct)
        assert len(pa) == len(pm) == len(pt)
        for a, b, c in zip(pa, pm, pt):
            assert bytes(a) == bytes(b) == bytes(c)


def test_request_larger_than_the_budget_fails_cleanly():
    p, x = _requests()[2]
    with A.Server(budget_bytes=A.footprint(p) // 2, workers=2) as sv:
        t = sv.submit(p, x)
        st, _, _ = sv.wait(t)
        assert st == -1
        assert sv.memory() == 0


def test_request_array_growth_while_workers_run():
    """More requests than the server's first request-array allocation (64),
    submitted while workers are busy, so the array is reallocated under
    running requests; each result equals the same request served alone."""
    print("case: 150 small requests (24 distinct) through 4 workers", flush=True)
    distinct = []
    for i in range(24):
        N, dt = (1 << 14) if i % 2 else (1 << 15), 1e-3
        ps = [dict(f0=31.7 + 2.1 * i, z=float(i % 5) - 2.0, duty=0.08, amp=0.5)] if i % 4 else []
        x = synth.make_series(N, dt, 7000 + i, ps)
        p = A.make_params(N // 2, N, dt, zmax=(0, 10, 20)[i % 3], numharm=(2, 4)[i % 2], sigma=2.0)
        distinct.append((p, x))
    alone = _serve(distinct, 1, 16 << 30, together=False)
    order = [k % len(distinct) for k in range(150)]
    with A.Server(budget_bytes=16 << 30, workers=4) as sv:
        tickets = [sv.submit(*distinct[k]) for k in order]
        for k, t in zip(order, tickets):
            st, _, _ = sv.wait(t)
            assert st == 0
            ca, pa = alone[k]
            assert np.array_equal(sv.candidates(t), ca)
            assert [bytes(v) for v in sv.props(t)] == [bytes(v) for v in pa]


def test_small_request_backfills_beside_a_waiting_large_one():
    """Two large requests that cannot share the budget and a small one
    behind them: the small one starts beside the first while the second
    waits for memory (backfill), and its results are those it has alone."""
    N, dt = 1 << 21, 64e-6
    big = A.make_params(N // 2, N, dt, zmax=200, numharm=16, sigma=2.0)
    xa = synth.make_series(N, dt, 931, [dict(f0=211.7, z=12.0, duty=0.05, amp=0.05)])
    xh = synth.make_series(N, dt, 932, [])
    n, sdt = 1 << 15, 1e-3
    small = A.make_params(n // 2, n, sdt, zmax=0, numharm=2, sigma=2.0)
    xs = synth.make_series(n, sdt, 933, [dict(f0=41.3, z=0.0, duty=0.06, amp=0.4)])
    fb, fs = A.footprint(big), A.footprint(small)
    budget = fb + fs + (64 << 20)
    print(f"case: backfill, big {fb >> 20} MiB x 2, small {fs >> 20} MiB, budget "
          f"{budget >> 20} MiB", flush=True)
    assert 2 * fb > budget
    with A.Server(budget_bytes=budget, workers=4) as sv:
        ta = sv.submit(big, xa)
        th = sv.submit(big, xh)
        ts = sv.submit(small, xs)
        for t in (ta, th, ts):
            assert sv.wait(t)[0] == 0
        qh, qs = sv.timing(th)["queued"], sv.timing(ts)["queued"]
        cs, ps = sv.candidates(ts), sv.props(ts)
        assert sv.memory() <= budget
    print(f"queued: second large {qh:.1f} ms, small {qs:.1f} ms", flush=True)
    assert qs < qh
    with A.Server(budget_bytes=budget, workers=1) as sv:
        t = sv.submit(small, xs)
        assert sv.wait(t)[0] == 0
        assert np.array_equal(sv.candidates(t), cs)
        assert [bytes(v) for v in sv.props(t)] == [bytes(v) for v in ps]
