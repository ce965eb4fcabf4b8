"""The streaming-read probe (presto_amd/csrc/bw_probe.hip) that measures the
read ceiling reported beside the harmonic sums' roofline: it loads on any
machine, and on the GPU it reads a buffer larger than the caches at a rate
between a tenth of the 8 TB/s peak and the peak."""
import ctypes as C
import os

import pytest

from presto_amd import profiling

LIB = os.path.join(os.path.dirname(os.path.abspath(profiling.__file__)), "libpresto_probe.so")


def test_probe_library_exports_its_entry_point():
    lib = C.CDLL(LIB)
    assert hasattr(lib, "probe_read_bandwidth")


@pytest.mark.gpu
def test_read_probe_rate_is_plausible():
    print("case: read probe over 1 GiB", flush=True)
    gbs, grid, ms = profiling.read_probe(1 << 30, reps=3)
    print(f"read probe: {gbs:.0f} GB/s at {grid} workgroups, {ms:.3f} ms", flush=True)
    assert 800.0 < gbs < 8000.0
    assert 1024 <= grid <= 1048576
    assert ms == pytest.approx((1 << 30) / (gbs * 1e9) * 1e3, rel=1e-6)
