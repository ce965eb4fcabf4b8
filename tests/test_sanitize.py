"""The library's host C code under AddressSanitizer, LeakSanitizer and
UndefinedBehaviorSanitizer: geometry, statistics, de-reddening, birdie
zapping, the ACCEL value(error) formatting and fourierprops, the sifter
over the sifting fixtures (read, reject, duplicates, DM problems,
harmonics, the three text outputs) and the .cand reader, built with gcc
into one host-only executable (tests/sanitize/host_driver.c; sections the
driver does not reach, and with them every device call, are dropped at
link time) and run.  Any sanitizer report fails the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "presto_amd", "csrc")
HOST_SOURCES = ["accel_host.c", "accel_stats.c", "accel_dered.c", "accel_output.c",
                "accel_sift.c", "accel_fold.c"]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_host_code_under_asan_ubsan(tmp_path):
    exe = tmp_path / "host_driver"
    cmd = ["gcc", "-O1", "-g", "-std=gnu11", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-ffunction-sections",
           "-fdata-sections", "-I", CSRC, "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "sanitize", "host_driver.c")] + \
          [os.path.join(CSRC, f) for f in HOST_SOURCES] + \
          ["-Wl,--gc-sections", "-lm", "-lpthread", "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    scratch = tmp_path / "scratch"
    scratch.mkdir()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe), os.path.join(ROOT, "tests", "golden", "sift"), str(scratch)],
                       capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert "sanitized host checks passed" in r.stdout
    assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr
    assert (scratch / "sifted.txt").stat().st_size > 0


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
@pytest.mark.parametrize("sanitizer", ["thread", "address,undefined"])
def test_request_loop_under_sanitizer(tmp_path, sanitizer):
    """The request loop (accel_server.c) with test doubles of the search
    handle (tests/sanitize/server_driver.c): 200 mixed requests through 4
    workers within a tight budget, a request above the budget, and a small
    request backfilled beside a large one that waits for memory; under
    ThreadSanitizer (data races on the queue, the handle cache and the
    request array) and under AddressSanitizer/UBSan."""
    exe = tmp_path / "server_driver"
    cmd = ["gcc", "-O1", "-g", "-std=gnu11", f"-fsanitize={sanitizer}",
           "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-I", CSRC,
           "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "sanitize", "server_driver.c"),
           os.path.join(CSRC, "accel_server.c"), "-lpthread", "-lm", "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1", ASAN_OPTIONS="detect_leaks=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert "request loop checks passed" in r.stdout
    assert "ThreadSanitizer" not in r.stderr and "runtime error" not in r.stderr
