"""Builds libpresto_accel.so in-tree: the HIP kernels (hipcc, gfx950) and
the C host (gcc), linked into one shared library next to this file."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpresto_accel.so")
CLI = os.path.join(HERE, "accelsearch")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I", CSRC, "-I", os.path.join(ROOT, "include")]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-ffp-contract=off", "-Wall",
           "-I", CSRC, "-I", os.path.join(ROOT, "include")]
HIP_SRC = ["accel_gpu.hip", "refine_gpu.hip", "fft_gpu.hip", "surface_gpu.hip", "dedisp_gpu.hip",
           "fold_gpu.hip"]
C_SRC = ["accel_host.c", "accel_stats.c", "accel_dered.c", "accel_output.c", "accel_batch.c",
         "accel_server.c", "accel_surface.c", "accel_dedisp.c", "accel_sift.c", "accel_fold.c"]


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in HIP_SRC:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in C_SRC:
        obj = os.path.join(BUILD, src.replace(".c", ".o"))
        run(["gcc"] + C_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs +
        ["-lm", "-lpthread", "-ldl"])
    os.replace(tmp, LIB)
    # the accelsearch command, linked against the library next to it
    cli_obj = os.path.join(BUILD, "accelsearch_cli.o")
    run(["gcc"] + C_FLAGS + ["-c", os.path.join(CSRC, "accelsearch_cli.c"), "-o", cli_obj])
    tmp = CLI + ".tmp"
    run(["gcc", "-o", tmp, cli_obj, "-L", HERE, "-lpresto_accel", "-Wl,-rpath,$ORIGIN", "-lm"])
    os.replace(tmp, CLI)
    # accel_sift: the sifting flow over a directory's ACCEL files
    obj = os.path.join(BUILD, "accel_sift_cli.o")
    run(["gcc"] + C_FLAGS + ["-c", os.path.join(CSRC, "accel_sift_cli.c"), "-o", obj])
    out = os.path.join(HERE, "accel_sift")
    run(["gcc", "-o", out + ".tmp", obj, "-L", HERE, "-lpresto_accel", "-Wl,-rpath,$ORIGIN",
         "-lm"])
    os.replace(out + ".tmp", out)
    # accelfold: prepfold -accelcand on a .dat
    obj = os.path.join(BUILD, "accelfold_cli.o")
    run(["gcc"] + C_FLAGS + ["-c", os.path.join(CSRC, "accelfold_cli.c"), "-o", obj])
    out = os.path.join(HERE, "accelfold")
    run(["gcc", "-o", out + ".tmp", obj, "-L", HERE, "-lpresto_accel", "-Wl,-rpath,$ORIGIN",
         "-lm"])
    os.replace(out + ".tmp", out)
    # realfft, rednoise, zapbirds: one source, three commands
    for tool in ("realfft", "rednoise", "zapbirds"):
        obj = os.path.join(BUILD, tool + ".o")
        run(["gcc"] + C_FLAGS + ["-DTOOL_" + tool.upper(), "-c",
                                 os.path.join(CSRC, "fft_tools.c"), "-o", obj])
        out = os.path.join(HERE, tool)
        run(["gcc", "-o", out + ".tmp", obj, "-L", HERE, "-lpresto_accel",
             "-Wl,-rpath,$ORIGIN", "-lm"])
        os.replace(out + ".tmp", out)
    # the streaming-read probe (measurement only; nothing in the search
    # links it)
    obj = os.path.join(BUILD, "bw_probe.o")
    run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, "bw_probe.hip"), "-o", obj])
    out = os.path.join(HERE, "libpresto_probe.so")
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out + ".tmp", obj])
    os.replace(out + ".tmp", out)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
