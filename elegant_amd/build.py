"""Builds libelegant_hip.so in-tree: HIP kernels with hipcc for gfx950 and the
C host layer with gcc (ISO C11, no FP contraction), linked into one shared
library with a C ABI (include/elegant_hip.h)."""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
OUT = os.path.join(HERE, "libelegant_hip.so")
BUILD = os.path.join(HERE, "_build")

HIP_SOURCES = ["eh_track.hip", "eh_collective.hip", "eh_mlp.hip"]
C_SOURCES = ["eh_compile.c", "eh_driver.c", "eh_collective.c", "eh_batch.c", "eh_da.c", "eh_learn.c", "eh_gen.c",
             "eh_runs.c"]

# FP64 parity needs IEEE division/sqrt and no contraction on both sides.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-ffp-contract=off", "-fno-fast-math", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + CSRC, "-I" + os.path.join(ROOT, "include"),
]
GCC_FLAGS = ["-O2", "-std=c11", "-ffp-contract=off", "-fPIC", "-Wall", "-Wno-unused-function",
             "-I" + CSRC, "-I" + HOST, "-I" + os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError("build step failed: %s" % cmd[0])
    return r.stdout


def build(verbose=False):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        _run([hipcc] + HIPCC_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in C_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        _run(["gcc"] + GCC_FLAGS + ["-c", os.path.join(HOST, src), "-o", obj])
        objs.append(obj)
    tmp = OUT + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lm"])
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
