"""In-tree build of libnvq.so for gfx950 (hipcc).

Flags: -ffp-contract=off keeps every float operation of the reference
unfused (the kernels write fmaf where the reference does); packed FP32 ops are
disabled.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libnvq.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_SOURCES = ["nvq_kernels.hip"]
HOST_SOURCES = []

HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I", os.path.join(ROOT, "include")]


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES + HOST_SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + HIP_FLAGS + ["-shared", "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
