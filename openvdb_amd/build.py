"""In-tree build of libnvq.so for gfx950 (hipcc).

Flags: -ffp-contract=off keeps every float operation of the reference
unfused (the kernels write fmaf where the reference does); packed FP32 ops are
disabled.
"""
import concurrent.futures
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libnvq.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_SOURCES = ["nvq_kernels.hip", "nvq_binned.hip", "nvq_adjoint.hip", "nvq_multi.hip", "nvq_profile.hip", "nvq_iso.hip",
               "nvq_voxels.hip", "nvq_stats.hip",
               "nvq_flood.hip"]
HOST_SOURCES = ["nvq_build.cpp", "nvq_io.cpp"]  # host-only C++ (grid builder), compiled by g++
OBJ_DIR = os.path.join(HERE, "_build")
HOST_FLAGS = ["-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-pthread", "-Wall", "-Wextra",
              "-I", os.path.join(ROOT, "include")]

HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I", os.path.join(ROOT, "include")]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _stale(obj, src):
    if not os.path.exists(obj):
        return True
    deps = [src] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".h")] + \
        [os.path.join(ROOT, "include", h) for h in os.listdir(os.path.join(ROOT, "include"))]
    return any(os.path.getmtime(d) > os.path.getmtime(obj) for d in deps)


def build(verbose=False, incremental=False):
    """Compile every source (incremental=True skips objects newer than their sources and headers)."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    objs, jobs = [], []
    for src in HOST_SOURCES + HIP_SOURCES:
        obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        if incremental and not _stale(obj, os.path.join(CSRC, src)):
            continue
        cc = [HIPCC] + HIP_FLAGS if src.endswith(".hip") else ["g++"] + HOST_FLAGS
        jobs.append(cc + ["-c", os.path.join(CSRC, src), "-o", obj])
    # one compiler per source, at most MAX_JOBS (or the CPU count, capped at 8) at a time
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", 0)) or min(8, os.cpu_count() or 1)))
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        for f in [pool.submit(_run, cmd, verbose) for cmd in jobs]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-o", tmp] + objs + ["-pthread"], verbose)
    os.replace(tmp, LIB)
    build_api_test(verbose)
    return LIB


# Client kernels of the device-API tests: written against include/nvq_device.h alone (plus the
# C ABI of libnvq.so for nvq_grid_get_view), as a NanoVDB user's own kernels would be.
API_TEST_SRC = os.path.join(ROOT, "tests", "device_api", "api_kernels.hip")
API_TEST_LIB = os.path.join(ROOT, "tests", "device_api", "libnvq_api_test.so")


def build_api_test(verbose=False):
    if not os.path.exists(API_TEST_SRC):
        return None
    tmp = API_TEST_LIB + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-shared", API_TEST_SRC, "-o", tmp, "-L", HERE, "-l:libnvq.so",
                                "-Wl,-rpath,$ORIGIN/../../openvdb_amd"], verbose)
    os.replace(tmp, API_TEST_LIB)
    return API_TEST_LIB


if __name__ == "__main__":
    print(build(verbose=True))
