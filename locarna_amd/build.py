"""Build step for the native library ``liblocarna_amd.so`` (gfx950).

Host C++ (input handling, score tables, trace band, arc-match lists, C ABI)
is compiled with g++ -O2 without -march flags, so that its floating point is
the same as a default build of the reference (x87 long double, no FMA
contraction).  HIP sources are compiled with hipcc for gfx950 only.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")

HOST_SOURCES = [
    "synth.cpp",
    "host.cpp",
]
# host code that calls the HIP runtime API (no kernels): compiled by g++
HIPAPI_SOURCES = [
    "batch.cpp",
]
HIP_SOURCES = [
    "align_kernels.hip",
    "pf_kernels.hip",
    "compact_kernels.hip",
    "band_kernels.hip",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-O2", "-g", "-std=c++17", "-fPIC", "-Wall", "-Wno-sign-compare",
              "-I" + INCLUDE, "-I" + CSRC]
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + INCLUDE, "-I" + CSRC]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError("build step failed: " + os.path.basename(cmd[-1]))
    return r.stdout


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(OUT, os.path.splitext(src)[0] + ".o")
        _run([CXX] + HOST_FLAGS + ["-c", "-o", obj, os.path.join(CSRC, src)])
        objs.append(obj)
    for src in HIPAPI_SOURCES:
        obj = os.path.join(OUT, os.path.splitext(src)[0] + ".o")
        _run([CXX] + HOST_FLAGS + ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-c", "-o", obj,
                                   os.path.join(CSRC, src)])
        objs.append(obj)
    for src in HIP_SOURCES:
        obj = os.path.join(OUT, os.path.splitext(src)[0] + ".hip.o")
        _run([HIPCC] + HIP_FLAGS + ["-c", "-o", obj, os.path.join(CSRC, src)])
        objs.append(obj)
    lib = os.path.join(OUT, "liblocarna_amd.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, lib)
    # locarna_gpu: command-line program on the C++ facade (include/locarna_amd.hpp)
    exe = os.path.join(OUT, "locarna_gpu")
    _run([CXX, "-O2", "-std=c++17", "-Wall", "-I" + INCLUDE, "-o", exe + ".tmp", os.path.join(CSRC, "locarna_gpu.cpp"),
          "-L" + OUT, "-llocarna_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(exe + ".tmp", exe)
    if verbose:
        print("built", lib, exe)
    return lib


if __name__ == "__main__":
    build(verbose=True)
