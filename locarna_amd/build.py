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
    "progressive.cpp",
    "mlocarna.cpp",
]
# host code that calls the HIP runtime API (no kernels): compiled by g++
HIPAPI_SOURCES = [
    "batch.cpp",
    "fold.cpp",
]
HIP_SOURCES = [
    "align_kernels.hip",
    "pf_kernels.hip",
    "compact_kernels.hip",
    "band_kernels.hip",
    "fold_kernels.hip",
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


# host code with AddressSanitizer and UndefinedBehaviorSanitizer (CPU checks
# of the packing, scheduling, sampling and input code; the device code is
# built as usual); loaded by tests run with LOCARNA_AMD_LIB pointing at it
# and the sanitizer runtimes preloaded (see sanitize_env)
SAN_OUT = os.path.join(HERE, "_build_san")
SAN_FLAGS = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined"]


def sanitize_env():
    """Environment for running the CPU tests against the sanitizer build."""
    libs = [subprocess.run([CXX, "-print-file-name=" + n], stdout=subprocess.PIPE, text=True).stdout.strip()
            for n in ("libasan.so", "libubsan.so")]
    env = dict(os.environ)
    env["LD_PRELOAD"] = ":".join(libs)
    env["LOCARNA_AMD_LIB"] = os.path.join(SAN_OUT, "liblocarna_amd.so")
    # leaks of the interpreter and of the HIP runtime are not this library's
    env["ASAN_OPTIONS"] = "detect_leaks=0:verify_asan_link_order=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    return env


def build(verbose=False, sanitize=False):
    out = SAN_OUT if sanitize else OUT
    host_flags = HOST_FLAGS + (SAN_FLAGS if sanitize else [])
    os.makedirs(out, exist_ok=True)
    objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(out, os.path.splitext(src)[0] + ".o")
        _run([CXX] + host_flags + ["-c", "-o", obj, os.path.join(CSRC, src)])
        objs.append(obj)
    for src in HIPAPI_SOURCES:
        obj = os.path.join(out, os.path.splitext(src)[0] + ".o")
        _run([CXX] + host_flags + ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-c", "-o", obj,
                                   os.path.join(CSRC, src)])
        objs.append(obj)
    for src in HIP_SOURCES:
        obj = os.path.join(out, os.path.splitext(src)[0] + ".hip.o")
        _run([HIPCC] + HIP_FLAGS + ["-c", "-o", obj, os.path.join(CSRC, src)])
        objs.append(obj)
    lib = os.path.join(out, "liblocarna_amd.so")
    tmp = lib + ".tmp"
    link_san = []
    if sanitize:
        for f in SAN_FLAGS[:1]:
            link_san += ["-Xarch_host", f]
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + link_san + ["-lpthread"])
    os.replace(tmp, lib)
    if sanitize:
        if verbose:
            print("built", lib)
        return lib
    # locarna_gpu: command-line program on the C++ facade (include/locarna_amd.hpp)
    exe = os.path.join(OUT, "locarna_gpu")
    _run([CXX, "-O2", "-std=c++17", "-Wall", "-I" + INCLUDE, "-o", exe + ".tmp", os.path.join(CSRC, "locarna_gpu.cpp"),
          "-L" + OUT, "-llocarna_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(exe + ".tmp", exe)
    # locarna_rnafold_pp: the folding program mlocarna calls per sequence
    exe2 = os.path.join(OUT, "locarna_rnafold_pp")
    _run([CXX, "-O2", "-std=c++17", "-Wall", "-I" + INCLUDE, "-o", exe2 + ".tmp", os.path.join(CSRC, "rnafold_pp.cpp"),
          "-L" + OUT, "-llocarna_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(exe2 + ".tmp", exe2)
    if verbose:
        print("built", lib, exe)
    return lib


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "sanitize":
        # python -m locarna_amd.build sanitize [pytest args]: builds the
        # sanitizer library and runs the CPU tests against it
        build(verbose=True, sanitize=True)
        args = sys.argv[2:] or ["tests", "-m", "not gpu", "-q", "-x"]
        sys.exit(subprocess.run([sys.executable, "-m", "pytest"] + args, cwd=ROOT, env=sanitize_env()).returncode)
    build(verbose=True)
