"""Builds libbnamd.so (the HIP kernels + C ABI) for gfx950, in-tree, and the
C++ host layer over it (librtcore_amd.so, plain g++) with its self-test
program (tests/cpp/rtcore_amd_selftest).

Each .hip translation unit is compiled separately (in parallel) and linked
with hipcc into barney_amd/libbnamd.so. Usage: python -m barney_amd.build
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libbnamd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
RTC_SRC = os.path.join(HERE, "rtcore_amd", "rtcore_amd.cpp")
RTC_LIB = os.path.join(HERE, "librtcore_amd.so")
SELFTEST_SRC = os.path.join(os.path.dirname(HERE), "tests", "cpp", "rtcore_amd_selftest.cpp")
SELFTEST = os.path.join(os.path.dirname(HERE), "tests", "cpp", "rtcore_amd_selftest")
BN_SRC = os.path.join(HERE, "bn", "bn_api.cpp")
BN_LIB = os.path.join(HERE, "libbn_api.so")
BN_TEST_SRC = os.path.join(os.path.dirname(HERE), "tests", "cpp", "bn_api_selftest.cpp")
BN_TEST = os.path.join(os.path.dirname(HERE), "tests", "cpp", "bn_api_selftest")
PROG_TEST_SRC = os.path.join(os.path.dirname(HERE), "tests", "cpp", "rtcore_amd_programs.hip")
PROG_TEST = os.path.join(os.path.dirname(HERE), "tests", "cpp", "rtcore_amd_programs")

SOURCES = [
    "runtime.hip",
    "bvh_build.hip",
    "trace.hip",
    "render_surface.hip",
    "volume.hip",
    "nanovdb.hip",
    "lights.hip",
    "ea.hip",
    "render_batch.hip",
    "umesh.hip",
    "umesh_count.hip",
    "fb.hip",
    "envmap.hip",
]

# -packed-fp32-ops: packed FP32 forms are kept out of the generated code.
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result",
    # every fp32 expression rounds as written (bit-comparable with the
    # oracle); hot loops use explicit fmaf where fusion is wanted
    "-ffp-contract=off",
]


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_host_layer(verbose)
    return LIB


def _run(cmd, what):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{what} failed:\n{' '.join(cmd)}\n{r.stderr}")


def build_host_layer(verbose=False):
    """The rtc::amd C++ layer is plain host C++ over the C ABI."""
    common = ["-std=c++17", "-O2", "-Wall", "-Wextra", "-fPIC"]
    tmp = RTC_LIB + ".tmp"
    _run([CXX, *common, "-shared", RTC_SRC, "-o", tmp, "-L" + HERE, "-lbnamd", "-Wl,-rpath,$ORIGIN"],
         "g++ (rtcore_amd)")
    os.replace(tmp, RTC_LIB)
    tmp = SELFTEST + ".tmp"
    _run([CXX, *common, SELFTEST_SRC, "-o", tmp, "-L" + HERE, "-lrtcore_amd", "-lbnamd",
          "-Wl,-rpath,$ORIGIN/../../barney_amd"], "g++ (selftest)")
    os.replace(tmp, SELFTEST)
    # the bn* API (barney.h subset) over the C ABI, and its self-test
    tmp = BN_LIB + ".tmp"
    inc = "-I" + os.path.join(os.path.dirname(HERE), "include")
    _run([CXX, *common, inc, "-shared", BN_SRC, "-o", tmp, "-L" + HERE, "-lbnamd", "-Wl,-rpath,$ORIGIN"],
         "g++ (bn_api)")
    os.replace(tmp, BN_LIB)
    tmp = BN_TEST + ".tmp"
    _run([CXX, *common, inc, BN_TEST_SRC, "-o", tmp, "-L" + HERE, "-lbn_api", "-lbnamd",
          "-Wl,-rpath,$ORIGIN/../../barney_amd"], "g++ (bn_api selftest)")
    os.replace(tmp, BN_TEST)
    # programs and kernels against rtc::amd's device interface (hipcc: the
    # traversal and the programs share the program's code object)
    tmp = PROG_TEST + ".tmp"
    _run([HIPCC, *FLAGS, PROG_TEST_SRC, "-o", tmp, "-L" + HERE, "-lrtcore_amd", "-lbnamd",
          "-Wl,-rpath,$ORIGIN/../../barney_amd"], "hipcc (rtcore_amd programs)")
    os.replace(tmp, PROG_TEST)
    if verbose:
        print(f"built {RTC_LIB}, {SELFTEST}, {BN_LIB}, {BN_TEST} and {PROG_TEST}")


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
