"""The rtc::amd C++ host layer (barney's rtcore backend classes over the C
ABI), exercised by a g++-built program: without a GPU it must fail loudly
with the runtime's message; on the GPU it checks memory, groups, geometry
types, tracing against analytic hits, and error reporting. A hipcc-built
program checks the device side (programs, trace and compute kernels)."""
import os
import subprocess

import pytest

from conftest import gpu_available

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "rtcore_amd_selftest")
LIB = os.path.join(ROOT, "barney_amd", "librtcore_amd.so")


def _run():
    if not os.path.exists(EXE):
        pytest.skip("self-test not built (python -m barney_amd.build)")
    return subprocess.run([EXE], capture_output=True, text=True, timeout=120)


def test_host_layer_exports():
    if not os.path.exists(LIB):
        pytest.skip("librtcore_amd.so not built")
    out = subprocess.run(["nm", "-DC", LIB], capture_output=True, text=True).stdout
    for name in ("rtc::amd::Device::Device(int)", "rtc::amd::Device::createInstanceGroup",
                 "rtc::amd::InstanceGroup::buildAccel()", "rtc::amd::InstanceGroup::traceRays",
                 "rtc::amd::createGeomType_Spheres", "rtc::amd::Buffer::upload"):
        assert name in out, name


def test_without_gpu_fails_loudly():
    if gpu_available():
        pytest.skip("a GPU is present")
    r = _run()
    assert r.returncode == 3, r.stdout + r.stderr
    assert r.stdout.startswith("NO_GPU: rtc::amd::Device")


@pytest.mark.gpu
def test_selftest_on_gpu(gpu):
    r = _run()
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK")


PROGRAMS = os.path.join(ROOT, "tests", "cpp", "rtcore_amd_programs")


@pytest.mark.gpu
def test_programmable_traversal_on_gpu(gpu):
    """Device programs and kernels written against rtcore_amd_device.h
    (RTC_EXPORT_TRIANGLES_GEOM / USER_GEOM / TRACE2D / COMPUTE1D), launched
    through TraceKernel2D / ComputeKernel1D: every ray's closest hit through
    TraceInterface::traceRay (anyHit ignoring a fifth of the triangles, user
    spheres by their intersect program, a two-level world) equals the native
    trace of the kept triangles and a float64 brute force over the spheres,
    t bit for bit, with the IDs and instance transforms the closestHit
    program reads."""
    if not os.path.exists(PROGRAMS):
        pytest.skip("programs test not built (python -m barney_amd.build)")
    r = subprocess.run([PROGRAMS], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK")
