"""The rtc::amd C++ host layer (barney's rtcore backend classes over the C
ABI), exercised by a g++-built program: without a GPU it must fail loudly
with the runtime's message; on the GPU it checks memory, groups, geometry
types, tracing against analytic hits, and error reporting."""
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
