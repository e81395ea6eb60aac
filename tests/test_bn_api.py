"""The bn* API (include/bn_api.h, barney.h's surface subset) as a C++
program uses it: libbn_api.so exports every declared function; without a
GPU bnContextCreate throws with the runtime's message; on the GPU a scene
built and rendered through bn* equals, bit for bit, the same scene through
the bnamd C ABI (color over two progressive frames, depth and ID channels,
bnAccumReset; the volume field kinds, NanoVDB among them), and unknown types
throw."""
import os
import re
import subprocess
import tempfile

import pytest

from conftest import gpu_available

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "bn_api_selftest")
LIB = os.path.join(ROOT, "barney_amd", "libbn_api.so")
HDR = os.path.join(ROOT, "include", "bn_api.h")


def _run():
    if not os.path.exists(EXE):
        pytest.skip("self-test not built (python -m barney_amd.build)")
    # a NanoVDB grid around the origin for the "NanoVDB" scalar field case
    from barney_amd import nanovdb, scenes
    f = scenes.structured_field(24, blobs=4, sigma=(3.0, 8.0), seed=0x5EED0A0A)
    grid = nanovdb.from_dense(f, voxel_size=0.125, origin=(-1.5, -1.5, -1.5))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "grid.nvdb")
        grid.tofile(path)
        return subprocess.run([EXE, path], capture_output=True, text=True, timeout=120)


def test_exports_every_declared_function():
    if not os.path.exists(LIB):
        pytest.skip("libbn_api.so not built")
    names = set(re.findall(r"^BN_API\s+[\w\s\*]+?\b(bn\w+)\(", open(HDR).read(), re.M))
    assert len(names) >= 25
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    assert names <= exported, names - exported


def test_without_gpu_fails_loudly():
    if gpu_available():
        pytest.skip("a GPU is present")
    r = _run()
    assert r.returncode == 3, r.stdout + r.stderr
    assert r.stdout.startswith("NO_GPU: bnContextCreate")


@pytest.mark.gpu
def test_bn_api_equals_c_abi_on_gpu(gpu):
    r = _run()
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK")
